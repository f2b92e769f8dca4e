// ba_certify.cpp — host side of the post-loop certification (see ba_certify.hpp for the
// reference functions each piece follows).
#include <hip/hip_runtime.h>

#include "ba_certify.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

#include "ba_math.hpp"

namespace ba {

void OwnedProblem::bind() {
    p.n_poses = var_rot.size();
    p.pose_rotation = rot.data();
    p.pose_center = cen.data();
    p.pose_var_rotation = var_rot.data();
    p.pose_var_translation = var_trans.data();
    p.n_intrinsics = intr.size();
    p.intrinsics = intr.data();
    p.n_points = point_fixed.size();
    p.point_xyz = xyz.data();
    p.point_fixed = point_fixed.data();
    p.n_observations = obs_point.size();
    p.obs_pose = obs_pose.data();
    p.obs_intrinsic = obs_intr.data();
    p.obs_point = obs_point.data();
    p.obs_pixel = pix.data();
}

void HostSolution::alloc(const ba_problem* p) {
    rot.assign(9 * p->n_poses, 0.0);
    cen.assign(3 * p->n_poses, 0.0);
    focal.assign(p->n_intrinsics, 0.0);
    k1.assign(p->n_intrinsics, 0.0);
    k2.assign(p->n_intrinsics, 0.0);
    xyz.assign(3 * p->n_points, 0.0);
    cov.assign(9 * p->n_intrinsics, NAN);
    quarantined.assign(p->n_points, 0);
    unconstrained.assign(p->n_intrinsics, 0);
    std::memset(&rep, 0, sizeof(rep));
}

ba_solution HostSolution::view() {
    ba_solution s{};
    s.pose_rotation = rot.data();
    s.pose_center = cen.data();
    s.intrinsic_focal = focal.data();
    s.intrinsic_k1 = k1.data();
    s.intrinsic_k2 = k2.data();
    s.point_xyz = xyz.data();
    s.point_quarantined = quarantined.data();
    s.intrinsic_unconstrained = unconstrained.data();
    s.intrinsic_covariance = cov.data();
    return s;
}

void HostSolution::copy_to(ba_solution* out) const {
    auto cp = [](void* dst, const void* src, size_t bytes) {
        if (dst && bytes) std::memcpy(dst, src, bytes);
    };
    cp(out->pose_rotation, rot.data(), rot.size() * 8);
    cp(out->pose_center, cen.data(), cen.size() * 8);
    cp(out->intrinsic_focal, focal.data(), focal.size() * 8);
    cp(out->intrinsic_k1, k1.data(), k1.size() * 8);
    cp(out->intrinsic_k2, k2.data(), k2.size() * 8);
    cp(out->point_xyz, xyz.data(), xyz.size() * 8);
    cp(out->point_quarantined, quarantined.data(), quarantined.size());
    cp(out->intrinsic_unconstrained, unconstrained.data(), unconstrained.size());
    cp(out->intrinsic_covariance, cov.data(), cov.size() * 8);
}

void without_points(const ba_problem* p, const HostSolution* state, const std::vector<uint8_t>& drop,
                    OwnedProblem& out, std::vector<uint32_t>& kept) {
    const size_t P = p->n_poses, G = p->n_intrinsics, N = p->n_points, O = p->n_observations;
    if (state) {
        out.rot = state->rot;
        out.cen = state->cen;
    } else {
        out.rot.assign(p->pose_rotation, p->pose_rotation + 9 * P);
        out.cen.assign(p->pose_center, p->pose_center + 3 * P);
    }
    out.var_rot.assign(p->pose_var_rotation, p->pose_var_rotation + P);
    out.var_trans.assign(p->pose_var_translation, p->pose_var_translation + P);
    out.intr.assign(p->intrinsics, p->intrinsics + G);
    if (state)
        for (size_t g = 0; g < G; ++g) {  // the solution's radial camera (fx = fy = f)
            ba_intrinsic& in = out.intr[g];
            in.fx = in.fy = state->focal[g];
            if (in.model >= BA_CAMERA_BROWN_K1 && in.model <= BA_CAMERA_BROWN_K1K2K3) in.k1 = state->k1[g];
            if (in.model >= BA_CAMERA_BROWN_K1K2 && in.model <= BA_CAMERA_BROWN_K1K2K3) in.k2 = state->k2[g];
        }
    std::vector<uint32_t> new_index(N, UINT32_MAX);
    kept.clear();
    out.xyz.clear();
    out.point_fixed.clear();
    for (size_t i = 0; i < N; ++i) {
        if (drop[i]) continue;
        new_index[i] = (uint32_t)kept.size();
        kept.push_back((uint32_t)i);
        const double* X = state ? &state->xyz[3 * i] : p->point_xyz + 3 * i;
        out.xyz.insert(out.xyz.end(), X, X + 3);
        out.point_fixed.push_back(p->point_fixed ? p->point_fixed[i] : 0);
    }
    out.obs_pose.clear();
    out.obs_intr.clear();
    out.obs_point.clear();
    out.pix.clear();
    for (size_t o = 0; o < O; ++o) {
        const uint32_t q = new_index[p->obs_point[o]];
        if (q == UINT32_MAX) continue;
        out.obs_pose.push_back(p->obs_pose[o]);
        out.obs_intr.push_back(p->obs_intrinsic[o]);
        out.obs_point.push_back(q);
        out.pix.push_back(p->obs_pixel[2 * o]);
        out.pix.push_back(p->obs_pixel[2 * o + 1]);
    }
    out.p = *p;  // residual space, loss and gauge carry over
    out.bind();
}

void symmetric_eigenvalues3(const double m[9], double out[3]) {
    double a[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) a[i][j] = 0.5 * (m[3 * i + j] + m[3 * j + i]);
    for (int sweep = 0; sweep < 64; ++sweep) {
        const double off = std::fabs(a[0][1]) + std::fabs(a[0][2]) + std::fabs(a[1][2]);
        if (!(off > 0.0)) break;  // diagonal (or NaN: the eigenvalues stay NaN below)
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                if (a[p][q] == 0.0) continue;
                const double theta = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
                const double t = (theta >= 0.0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < 3; ++k) {  // columns p, q
                    const double akp = a[k][p], akq = a[k][q];
                    a[k][p] = c * akp - s * akq;
                    a[k][q] = s * akp + c * akq;
                }
                for (int k = 0; k < 3; ++k) {  // rows p, q
                    const double apk = a[p][k], aqk = a[q][k];
                    a[p][k] = c * apk - s * aqk;
                    a[q][k] = s * apk + c * aqk;
                }
            }
    }
    out[0] = a[0][0];
    out[1] = a[1][1];
    out[2] = a[2][2];
    std::sort(out, out + 3);
}

namespace {

// first_radial_turn_squared (openmvg-core camera/distortion.rs): the first positive zero of
// 1 + 3 k1 q + 5 k2 q^2.
bool radial_turn_squared(double k1, double k2, double* root) { return first_turn(k1, k2, root); }

// CameraModel::project_with_jacobian of a pinhole or radial (k1, k2) Brown camera with fx = fy
// (openmvg-core camera/pinhole.rs:301-322, brown_conrady.rs:240-267, 334-360,
// distortion.rs:613-633): the pixel and d pixel / d camera point.
bool project_with_jacobian(const ba_intrinsic& in, double f, double k1, double k2, const double P[3], double px[2],
                           double J[2][3]) {
    for (int k = 0; k < 3; ++k)
        if (!std::isfinite(P[k])) return false;
    if (P[2] <= 0.0) return false;
    const bool brown = in.model != BA_CAMERA_PINHOLE;
    const double inv = 1.0 / P[2];
    if (!brown) {
        px[0] = std::fma(f, P[0] * inv, in.cx);
        px[1] = std::fma(f, P[1] * inv, in.cy);
        const double inv2 = inv * inv;
        J[0][0] = f * inv;
        J[0][1] = 0.0;
        J[0][2] = -f * P[0] * inv2;
        J[1][0] = 0.0;
        J[1][1] = f * inv;
        J[1][2] = -f * P[1] * inv2;
    } else {
        const double x = P[0] / P[2], y = P[1] / P[2];
        if (!std::isfinite(x) || !std::isfinite(y)) return false;
        auto jac = [&](double& j11, double& j12, double& j22, double& det) {
            const double r2 = std::fma(x, x, y * y);
            const double factor = 1.0 + r2 * (k1 + r2 * k2);
            const double deriv = k1 + 2.0 * (k2 * r2);
            j11 = factor + 2.0 * x * x * deriv;
            j12 = 2.0 * x * y * deriv;
            j22 = factor + 2.0 * y * y * deriv;
            det = std::fma(j11, j22, -j12 * j12);
            return std::isfinite(j11) && std::isfinite(j12) && std::isfinite(j22) && std::isfinite(det);
        };
        double j11, j12, j22, det;
        if (k1 != 0.0 || k2 != 0.0) {  // validate_forward_domain of a distorting camera
            const double r2 = std::fma(x, x, y * y);
            if (!std::isfinite(r2)) return false;
            double turn2;
            if (radial_turn_squared(k1, k2, &turn2) && r2 >= turn2) return false;
            if (!jac(j11, j12, j22, det) || det <= 1e-12) return false;
        }
        const double r2 = x * x + y * y;
        const double factor = 1.0 + r2 * (k1 + r2 * k2);
        const double xd = x * factor, yd = y * factor;
        px[0] = std::fma(f, xd, in.cx);
        px[1] = std::fma(f, yd, in.cy);
        if (!jac(j11, j12, j22, det)) return false;
        J[0][0] = f * j11 * inv;
        J[0][1] = f * j12 * inv;
        J[0][2] = -f * (j11 * x + j12 * y) * inv;
        J[1][0] = f * j12 * inv;
        J[1][1] = f * j22 * inv;
        J[1][2] = -f * (j12 * x + j22 * y) * inv;
    }
    if (!std::isfinite(px[0]) || !std::isfinite(px[1])) return false;
    for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 3; ++c)
            if (!std::isfinite(J[r][c])) return false;
    return true;
}

}  // namespace

std::vector<uint8_t> weak_points(const ba_problem* p, const HostSolution& s) {
    const size_t N = p->n_points, O = p->n_observations;
    const int loss = p->robust_loss;
    const double scale = loss == BA_LOSS_CAUCHY ? p->loss_scale / std::sqrt(2.0) : p->loss_scale;
    std::vector<double> block(9 * N, 0.0);
    std::vector<uint8_t> none(N, 0);
    for (size_t o = 0; o < O; ++o) {
        const uint32_t q = p->obs_point[o], c = p->obs_pose[o], g = p->obs_intrinsic[o];
        if (none[q]) continue;
        const double* R = &s.rot[9 * c];
        const double* C = &s.cen[3 * c];
        const double* X = &s.xyz[3 * q];
        const double d[3] = {X[0] - C[0], X[1] - C[1], X[2] - C[2]};
        double P[3];
        for (int i = 0; i < 3; ++i) P[i] = (R[3 * i] * d[0] + R[3 * i + 1] * d[1]) + R[3 * i + 2] * d[2];
        double px[2], J[2][3];
        if (!project_with_jacobian(p->intrinsics[g], s.focal[g], s.k1[g], s.k2[g], P, px, J)) {
            none[q] = 1;
            continue;
        }
        const double e0 = px[0] - p->obs_pixel[2 * o], e1 = px[1] - p->obs_pixel[2 * o + 1];
        const double r = std::sqrt(e0 * e0 + e1 * e1);
        if (!std::isfinite(r)) {
            none[q] = 1;
            continue;
        }
        double JR[2][3];  // J R
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 3; ++b) JR[a][b] = (J[a][0] * R[b] + J[a][1] * R[3 + b]) + J[a][2] * R[6 + b];
        const double w = loss_sqrt_weight(loss, scale, r);
        const double w2 = w * w;
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) block[9 * q + 3 * a + b] += w2 * (JR[0][a] * JR[0][b] + JR[1][a] * JR[1][b]);
    }
    std::vector<uint8_t> weak(N, 0);
    const std::vector<uint8_t> degenerate = degenerate_points(p, s.cen.data(), s.xyz.data());
    for (size_t i = 0; i < N; ++i) {
        const bool fixed = p->point_fixed && p->point_fixed[i];
        if (s.quarantined.size() == N && s.quarantined[i]) {
            weak[i] = 1;
            continue;
        }
        if (fixed) continue;
        if (degenerate[i] || none[i]) {
            weak[i] = 1;
            continue;
        }
        double ev[3];
        symmetric_eigenvalues3(&block[9 * i], ev);
        const double mx = std::max(std::max(ev[0], ev[1]), ev[2]);
        const double mn = std::min(std::min(ev[0], ev[1]), ev[2]);
        const double threshold = 1.0e-5 * mx;  // POINT_BLOCK_QUARANTINE_RCOND (point_block.rs:20)
        weak[i] = !(mn > threshold);           // NaN counts as weak
    }
    return weak;
}

bool needs_resolve(const ba_solve_report& r) {
    return r.weak_point_count > 0 || r.observability == BA_OBSERVABILITY_RANK_DEFICIENT ||
           r.status == BA_STATUS_DEGENERATE;
}

std::vector<uint32_t> unconstrained_groups(const ba_problem* p, const Lowered& L, const double* final_gp,
                                           const double* cov) {
    std::vector<uint32_t> out;
    for (size_t g = 0; g < L.n_groups(); ++g) {
        const uint32_t d = L.grp_dim[g];
        if (!d) continue;
        if (!cov) {
            out.push_back((uint32_t)g);
            continue;
        }
        const ba_intrinsic& in = p->intrinsics[g];
        const double f = final_gp[3 * g];
        const double rx = std::max(in.cx, (double)in.width - in.cx) / f;
        const double ry = std::max(in.cy, (double)in.height - in.cy) / f;
        const double r2 = rx * rx + ry * ry;
        const double scales[3] = {1.0, r2, r2 * r2};
        bool bad = false;
        for (uint32_t k = 0; k < d; ++k) {
            const double dev = std::sqrt(cov[9 * g + 4 * k]) * scales[k];
            bad |= std::isnan(dev) || (!in.certification_waived && dev > MAX_INTRINSIC_CORNER_DISPLACEMENT_STD);
        }
        if (bad) out.push_back((uint32_t)g);
    }
    return out;
}

void continue_report(ba_solve_report& later, const ba_solve_report& first) {
    later.initial_cost = first.initial_cost;
    later.initial_rmse = first.initial_rmse;
    later.iterations += first.iterations;
    later.accepted_steps += first.accepted_steps;
    later.rejected_steps += first.rejected_steps;
    later.rejected_linearization += first.rejected_linearization;
    later.rejected_step_solve += first.rejected_step_solve;
    later.rejected_candidate += first.rejected_candidate;
    later.rejected_insufficient_reduction += first.rejected_insufficient_reduction;
    later.pcg_iterations += first.pcg_iterations;
    later.seconds_setup += first.seconds_setup;
    later.seconds_iterations += first.seconds_iterations;
    later.seconds_final += first.seconds_final;
    later.weak_point_resolves += first.weak_point_resolves;
}

}  // namespace ba
