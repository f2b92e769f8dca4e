// ba_host.cpp — problem checks, lowering and index maps of the MI355X bundle adjuster.
//
// Follows crates/openmvg-ba/src:
//   problem.rs:492-576            BundleProblemBuilder::build checks
//   solver.rs:122-141             SolveConfig::validate
//   quarantine.rs:33-96           degenerate input points
//   native/mod.rs:260-340,457-497 quarantine pass, free-intrinsic test, one intrinsic per pose
//   native/gauge.rs:9-330         pose-variable mask, gauge anchor and scale
//   native/world_rebase.rs        anchor-frame conversion
//   native/self_calibration.rs    PreparedSelfCalibratingProblem::try_new, radial_camera,
//                                 free_radial_dimension
//   kernel/multi_view/sparse/workspace/topology.rs:351-397   ReducedLayout::try_radial
//   kernel/multi_view/sparse/workspace/pattern.rs:306-468    point parameters, block pattern
#include "ba_host.hpp"
#include "ba_kernels.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <numeric>
#include <thread>

namespace ba {
namespace {

struct M3 {
    double a[9];
};
inline M3 rot_of(const ba_problem* p, size_t i) {
    M3 m;
    std::memcpy(m.a, p->pose_rotation + 9 * i, sizeof(m.a));
    return m;
}
inline void center_of(const ba_problem* p, size_t i, double c[3]) {
    std::memcpy(c, p->pose_center + 3 * i, 3 * sizeof(double));
}
inline bool is_fixed_pose(const ba_problem* p, size_t i) {
    return !p->pose_var_rotation[i] && !p->pose_var_translation[i];
}
inline bool is_fixed_point(const ba_problem* p, size_t i) { return p->point_fixed && p->point_fixed[i]; }
inline double sq_dist(const double a[3], const double b[3]) {
    double d0 = a[0] - b[0], d1 = a[1] - b[1], d2 = a[2] - b[2];
    return d0 * d0 + d1 * d1 + d2 * d2;
}
// row-major a * b^T, with the product sums in index order
inline void mul_abt(const double* a, const double* b, double* o) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += a[3 * i + k] * b[3 * j + k];
            o[3 * i + j] = s;
        }
}
inline void mul_ab(const double* a, const double* b, double* o) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += a[3 * i + k] * b[3 * k + j];
            o[3 * i + j] = s;
        }
}
inline void mul_av(const double* a, const double* v, double* o) {
    for (int i = 0; i < 3; ++i) o[i] = (a[3 * i] * v[0] + a[3 * i + 1] * v[1]) + a[3 * i + 2] * v[2];
}
inline void mul_atv(const double* a, const double* v, double* o) {
    for (int i = 0; i < 3; ++i) o[i] = (a[i] * v[0] + a[3 + i] * v[1]) + a[6 + i] * v[2];
}

// Threads for n items of at least `grain` each, at most 16.
unsigned range_count(size_t n, size_t grain) {
    const size_t hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    return (unsigned)std::max<size_t>(1, std::min(hw, n / std::max<size_t>(grain, 1)));
}
// f(t, begin, end) over T contiguous ranges of [0, n), range t on thread t.  Callers combine
// per-range results in range order, so what they build does not depend on T.
template <class F>
void for_ranges(size_t n, unsigned T, F&& f) {
    if (T <= 1) {
        f(0u, size_t(0), n);
        return;
    }
    std::vector<std::thread> th;
    th.reserve(T);
    for (unsigned t = 0; t < T; ++t) th.emplace_back([&f, t, n, T] { f(t, n * t / T, n * (t + 1) / T); });
    for (auto& x : th) x.join();
}

// TrackGeometry of kernel/point_geometry.rs:29-101, the input-quarantine subset.
struct Geometry {
    double near2, far2;
    double first[3];
    size_t count = 0;
    bool has_first = false, parallax = false, invalid = false;
    explicit Geometry(double scene) : near2(INFINITY), far2(scene * scene) {}
    void push(const double o[3]) {
        double sq = o[0] * o[0] + o[1] * o[1] + o[2] * o[2];
        if (!(std::isfinite(sq) && sq > 0.0)) {
            invalid = true;
            return;
        }
        count += 1;
        near2 = std::fmin(near2, sq);
        far2 = std::fmax(far2, sq);
        double inv = 1.0 / std::sqrt(sq);
        double ray[3] = {o[0] * inv, o[1] * inv, o[2] * inv};
        if (!has_first) {
            has_first = true;
            std::memcpy(first, ray, sizeof(ray));
            return;
        }
        if (parallax) return;
        double dot = ray[0] * first[0] + ray[1] * first[1] + ray[2] * first[2];
        double c0 = ray[1] * first[2] - ray[2] * first[1], c1 = ray[2] * first[0] - ray[0] * first[2],
               c2 = ray[0] * first[1] - ray[1] * first[0];
        parallax = dot < 0.0 || c0 * c0 + c1 * c1 + c2 * c2 >= 1.7e-5 * 1.7e-5;
    }
    bool degenerate() const { return invalid || (count >= 2 && (!parallax || near2 < 1e-4 * 1e-4 * far2)); }
};

}  // namespace

std::vector<uint8_t> degenerate_points(const ba_problem* p, const double* centers, const double* xyz) {
    const size_t N = p->n_points, O = p->n_observations;
    std::vector<double> dist(O);
    auto offset = [&](size_t o, double v[3]) {
        const double* X = xyz + 3 * p->obs_point[o];
        const double* C = centers + 3 * p->obs_pose[o];
        for (int k = 0; k < 3; ++k) v[k] = X[k] - C[k];
    };
    for (size_t o = 0; o < O; ++o) {
        double v[3];
        offset(o, v);
        dist[o] = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    }
    double scene = 0.0;
    if (O) {
        std::nth_element(dist.begin(), dist.begin() + O / 2, dist.end());
        scene = dist[O / 2];
    }
    std::vector<Geometry> geo(N, Geometry(scene));
    for (size_t o = 0; o < O; ++o) {
        double v[3];
        offset(o, v);
        geo[p->obs_point[o]].push(v);
    }
    std::vector<uint8_t> out(N, 0);
    for (size_t i = 0; i < N; ++i) out[i] = geo[i].degenerate();
    return out;
}

Status validate_config(const ba_solve_config* c) {
    if (c->max_iterations == 0) return error(BA_ERR_INVALID_CONFIG, "max_iterations must be positive");
    if (!std::isfinite(c->function_tolerance) || c->function_tolerance <= 0.0)
        return error(BA_ERR_INVALID_CONFIG, "function_tolerance");
    if (!std::isfinite(c->gradient_tolerance) || c->gradient_tolerance <= 0.0)
        return error(BA_ERR_INVALID_CONFIG, "gradient_tolerance");
    if (!std::isfinite(c->parameter_tolerance) || c->parameter_tolerance <= 0.0)
        return error(BA_ERR_INVALID_CONFIG, "parameter_tolerance");
    if (c->linear_solver < BA_SOLVER_AUTO || c->linear_solver > BA_SOLVER_ITERATIVE)
        return error(BA_ERR_INVALID_CONFIG, "linear_solver");
    if (c->precision != BA_PRECISION_FP64 && c->precision != BA_PRECISION_MIXED)
        return error(BA_ERR_INVALID_CONFIG, "precision");
    if (!(c->pcg_relative_tolerance >= 0.0 && c->pcg_relative_tolerance < 1.0))
        return error(BA_ERR_INVALID_CONFIG, "pcg_relative_tolerance");
    return Status{};
}

Status validate_problem(const ba_problem* p) {
    if (!p) return error(BA_ERR_INVALID_PROBLEM, "null problem");
    if (p->n_poses == 0) return error(BA_ERR_INVALID_PROBLEM, "empty poses");
    if (p->n_intrinsics == 0) return error(BA_ERR_INVALID_PROBLEM, "empty intrinsics");
    if (p->n_points == 0) return error(BA_ERR_INVALID_PROBLEM, "empty points");
    if (p->n_observations == 0) return error(BA_ERR_INVALID_PROBLEM, "empty observations");
    for (size_t i = 0; i < p->n_poses; ++i) {
        const double* R = p->pose_rotation + 9 * i;
        const double* C = p->pose_center + 3 * i;
        bool ok = std::isfinite(C[0]) && std::isfinite(C[1]) && std::isfinite(C[2]);
        double rrt[9];
        mul_abt(R, R, rrt);
        for (int k = 0; k < 9 && ok; ++k)
            ok = std::fabs(rrt[k] - ((k % 4 == 0) ? 1.0 : 0.0)) <= 9.5367431640625e-7;
        if (!ok) return error(BA_ERR_INVALID_PROBLEM, "invalid pose " + std::to_string(i));
    }
    for (size_t i = 0; i < 3 * p->n_points; ++i)
        if (!std::isfinite(p->point_xyz[i])) return error(BA_ERR_INVALID_PROBLEM, "non-finite point");
    for (size_t o = 0; o < p->n_observations; ++o) {
        if (p->obs_pose[o] >= p->n_poses) return error(BA_ERR_INVALID_PROBLEM, "observation pose out of bounds");
        if (p->obs_point[o] >= p->n_points) return error(BA_ERR_INVALID_PROBLEM, "observation point out of bounds");
        if (p->obs_intrinsic[o] >= p->n_intrinsics)
            return error(BA_ERR_INVALID_PROBLEM, "observation intrinsic out of bounds");
        if (!std::isfinite(p->obs_pixel[2 * o]) || !std::isfinite(p->obs_pixel[2 * o + 1]))
            return error(BA_ERR_INVALID_PROBLEM, "non-finite observation");
    }
    if (p->robust_loss != BA_LOSS_SQUARED && p->residual_space != BA_RESIDUAL_PIXELS)
        return error(BA_ERR_INVALID_PROBLEM, "loss unit mismatch");
    if (p->robust_loss != BA_LOSS_SQUARED && !(std::isfinite(p->loss_scale) && p->loss_scale > 0.0))
        return error(BA_ERR_INVALID_PROBLEM, "invalid loss scale");
    if (p->gauge == BA_GAUGE_FIXED_BASELINE) {
        if (p->gauge_first >= p->n_poses || p->gauge_second >= p->n_poses)
            return error(BA_ERR_INVALID_PROBLEM, "gauge pose out of bounds");
        if (p->gauge_first == p->gauge_second) return error(BA_ERR_INVALID_PROBLEM, "gauge uses one pose");
        if (!std::isfinite(p->gauge_distance) || p->gauge_distance <= 0.0)
            return error(BA_ERR_INVALID_PROBLEM, "invalid baseline distance");
    }
    return Status{};
}

// Camera and intrinsic blocks of the reduced system (topology.rs:351-397) and the number of
// point parameters (each free point's distinct active camera parameters), from the cameras'
// and groups' fixed flags.
void build_layout(Lowered& L) {
    const size_t P = L.n_cameras(), G = L.n_groups(), NP = L.n_points();
    L.cam_off.assign(P, 0);
    L.cam_dim.assign(P, 0);
    size_t D = 0;
    for (size_t c = 0; c < P; ++c) {
        uint32_t local = ((c == 0 && !L.anchor_free) || L.cam_fixed[c]) ? 0u : (L.gauge_camera == (int)c ? 5u : 6u);
        L.cam_off[c] = (uint32_t)D;
        L.cam_dim[c] = local;
        D += local;
    }
    L.grp_off.assign(G, 0);
    L.grp_dim.assign(G, 0);
    for (size_t g = 0; g < G; ++g) {
        uint32_t local = L.grp_fixed[g] ? 0u : L.grp_free[g];
        L.grp_off[g] = (uint32_t)D;
        L.grp_dim[g] = local;
        D += local;
    }
    L.D = D;
    {
        const size_t NB = P + G;
        const unsigned T = range_count(NP, 1 << 14);
        std::vector<uint64_t> part(T, 0);
        for_ranges(NP, T, [&](unsigned t, size_t b, size_t e) {
            std::vector<uint32_t> seen(NB, UINT32_MAX);  // last point that counted a block
            uint64_t n = 0;
            for (size_t q = b; q < e; ++q) {
                if (L.point_fixed[q]) continue;
                for (uint32_t o = L.point_off[q]; o < L.point_off[q + 1]; ++o) {
                    const uint32_t c = L.obs_cam[o], g = L.cam_group[c];
                    if (seen[c] != (uint32_t)q) n += L.cam_dim[c];
                    seen[c] = (uint32_t)q;
                    if (seen[P + g] != (uint32_t)q) n += L.grp_dim[g];
                    seen[P + g] = (uint32_t)q;
                }
            }
            part[t] = n;
        });
        L.n_cross = 0;
        for (uint64_t n : part) L.n_cross += n;
    }
}

void hold_groups(Lowered& L, const std::vector<uint32_t>& groups) {
    for (uint32_t g : groups)
        if (g < L.n_groups()) L.grp_fixed[g] = 1;
    build_layout(L);
}

Status lower_problem(const ba_problem* p, Lowered& L, bool build_pattern, bool quarantine) {
    const size_t P = p->n_poses, N = p->n_points, O = p->n_observations, G = p->n_intrinsics;
    // ---- input quarantine (quarantine.rs:33-96) ----
    L.n_input_points = N;
    L.quarantined.assign(N, 0);
    // observations grouped by input point, in input order inside each point (a counting sort)
    std::vector<uint32_t> by_off(N + 1, 0), by_pt(O);
    for (size_t o = 0; o < O; ++o) by_off[p->obs_point[o] + 1]++;
    for (size_t i = 0; i < N; ++i) by_off[i + 1] += by_off[i];
    {
        std::vector<uint32_t> cur(by_off.begin(), by_off.end() - 1);
        for (size_t o = 0; o < O; ++o) by_pt[cur[p->obs_point[o]]++] = (uint32_t)o;
    }
    if (quarantine) {
        auto offset = [&](size_t o, double v[3]) {
            const double* X = p->point_xyz + 3 * p->obs_point[o];
            const double* C = p->pose_center + 3 * p->obs_pose[o];
            for (int k = 0; k < 3; ++k) v[k] = X[k] - C[k];
        };
        std::vector<double> dist(O);
        for_ranges(O, range_count(O, 1 << 16), [&](unsigned, size_t b, size_t e) {
            for (size_t o = b; o < e; ++o) {
                double v[3];
                offset(o, v);
                dist[o] = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
            }
        });
        size_t mid = O / 2;
        std::nth_element(dist.begin(), dist.begin() + mid, dist.end());
        const double scene = dist[mid];
        // each point's rays pushed in input order, as a pass over the observations would
        for_ranges(N, range_count(N, 1 << 14), [&](unsigned, size_t b, size_t e) {
            for (size_t i = b; i < e; ++i) {
                Geometry geo(scene);
                for (uint32_t k = by_off[i]; k < by_off[i + 1]; ++k) {
                    double v[3];
                    offset(by_pt[k], v);
                    geo.push(v);
                }
                L.quarantined[i] = (by_off[i + 1] - by_off[i] < 2 || geo.degenerate()) && !is_fixed_point(p, i);
            }
        });
    }
    L.n_quarantined = 0;
    for (auto q : L.quarantined) L.n_quarantined += q;
    if (L.n_quarantined == N)
        return error(BA_ERR_UNSUPPORTED,
                     "no point can be adjusted: every point is seen fewer than twice or without parallax");
    if (p->residual_space != BA_RESIDUAL_PIXELS)
        return error(BA_ERR_UNSUPPORTED, "self-calibration requires pixel residuals");
    std::vector<uint32_t> kept_obs;
    kept_obs.reserve(O);
    for (size_t o = 0; o < O; ++o)
        if (!L.quarantined[p->obs_point[o]]) kept_obs.push_back((uint32_t)o);
    // ---- native/mod.rs:318-335, 457-497: the self-calibrating or the calibrated path ----
    bool free_intrinsic = false;
    for (uint32_t o : kept_obs) free_intrinsic |= !is_fixed_intrinsic(p->intrinsics[p->obs_intrinsic[o]]);
    L.calibrated = !free_intrinsic;
    L.anchor_free = false;
    std::vector<int64_t> pose_intrinsic(P, -1);
    std::vector<uint8_t> observed(P, 0);
    for (uint32_t o : kept_obs) {
        int64_t& slot = pose_intrinsic[p->obs_pose[o]];
        if (slot >= 0 && slot != (int64_t)p->obs_intrinsic[o])
            return error(BA_ERR_UNSUPPORTED, L.calibrated
                                                 ? "calibrated pixel BA here requires one camera model per observed pose"
                                                 : "native self-calibration requires one intrinsic group per observed pose");
        slot = p->obs_intrinsic[o];
        observed[p->obs_pose[o]] = 1;
    }
    for (auto& s : pose_intrinsic)
        if (s < 0) s = 0;
    if (L.calibrated) {
        // PreparedCalibratedProblem::try_new (native/calibrated.rs:200-226): fixed pinhole and
        // Brown-Conrady cameras; this path covers the radial k1/k2 models with one focal length
        for (uint32_t o : kept_obs) {
            const int m = p->intrinsics[p->obs_intrinsic[o]].model;
            if (m == BA_CAMERA_FISHEYE || m == BA_CAMERA_SPHERICAL)
                return error(BA_ERR_UNSUPPORTED,
                             "pixel BA supports fixed pinhole and Brown-Conrady cameras; use angular residuals for "
                             "wide-angle cameras");
        }
        // can_use_two_view (native/two_view.rs:14-50): the reference's two-view solver is declined
        if (P == 2) {
            size_t kept_points = 0;
            bool any_fixed_point = false;
            for (size_t i = 0; i < N; ++i)
                if (!L.quarantined[i]) {
                    kept_points += 1;
                    any_fixed_point |= is_fixed_point(p, i);
                }
            bool pinholes = true;
            for (size_t i = 0; i < P; ++i) {
                const ba_intrinsic& in = p->intrinsics[pose_intrinsic[i]];
                const bool distortion = in.model != BA_CAMERA_PINHOLE &&
                                        ((in.model >= BA_CAMERA_BROWN_K1 && in.k1 != 0.0) ||
                                         (in.model >= BA_CAMERA_BROWN_K1K2 && in.k2 != 0.0) ||
                                         (in.model >= BA_CAMERA_BROWN_K1K2K3 && in.k3 != 0.0) || in.p1 != 0.0 ||
                                         in.p2 != 0.0);
                pinholes &= in.model != BA_CAMERA_FISHEYE && in.model != BA_CAMERA_SPHERICAL && !distortion;
            }
            std::vector<std::pair<uint32_t, uint32_t>> pp;
            pp.reserve(kept_obs.size());
            for (uint32_t o : kept_obs) pp.emplace_back(p->obs_pose[o], p->obs_point[o]);
            std::sort(pp.begin(), pp.end());
            const bool repeated = std::adjacent_find(pp.begin(), pp.end()) != pp.end();
            if (!any_fixed_point && kept_obs.size() == 2 * kept_points && !repeated && pinholes) {
                int64_t a = -1;
                if (p->gauge == BA_GAUGE_FIXED_BASELINE)
                    a = p->gauge_first;
                else
                    for (size_t i = 0; i < 2 && a < 0; ++i)
                        if (is_fixed_pose(p, i)) a = (int64_t)i;
                if (a >= 0 && is_fixed_pose(p, (size_t)a) && p->pose_var_rotation[1 - a] && p->pose_var_translation[1 - a])
                    return error(BA_ERR_UNSUPPORTED, "two-view refinement (native/two_view.rs) is not this adjuster's");
            }
        }
        for (size_t i = 0; i < P; ++i)
            if (observed[i] && !is_fixed_pose(p, i) && !(p->pose_var_rotation[i] && p->pose_var_translation[i]))
                return error(BA_ERR_UNSUPPORTED, "calibrated pixel BA here requires full-pose or fixed camera blocks");
    } else {
        // ---- common_pose_variables (gauge.rs:9-40) ----
        for (size_t i = 0; i < P; ++i) {
            if (!observed[i] || is_fixed_pose(p, i)) continue;
            if (!p->pose_var_translation[i])
                return error(BA_ERR_UNSUPPORTED, "direct native lowering cannot optimize rotation without translation");
            if (!p->pose_var_rotation[i])
                return error(BA_ERR_UNSUPPORTED, "self-calibration requires full-pose or fixed camera blocks");
        }
        // ---- intrinsic checks (native/self_calibration.rs:63-85) ----
        for (size_t g = 0; g < G; ++g) {
            const ba_intrinsic& in = p->intrinsics[g];
            if (in.var_radial && in.model == BA_CAMERA_BROWN_K1K2K3)
                return error(BA_ERR_UNSUPPORTED, "native self-calibration refines at most two declared radial coefficients");
            if (!is_fixed_intrinsic(in) && (!in.var_focal || in.var_principal_point || in.var_tangential))
                return error(BA_ERR_UNSUPPORTED,
                             "self-calibration refines the focal length and radial coefficients with a fixed principal point");
        }
    }
    // ---- gauge_plan (gauge.rs:84-107, 231-330) ----
    bool observes_fixed_point = false;
    for (uint32_t o : kept_obs) observes_fixed_point |= is_fixed_point(p, p->obs_point[o]);
    size_t anchor = 0;
    double scale = 1.0;
    int64_t gauge_pose = -1;
    if (L.calibrated && p->gauge == BA_GAUGE_FREE) {
        // calibrated_gauge_plan (gauge.rs:123-229) for full-pose or fixed cameras: its held
        // centres and known rotations are then exactly the fixed observed cameras
        int64_t first_observed = -1, fixed_anchor = -1;
        for (size_t i = 0; i < P; ++i) {
            if (observed[i] && first_observed < 0) first_observed = (int64_t)i;
            if (observed[i] && is_fixed_pose(p, i) && fixed_anchor < 0) fixed_anchor = (int64_t)i;
        }
        if (first_observed < 0) return error(BA_ERR_INTERNAL, "checked bundle has no observed pose");
        anchor = (size_t)(fixed_anchor >= 0 ? fixed_anchor : first_observed);
        double ca[3];
        center_of(p, anchor, ca);
        if (observes_fixed_point && fixed_anchor < 0) {
            L.anchor_free = true;  // survey controls fix the frame; the anchor only sets the native frame
        } else {
            bool scale_set = false;
            for (size_t i = 0; i < P && !scale_set; ++i) {
                double c[3];
                center_of(p, i, c);
                scale_set = observed[i] && is_fixed_pose(p, i) && sq_dist(c, ca) > 0.0;
            }
            for (size_t k = 0; k < kept_obs.size() && !scale_set; ++k) {
                size_t q = p->obs_point[kept_obs[k]];
                scale_set = is_fixed_point(p, q) && sq_dist(p->point_xyz + 3 * q, ca) > 0.0;
            }
            if (!scale_set) {
                double best = -1.0;
                for (size_t i = 0; i < P; ++i) {
                    if (i == anchor || !observed[i] || is_fixed_pose(p, i)) continue;
                    double c[3];
                    center_of(p, i, c);
                    double d = sq_dist(c, ca);
                    if (std::isfinite(d) && d > best) {
                        best = d;
                        gauge_pose = (int64_t)i;
                    }
                }
                if (gauge_pose < 0 || !(best > 0.0))
                    return error(BA_ERR_UNSUPPORTED, "free-scale BA has no nonzero active baseline");
            }
        }
    } else if (p->gauge == BA_GAUGE_FREE) {
        if (observes_fixed_point) {
            bool fixed_observed = false;
            for (size_t i = 0; i < P; ++i) fixed_observed |= observed[i] && is_fixed_pose(p, i);
            if (!fixed_observed)
                return error(BA_ERR_UNSUPPORTED,
                             "native BA with fixed world points requires an explicitly fixed observed pose or fixed-baseline anchor");
        }
        int64_t first_observed = -1, first_fixed = -1;
        for (size_t i = 0; i < P && first_fixed < 0; ++i) {
            if (observed[i] && is_fixed_pose(p, i)) first_fixed = (int64_t)i;
            if (observed[i] && first_observed < 0) first_observed = (int64_t)i;
        }
        anchor = (size_t)(first_fixed >= 0 ? first_fixed : (first_observed >= 0 ? first_observed : 0));
        double ca[3];
        center_of(p, anchor, ca);
        bool scale_set = false;
        for (size_t i = 0; i < P && !scale_set; ++i) {
            double c[3];
            center_of(p, i, c);
            scale_set = i != anchor && observed[i] && is_fixed_pose(p, i) && sq_dist(c, ca) > 0.0;
        }
        for (size_t k = 0; k < kept_obs.size() && !scale_set; ++k) {
            size_t q = p->obs_point[kept_obs[k]];
            scale_set = is_fixed_point(p, q) && sq_dist(p->point_xyz + 3 * q, ca) > 0.0;
        }
        if (!scale_set) {  // farthest_camera: largest squared distance, ties to the lower index
            double best = -1.0;
            for (size_t i = 0; i < P; ++i) {
                if (i == anchor || !observed[i] || is_fixed_pose(p, i)) continue;
                double c[3];
                center_of(p, i, c);
                double d = sq_dist(c, ca);
                if (std::isfinite(d) && d > best) {
                    best = d;
                    gauge_pose = (int64_t)i;
                }
            }
            if (gauge_pose < 0 || !(best > 0.0))
                return error(BA_ERR_UNSUPPORTED, "free-scale BA has no nonzero active baseline");
        }
    } else {
        size_t first = p->gauge_first, second = p->gauge_second;
        double c1[3], c2[3];
        center_of(p, first, c1);
        center_of(p, second, c2);
        double current = std::sqrt(sq_dist(c2, c1));
        if (!std::isfinite(current) || current <= 0.0)
            return error(BA_ERR_UNSUPPORTED, "fixed-baseline cameras have no usable separation");
        double raw = p->gauge_distance / current;
        double tol = 256.0 * DBL_EPSILON * std::fmax(std::fmax(std::fabs(current), std::fabs(p->gauge_distance)), 1.0);
        scale = std::fabs(current - p->gauge_distance) <= tol ? 1.0 : raw;
        if (scale != 1.0) {
            for (size_t i = 0; i < P; ++i)
                if (i != first && !p->pose_var_translation[i])
                    return error(BA_ERR_UNSUPPORTED, "fixed-baseline scaling would move fixed state");
            for (size_t i = 0; i < N; ++i)
                if (is_fixed_point(p, i)) return error(BA_ERR_UNSUPPORTED, "fixed-baseline scaling would move fixed state");
            if (L.n_quarantined > 0)
                return error(BA_ERR_UNSUPPORTED, "native BA cannot change its physical baseline while holding out points");
        }
        gauge_pose = p->pose_var_translation[second] ? (int64_t)second : -1;
        anchor = first;
    }
    // ---- world rebase (world_rebase.rs:19-59) ----
    std::memcpy(L.anchor_rotation, p->pose_rotation + 9 * anchor, sizeof(L.anchor_rotation));
    center_of(p, anchor, L.anchor_center);
    L.rebase_scale = scale;
    L.camera_source.assign(1, (uint32_t)anchor);
    for (size_t i = 0; i < P; ++i)
        if (i != anchor) L.camera_source.push_back((uint32_t)i);
    std::vector<uint32_t> dest(P);
    for (size_t d = 0; d < P; ++d) dest[L.camera_source[d]] = (uint32_t)d;
    L.R.assign(9 * P, 0.0);
    L.t.assign(3 * P, 0.0);
    for (size_t d = 0; d < P; ++d) {
        size_t s = L.camera_source[d];
        const double* R = p->pose_rotation + 9 * s;
        const double* C = p->pose_center + 3 * s;
        if (std::memcmp(R, L.anchor_rotation, sizeof(L.anchor_rotation)) == 0 && C[0] == L.anchor_center[0] &&
            C[1] == L.anchor_center[1] && C[2] == L.anchor_center[2]) {
            const double I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
            std::memcpy(&L.R[9 * d], I, sizeof(I));
            continue;
        }
        mul_abt(R, L.anchor_rotation, &L.R[9 * d]);
        double diff[3] = {L.anchor_center[0] - C[0], L.anchor_center[1] - C[1], L.anchor_center[2] - C[2]};
        double rd[3];
        mul_av(R, diff, rd);
        for (int k = 0; k < 3; ++k) L.t[3 * d + k] = rd[k] * scale;
    }
    // ---- points without the quarantined (problem.rs:650-701) ----
    std::vector<uint32_t> new_index(N, UINT32_MAX);
    L.point_source.clear();
    for (size_t i = 0; i < N; ++i)
        if (!L.quarantined[i]) {
            new_index[i] = (uint32_t)L.point_source.size();
            L.point_source.push_back((uint32_t)i);
        }
    const size_t NP = L.point_source.size();
    L.X.assign(3 * NP, 0.0);
    L.point_fixed.assign(NP, 0);
    for (size_t q = 0; q < NP; ++q) {
        const double* Xi = p->point_xyz + 3 * L.point_source[q];
        double diff[3] = {Xi[0] - L.anchor_center[0], Xi[1] - L.anchor_center[1], Xi[2] - L.anchor_center[2]};
        double rd[3];
        mul_av(L.anchor_rotation, diff, rd);
        for (int k = 0; k < 3; ++k) L.X[3 * q + k] = rd[k] * scale;
        L.point_fixed[q] = is_fixed_point(p, L.point_source[q]);
    }
    // ---- radial groups (radial_camera :483-512, free_radial_dimension :442-451) ----
    L.gp.assign(3 * G, 0.0);
    L.grp_fixed.assign(G, 0);
    L.grp_free.assign(G, 1);
    std::vector<uint8_t> used(G, 0);
    for (uint32_t o : kept_obs) used[p->obs_intrinsic[o]] = 1;
    for (size_t g = 0; g < G; ++g) {
        const ba_intrinsic& in = p->intrinsics[g];
        if (L.calibrated) {
            // fixed cameras: the radial projection with f = fx and the declared k1, k2; the models
            // of unobserved intrinsics never enter the objective
            L.grp_fixed[g] = 1;
            L.gp[3 * g] = 1.0;
            if (!used[g]) continue;
            const bool brown = in.model != BA_CAMERA_PINHOLE;
            const double k3 = brown && in.model >= BA_CAMERA_BROWN_K1K2K3 ? in.k3 : 0.0;
            if (in.fx != in.fy || k3 != 0.0 || (brown && (in.p1 != 0.0 || in.p2 != 0.0)))
                return error(BA_ERR_UNSUPPORTED,
                             "calibrated pixel BA here supports one focal length and k1+k2 radial distortion without "
                             "tangential terms");
            L.gp[3 * g] = in.fx;
            L.gp[3 * g + 1] = brown && in.model >= BA_CAMERA_BROWN_K1 ? in.k1 : 0.0;
            L.gp[3 * g + 2] = brown && in.model >= BA_CAMERA_BROWN_K1K2 ? in.k2 : 0.0;
            continue;
        }
        if (in.model == BA_CAMERA_FISHEYE || in.model == BA_CAMERA_SPHERICAL)
            return error(BA_ERR_UNSUPPORTED, "self-calibration supports perspective cameras only");
        if (in.model < BA_CAMERA_PINHOLE || in.model > BA_CAMERA_SPHERICAL)
            return error(BA_ERR_INVALID_PROBLEM, "unknown camera model");
        bool brown = in.model != BA_CAMERA_PINHOLE;
        double k1 = brown && in.model >= BA_CAMERA_BROWN_K1 ? in.k1 : 0.0;
        double k2 = brown && in.model >= BA_CAMERA_BROWN_K1K2 ? in.k2 : 0.0;
        double k3 = brown && in.model >= BA_CAMERA_BROWN_K1K2K3 ? in.k3 : 0.0;
        double p1 = brown ? in.p1 : 0.0, p2 = brown ? in.p2 : 0.0;
        if (in.fx != in.fy || k3 != 0.0 || p1 != 0.0 || p2 != 0.0)
            return error(BA_ERR_UNSUPPORTED,
                         "self-calibration requires isotropic focal length, k1+k2 radial distortion, and no tangential distortion");
        L.gp[3 * g] = in.fx;
        L.gp[3 * g + 1] = k1;
        L.gp[3 * g + 2] = k2;
        L.grp_fixed[g] = is_fixed_intrinsic(in);
        if (brown && in.var_radial)
            L.grp_free[g] = in.model == BA_CAMERA_BROWN_NONE ? 1 : (in.model == BA_CAMERA_BROWN_K1 ? 2 : 3);
    }
    L.cam_group.assign(P, 0);
    L.cam_fixed.assign(P, 0);
    for (size_t d = 0; d < P; ++d) {
        size_t s = L.camera_source[d];
        L.cam_group[d] = (uint32_t)pose_intrinsic[s];
        L.cam_fixed[d] = (s == anchor && !L.anchor_free) || is_fixed_pose(p, s);
    }
    L.gauge_camera = gauge_pose < 0 ? -1 : (int)dest[gauge_pose];
    // ---- observations sorted by (point, input index), principal point subtracted ----
    // kept observations point by point (new point indices ascend with the input's), input order
    // inside each point: the stable order by point
    std::vector<uint32_t> order;
    order.reserve(kept_obs.size());
    for (size_t i = 0; i < N; ++i)
        if (!L.quarantined[i]) order.insert(order.end(), by_pt.begin() + by_off[i], by_pt.begin() + by_off[i + 1]);
    const size_t NO = order.size();
    L.obs_cam.resize(NO);
    L.obs_point.resize(NO);
    L.uv.resize(2 * NO);
    for_ranges(NO, range_count(NO, 1 << 16), [&](unsigned, size_t b, size_t e) {
        for (size_t k = b; k < e; ++k) {
            uint32_t s = order[k];
            const ba_intrinsic& in = p->intrinsics[p->obs_intrinsic[s]];
            L.obs_cam[k] = dest[p->obs_pose[s]];
            L.obs_point[k] = new_index[p->obs_point[s]];
            L.uv[2 * k] = p->obs_pixel[2 * s] - in.cx;
            L.uv[2 * k + 1] = p->obs_pixel[2 * s + 1] - in.cy;
        }
    });
    L.obs_source = std::move(order);
    L.point_off.assign(NP + 1, 0);
    for (size_t k = 0; k < NO; ++k) L.point_off[L.obs_point[k] + 1] += 1;
    for (size_t q = 0; q < NP; ++q) L.point_off[q + 1] += L.point_off[q];
    // ---- loss (problem.rs:372-382) ----
    L.loss = p->robust_loss;
    L.loss_scale = p->robust_loss == BA_LOSS_CAUCHY ? p->loss_scale / std::sqrt(2.0) : p->loss_scale;
    // ---- scene scale at the initial state (optimizer.rs:620-627) ----
    if (L.gauge_camera >= 0) {
        const double* tg = &L.t[3 * L.gauge_camera];
        L.scene_scale = std::sqrt(tg[0] * tg[0] + tg[1] * tg[1] + tg[2] * tg[2]);
    } else if (L.calibrated) {
        // observed_scene_scale (optimizer.rs:629-668): the calibrated workspace carries explicit
        // pose masks, so the length comes from observed points, by a scaled running sum of squares
        double sc = 0.0, sqs = 1.0;
        size_t count = 0;
        for (size_t q = 0; q < NP; ++q) {
            if (L.point_off[q] == L.point_off[q + 1]) continue;
            count += 1;
            for (int k = 0; k < 3; ++k) {
                const double m = std::fabs(L.X[3 * q + k]);
                if (m == 0.0) continue;
                if (sc < m) {
                    const double ratio = sc / m;
                    sqs = 1.0 + sqs * ratio * ratio;
                    sc = m;
                } else {
                    const double ratio = m / sc;
                    sqs += ratio * ratio;
                }
            }
        }
        if (sc == 0.0 || count == 0) {
            L.scene_scale = DBL_MIN;
        } else {
            const double rms = sc * std::sqrt(sqs / (double)count);
            L.scene_scale = std::isfinite(rms) ? std::fmax(rms, DBL_MIN) : sc;
        }
    } else {
        double s = 0.0;
        for (size_t q = 0; q < NP; ++q)
            s += L.X[3 * q] * L.X[3 * q] + L.X[3 * q + 1] * L.X[3 * q + 1] + L.X[3 * q + 2] * L.X[3 * q + 2];
        L.scene_scale = std::fmax(std::sqrt(s / (double)NP), DBL_MIN);
    }
    build_layout(L);
    if (!build_pattern) return Status{};
    // ---- per-point active parameters, first-encounter order (pattern.rs:306-379) ----
    L.cross_off.assign(NP + 1, 0);
    L.params.clear();
    {
        std::vector<uint8_t> mark(L.D, 0);
        size_t start = 0;
        for (size_t q = 0; q < NP; ++q) {
            if (!L.point_fixed[q])
                for (uint32_t o = L.point_off[q]; o < L.point_off[q + 1]; ++o) {
                    uint32_t c = L.obs_cam[o], g = L.cam_group[c];
                    for (uint32_t k = 0; k < L.cam_dim[c]; ++k)
                        if (!mark[L.cam_off[c] + k]) {
                            mark[L.cam_off[c] + k] = 1;
                            L.params.push_back(L.cam_off[c] + k);
                        }
                    for (uint32_t k = 0; k < L.grp_dim[g]; ++k)
                        if (!mark[L.grp_off[g] + k]) {
                            mark[L.grp_off[g] + k] = 1;
                            L.params.push_back(L.grp_off[g] + k);
                        }
                }
            for (size_t k = start; k < L.params.size(); ++k) mark[L.params[k]] = 0;
            start = L.params.size();
            L.cross_off[q + 1] = (uint32_t)start;
        }
    }
    // ---- reduced block pattern (pattern.rs:381-468) ----
    const size_t NB = P + G;
    auto bdim = [&](size_t b) { return b < P ? L.cam_dim[b] : L.grp_dim[b - P]; };
    auto boff = [&](size_t b) { return b < P ? L.cam_off[b] : L.grp_off[b - P]; };
    std::vector<std::vector<uint32_t>> rows(NB);
    for (size_t b = 0; b < NB; ++b)
        if (bdim(b)) rows[b].push_back((uint32_t)b);
    std::vector<uint32_t> active;
    for (size_t q = 0; q < NP; ++q) {
        active.clear();
        for (uint32_t o = L.point_off[q]; o < L.point_off[q + 1]; ++o) {
            uint32_t c = L.obs_cam[o], g = L.cam_group[c];
            if (L.point_fixed[q]) {
                if (L.cam_dim[c] && L.grp_dim[g]) rows[c].push_back((uint32_t)(P + g));
                continue;
            }
            if (L.cam_dim[c]) active.push_back(c);
            if (L.grp_dim[g]) active.push_back((uint32_t)(P + g));
        }
        std::sort(active.begin(), active.end());
        active.erase(std::unique(active.begin(), active.end()), active.end());
        for (size_t r = 0; r < active.size(); ++r)
            for (size_t c = 0; c <= r; ++c) rows[active[c]].push_back(active[r]);
    }
    L.col_ptr.assign(1, 0);
    L.row_idx.clear();
    for (size_t cb = 0; cb < NB; ++cb) {
        auto& rb = rows[cb];
        std::sort(rb.begin(), rb.end());
        rb.erase(std::unique(rb.begin(), rb.end()), rb.end());
        for (uint32_t cl = 0; cl < bdim(cb); ++cl) {
            for (uint32_t r : rb) {
                for (uint32_t k = (r == cb ? cl : 0u); k < bdim(r); ++k) L.row_idx.push_back(boff(r) + k);
            }
            L.col_ptr.push_back((uint32_t)L.row_idx.size());
        }
    }
    return Status{};
}

BlockLists build_block_lists(const Lowered& L) {
    BlockLists B;
    const size_t P = L.n_cameras(), G = L.n_groups();
    B.n_blocks = P + G;
    B.block_off.resize(B.n_blocks);
    B.block_dim.resize(B.n_blocks);
    for (size_t c = 0; c < P; ++c) {
        B.block_off[c] = L.cam_off[c];
        B.block_dim[c] = L.cam_dim[c];
    }
    for (size_t g = 0; g < G; ++g) {
        B.block_off[P + g] = L.grp_off[g];
        B.block_dim[P + g] = L.grp_dim[g];
    }
    // counting sort by block keeps the observation index ascending inside each block
    std::vector<uint32_t> count(B.n_blocks + 1, 0);
    const size_t NO = L.n_obs();
    for (size_t o = 0; o < NO; ++o) {
        uint32_t c = L.obs_cam[o];
        if (L.cam_dim[c]) count[c + 1]++;
        uint32_t g = L.cam_group[c];
        if (L.grp_dim[g]) count[P + g + 1]++;
    }
    for (size_t b = 0; b < B.n_blocks; ++b) count[b + 1] += count[b];
    B.obs_off = count;
    B.obs.resize(count[B.n_blocks]);
    std::vector<uint32_t> cursor(count.begin(), count.end() - 1);
    for (size_t o = 0; o < NO; ++o) {
        uint32_t c = L.obs_cam[o];
        if (L.cam_dim[c]) B.obs[cursor[c]++] = (uint32_t)o;
        uint32_t g = L.cam_group[c];
        if (L.grp_dim[g]) B.obs[cursor[P + g]++] = (uint32_t)o;
    }
    return B;
}

namespace {

// The terms of point q in (i, j, x, y) order: emit(A, B, i, j | flags) for block A of observation
// i (x-th of its blocks) and block B of observation j (y-th), A >= B.
template <class Emit>
void point_terms(const Lowered& L, size_t q, bool all, Emit&& emit) {
    const size_t P = L.n_cameras();
    auto blocks_of = [&](uint32_t o, uint32_t out_b[2]) {
        uint32_t c = L.obs_cam[o], g = L.cam_group[c];
        int n = 0;
        if (L.cam_dim[c]) out_b[n++] = c;
        if (L.grp_dim[g]) out_b[n++] = (uint32_t)(P + g);
        return n;
    };
    const uint32_t b0 = L.point_off[q], b1 = L.point_off[q + 1];
    const bool fixed = L.point_fixed[q];
    for (uint32_t i = b0; i < b1; ++i) {
        uint32_t bi[2];
        const int ni = blocks_of(i, bi);
        for (uint32_t j = b0; j < b1; ++j) {
            if (fixed && i != j) continue;
            uint32_t bj[2];
            const int nj = blocks_of(j, bj);
            for (int x = 0; x < ni; ++x)
                for (int y = 0; y < nj; ++y) {
                    const uint32_t A = bi[x], Bb = bj[y];
                    if (A < Bb) continue;
                    if (!all) {
                        const bool diag = A == Bb;
                        const bool edge = i == j && A >= P && Bb < P && L.cam_group[Bb] == A - P;
                        if (!diag && !edge) continue;
                        if (edge && !diag) {  // edges carry only the U Gram term
                            emit(A, Bb, i, j | TERM_U);
                            continue;
                        }
                    }
                    const uint32_t flags = (fixed ? 0u : TERM_SCHUR) | (i == j ? TERM_U : 0u);
                    if (flags) emit(A, Bb, i, j | flags);
                }
        }
    }
}

// Every block pair of the pattern: terms keyed by A * NB + B, bucketed by a stable counting
// sort over the distinct keys ascending.
PairLists all_pairs(const Lowered& L, const BlockLists& B) {
    PairLists out;
    const size_t NB = B.n_blocks, NP = L.n_points();
    struct Term {
        uint64_t key;
        uint32_t i, j;
    };
    std::vector<Term> terms;
    for (size_t q = 0; q < NP; ++q)
        point_terms(L, q, true, [&](uint32_t A, uint32_t Bb, uint32_t i, uint32_t j) {
            terms.push_back(Term{(uint64_t)A * NB + Bb, i, j});
        });
    std::vector<uint64_t> keys;
    keys.reserve(terms.size());
    for (auto& t : terms) keys.push_back(t.key);
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    const size_t NPAIR = keys.size();
    out.block_a.resize(NPAIR);
    out.block_b.resize(NPAIR);
    out.term_off.assign(NPAIR + 1, 0);
    out.value_off.assign(NPAIR + 1, 0);
    std::vector<uint32_t> pair_of(terms.size());
    for (size_t k = 0; k < terms.size(); ++k) {
        size_t idx = (size_t)(std::lower_bound(keys.begin(), keys.end(), terms[k].key) - keys.begin());
        pair_of[k] = (uint32_t)idx;
        out.term_off[idx + 1]++;
    }
    for (size_t k = 0; k < NPAIR; ++k) {
        out.block_a[k] = (uint32_t)(keys[k] / NB);
        out.block_b[k] = (uint32_t)(keys[k] % NB);
        out.term_off[k + 1] += out.term_off[k];
        out.value_off[k + 1] = out.value_off[k] + (uint64_t)B.block_dim[out.block_a[k]] * B.block_dim[out.block_b[k]];
    }
    out.term_i.resize(terms.size());
    out.term_j.resize(terms.size());
    std::vector<uint32_t> cursor(out.term_off.begin(), out.term_off.end() - 1);
    for (size_t k = 0; k < terms.size(); ++k) {
        uint32_t at = cursor[pair_of[k]]++;
        out.term_i[at] = terms[k].i;
        out.term_j[at] = terms[k].j;
    }
    return out;
}

// The diagonal blocks and pose-intrinsic edges only: one slot per diagonal block A (slot A) and
// per camera c's edge (P + group(c), c) (slot NB + c), so no sort is needed.  Terms are counted
// per point range on each thread, and each range's terms start after those of the ranges before
// it, which keeps every pair's terms in point order.
PairLists local_pairs(const Lowered& L, const BlockLists& B, bool with_terms) {
    PairLists out;
    const size_t P = L.n_cameras(), G = L.n_groups(), NB = B.n_blocks, NS = NB + P, NP = L.n_points();
    auto slot = [NB](uint32_t A, uint32_t Bb) { return A == Bb ? (size_t)A : NB + Bb; };
    const unsigned T = range_count(NP, 1 << 14);
    std::vector<uint32_t> cnt((size_t)T * NS, 0);
    for_ranges(NP, T, [&](unsigned t, size_t b, size_t e) {
        uint32_t* c = cnt.data() + (size_t)t * NS;
        for (size_t q = b; q < e; ++q)
            point_terms(L, q, false, [&](uint32_t A, uint32_t Bb, uint32_t, uint32_t) { c[slot(A, Bb)]++; });
    });
    std::vector<uint64_t> total(NS, 0);
    for (unsigned t = 0; t < T; ++t)
        for (size_t k = 0; k < NS; ++k) total[k] += cnt[(size_t)t * NS + k];
    // cameras of each group, ascending
    std::vector<uint32_t> goff(G + 1, 0), gcam(P);
    for (size_t c = 0; c < P; ++c) goff[L.cam_group[c] + 1]++;
    for (size_t g = 0; g < G; ++g) goff[g + 1] += goff[g];
    {
        std::vector<uint32_t> cur(goff.begin(), goff.end() - 1);
        for (size_t c = 0; c < P; ++c) gcam[cur[L.cam_group[c]]++] = (uint32_t)c;
    }
    // pairs ascending in (A, B): block A's edges (B = its group's cameras, all < P <= A), then (A, A)
    std::vector<uint32_t> pair_of(NS, UINT32_MAX);
    out.term_off.assign(1, 0);
    out.value_off.assign(1, 0);
    auto add = [&](uint32_t A, uint32_t Bb) {
        const size_t s = slot(A, Bb);
        if (!total[s]) return;
        pair_of[s] = (uint32_t)out.block_a.size();
        out.block_a.push_back(A);
        out.block_b.push_back(Bb);
        out.term_off.push_back(out.term_off.back() + (uint32_t)total[s]);
        out.value_off.push_back(out.value_off.back() + (uint64_t)B.block_dim[A] * B.block_dim[Bb]);
    };
    for (size_t A = 0; A < NB; ++A) {
        if (A >= P)
            for (uint32_t k = goff[A - P]; k < goff[A - P + 1]; ++k) add((uint32_t)A, gcam[k]);
        add((uint32_t)A, (uint32_t)A);
    }
    if (!with_terms) return out;
    for (size_t k = 0; k < NS; ++k) {  // per range: where its terms of slot k start
        uint32_t at = pair_of[k] == UINT32_MAX ? 0u : out.term_off[pair_of[k]];
        for (unsigned t = 0; t < T; ++t) {
            const uint32_t n = cnt[(size_t)t * NS + k];
            cnt[(size_t)t * NS + k] = at;
            at += n;
        }
    }
    out.term_i.resize(out.term_off.back());
    out.term_j.resize(out.term_off.back());
    for_ranges(NP, T, [&](unsigned t, size_t b, size_t e) {
        uint32_t* cur = cnt.data() + (size_t)t * NS;
        for (size_t q = b; q < e; ++q)
            point_terms(L, q, false, [&](uint32_t A, uint32_t Bb, uint32_t i, uint32_t j) {
                const uint32_t at = cur[slot(A, Bb)]++;
                out.term_i[at] = i;
                out.term_j[at] = j;
            });
    });
    return out;
}

}  // namespace

PairLists build_pair_lists(const Lowered& L, const BlockLists& B, bool all, bool with_terms) {
    return all ? all_pairs(L, B) : local_pairs(L, B, with_terms);
}

void build_pair_groups(const std::vector<uint32_t>& pair_cam, uint32_t nc, std::vector<uint32_t>& group,
                       std::vector<uint8_t>& group_n) {
    constexpr uint32_t NX = 8;
    const size_t np = pair_cam.size() / 2;
    std::vector<std::vector<uint32_t>> qs(NX), qn(NX);  // per residue: group starts and counts
    size_t k = 0;
    for (uint32_t c = 0; c < nc; ++c) {
        size_t e = k;
        while (e < np && pair_cam[2 * e] == c) ++e;
        for (size_t g = k; g < e; g += SP_GROUP) {
            qs[c % NX].push_back((uint32_t)g);
            qn[c % NX].push_back((uint32_t)std::min<size_t>(SP_GROUP, e - g));
        }
        k = e;
    }
    size_t depth = 0;
    for (auto& q : qs) depth = std::max(depth, q.size());
    group.assign(NX * depth, 0);
    group_n.assign(NX * depth, 0);
    for (uint32_t x = 0; x < NX; ++x)
        for (size_t i = 0; i < qs[x].size(); ++i) {
            group[NX * i + x] = qs[x][i];
            group_n[NX * i + x] = (uint8_t)qn[x][i];
        }
}

uint64_t count_schur_terms(const Lowered& L) {
    uint64_t n = 0;
    for (size_t p = 0; p < L.n_points(); ++p) {
        if (L.point_fixed[p]) continue;
        const uint64_t k = L.point_off[p + 1] - L.point_off[p];
        n += k * (k - 1) / 2;
    }
    return n;
}

SchurLists build_schur_lists(const Lowered& L, const std::vector<uint32_t>& pos) {
    const size_t nc = L.n_cameras(), np = L.n_points();
    SchurLists S;
    // 1. observation pairs bucketed by their lower camera, in point order: counted per point
    //    range, each range's entries placed after those of the ranges before it
    const unsigned T = range_count(np, 1 << 13);
    std::vector<uint64_t> cnt((size_t)T * nc, 0);
    auto each_pair = [&](size_t b, size_t e, auto&& f) {
        for (size_t p = b; p < e; ++p) {
            if (L.point_fixed[p]) continue;
            for (uint32_t a = L.point_off[p]; a < L.point_off[p + 1]; ++a)
                for (uint32_t c = a + 1; c < L.point_off[p + 1]; ++c) f(a, c);
        }
    };
    for_ranges(np, T, [&](unsigned t, size_t b, size_t e) {
        uint64_t* n = cnt.data() + (size_t)t * nc;
        each_pair(b, e, [&](uint32_t a, uint32_t c) { n[std::min(L.obs_cam[a], L.obs_cam[c])]++; });
    });
    std::vector<uint64_t> boff(nc + 1, 0);
    for (size_t c = 0; c < nc; ++c) {
        uint64_t at = boff[c];
        for (unsigned t = 0; t < T; ++t) {
            const uint64_t n = cnt[(size_t)t * nc + c];
            cnt[(size_t)t * nc + c] = at;
            at += n;
        }
        boff[c + 1] = at;
    }
    struct Entry {
        uint32_t j, pc, pj;
    };
    std::vector<Entry, NoInitAlloc<Entry>> ent(boff[nc]);
    for_ranges(np, T, [&](unsigned t, size_t b, size_t e) {
        uint64_t* cur = cnt.data() + (size_t)t * nc;
        each_pair(b, e, [&](uint32_t a, uint32_t c) {
            const uint32_t ca = L.obs_cam[a], cb = L.obs_cam[c];
            if (ca < cb) ent[cur[ca]++] = Entry{cb, pos[a], pos[c]};
            else ent[cur[cb]++] = Entry{ca, pos[c], pos[a]};
        });
    });
    // 2. per camera: its partner cameras ascending (all above it), the entries grouped by
    //    partner in their order
    std::vector<std::vector<uint32_t>> js(nc), jn(nc);
    for_ranges(nc, range_count(nc, 16), [&](unsigned, size_t c0, size_t c1) {
        std::vector<uint64_t> count(nc, 0);
        std::vector<Entry> tmp;
        for (size_t c = c0; c < c1; ++c) {
            const uint64_t b = boff[c], e = boff[c + 1];
            if (b == e) continue;
            for (uint64_t i = b; i < e; ++i) count[ent[i].j]++;
            uint64_t at = 0;
            for (size_t j = c + 1; j < nc; ++j) {
                if (!count[j]) continue;
                js[c].push_back((uint32_t)j);
                jn[c].push_back((uint32_t)count[j]);
                const uint64_t n = count[j];
                count[j] = at;  // now the partner's first slot
                at += n;
            }
            tmp.resize(e - b);
            for (uint64_t i = b; i < e; ++i) tmp[count[ent[i].j]++] = ent[i];
            std::copy(tmp.begin(), tmp.end(), ent.begin() + b);
            for (uint32_t j : js[c]) count[j] = 0;
        }
    });
    // 3. upper rows: the camera, then its partners j > c; pairs numbered in (c, j) order
    S.nbr_off.assign(nc + 1, 0);
    S.pair_off.assign(nc + 1, 0);
    S.row_base.resize(nc);
    uint64_t base = 0;
    for (size_t c = 0; c < nc; ++c) {
        const uint32_t nn = (uint32_t)(1 + js[c].size());
        S.nbr_off[c + 1] = S.nbr_off[c] + nn;
        S.pair_off[c + 1] = S.pair_off[c] + (uint32_t)js[c].size();
        S.row_base[c] = base;
        base += 9ull * schur_row_len(nn);
    }
    S.n_values = base;
    S.nbr.resize(S.nbr_off[nc]);
    for (size_t c = 0; c < nc; ++c) {
        uint32_t* out = S.nbr.data() + S.nbr_off[c];
        *out++ = (uint32_t)c;
        for (uint32_t j : js[c]) *out++ = j;
    }
    // 4. pairs and their terms (the entries in their grouped order); per camera j the pairs
    //    (c < j, j) in ascending c
    const uint32_t npairs = S.pair_off[nc];
    S.rev_off.assign(nc + 1, 0);
    for (size_t c = 0; c < nc; ++c)
        for (uint32_t j : js[c]) S.rev_off[j + 1]++;
    for (size_t c = 0; c < nc; ++c) S.rev_off[c + 1] += S.rev_off[c];
    S.rev_pair.resize(npairs);
    std::vector<uint32_t> rcur(S.rev_off.begin(), S.rev_off.end() - 1);
    S.term_off.assign(1, 0);
    S.term_off.reserve((size_t)npairs + 1);
    S.pair_cam.reserve(2 * (size_t)npairs);
    uint32_t k = 0;
    for (size_t c = 0; c < nc; ++c)
        for (size_t m = 0; m < js[c].size(); ++m, ++k) {
            S.pair_cam.push_back((uint32_t)c);
            S.pair_cam.push_back(js[c][m]);
            S.rev_pair[rcur[js[c][m]]++] = k;
            S.term_off.push_back(S.term_off.back() + jn[c][m]);
        }
    S.term.resize(2 * ent.size());
    for_ranges(ent.size(), range_count(ent.size(), 1 << 16), [&](unsigned, size_t b, size_t e) {
        for (size_t i = b; i < e; ++i) {
            S.term[2 * i] = ent[i].pc;
            S.term[2 * i + 1] = ent[i].pj;
        }
    });
    return S;
}

void raise_solution(const ba_problem* p, const Lowered& L, const double* R, const double* t, const double* gp,
                    const double* X, ba_solution* out) {
    const size_t P = p->n_poses;
    std::memcpy(out->pose_rotation, p->pose_rotation, 9 * P * sizeof(double));
    std::memcpy(out->pose_center, p->pose_center, 3 * P * sizeof(double));
    for (size_t d = 0; d < L.n_cameras(); ++d) {
        size_t s = L.camera_source[d];
        if (is_fixed_pose(p, s)) continue;
        const double* Rn = R + 9 * d;
        mul_ab(Rn, L.anchor_rotation, out->pose_rotation + 9 * s);
        double rc[3], rel[3], w[3];
        mul_atv(Rn, t + 3 * d, rc);
        for (int k = 0; k < 3; ++k) rel[k] = -rc[k];
        mul_atv(L.anchor_rotation, rel, w);
        for (int k = 0; k < 3; ++k) out->pose_center[3 * s + k] = L.anchor_center[k] + w[k];
    }
    std::memcpy(out->point_xyz, p->point_xyz, 3 * p->n_points * sizeof(double));
    for (size_t q = 0; q < L.n_points(); ++q) {
        size_t s = L.point_source[q];
        if (is_fixed_point(p, s)) continue;
        double w[3];
        mul_atv(L.anchor_rotation, X + 3 * q, w);
        for (int k = 0; k < 3; ++k) out->point_xyz[3 * s + k] = L.anchor_center[k] + w[k];
    }
    if (out->point_quarantined)
        for (size_t i = 0; i < p->n_points; ++i) out->point_quarantined[i] = L.quarantined[i];
    for (size_t g = 0; g < p->n_intrinsics; ++g) {
        const ba_intrinsic& in = p->intrinsics[g];
        if (L.grp_fixed[g]) {
            out->intrinsic_focal[g] = in.fx;
            out->intrinsic_k1[g] = in.model >= BA_CAMERA_BROWN_K1 && in.model <= BA_CAMERA_BROWN_K1K2K3 ? in.k1 : 0.0;
            out->intrinsic_k2[g] = in.model >= BA_CAMERA_BROWN_K1K2 && in.model <= BA_CAMERA_BROWN_K1K2K3 ? in.k2 : 0.0;
        } else {
            out->intrinsic_focal[g] = gp[3 * g];
            out->intrinsic_k1[g] = gp[3 * g + 1];
            out->intrinsic_k2[g] = gp[3 * g + 2];
        }
        if (out->intrinsic_unconstrained) out->intrinsic_unconstrained[g] = 0;
    }
}

// ---- LinearSolver::Auto (kernel/multi_view/sparse/execution_policy.rs:98-227,
//      workspace/self_calibration.rs:466-535, self_calibration_retained.rs:233-290) ----
namespace {
// SparseWorkspaceMemoryBudget::AUTO_BYTES (kernel/multi_view/sparse/api.rs:43)
constexpr uint64_t AUTO_BUDGET_BYTES = 1024ull * 1024ull * 1024ull;
// direct_fits_budget: at most 4/5 of the limit, without overflow (execution_policy.rs:280-286)
bool direct_fits(uint64_t estimate, uint64_t limit) {
    const uint64_t guarded = limit / 5 * 4 + limit % 5 * 4 / 5;
    return estimate <= guarded;
}
// Lower-triangle nonzeros of the reduced pattern (pattern.rs:381-491): every block pair that a
// free point couples, plus each camera's pose-intrinsic pair for fixed points, counted element
// by element.  Block pairs are marked in a bitmap of (blocks x blocks) bits.
uint64_t reduced_nonzeros(const Lowered& L) {
    const size_t P = L.n_cameras(), G = L.n_groups(), NB = P + G, NP = L.n_points();
    auto bdim = [&](size_t b) -> uint64_t { return b < P ? L.cam_dim[b] : L.grp_dim[b - P]; };
    const size_t words = (NB + 63) / 64;
    std::vector<uint64_t> bits(words * NB, 0);
    auto mark = [&](size_t row, size_t col) { bits[col * words + row / 64] |= 1ull << (row % 64); };
    for (size_t b = 0; b < NB; ++b)
        if (bdim(b)) mark(b, b);
    std::vector<uint32_t> active;
    for (size_t q = 0; q < NP; ++q) {
        const uint32_t o0 = L.point_off[q], o1 = L.point_off[q + 1];
        if (L.point_fixed[q]) {
            for (uint32_t o = o0; o < o1; ++o) {
                const uint32_t c = L.obs_cam[o], g = L.cam_group[c];
                if (L.cam_dim[c] && L.grp_dim[g]) mark(P + g, c);
            }
            continue;
        }
        active.clear();
        for (uint32_t o = o0; o < o1; ++o) {
            const uint32_t c = L.obs_cam[o], g = L.cam_group[c];
            if (L.cam_dim[c]) active.push_back(c);
            if (L.grp_dim[g]) active.push_back((uint32_t)(P + g));
        }
        std::sort(active.begin(), active.end());
        active.erase(std::unique(active.begin(), active.end()), active.end());
        for (size_t i = 0; i < active.size(); ++i)
            for (size_t j = 0; j <= i; ++j) mark(active[i], active[j]);
    }
    uint64_t count = 0;
    for (size_t col = 0; col < NB; ++col) {
        const uint64_t dc = bdim(col);
        if (!dc) continue;
        for (size_t w = 0; w < words; ++w) {
            uint64_t m = bits[col * words + w];
            while (m) {
                const size_t row = w * 64 + (size_t)__builtin_ctzll(m);
                m &= m - 1;
                count += row == col ? dc * (dc + 1) / 2 : bdim(row) * dc;
            }
        }
    }
    return count;
}
}  // namespace

bool rank_certificate_fits(const Lowered& L, const ba_execution_plan& plan) {
    const uint64_t nnz = plan.reduced_nonzeros ? plan.reduced_nonzeros : reduced_nonzeros(L);
    const uint64_t cert = nnz * 8 * 4;  // estimated_direct_bytes(nnz, D, 0, 0, false)
    const uint64_t total = plan.estimated_retained_bytes > UINT64_MAX - cert ? UINT64_MAX
                                                                              : plan.estimated_retained_bytes + cert;
    return direct_fits(total, plan.budget_bytes);
}

void plan_execution(const Lowered& L, const ba_solve_config* c, ba_execution_plan& out) {
    std::memset(&out, 0, sizeof(out));
    out.reduced_dimension = (uint32_t)L.D;
    if (c->linear_solver != BA_SOLVER_AUTO) {  // explicit_decision (execution_policy.rs:252-268)
        out.linear_solver = c->linear_solver;
        out.reason = BA_PLAN_EXPLICIT_SOLVER;
        return;
    }
    const uint64_t P = L.n_cameras(), G = L.n_groups(), NP = L.n_points(), D = L.D;
    const uint64_t cross = L.n_cross;
    const uint64_t limit = c->workspace_budget_bytes ? c->workspace_budget_bytes : AUTO_BUDGET_BYTES;
    out.budget_bytes = limit;
    if (L.calibrated) {
        // select_calibrated (execution_policy.rs:123-183) over the serial estimates of
        // workspace/calibrated.rs:480-530: the cached point cache keeps point steps and f64 cross
        // rows (execution_policy.rs:68-96), the direct estimate adds the dense rank certificate up
        // to dimension 2048 (workspace.rs:126-132), the retained estimate is
        // estimated_calibrated_retained_bytes (retained/mod.rs:156-191) without worker shards
        const uint64_t cached_points = (NP + 1) * 8 + NP * (1 + 2 * 24) + cross * 24;
        const uint64_t dense_rank = D <= 2048 ? (D + 2) * D * 8 : 0;
        out.estimated_retained_bytes = P * (2 * 168 + 216) + NP * 48 + D * 48 + cached_points;
        const bool prefers_cached = L.n_obs() >= 250000 && D >= 768;  // execution_policy.rs:14-22
        // the exact pattern count only where the bounds (diagonal blocks, full lower triangle) do
        // not decide
        uint64_t diagonal_nnz = 0;
        for (size_t k = 0; k < P; ++k) diagonal_nnz += (uint64_t)L.cam_dim[k] * (L.cam_dim[k] + 1) / 2;
        const uint64_t lower = diagonal_nnz * 32 + dense_rank + cached_points;
        const uint64_t upper = D * (D + 1) / 2 * 32 + dense_rank + cached_points;
        if ((prefers_cached && out.estimated_retained_bytes <= limit) || !direct_fits(lower, limit)) {
            out.estimated_direct_bytes = lower;
        } else if (direct_fits(upper, limit)) {
            out.estimated_direct_bytes = upper;
        } else {
            out.reduced_nonzeros = reduced_nonzeros(L);
            out.estimated_direct_bytes = out.reduced_nonzeros * 32 + dense_rank + cached_points;
        }
        if (prefers_cached && out.estimated_retained_bytes <= limit) {
            out.linear_solver = BA_SOLVER_ITERATIVE;
            out.reason = BA_PLAN_WORKLOAD_PREFERS_RETAINED;
        } else if (direct_fits(out.estimated_direct_bytes, limit)) {
            out.linear_solver = BA_SOLVER_DIRECT;
            out.reason = BA_PLAN_DIRECT_FITS_BUDGET;
        } else {
            out.linear_solver = BA_SOLVER_ITERATIVE;
            out.reason = out.estimated_retained_bytes <= limit ? BA_PLAN_DIRECT_EXCEEDS_BUDGET : BA_PLAN_LOWEST_MEMORY;
        }
        return;
    }
    // estimated_point_cache_bytes (execution_policy.rs:68-96): usize offsets, no retained point
    // steps (serial direct assembly), one [f64; 3] row per point parameter
    const uint64_t direct_points = (NP + 1) * 8 + cross * 24;
    // estimated_direct_bytes (execution_policy.rs:98-120): 4 x 8 bytes per lower nonzero
    auto direct_bytes = [&](uint64_t nnz) { return nnz * 8 * 4 + direct_points; };
    uint64_t diagonal_nnz = 0;
    for (size_t k = 0; k < P; ++k) diagonal_nnz += (uint64_t)L.cam_dim[k] * (L.cam_dim[k] + 1) / 2;
    for (size_t g = 0; g < G; ++g) diagonal_nnz += (uint64_t)L.grp_dim[g] * (L.grp_dim[g] + 1) / 2;
    const uint64_t lower = direct_bytes(diagonal_nnz), upper = direct_bytes(D * (D + 1) / 2);
    if (!direct_fits(lower, limit)) {
        out.estimated_direct_bytes = lower;  // at least this much: the exact count cannot fit
    } else if (direct_fits(upper, limit)) {
        out.estimated_direct_bytes = upper;  // at most this much: the exact count fits
    } else {
        out.reduced_nonzeros = reduced_nonzeros(L);
        out.estimated_direct_bytes = direct_bytes(out.reduced_nonzeros);
    }
    // estimated_self_calibrating_retained_bytes (self_calibration_retained.rs:233-290) without the
    // worker term: per block two packed 6x6 blocks and a factor (2*168 + 216 bytes) and three
    // usize; per pose-intrinsic edge 160 bytes; per point its packed 3x3 block (48 bytes); per
    // point parameter a u32; per dimension five f64 and a usize; and the retained point cache
    // (offsets, point steps of 1 + 48 bytes, [f64; 3] rows)
    uint64_t edges = 0;
    for (size_t k = 0; k < P; ++k) edges += (L.cam_dim[k] > 0 && L.grp_dim[L.cam_group[k]] > 0) ? 1 : 0;
    const uint64_t NB = P + G;
    out.estimated_retained_bytes = NB * (2 * 168 + 216) + edges * 160 + NB * 3 * 8 + NP * 48 + cross * 4 +
                                   D * (5 * 8 + 8) + (NP + 1) * 8 + NP * 49 + cross * 24;
    if (direct_fits(out.estimated_direct_bytes, limit)) {
        out.linear_solver = BA_SOLVER_DIRECT;
        out.reason = BA_PLAN_DIRECT_FITS_BUDGET;
    } else {
        out.linear_solver = BA_SOLVER_ITERATIVE;
        out.reason = out.estimated_retained_bytes <= limit ? BA_PLAN_DIRECT_EXCEEDS_BUDGET : BA_PLAN_LOWEST_MEMORY;
    }
}

}  // namespace ba
