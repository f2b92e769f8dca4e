// ba_solver.cpp — the LM driver and the C ABI of the MI355X bundle adjuster.
//
// The loop restates crates/openmvg-ba/src/kernel/multi_view/sparse/self_calibration/mod.rs:212-639:
// stages (numerics/refine/nonlinear/continuation.rs:26-74), Nielsen damping (controller.rs:6-91),
// run bookkeeping and stop tests (run_state.rs:15-181), the camera step by dense Cholesky or
// retained PCG (camera_step.rs:69-128), retraction, candidate evaluation, model reduction
// (camera_step.rs:33-67) and the final zero-damping assembly.  The scalar LM control is FP64 on
// the host; every pass over observations, points and the reduced system runs on the GPU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/ba_hip.h"
#include "ba_certify.hpp"
#include "ba_handle.hpp"
#include "ba_host.hpp"
#include "ba_kernels.hpp"
#include "ba_math.hpp"

#include "ba_engine.hpp"

namespace ba {
namespace {

// A Lowered copy of a device-resident lowered problem (the per-kernel hooks).
Status lowered_from_device(const ba_device_problem* dp, Lowered& L, hipStream_t s) {
    if (!dp) return error(BA_ERR_INVALID_CONFIG, "null problem");
    const size_t nc = dp->n_cameras, ng = dp->n_groups, np = dp->n_points, no = dp->n_observations;
    auto down = [&](auto& vec, const void* src, size_t count) {
        vec.resize(count);
        if (count) HIP_OK(hipMemcpyAsync(vec.data(), src, count * sizeof(vec[0]), hipMemcpyDeviceToHost, s));
    };
    down(L.R, dp->camera_rotation, 9 * nc);
    down(L.t, dp->camera_translation, 3 * nc);
    down(L.gp, dp->group_params, 3 * ng);
    down(L.cam_group, dp->camera_group, nc);
    down(L.X, dp->point_xyz, 3 * np);
    down(L.obs_cam, dp->obs_camera, no);
    down(L.obs_point, dp->obs_point, no);
    down(L.uv, dp->obs_uv, 2 * no);
    down(L.point_off, dp->point_offsets, np + 1);
    down(L.point_fixed, dp->point_fixed, np);
    down(L.cam_off, dp->camera_offset, nc);
    down(L.cam_dim, dp->camera_dim, nc);
    down(L.grp_off, dp->group_offset, ng);
    down(L.grp_dim, dp->group_dim, ng);
    HIP_OK(hipStreamSynchronize(s));
    // topology checks the kernels rely on
    if (L.point_off.empty() || L.point_off[0] != 0 || L.point_off[np] != no)
        return error(BA_ERR_INVALID_PROBLEM, "point offsets do not cover the observations");
    for (size_t q = 0; q < np; ++q) {
        if (L.point_off[q] > L.point_off[q + 1]) return error(BA_ERR_INVALID_PROBLEM, "point offsets decrease");
        for (uint32_t o = L.point_off[q]; o < L.point_off[q + 1]; ++o)
            if (L.obs_point[o] != q) return error(BA_ERR_INVALID_PROBLEM, "obs_point disagrees with the offsets");
    }
    for (size_t o = 0; o < no; ++o)
        if (L.obs_cam[o] >= nc) return error(BA_ERR_INVALID_PROBLEM, "observation camera out of bounds");
    size_t D = 0;
    for (size_t c = 0; c < nc; ++c) {
        if (L.cam_group[c] >= ng) return error(BA_ERR_INVALID_PROBLEM, "camera group out of bounds");
        if (L.cam_dim[c] > 6 || (L.cam_dim[c] && L.cam_off[c] + L.cam_dim[c] > dp->dimension))
            return error(BA_ERR_INVALID_PROBLEM, "camera block outside the layout");
        D += L.cam_dim[c];
    }
    for (size_t g = 0; g < ng; ++g) {
        if (L.grp_dim[g] > 3 || (L.grp_dim[g] && L.grp_off[g] + L.grp_dim[g] > dp->dimension))
            return error(BA_ERR_INVALID_PROBLEM, "intrinsic block outside the layout");
        D += L.grp_dim[g];
    }
    if (D != dp->dimension) return error(BA_ERR_INVALID_PROBLEM, "block dimensions do not sum to the dimension");
    if (dp->gauge_camera >= (int32_t)nc) return error(BA_ERR_INVALID_PROBLEM, "gauge camera out of bounds");
    L.D = D;
    L.loss = dp->loss;
    L.loss_scale = dp->loss_scale;
    L.gauge_camera = dp->gauge_camera;
    L.scene_scale = dp->scene_scale;
    L.grp_fixed.assign(ng, 0);
    L.cam_fixed.assign(nc, 0);
    return Status{};
}

GaugeDev gauge_from(const ba_device_problem* dp) {
    GaugeDev g{};
    g.camera = dp->gauge_camera;
    g.scene_scale = dp->scene_scale;
    g.baseline = dp->baseline;
    for (int k = 0; k < 6; ++k) g.basis[k] = dp->basis[k];
    return g;
}

// Per-iteration trace rows; `loop_start` is the host clock when the LM loop began.
struct TraceSink {
    ba_iteration_trace* tr = nullptr;
    double loop_start = 0.0;
};

void trace_push(const TraceSink& sink, double cost, double damping, double cand, double pred, double gain, bool acc,
                uint32_t pcg) {
    ba_iteration_trace* tr = sink.tr;
    if (!tr || tr->count >= tr->capacity) return;
    uint32_t i = tr->count++;
    if (tr->seconds) tr->seconds[i] = now() - sink.loop_start;
    if (tr->cost) tr->cost[i] = cost;
    if (tr->damping) tr->damping[i] = damping;
    if (tr->candidate_cost) tr->candidate_cost[i] = cand;
    if (tr->predicted_reduction) tr->predicted_reduction[i] = pred;
    if (tr->gain_ratio) tr->gain_ratio[i] = gain;
    if (tr->accepted) tr->accepted[i] = acc;
    if (tr->pcg_iterations) tr->pcg_iterations[i] = pcg;
}

struct StageDef {
    int loss;
    double scale;
    uint32_t iters;
};

// The LM loop (sparse/self_calibration/mod.rs:212-639) on the GPU engine.
// `loop_seconds` receives the host time from the first iteration's start to the end of the last one;
// every iteration ends in a synchronous scalar readback, so this brackets completed device work.
// `inlier_limit` and `initial_inliers` (count, RMSE): the inlier statistics at the initial state.
template <class T>
Status refine(Engine<T>& E, const Lowered& L, const ba_solve_config* cfg, ba_solve_report& rep,
              ba_iteration_trace* trace_rows, int* final_state, double* loop_seconds, double inlier_limit,
              uint64_t* initial_inlier_count, double* initial_inlier_rmse) {
    const size_t D = L.D;
    const bool iterative = E.iterative();
    const double no = (double)L.n_obs();
    const double npts = (double)L.n_points();
    StageDef sched[3];
    if (L.loss == BA_LOSS_CAUCHY) {
        uint32_t l2 = std::min<uint32_t>(cfg->max_iterations / 4, 5);
        sched[0] = {BA_LOSS_SQUARED, 0.0, l2};
        sched[1] = {L.loss, L.loss_scale, cfg->max_iterations - l2};
        sched[2] = {BA_LOSS_SQUARED, 0.0, 0};
    } else {
        sched[0] = {L.loss, L.loss_scale, cfg->max_iterations};
        sched[1] = {BA_LOSS_SQUARED, 0.0, 0};
        sched[2] = {BA_LOSS_SQUARED, 0.0, 0};
    }
    int target = 0;
    for (int i = 2; i >= 0; --i)
        if (sched[i].iters > 0) {
            target = i;
            break;
        }
    int state = 0;
    // validation::projection_domain: every observation projects at the initial state
    {
        AsmScalars a0 = E.assemble(state, 0.0, L.loss, L.loss_scale, false);
        if (a0.unprojectable > 0.0) return error(BA_ERR_NUMERICAL, "point not projectable at the initial state");
        E.inliers(inlier_limit, initial_inlier_count, initial_inlier_rmse);
    }
    TraceSink trace{trace_rows, now()};
    const double initial_damping = iterative ? 1e-1 : 1e-4;  // sparse/api.rs:300-305
    Nielsen damping{initial_damping};
    uint32_t iterations = 0, accepted = 0, target_accepted = 0;
    uint32_t rej[4] = {0, 0, 0, 0};
    double step_norm = 0.0;
    int conv = -1;
    bool have_initial = false;
    double initial_cost = INFINITY, initial_rmse = INFINITY;
    bool reuse = false;
    for (int i = 0; i < 3; ++i)
        if (sched[i].iters > 0) {
            reuse = sched[i].loss == L.loss;
            break;
        }
    if (!reuse) {
        AsmScalars a0 = E.assemble(state, 0.0, L.loss, L.loss_scale, false);
        initial_cost = a0.cost;
        initial_rmse = std::sqrt(a0.sq / (2.0 * no));
        have_initial = true;
    }
    const double conv_tol = std::fmin(std::sqrt(cfg->function_tolerance), 1e-6);
    const double base_tol = cfg->pcg_relative_tolerance > 0.0 ? cfg->pcg_relative_tolerance : 1e-1;
    const uint32_t pcg_cap = cfg->pcg_max_iterations ? cfg->pcg_max_iterations : 600u;
    uint32_t pcg_total = 0;
    double pcg_max_rel = NAN;
    bool stop = false;
    bool same_state = true;  // state `state` unchanged since the last assembly
    uint32_t last_terms = 0;  // PCG iterations of the last solve: picks the next product form
    for (int si = 0; si < 3 && !stop; ++si) {
        bool confirm = false;
        for (uint32_t it = 0; it < sched[si].iters; ++it) {
            iterations += 1;
            const double lambda = damping.value;
            if (iterative) E.plan_product(last_terms);
            AsmScalars a = E.assemble(state, lambda, sched[si].loss, sched[si].scale, true, same_state);
            same_state = true;
            if (!have_initial && std::isfinite(a.cost) && std::isfinite(a.sq) && no > 0) {
                initial_cost = a.cost;
                initial_rmse = std::sqrt(a.sq / (2.0 * no));
                have_initial = true;
            }
            const bool valid = a.invalid == 0.0 && a.nonfinite == 0.0;
            if (!valid) {
                rej[0]++;
                damping.reject();
                trace_push(trace, a.cost, lambda, NAN, NAN, NAN, false, 0);
                continue;
            }
            const double current = a.cost;
            const double sgin = std::fmax(a.sgin, a.sgin_prepared);
            if (si == target && conv < 0 && std::isfinite(sgin) && sgin >= 0.0 && std::isfinite(current) &&
                sgin <= cfg->gradient_tolerance * (1.0 + std::fabs(current))) {
                conv = BA_STATUS_CONVERGED_GRADIENT;
                trace_push(trace, current, lambda, NAN, NAN, NAN, false, 0);
                stop = true;
                break;
            }
            bool solved;
            uint32_t terms = 0;
            double rel = NAN;
            if (!iterative) {
                solved = E.solve_direct();
            } else {
                double tol = confirm ? std::fmin(base_tol, conv_tol) : base_tol;
                uint32_t limit = confirm ? pcg_cap : std::min<uint32_t>(pcg_cap, (uint32_t)D);
                solved = E.solve_pcg(tol, (int)limit, &terms, &rel);
                pcg_total += terms;
                last_terms = terms;
                if (std::isfinite(rel)) pcg_max_rel = std::isfinite(pcg_max_rel) ? std::fmax(pcg_max_rel, rel) : rel;
            }
            if (!solved) {
                rej[1]++;
                damping.reject();
                trace_push(trace, current, lambda, NAN, NAN, NAN, false, terms);
                continue;
            }
            if (L.calibrated) E.limit_step(0.5, iterative);  // calibrated/candidate.rs:36-44
            auto c = E.candidate(state, sched[si].loss, sched[si].scale, iterative);
            if (c.c.rejected != 0.0) {
                rej[2]++;
                damping.reject();
                trace_push(trace, current, lambda, NAN, NAN, NAN, false, terms);
                continue;
            }
            const double point_rms = std::sqrt(c.c.step2 / npts);
            const double gdot = c.dots[0], cdn = c.dots[1], xx = c.dots[2], xq = c.dots[3];
            double predicted;
            if (!iterative && L.calibrated) {
                // calibrated_model_reduction (calibrated/camera_step.rs:40-62): pmr - g.x - x.S x / 2
                // with the damped S; the limited step x = alpha x0 of the direct solve S x0 = -g
                // gives x.S x = -alpha g.x
                predicted = (a.pmr - gdot + 0.5 * (c.alpha * gdot)) + 0.5 * lambda * (cdn + c.c.damping_norm);
            } else if (!iterative) {  // exact_model_reduction (self_calibration/optimizer.rs:30-52)
                predicted = (a.pmr - 0.5 * gdot) + 0.5 * lambda * (cdn + c.c.damping_norm);
            } else {  // predicted_reduction (self_calibration_retained.rs:585-598) + undamped
                double pr = a.pmr + (-gdot - 0.5 * xq);
                predicted = (std::isfinite(pr) && pr > 0.0) ? pr + 0.5 * lambda * (cdn + c.c.damping_norm) : NAN;
            }
            const double reduction = current - c.c.cost;
            const bool acc = damping.accept(reduction, predicted);
            trace_push(trace, current, lambda, c.c.cost, predicted, damping.gain, acc, terms);
            if (!acc) {
                rej[3]++;
                continue;
            }
            state = 1 - state;
            same_state = false;
            accepted += 1;
            const double rms = std::sqrt(xx / (double)std::max<size_t>(D, 1));
            const double sn = std::sqrt(rms * rms + point_rms * point_rms);
            step_norm = sn;
            if (si == target) {
                target_accepted += 1;
                const double relred = reduction / (1.0 + std::fabs(current));
                int cand_conv = -1;
                if (sn <= cfg->parameter_tolerance)
                    cand_conv = BA_STATUS_CONVERGED_STEP;
                else if (relred <= cfg->function_tolerance)
                    cand_conv = BA_STATUS_CONVERGED_COST;
                if (iterative && cand_conv >= 0 && rel > conv_tol) {
                    confirm = true;
                    conv = -1;
                } else {
                    confirm = false;
                    conv = cand_conv;
                    if (conv >= 0) {
                        stop = true;
                        break;
                    }
                }
            }
        }
    }
    *loop_seconds = now() - trace.loop_start;
    AsmScalars fa = E.assemble(state, 0.0, L.loss, L.loss_scale, false);
    std::memset(&rep, 0, sizeof(rep));
    if (conv >= 0)
        rep.status = conv;
    else if (iterations > 0 && target_accepted == 0)
        rep.status = BA_STATUS_STALLED;
    else
        rep.status = BA_STATUS_MAX_ITERATIONS;
    rep.linear_solver_used = iterative ? BA_SOLVER_ITERATIVE : BA_SOLVER_DIRECT;
    rep.iterations = iterations;
    rep.accepted_steps = accepted;
    rep.rejected_linearization = rej[0];
    rep.rejected_step_solve = rej[1];
    rep.rejected_candidate = rej[2];
    rep.rejected_insufficient_reduction = rej[3];
    rep.rejected_steps = rej[0] + rej[1] + rej[2] + rej[3];
    rep.initial_cost = initial_cost;
    rep.final_cost = std::isfinite(fa.cost) ? fa.cost : INFINITY;
    rep.initial_rmse = initial_rmse * std::sqrt(2.0);
    rep.final_rmse = std::sqrt(fa.sq / (2.0 * no)) * std::sqrt(2.0);
    rep.gradient_inf_norm = std::fmax(fa.sgin, fa.sgin_prepared);
    rep.step_norm = step_norm;
    rep.initial_damping = initial_damping;
    rep.final_damping = damping.value;
    rep.last_gain_ratio = damping.has_gain ? damping.gain : NAN;
    rep.reduced_dimension = (uint32_t)D;
    rep.weak_point_count = (uint32_t)fa.weak;
    rep.minimum_point_rcond = fa.min_rcond;
    rep.pcg_iterations = pcg_total;
    rep.maximum_pcg_relative_residual = pcg_max_rel;
    rep.normalization_scale = L.scene_scale;
    rep.observation_count = L.n_obs();
    rep.point_count = L.n_points();
    rep.pose_count = L.n_cameras();
    rep.quarantined_points = (uint32_t)L.n_quarantined;
    *final_state = state;
    return Status{};
}

int finish(const Status& st, char* msg, size_t len) {
    if (msg && len) std::snprintf(msg, len, "%s", st.message.c_str());
    return st.code;
}

}  // namespace
}  // namespace ba

namespace ba {
namespace {

// The native final state and report of one LM solve.
struct Refined {
    ba_solve_report r{};
    std::vector<double> R, t, gp, X;
    uint64_t initial_inliers = 0;
    double initial_inlier_rmse = NAN;
};

// One LM solve of a lowered problem on the device (the plan's camera step), with the inlier
// statistics of its first and last linearisations.
template <class T>
Status run_refine(ba_hip_handle* h, const ba_problem* p, const ba_solve_config* cfg, const Lowered& L,
                  const ba_execution_plan& plan, ba_iteration_trace* trace, Refined& out) {
    const double t0 = now();
    const bool iterative = plan.linear_solver == BA_SOLVER_ITERATIVE;
    Engine<T> E(h->stream, L, iterative, cfg->profile_kernel);
    E.set_schur_product(h->schur_product);
    E.set_budget(cfg->workspace_budget_bytes);
    E.setup();
    // check_workspace_budget (native/mod.rs:557-570): an explicit budget refuses a larger
    // workspace before the first iteration; the automatic budget only guides the plan
    if (cfg->workspace_budget_bytes && E.device_bytes() > cfg->workspace_budget_bytes)
        return error(BA_ERR_RESOURCE_EXHAUSTED,
                     "native self-calibration workspace " + std::to_string(E.device_bytes()) +
                         " bytes exceeds the workspace budget " + std::to_string(cfg->workspace_budget_bytes));
    const double t1 = now();
    int state = 0;
    double loop_seconds = 0.0;
    const double limit = p->robust_loss == BA_LOSS_SQUARED ? INFINITY : p->loss_scale;
    ba_solve_report& r = out.r;
    Status st = refine(E, L, cfg, r, trace, &state, &loop_seconds, limit, &out.initial_inliers,
                       &out.initial_inlier_rmse);
    if (iterative) h->last_schur_product = E.explicit_product() ? BA_SCHUR_EXPLICIT : BA_SCHUR_IMPLICIT;
    if (!st.ok()) return st;
    E.inliers(limit, &r.inlier_count, &r.final_inlier_rmse);  // at refine's final zero-damping assembly
    r.plan_reason = plan.reason;
    r.workspace_bytes = E.device_bytes();
    E.download(state, out.R, out.t, out.gp, out.X);
    const double t3 = now();
    r.seconds_setup = t1 - t0;
    r.seconds_iterations = loop_seconds;
    r.seconds_final = (t3 - t1) - loop_seconds;
    E.profile_result(&r.profile_launches, &r.profile_seconds);
    r.precision_fallbacks = E.mixed_fallbacks();
    r.refinement_residual = E.refinement_residual();
    r.observability = BA_OBSERVABILITY_NOT_ASSESSED;
    return st;
}

Status refine_lowered(ba_hip_handle* h, const ba_problem* p, const ba_solve_config* cfg, const Lowered& L,
                      const ba_execution_plan& plan, ba_iteration_trace* trace, Refined& out) {
    return cfg->precision == BA_PRECISION_MIXED ? run_refine<float>(h, p, cfg, L, plan, trace, out)
                                                : run_refine<double>(h, p, cfg, L, plan, trace, out);
}

// The final rank certificate and the intrinsic marginals of a state, on the device in FP64: the
// undamped reduced system of the direct form (a second engine on that state), factored in place
// with every pivot required positive (the sparse LLT of direct.rs:259-264 without
// regularisation), and (S^-1)_gg of every free group from the factor's trailing columns
// (sparse/self_calibration/mod.rs:110-209, covariance.rs:562-587).
struct Certificate {
    bool formed = false;      // within this library's limits (BA_CERTIFY_MAX_*)
    int32_t observability = BA_OBSERVABILITY_NOT_ASSESSED;
    bool marginals = false;   // cov holds every free group's marginal
    std::vector<double> cov;  // 9 per group, NaN where none
};

// Whether the dense certificate of a lowering stays within BA_CERTIFY_MAX_DIMENSION and
// BA_CERTIFY_MAX_TERMS.
bool certificate_in_reach(const Lowered& L) {
    if (L.D > BA_CERTIFY_MAX_DIMENSION) return false;
    uint64_t terms = 0;
    for (size_t q = 0; q < L.n_points(); ++q) {
        const uint64_t k = 2ull * (L.point_off[q + 1] - L.point_off[q]);
        terms += L.point_fixed[q] ? k : k * k;
    }
    return terms <= BA_CERTIFY_MAX_TERMS;
}

Certificate certify(hipStream_t s, const Lowered& L, const Refined& f, bool exact_rank, bool want_marginals) {
    Certificate c;
    c.cov.assign(9 * L.n_groups(), NAN);
    const size_t D = L.D;
    if (!certificate_in_reach(L)) return c;
    c.formed = true;
    if (!exact_rank && !want_marginals) return c;
    if (D == 0) {  // nothing free: full rank, and no group to certify
        if (exact_rank) c.observability = BA_OBSERVABILITY_FULL_RANK;
        c.marginals = want_marginals;
        return c;
    }
    Lowered F = L;
    F.R = f.R;
    F.t = f.t;
    F.gp = f.gp;
    F.X = f.X;
    Engine<double> E(s, F, false);
    // The rank verdict is a threshold test at the rounding level of the undamped S: it is formed
    // as U - sum W V^-1 W^T, as the reference forms it, rather than from the explicit rows (whose
    // J^T (I - M) J diagonal rounds differently when a point block is nearly singular).
    E.keep_pair_terms();
    E.setup();
    const AsmScalars a = E.assemble(0, 0.0, L.loss, L.loss_scale, true);
    const bool valid = a.invalid == 0.0 && a.nonfinite == 0.0 && std::isfinite(a.cost);
    bool factored = false;
    if (valid) {
        Buf<double> work, cov, scratch;
        Buf<int> fail;
        work.alloc(cholesky_work_doubles((uint32_t)D));
        fail.alloc(1);
        Buf<double> tols;
        tols.alloc(3 * D);
        launch_rank_tolerances(s, E.S(), (uint32_t)D, (uint32_t)D, tols.p);
        launch_cholesky_factor(s, E.S_mut(), (uint32_t)D, (uint32_t)D, work.p, fail.p, 0.0, tols.p + 2 * D);
        int failed = 1;
        HIP_OK(hipMemcpyAsync(&failed, fail.p, sizeof(int), hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
        factored = failed == 0;
        size_t m = 0;
        for (uint32_t d : L.grp_dim) m += d;
        if (factored && want_marginals && m > 0) {
            const uint32_t j0 = (uint32_t)(D - m);  // the free intrinsic coordinates come last
            const uint32_t G = (uint32_t)L.n_groups();
            Buf<uint32_t> goff, gdim;
            goff.upload(L.grp_off, s);
            gdim.upload(L.grp_dim, s);
            scratch.alloc(trailing_inverse_doubles((uint32_t)D, j0));
            cov.alloc(9 * (size_t)G);
            const double w0 = loss_sqrt_weight(L.loss, L.loss_scale, 0.0);
            launch_group_marginals(s, E.S(), (uint32_t)D, (uint32_t)D, work.p, fail.p, j0, G, goff.p, gdim.p,
                                   w0 * w0, scratch.p, cov.p);
            std::vector<double> host(9 * (size_t)G);
            HIP_OK(hipMemcpyAsync(host.data(), cov.p, host.size() * sizeof(double), hipMemcpyDeviceToHost, s));
            HIP_OK(hipStreamSynchronize(s));
            bool finite = true;
            for (size_t g = 0; g < G; ++g)
                for (uint32_t k = 0; k < 9; ++k)
                    if (L.grp_dim[g]) finite &= std::isfinite(host[9 * g + k]);
            if (finite) {
                for (size_t g = 0; g < G; ++g)
                    if (L.grp_dim[g]) std::copy(host.begin() + 9 * g, host.begin() + 9 * g + 9, c.cov.begin() + 9 * g);
                c.marginals = true;
            }
        } else if (factored && want_marginals) {
            c.marginals = true;  // no free group
        }
    }
    if (exact_rank) c.observability = factored ? BA_OBSERVABILITY_FULL_RANK : BA_OBSERVABILITY_RANK_DEFICIENT;
    return c;
}

// Whether the exact rank certificate runs (workspace/self_calibration.rs:515-531): always, except
// when Auto chose the retained PCG step and its estimate plus the certificate's reduced values do
// not fit 4/5 of the budget.  The marginals need the direct storage, which Auto's budget
// excludes whenever it chose PCG (workspace.rs:251-265).
void certificate_policy(const Lowered& L, const ba_solve_config* cfg, const ba_execution_plan& plan, bool* exact,
                        bool* marginals) {
    const bool auto_pcg = cfg->linear_solver == BA_SOLVER_AUTO && plan.linear_solver == BA_SOLVER_ITERATIVE;
    *marginals = !auto_pcg;
    *exact = !auto_pcg || rank_certificate_fits(L, plan);
}

// The certificate's verdict in the report (native/reporting.rs:24-100, run_state.rs:170-180).
void apply_certificate(ba_solve_report& r, const Certificate& c, size_t D) {
    r.observability = c.observability;
    if (c.observability == BA_OBSERVABILITY_FULL_RANK) {
        r.pose_rank = r.pose_dimension = (uint32_t)D;
    } else if (c.observability == BA_OBSERVABILITY_RANK_DEFICIENT) {
        r.pose_rank = 0;
        r.pose_dimension = (uint32_t)D;
    } else {
        r.pose_rank = r.pose_dimension = 0;
    }
    // ObservabilityDiagnostics needs a positive dimension
    if (r.pose_dimension == 0) r.pose_rank = 0;
    const bool well_constrained = c.observability != BA_OBSERVABILITY_RANK_DEFICIENT && r.weak_point_count == 0;
    if (!well_constrained) r.status = BA_STATUS_DEGENERATE;
}

// Solution::canonicalize_roundoff_non_descent (solver.rs:1030-1101): a converged solve that only
// raised the objective by round-off returns its input state, with status Skipped.
void canonicalize_roundoff(const ba_problem* p, HostSolution& s, const Refined& first) {
    ba_solve_report& r = s.rep;
    if (p->gauge == BA_GAUGE_FIXED_BASELINE) {
        const double* a = p->pose_center + 3 * p->gauge_first;
        const double* b = p->pose_center + 3 * p->gauge_second;
        const double d0 = a[0] - b[0], d1 = a[1] - b[1], d2 = a[2] - b[2];
        const double initial = std::sqrt(d0 * d0 + d1 * d1 + d2 * d2);
        if (!std::isfinite(initial) ||
            std::fabs(initial - p->gauge_distance) > 512.0 * DBL_EPSILON * std::fmax(initial, p->gauge_distance))
            return;
    }
    const double increase = r.final_cost - r.initial_cost;
    const double tol = 2.0 * std::sqrt(DBL_EPSILON) * std::fmax(std::fabs(r.initial_cost), 1.0);
    const bool converged = r.status == BA_STATUS_CONVERGED_GRADIENT || r.status == BA_STATUS_CONVERGED_STEP ||
                           r.status == BA_STATUS_CONVERGED_COST;
    if (!(increase > 0.0 && increase <= tol && converged && r.observability != BA_OBSERVABILITY_RANK_DEFICIENT)) return;
    bool same = true;
    for (size_t k = 0; k < 9 * p->n_poses && same; ++k) same = s.rot[k] == p->pose_rotation[k];
    for (size_t k = 0; k < 3 * p->n_poses && same; ++k) same = s.cen[k] == p->pose_center[k];
    for (size_t g = 0; g < p->n_intrinsics && same; ++g) {
        const ba_intrinsic& in = p->intrinsics[g];
        const bool brown = in.model != BA_CAMERA_PINHOLE;
        same = s.focal[g] == in.fx && s.k1[g] == (brown && in.model >= BA_CAMERA_BROWN_K1 ? in.k1 : 0.0) &&
               s.k2[g] == (brown && in.model >= BA_CAMERA_BROWN_K1K2 ? in.k2 : 0.0);
    }
    for (size_t k = 0; k < 3 * p->n_points && same; ++k) same = s.xyz[k] == p->point_xyz[k];
    if (!same) {
        std::copy(p->pose_rotation, p->pose_rotation + 9 * p->n_poses, s.rot.begin());
        std::copy(p->pose_center, p->pose_center + 3 * p->n_poses, s.cen.begin());
        for (size_t g = 0; g < p->n_intrinsics; ++g) {
            const ba_intrinsic& in = p->intrinsics[g];
            const bool brown = in.model != BA_CAMERA_PINHOLE;
            s.focal[g] = in.fx;
            s.k1[g] = brown && in.model >= BA_CAMERA_BROWN_K1 ? in.k1 : 0.0;
            s.k2[g] = brown && in.model >= BA_CAMERA_BROWN_K1K2 ? in.k2 : 0.0;
        }
        std::copy(p->point_xyz, p->point_xyz + 3 * p->n_points, s.xyz.begin());
        // the input's rank and gradient need not match the accepted proposal
        r.observability = BA_OBSERVABILITY_NOT_ASSESSED;
        r.pose_rank = r.pose_dimension = 0;
        r.weak_point_count = 0;
    }
    r.final_cost = r.initial_cost;
    r.final_rmse = r.initial_rmse;
    r.status = BA_STATUS_SKIPPED;
    r.inlier_count = first.initial_inliers;
    r.final_inlier_rmse = first.initial_inlier_rmse;
}

// The native solve of one problem (solve_once of native/mod.rs:318-335 with
// solve_self_calibrating, native/self_calibration.rs:223-249): the LM solve; under FULL the
// certificate, and when the marginals leave groups unconstrained a second solve from the input
// state with those groups held; then the round-off canonicalisation.  `pre`: the problem's
// lowering when it is already at hand.
Status solve_once(ba_hip_handle* h, const ba_problem* p, const ba_solve_config* cfg, ba_iteration_trace* trace,
                  bool quarantine, const Lowered* pre, HostSolution& out) {
    const bool full = cfg->certification == BA_CERTIFY_FULL;
    Lowered own;
    if (!pre) {
        Status st = lower_problem(p, own, false, quarantine);
        if (!st.ok()) return st;
    }
    const Lowered& L = pre ? *pre : own;
    ba_execution_plan plan;
    plan_execution(L, cfg, plan);
    const uint32_t trace_start = trace ? trace->count : 0;
    Refined first;
    Status st = refine_lowered(h, p, cfg, L, plan, trace, first);
    if (!st.ok()) return st;
    out.alloc(p);
    // the calibrated path reports the unobserved free intrinsics as unconstrained
    // (native/calibrated.rs:485-487, native/mod.rs:465-477) and re-solves nothing
    std::vector<uint32_t> unobserved_free;
    if (L.calibrated) {
        std::vector<uint8_t> seen(p->n_intrinsics, 0);
        for (size_t o = 0; o < p->n_observations; ++o) seen[p->obs_intrinsic[o]] = 1;
        for (size_t g = 0; g < p->n_intrinsics; ++g)
            if (!seen[g] && !is_fixed_intrinsic(p->intrinsics[g])) unobserved_free.push_back((uint32_t)g);
        for (uint32_t g : unobserved_free) out.unconstrained[g] = 1;
    }
    if (!full) {
        ba_solution v = out.view();
        raise_solution(p, L, first.R.data(), first.t.data(), first.gp.data(), first.X.data(), &v);
        out.rep = first.r;
        out.rep.unconstrained_intrinsics = (uint32_t)unobserved_free.size();
        return st;
    }
    bool exact, marginals;
    certificate_policy(L, cfg, plan, &exact, &marginals);
    const Certificate c = certify(h->stream, L, first, exact, marginals);
    apply_certificate(first.r, c, L.D);
    out.cov = c.cov;
    // beyond this library's limits nothing is certified and nothing is held
    const std::vector<uint32_t> held =
        c.formed ? unconstrained_groups(p, L, first.gp.data(), marginals && c.marginals ? c.cov.data() : nullptr)
                 : std::vector<uint32_t>{};
    if (held.empty()) {
        ba_solution v = out.view();
        raise_solution(p, L, first.R.data(), first.t.data(), first.gp.data(), first.X.data(), &v);
        out.rep = first.r;
    } else {
        // re-solve from the input state with the unconstrained groups at their input values
        Lowered L2;
        st = lower_problem(p, L2, false, quarantine);
        if (!st.ok()) return st;
        hold_groups(L2, held);
        ba_execution_plan plan2;
        plan_execution(L2, cfg, plan2);
        if (trace) trace->count = trace_start;
        Refined second;
        st = refine_lowered(h, p, cfg, L2, plan2, trace, second);
        if (!st.ok()) return st;
        certificate_policy(L2, cfg, plan2, &exact, &marginals);
        apply_certificate(second.r, certify(h->stream, L2, second, exact, false), L2.D);
        ba_solution v = out.view();
        raise_solution(p, L2, second.R.data(), second.t.data(), second.gp.data(), second.X.data(), &v);
        out.rep = second.r;
        first.initial_inliers = second.initial_inliers;
        first.initial_inlier_rmse = second.initial_inlier_rmse;
    }
    for (uint32_t g : held) out.unconstrained[g] = 1;
    out.rep.unconstrained_intrinsics = (uint32_t)(held.size() + unobserved_free.size());
    canonicalize_roundoff(p, out, first);
    return Status{};
}

// solve_with_quarantine (native/mod.rs:252-316): the input-degenerate points held out, the solve,
// and up to four re-solves without the weak points of the previous solution, each warm-started
// from it with at most 8 iterations; held-out points keep their input positions.
Status solve_native(ba_hip_handle* h, const ba_problem* p, const ba_solve_config* cfg, ba_iteration_trace* trace,
                    HostSolution& result) {
    Lowered L;
    Status st = lower_problem(p, L, false, true);  // input quarantine and every lowering check
    if (!st.ok()) return st;
    const size_t N = p->n_points;
    OwnedProblem working_own;
    const ba_problem* working = p;
    std::vector<uint32_t> kept(N);
    for (size_t i = 0; i < N; ++i) kept[i] = (uint32_t)i;
    size_t quarantined = L.n_quarantined;
    if (quarantined) {
        without_points(p, nullptr, L.quarantined, working_own, kept);
        working = &working_own.p;
    }
    const bool changes_scale = p->gauge == BA_GAUGE_FIXED_BASELINE && std::fabs(L.rebase_scale - 1.0) > 1e-9;
    HostSolution sol;
    st = solve_once(h, working, cfg, trace, false, quarantined ? nullptr : &L, sol);
    if (!st.ok()) return st;
    if (cfg->certification == BA_CERTIFY_FULL) {
        ba_solve_config retry = *cfg;
        retry.max_iterations = std::min<uint32_t>(cfg->max_iterations, RESOLVE_ITERATIONS);
        OwnedProblem next_own[2];
        const ba_problem* current = working;
        bool have_origin = false;
        ba_solve_report origin{};
        uint32_t rounds = 0;
        for (uint32_t round = 0; round < MAX_RESOLVES; ++round) {
            if (!needs_resolve(sol.rep)) break;
            std::vector<uint8_t> weak = weak_points(current, sol);
            size_t nweak = 0;
            for (size_t i = 0; i < weak.size(); ++i) {
                if (current->point_fixed && current->point_fixed[i]) weak[i] = 0;
                nweak += weak[i];
            }
            if (!nweak) break;
            OwnedProblem& next = next_own[round % 2];
            std::vector<uint32_t> next_kept;
            without_points(current, &sol, weak, next, next_kept);
            HostSolution resolved;
            if (!solve_once(h, &next.p, &retry, trace, false, nullptr, resolved).ok()) break;
            std::vector<uint32_t> composed(next_kept.size());
            for (size_t i = 0; i < next_kept.size(); ++i) composed[i] = kept[next_kept[i]];
            kept.swap(composed);
            quarantined += nweak;
            if (have_origin) continue_report(sol.rep, origin);
            origin = sol.rep;
            have_origin = true;
            sol = std::move(resolved);
            current = &next.p;
            ++rounds;
        }
        if (have_origin) continue_report(sol.rep, origin);
        sol.rep.weak_point_resolves = rounds;
    }
    if (changes_scale && quarantined > 0)
        return error(BA_ERR_UNSUPPORTED, "native BA cannot change its physical baseline while holding out points");
    if (kept.size() == N) {
        result = std::move(sol);
        result.rep.quarantined_points = 0;
        return Status{};
    }
    // Solution::lifted (solver.rs:891-935): back to the caller's points
    result.alloc(p);
    result.rot = sol.rot;
    result.cen = sol.cen;
    result.focal = sol.focal;
    result.k1 = sol.k1;
    result.k2 = sol.k2;
    result.cov = sol.cov;
    result.unconstrained = sol.unconstrained;
    result.rep = sol.rep;
    std::copy(p->point_xyz, p->point_xyz + 3 * N, result.xyz.begin());
    std::fill(result.quarantined.begin(), result.quarantined.end(), 1);
    for (size_t i = 0; i < kept.size(); ++i) {
        std::copy(sol.xyz.begin() + 3 * i, sol.xyz.begin() + 3 * i + 3, result.xyz.begin() + 3 * (size_t)kept[i]);
        result.quarantined[kept[i]] = 0;
    }
    result.rep.quarantined_points = (uint32_t)quarantined;
    result.rep.point_count = N;
    return Status{};
}

}  // namespace
}  // namespace ba

using namespace ba;

extern "C" {

const char* ba_hip_name(void) { return "openmvg-ba-hip"; }

int ba_hip_create(int device, ba_hip_handle** out) {
    if (!out) return BA_ERR_INVALID_CONFIG;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) return BA_ERR_INTERNAL;
    if (hipSetDevice(device) != hipSuccess) return BA_ERR_INTERNAL;
    auto* h = new ba_hip_handle();
    h->device = device;
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
        delete h;
        return BA_ERR_INTERNAL;
    }
    if (hipHostMalloc((void**)&h->pinned, 4 * sizeof(double)) != hipSuccess) {
        h->pinned = nullptr;
        (void)hipStreamDestroy(h->stream);
        delete h;
        return BA_ERR_INTERNAL;
    }
    ba::noalloc_prepare_device(device);  // with `device` current
    *out = h;
    return BA_OK;
}

int ba_hip_set_schur_product(ba_hip_handle* h, int32_t mode) {
    if (!h || mode < BA_SCHUR_AUTO || mode > BA_SCHUR_EXPLICIT) return BA_ERR_INVALID_CONFIG;
    h->schur_product = mode;
    return BA_OK;
}

int32_t ba_hip_last_schur_product(const ba_hip_handle* h) { return h ? h->last_schur_product : BA_SCHUR_AUTO; }

void ba_hip_destroy(ba_hip_handle* h) {
    if (!h) return;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    if (h->pinned) (void)hipHostFree(h->pinned);
    delete h;
}

int ba_hip_prepare(const ba_problem* p, ba_native_sizes* sizes, ba_native_problem* n, char* msg, size_t len) {
    Status st = validate_problem(p);
    if (!st.ok()) return finish(st, msg, len);
    Lowered L;
    st = lower_problem(p, L, true);
    if (!st.ok()) return finish(st, msg, len);
    if (L.n_cross != L.params.size())  // the count the planner reads, against the map itself
        return finish(error(BA_ERR_INTERNAL, "point parameter count differs from the point parameter map"), msg, len);
    if (sizes) {
        sizes->n_cameras = L.n_cameras();
        sizes->n_groups = L.n_groups();
        sizes->n_points = L.n_points();
        sizes->n_observations = L.n_obs();
        sizes->dimension = L.D;
        sizes->n_point_parameters = L.params.size();
        sizes->n_pattern_entries = L.row_idx.size();
        sizes->n_quarantined = L.n_quarantined;
    }
    if (!n) return BA_OK;
    auto cp = [](void* dst, const void* src, size_t bytes) {
        if (bytes) std::memcpy(dst, src, bytes);
    };
    cp(n->camera_rotation, L.R.data(), L.R.size() * 8);
    cp(n->camera_translation, L.t.data(), L.t.size() * 8);
    cp(n->group_params, L.gp.data(), L.gp.size() * 8);
    cp(n->camera_group, L.cam_group.data(), L.cam_group.size() * 4);
    cp(n->camera_source, L.camera_source.data(), L.camera_source.size() * 4);
    cp(n->point_xyz, L.X.data(), L.X.size() * 8);
    cp(n->point_source, L.point_source.data(), L.point_source.size() * 4);
    cp(n->obs_camera, L.obs_cam.data(), L.obs_cam.size() * 4);
    cp(n->obs_source, L.obs_source.data(), L.obs_source.size() * 4);
    cp(n->obs_uv, L.uv.data(), L.uv.size() * 8);
    cp(n->point_offsets, L.point_off.data(), L.point_off.size() * 4);
    cp(n->camera_fixed, L.cam_fixed.data(), L.cam_fixed.size());
    cp(n->point_fixed, L.point_fixed.data(), L.point_fixed.size());
    cp(n->camera_offset, L.cam_off.data(), L.cam_off.size() * 4);
    cp(n->camera_dim, L.cam_dim.data(), L.cam_dim.size() * 4);
    cp(n->group_offset, L.grp_off.data(), L.grp_off.size() * 4);
    cp(n->group_dim, L.grp_dim.data(), L.grp_dim.size() * 4);
    cp(n->cross_offsets, L.cross_off.data(), L.cross_off.size() * 4);
    cp(n->point_parameters, L.params.data(), L.params.size() * 4);
    cp(n->pattern_col_ptr, L.col_ptr.data(), L.col_ptr.size() * 4);
    cp(n->pattern_row_idx, L.row_idx.data(), L.row_idx.size() * 4);
    n->gauge_camera = L.gauge_camera;
    n->scene_scale = L.scene_scale;
    return BA_OK;
}


int ba_hip_plan(const ba_problem* p, const ba_solve_config* cfg, ba_execution_plan* plan, char* msg, size_t len) {
    if (!cfg || !plan) return finish(error(BA_ERR_INVALID_CONFIG, "null argument"), msg, len);
    Status st = validate_config(cfg);
    if (st.ok()) st = validate_problem(p);
    if (!st.ok()) return finish(st, msg, len);
    try {
        Lowered L;
        st = lower_problem(p, L, false);
        if (!st.ok()) return finish(st, msg, len);
        plan_execution(L, cfg, *plan);
        return BA_OK;
    } catch (const std::bad_alloc&) {
        return finish(error(BA_ERR_RESOURCE_EXHAUSTED, "host allocation failed"), msg, len);
    }
}

int ba_hip_solve(ba_hip_handle* h, const ba_problem* p, const ba_solve_config* cfg, ba_solution* sol,
                 ba_solve_report* rep, ba_iteration_trace* trace, char* msg, size_t len) {
    if (!h || !cfg || !sol) return finish(error(BA_ERR_INVALID_CONFIG, "null argument"), msg, len);
    Status st = validate_config(cfg);
    if (st.ok()) st = validate_problem(p);
    if (!st.ok()) return finish(st, msg, len);
    try {
        if (hipSetDevice(h->device) != hipSuccess) return finish(error(BA_ERR_INTERNAL, "hipSetDevice"), msg, len);
        if (trace) trace->count = 0;
        HostSolution out;
        if (cfg->certification == BA_CERTIFY_FULL) {
            st = solve_native(h, p, cfg, trace, out);
        } else {
            Lowered L;
            st = lower_problem(p, L, false, true);
            if (st.ok()) st = solve_once(h, p, cfg, trace, true, &L, out);
            if (st.ok()) out.rep.quarantined_points = (uint32_t)L.n_quarantined;
        }
        if (!st.ok()) return finish(st, msg, len);
        out.copy_to(sol);
        if (rep) *rep = out.rep;
        return BA_OK;
    } catch (const HipError& e) {
        return finish(error(BA_ERR_INTERNAL, std::string("HIP error: ") + hipGetErrorString(e.code) + " in " + e.what),
                      msg, len);
    } catch (const std::bad_alloc&) {
        return finish(error(BA_ERR_RESOURCE_EXHAUSTED, "host allocation failed"), msg, len);
    }
}

// The hooks read and write the caller's device buffers on the handle's stream, which does not
// synchronise with other streams: they first wait for all work queued on the device (the caller's
// writes to their inputs and outputs), and end with the handle's stream drained.
#define HOOK_BEGIN                                                       \
    if (!h) return BA_ERR_INVALID_CONFIG;                                \
    try {                                                                \
        if (hipSetDevice(h->device) != hipSuccess) return BA_ERR_INTERNAL; \
        if (hipDeviceSynchronize() != hipSuccess) return BA_ERR_INTERNAL;
#define HOOK_END                                                         \
    }                                                                    \
    catch (const HipError&) {                                            \
        return BA_ERR_INTERNAL;                                          \
    }                                                                    \
    catch (const std::bad_alloc&) {                                      \
        return BA_ERR_RESOURCE_EXHAUSTED;                                \
    }

int ba_hip_linearize(ba_hip_handle* h, const ba_device_problem* dp, double* residual, double* point_jacobian,
                     double* camera_jacobian, double* cost, uint8_t* valid) {
    HOOK_BEGIN
    Lowered L;
    Status st = lowered_from_device(dp, L, h->stream);
    if (!st.ok()) return st.code;
    Engine<double> E(h->stream, L, true);
    E.keep_pose_rows();
    E.setup();
    E.set_gauge(gauge_from(dp));
    E.assemble(0, 0.0, L.loss, L.loss_scale, false);
    const size_t no = L.n_obs();
    hipStream_t s = h->stream;
    if (no) {
        // tiled structure-of-arrays rows -> per-observation records (through the host; a test hook)
        std::vector<double> tiles(E.jpm_size()), rec;
        HIP_OK(hipMemcpyAsync(tiles.data(), E.jpm(), tiles.size() * sizeof(double), hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
        auto rows = [&](double* dst, int first, int count) {
            rec.resize((size_t)count * no);
            for (size_t o = 0; o < no; ++o)
                for (int k = 0; k < count; ++k) rec[o * count + k] = tiles[soa_at(first + k, o)];
            HIP_OK(hipMemcpyAsync(dst, rec.data(), rec.size() * sizeof(double), hipMemcpyHostToDevice, s));
            HIP_OK(hipStreamSynchronize(s));
        };
        rows(residual, SOA_R, 2);
        rows(point_jacobian, SOA_JX, 6);
        rows(camera_jacobian, SOA_JC, 18);
        HIP_OK(hipMemcpyAsync(cost, E.obs_cost(), no * 8, hipMemcpyDeviceToDevice, s));
        HIP_OK(hipMemcpyAsync(valid, E.valid(), no, hipMemcpyDeviceToDevice, s));
    }
    HIP_OK(hipStreamSynchronize(s));
    return BA_OK;
    HOOK_END
}

int ba_hip_schur(ba_hip_handle* h, const ba_device_problem* dp, double lambda, double* schur, double* gradient,
                 double* full_gradient, double* diagonal, ba_schur_scalars* sc) {
    HOOK_BEGIN
    Lowered L;
    Status st = lowered_from_device(dp, L, h->stream);
    if (!st.ok()) return st.code;
    Engine<double> E(h->stream, L, false);
    E.setup();
    E.set_gauge(gauge_from(dp));
    AsmScalars a = E.assemble(0, lambda, L.loss, L.loss_scale, true);
    const size_t D = L.D;
    hipStream_t s = h->stream;
    if (D) {
        HIP_OK(hipMemcpyAsync(schur, E.S(), D * D * 8, hipMemcpyDeviceToDevice, s));
        HIP_OK(hipMemcpyAsync(gradient, E.grad(), D * 8, hipMemcpyDeviceToDevice, s));
        HIP_OK(hipMemcpyAsync(full_gradient, E.full_grad(), D * 8, hipMemcpyDeviceToDevice, s));
        HIP_OK(hipMemcpyAsync(diagonal, E.diag(), D * 8, hipMemcpyDeviceToDevice, s));
    }
    HIP_OK(hipStreamSynchronize(s));
    if (sc) {
        sc->cost = a.cost;
        sc->squared_residual = a.sq;
        sc->point_model_reduction = a.pmr;
        sc->scaled_gradient_inf_norm = std::fmax(a.sgin, a.sgin_prepared);
        sc->minimum_point_rcond = a.min_rcond;
        sc->weak_point_count = (uint32_t)a.weak;
        sc->valid = a.invalid == 0.0 && a.nonfinite == 0.0;
    }
    return BA_OK;
    HOOK_END
}

int ba_hip_pcg(ba_hip_handle* h, const ba_device_problem* dp, double lambda, double tol, uint32_t max_iterations,
               double* step, uint32_t* iterations, double* relative_residual, int32_t* accepted) {
    HOOK_BEGIN
    Lowered L;
    Status st = lowered_from_device(dp, L, h->stream);
    if (!st.ok()) return st.code;
    Engine<double> E(h->stream, L, true);
    E.set_schur_product(h->schur_product);
    E.setup();
    E.set_gauge(gauge_from(dp));
    AsmScalars a = E.assemble(0, lambda, L.loss, L.loss_scale, true);  // the explicit system when set up
    uint32_t terms = 0;
    double rel = NAN;
    bool ok = (a.invalid == 0.0 && a.nonfinite == 0.0) && E.solve_pcg(tol, (int)max_iterations, &terms, &rel);
    h->last_schur_product = E.explicit_product() ? BA_SCHUR_EXPLICIT : BA_SCHUR_IMPLICIT;
    if (L.D) HIP_OK(hipMemcpyAsync(step, E.pcg_x(), L.D * 8, hipMemcpyDeviceToDevice, h->stream));
    HIP_OK(hipStreamSynchronize(h->stream));
    if (iterations) *iterations = terms;
    if (relative_residual) *relative_residual = rel;
    if (accepted) *accepted = ok ? 1 : 0;
    return BA_OK;
    HOOK_END
}

int ba_hip_cholesky_solve(ba_hip_handle* h, size_t n, const double* a, const double* b, double* x, int32_t precision) {
    HOOK_BEGIN
    if (precision != BA_PRECISION_FP64 && precision != BA_PRECISION_MIXED) return BA_ERR_INVALID_CONFIG;
    if (n == 0) return BA_OK;
    hipStream_t s = h->stream;
    Buf<double> work;
    Buf<int> fail;
    work.alloc(cholesky_work_doubles((uint32_t)n));
    fail.alloc(1);
    int f = 0;
    if (precision == BA_PRECISION_MIXED) {  // FP32 factor, FP64 refinement against the full matrix
        Buf<float> F;
        Buf<int> done;
        F.alloc(n * n);
        done.alloc(1);
        launch_cholesky_solve_mixed(s, a, F.p, (uint32_t)n, (uint32_t)n, b, x, work.p, fail.p, done.p, 8);
        double st[3];
        HIP_OK(hipMemcpyAsync(&f, fail.p, sizeof(int), hipMemcpyDeviceToHost, s));
        HIP_OK(hipMemcpyAsync(st, work.p + 8, 3 * sizeof(double), hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
        const double rel = std::sqrt(st[1] / st[2]);
        return (f || !(rel <= 1e-10)) ? BA_ERR_NUMERICAL : BA_OK;
    }
    Buf<double> A;
    A.alloc(n * n);
    HIP_OK(hipMemcpyAsync(A.p, a, n * n * 8, hipMemcpyDeviceToDevice, s));
    launch_cholesky_solve(s, A.p, (uint32_t)n, (uint32_t)n, b, x, work.p, fail.p);
    HIP_OK(hipMemcpyAsync(&f, fail.p, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    return f ? BA_ERR_NUMERICAL : BA_OK;
    HOOK_END
}

}  // extern "C"
