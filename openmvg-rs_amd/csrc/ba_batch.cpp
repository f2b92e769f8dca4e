// ba_batch.cpp — batched local bundle adjustment: host side and C entry point.
//
// ba_hip_solve_batched solves many independent windows (BundleProblems) with results equal to
// calling ba_hip_solve on each window with the direct camera step: each window keeps its own LM
// state (damping, acceptance, stop tests) on the device, while every pass runs over all windows
// in one launch.  One LM iteration is a fixed sequence of launches; it is captured once in a HIP
// graph and replayed for every iteration.  The per-window dense Cholesky tile edge is chosen by a
// small autotuner that times both candidates on the batch's own window sizes.
#include <hip/hip_runtime.h>

#include <map>

#include "ba_engine.hpp"

namespace ba {
namespace {

struct Windows {
    size_t W = 0;
    std::vector<Lowered> parts;
    Lowered L;  // the windows concatenated
    std::vector<uint32_t> cam_off{0}, grp_off{0}, pt_off{0}, obs_off{0}, param_off{0};
    std::vector<uint64_t> S_off{0}, linv_off{0};
    std::vector<uint32_t> cam_win, grp_win, pt_win, blk_win;
    std::vector<int32_t> gauge_cam;
    std::vector<double> scene;
};

template <class V>
void append(V& dst, const V& src) {
    dst.insert(dst.end(), src.begin(), src.end());
}

size_t linv_size(size_t D) {  // the larger of the two tile edges' needs
    const size_t a = (D + 31) / 32 * 32 * 32, b = (D + 63) / 64 * 64 * 64;
    return a > b ? a : b;
}

Status merge_windows(const ba_problem* probs, size_t W, Windows& B) {
    B.W = W;
    B.parts.resize(W);
    Lowered& M = B.L;
    M.point_off.push_back(0);
    for (size_t w = 0; w < W; ++w) {
        Lowered& P = B.parts[w];
        Status st = validate_problem(&probs[w]);
        if (st.ok()) st = lower_problem(&probs[w], P, false);
        if (!st.ok()) return error(st.code, "window " + std::to_string(w) + ": " + st.message);
        if (w > 0 && (P.loss != B.parts[0].loss || P.loss_scale != B.parts[0].loss_scale))
            return error(BA_ERR_UNSUPPORTED, "batched windows must share one robust loss");
        if (P.loss == BA_LOSS_CAUCHY)
            return error(BA_ERR_UNSUPPORTED, "batched windows run single-stage losses (squared, Huber)");
        const uint32_t cb = B.cam_off.back(), gb = B.grp_off.back(), pb = B.pt_off.back(), ob = B.obs_off.back();
        const uint32_t db = B.param_off.back();
        const size_t nc = P.n_cameras(), ng = P.n_groups(), np = P.n_points(), no = P.n_obs();
        append(M.R, P.R);
        append(M.t, P.t);
        append(M.gp, P.gp);
        append(M.X, P.X);
        append(M.uv, P.uv);
        append(M.cam_fixed, P.cam_fixed);
        append(M.point_fixed, P.point_fixed);
        append(M.grp_fixed, P.grp_fixed);
        for (size_t c = 0; c < nc; ++c) {
            M.cam_group.push_back(P.cam_group[c] + gb);
            M.cam_off.push_back(P.cam_off[c] + db);
            M.cam_dim.push_back(P.cam_dim[c]);
            B.cam_win.push_back((uint32_t)w);
        }
        for (size_t g = 0; g < ng; ++g) {
            M.grp_off.push_back(P.grp_off[g] + db);
            M.grp_dim.push_back(P.grp_dim[g]);
            B.grp_win.push_back((uint32_t)w);
        }
        for (size_t o = 0; o < no; ++o) {
            M.obs_cam.push_back(P.obs_cam[o] + cb);
            M.obs_point.push_back(P.obs_point[o] + pb);
        }
        for (size_t p = 0; p < np; ++p) {
            M.point_off.push_back(P.point_off[p + 1] + ob);
            B.pt_win.push_back((uint32_t)w);
        }
        B.gauge_cam.push_back(P.gauge_camera >= 0 ? (int32_t)(P.gauge_camera + cb) : -1);
        B.scene.push_back(P.scene_scale);
        B.cam_off.push_back(cb + (uint32_t)nc);
        B.grp_off.push_back(gb + (uint32_t)ng);
        B.pt_off.push_back(pb + (uint32_t)np);
        B.obs_off.push_back(ob + (uint32_t)no);
        B.param_off.push_back(db + (uint32_t)P.D);
        B.S_off.push_back(B.S_off.back() + (uint64_t)P.D * P.D);
        B.linv_off.push_back(B.linv_off.back() + linv_size(P.D));
    }
    M.D = B.param_off.back();
    M.loss = B.parts[0].loss;
    M.loss_scale = B.parts[0].loss_scale;
    M.gauge_camera = -1;
    M.scene_scale = 1.0;
    const size_t NC = M.n_cameras();
    B.blk_win.resize(NC + M.n_groups());
    for (size_t c = 0; c < NC; ++c) B.blk_win[c] = B.cam_win[c];
    for (size_t g = 0; g < M.n_groups(); ++g) B.blk_win[NC + g] = B.grp_win[g];
    return Status{};
}

template <class T>
class BatchEngine : public Engine<T> {
    using E = Engine<T>;

   public:
    BatchEngine(hipStream_t s, const Windows& B) : E(s, B.L, false), B_(B) { this->batched_ = true; }

    void setup_batch(const ba_solve_config* cfg) {
        this->setup();
        const size_t W = B_.W, D = this->L_.D;
        cam_off_.upload(B_.cam_off, this->s_);
        grp_off_.upload(B_.grp_off, this->s_);
        pt_off_.upload(B_.pt_off, this->s_);
        param_off_.upload(B_.param_off, this->s_);
        S_off_.upload(B_.S_off, this->s_);
        linv_off_.upload(B_.linv_off, this->s_);
        cam_win_.upload(B_.cam_win, this->s_);
        grp_win_.upload(B_.grp_win, this->s_);
        pt_win_.upload(B_.pt_win, this->s_);
        blk_win_.upload(B_.blk_win, this->s_);
        gauge_cam_.upload(B_.gauge_cam, this->s_);
        scene_.upload(B_.scene, this->s_);
        wt_.W = (uint32_t)W;
        wt_.cam_off = cam_off_.p;
        wt_.grp_off = grp_off_.p;
        wt_.pt_off = pt_off_.p;
        wt_.param_off = param_off_.p;
        wt_.S_off = S_off_.p;
        wt_.linv_off = linv_off_.p;
        wt_.grp_win = grp_win_.p;
        wt_.blk_win = blk_win_.p;
        wt_.gauge_cam = gauge_cam_.p;
        wt_.scene = scene_.p;
        this->gauge_.alloc(W);
        Sb_.alloc(B_.S_off.back());
        if (std::is_same<T, float>::value) {
            Fb_.alloc(B_.S_off.back());
            linv64_.alloc(B_.linv_off.back());
            fallbacks_.alloc(W);
            HIP_OK(hipMemsetAsync(fallbacks_.p, 0, W * sizeof(uint32_t), this->s_));
        }
        linv_.alloc(B_.linv_off.back());
        vecs_.alloc(4 * D + 1);
        fail_w_.alloc(W);
        rel_w_.alloc(W);
        lm_.alloc(W);
        lam_.alloc(W);
        lam0_.alloc(W);
        asm_w_.alloc(W);
        cand_w_.alloc(W);
        dots_w_.alloc(4 * W);
        this->tp_.cam_win = cam_win_.p;
        this->tp_.pt_win = pt_win_.p;
        this->tp_.win_lambda = lam_.p;
        // LM state: lambda_0 = 1e-4 for the direct step (sparse/api.rs:300-305), nu = 2
        std::vector<LmWin> init(W);
        for (auto& s : init) {
            std::memset(&s, 0, sizeof(s));
            s.nu = 2.0;
            s.gain = NAN;
            s.status = -1;
        }
        lm_.upload(init, this->s_);
        lam_.upload(std::vector<double>(W, 1e-4), this->s_);
        lam0_.upload(std::vector<double>(W, 0.0), this->s_);
        HIP_OK(hipMemsetAsync(this->step_.p, 0, D * sizeof(double), this->s_));
        params_.max_iterations = cfg->max_iterations;
        params_.ftol = cfg->function_tolerance;
        params_.gtol = cfg->gradient_tolerance;
        params_.ptol = cfg->parameter_tolerance;
        HIP_OK(hipStreamSynchronize(this->s_));
    }

    // Assembly of every window at its own damping (damp) or at zero damping (the final one).
    void assemble(bool damp) {
        Topo tp = this->tp_;
        tp.win_lambda = damp ? lam_.p : lam0_.p;
        hipStream_t s = this->s_;
        launch_win_gauge(s, wt_, this->t_[0].p, this->gauge_.p);
        launch_linearize<T>(s, tp, this->R_[0].p, this->t_[0].p, this->gp_[0].p, this->X_[0].p, this->uv_.p,
                            this->gauge_.p, this->lin());
        launch_point_elim<T>(s, tp, this->lin(), (T)0, !damp, this->pe());
        launch_win_reduce_points(s, wt_, this->pscal_.p, asm_w_.p);
        launch_block_gradient<T>(s, tp, this->lin(), this->pe(), this->grad_.p, this->fgrad_.p, this->diag_.p);
        if (damp) {
            launch_pairs<T>(s, tp, this->lin(), this->pe(), (uint32_t)this->PL_.block_a.size(), this->pair_a_.p,
                            this->pair_b_.p, this->term_off_.p, this->term_i_.p, this->term_j_.p, this->value_off_.p,
                            this->u_vals_.p, this->s_vals_.p);
            launch_win_zero(s, Sb_.p, Sb_.n);
            launch_win_scatter(s, tp, wt_, (uint32_t)this->PL_.block_a.size(), this->pair_a_.p, this->pair_b_.p,
                               this->value_off_.p, this->u_vals_.p, this->s_vals_.p, Sb_.p);
        }
        launch_win_finalize(s, tp, wt_, tp.win_lambda, this->diag_.p, this->fgrad_.p, this->grad_.p, Sb_.p, asm_w_.p,
                            damp);
    }

    void cholesky(int tb) {
        launch_win_cholesky<T>(this->s_, wt_, tb, Sb_.p, Fb_.p, linv_.p, this->grad_.p, this->step_.p, vecs_.p, lm_.p,
                               fail_w_.p, rel_w_.p, linv64_.p, fallbacks_.p);
    }

    // One LM iteration of every window: a fixed sequence of launches.
    void iteration(int tb, WinTrace tr) {
        hipStream_t s = this->s_;
        Topo tp = this->tp_;
        assemble(true);
        launch_lm_begin(s, wt_, lm_.p, lam_.p, asm_w_.p, params_, tr);
        cholesky(tb);
        launch_lm_solved(s, wt_, lm_.p, lam_.p, fail_w_.p, params_);
        launch_retract(s, tp, this->R_[0].p, this->t_[0].p, this->gp_[0].p, this->step_.p, this->gauge_.p,
                       this->R_[1].p, this->t_[1].p, this->gp_[1].p);
        launch_candidate<T>(s, tp, this->R_[0].p, this->t_[0].p, this->gp_[0].p, this->R_[1].p, this->t_[1].p,
                            this->gp_[1].p,
                            this->X_[0].p, this->uv_.p, this->lin(), this->pe(), this->step_.p, this->gauge_.p,
                            this->srec_.p, this->X_[1].p, this->cscal_.p);
        launch_win_reduce_cand(s, wt_, this->cscal_.p, cand_w_.p);
        launch_win_step_dots(s, wt_, this->grad_.p, this->diag_.p, this->step_.p, dots_w_.p);
        launch_lm_end(s, wt_, lm_.p, lam_.p, cand_w_.p, dots_w_.p, params_, tr);
        launch_win_accept(s, tp, wt_, lm_.p, this->R_[0].p, this->t_[0].p, this->gp_[0].p, this->X_[0].p,
                          this->R_[1].p, this->t_[1].p, this->gp_[1].p, this->X_[1].p);
    }

    // Tile edge of the per-window Cholesky: both candidates timed on this batch's window sizes with
    // a synthetic SPD system (identity plus a small symmetric perturbation); the faster wins.
    int autotune(double seconds_out[2]) {
        hipStream_t s = this->s_;
        std::vector<LmWin> go(B_.W);
        for (auto& g : go) {
            std::memset(&g, 0, sizeof(g));
            g.go = 1;
        }
        Buf<LmWin> lm_go;
        lm_go.upload(go, s);
        std::vector<double> Sh(B_.S_off.back());
        for (size_t w = 0; w < B_.W; ++w) {
            const size_t n = B_.param_off[w + 1] - B_.param_off[w];
            double* Sw = Sh.data() + B_.S_off[w];
            for (size_t i = 0; i < n; ++i)
                for (size_t j = 0; j < n; ++j) Sw[i * n + j] = (i == j) ? 2.0 : 1e-3 / (1.0 + (double)(i + j));
        }
        HIP_OK(hipMemsetAsync(this->grad_.p, 0, this->L_.D * sizeof(double), s));  // right-hand side 0
        int best = 64;
        double best_t = INFINITY;
        hipEvent_t a, b;
        HIP_OK(hipEventCreate(&a));
        HIP_OK(hipEventCreate(&b));
        const int tiles[2] = {32, 64};
        for (int k = 0; k < 2; ++k) {
            double tmin = INFINITY;
            for (int rep = 0; rep < 3; ++rep) {
                HIP_OK(hipMemcpyAsync(Sb_.p, Sh.data(), Sh.size() * sizeof(double), hipMemcpyHostToDevice, s));
                HIP_OK(hipEventRecord(a, s));
                launch_win_cholesky<T>(s, wt_, tiles[k], Sb_.p, Fb_.p, linv_.p, this->grad_.p, this->step_.p,
                                       vecs_.p, lm_go.p, fail_w_.p, rel_w_.p, nullptr, nullptr);
                HIP_OK(hipEventRecord(b, s));
                HIP_OK(hipEventSynchronize(b));
                float ms = 0.0f;
                HIP_OK(hipEventElapsedTime(&ms, a, b));
                tmin = std::fmin(tmin, ms * 1e-3);
            }
            seconds_out[k] = tmin;
            if (tmin < best_t) {
                best_t = tmin;
                best = tiles[k];
            }
        }
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        HIP_OK(hipMemsetAsync(this->step_.p, 0, this->L_.D * sizeof(double), s));
        HIP_OK(hipStreamSynchronize(s));
        return best;
    }

    std::vector<uint32_t> fallbacks() {
        std::vector<uint32_t> f(B_.W, 0);
        if (fallbacks_.p)
            HIP_OK(hipMemcpy(f.data(), fallbacks_.p, B_.W * sizeof(uint32_t), hipMemcpyDeviceToHost));
        return f;
    }

    void download_windows(std::vector<LmWin>& lm, std::vector<double>& lam, std::vector<AsmScalars>& fin,
                          std::vector<double>& rel) {
        const size_t W = B_.W;
        lm.resize(W);
        lam.resize(W);
        fin.resize(W);
        rel.resize(W);
        hipStream_t s = this->s_;
        HIP_OK(hipMemcpyAsync(lm.data(), lm_.p, W * sizeof(LmWin), hipMemcpyDeviceToHost, s));
        HIP_OK(hipMemcpyAsync(lam.data(), lam_.p, W * sizeof(double), hipMemcpyDeviceToHost, s));
        HIP_OK(hipMemcpyAsync(fin.data(), asm_w_.p, W * sizeof(AsmScalars), hipMemcpyDeviceToHost, s));
        HIP_OK(hipMemcpyAsync(rel.data(), rel_w_.p, W * sizeof(double), hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
    }

   private:
    const Windows& B_;
    WinTopo wt_{};
    LmParams params_{};
    Buf<uint32_t> cam_off_, grp_off_, pt_off_, param_off_, cam_win_, grp_win_, pt_win_, blk_win_;
    Buf<uint64_t> S_off_, linv_off_;
    Buf<int32_t> gauge_cam_;
    Buf<double> scene_, Sb_, vecs_, rel_w_, lam_, lam0_, dots_w_, linv64_;
    Buf<uint32_t> fallbacks_;
    Buf<T> Fb_, linv_;
    Buf<int> fail_w_;
    Buf<LmWin> lm_;
    Buf<AsmScalars> asm_w_;
    Buf<CandScalars> cand_w_;
};

template <class T>
Status run_batched(ba_hip_handle* h, const ba_problem* probs, size_t W, const ba_solve_config* cfg, ba_solution* sols,
                   ba_solve_report* reps, ba_batch_trace* trace, int mode) {
    double t0 = now();
    Windows B;
    Status st = merge_windows(probs, W, B);
    if (!st.ok()) return st;
    BatchEngine<T> E(h->stream, B);
    E.setup_batch(cfg);
    {  // validation::projection_domain: every observation projects at the initial state
        E.assemble(false);
        std::vector<LmWin> lm0;
        std::vector<double> lam0, rel0;
        std::vector<AsmScalars> a0;
        E.download_windows(lm0, lam0, a0, rel0);
        for (size_t w = 0; w < W; ++w)
            if (a0[w].unprojectable > 0.0)
                return error(BA_ERR_NUMERICAL,
                             "window " + std::to_string(w) + ": point not projectable at the initial state");
    }
    double tune[2] = {NAN, NAN};
    const int tb = E.autotune(tune);
    const uint32_t cap = cfg->max_iterations;
    Buf<double> tr_cost, tr_lam, tr_cand;
    Buf<uint8_t> tr_acc;
    WinTrace tr{nullptr, nullptr, nullptr, nullptr};
    if (trace && trace->capacity) {
        tr_cost.alloc((size_t)cap * W);
        tr_lam.alloc((size_t)cap * W);
        tr_cand.alloc((size_t)cap * W);
        tr_acc.alloc((size_t)cap * W);
        tr = WinTrace{tr_cost.p, tr_lam.p, tr_cand.p, tr_acc.p};
        HIP_OK(hipMemsetAsync(tr_acc.p, 0, (size_t)cap * W, h->stream));
    }
    HIP_OK(hipStreamSynchronize(h->stream));
    double t1 = now();
    if (mode == BA_BATCH_GRAPH) {
        hipGraph_t graph;
        hipGraphExec_t exec;
        HIP_OK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
        E.iteration(tb, tr);
        HIP_OK(hipStreamEndCapture(h->stream, &graph));
        HIP_OK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        for (uint32_t it = 0; it < cap; ++it) HIP_OK(hipGraphLaunch(exec, h->stream));
        HIP_OK(hipStreamSynchronize(h->stream));
        (void)hipGraphExecDestroy(exec);
        (void)hipGraphDestroy(graph);
    } else {
        for (uint32_t it = 0; it < cap; ++it) E.iteration(tb, tr);
        HIP_OK(hipStreamSynchronize(h->stream));
    }
    double t2 = now();
    E.assemble(false);  // final zero-damping assembly
    std::vector<LmWin> lm;
    std::vector<double> lam, rel;
    std::vector<AsmScalars> fin;
    E.download_windows(lm, lam, fin, rel);
    std::vector<double> R, t, gp, X;
    E.download(0, R, t, gp, X);
    const std::vector<uint32_t> fallbacks = E.fallbacks();
    for (size_t w = 0; w < W; ++w) {
        const Lowered& P = B.parts[w];
        const size_t c0 = B.cam_off[w], g0 = B.grp_off[w], p0 = B.pt_off[w];
        raise_solution(&probs[w], P, R.data() + 9 * c0, t.data() + 3 * c0, gp.data() + 3 * g0, X.data() + 3 * p0,
                       &sols[w]);
        if (!reps) continue;
        ba_solve_report& r = reps[w];
        std::memset(&r, 0, sizeof(r));
        const LmWin& s = lm[w];
        const double no = (double)P.n_obs();
        if (s.status >= 0)
            r.status = s.status;
        else
            r.status = (s.iterations > 0 && s.target_accepted == 0) ? BA_STATUS_STALLED : BA_STATUS_MAX_ITERATIONS;
        r.linear_solver_used = BA_SOLVER_DIRECT;
        r.iterations = (uint32_t)s.iterations;
        r.accepted_steps = (uint32_t)s.accepted;
        r.rejected_linearization = (uint32_t)s.rej[0];
        r.rejected_step_solve = (uint32_t)s.rej[1];
        r.rejected_candidate = (uint32_t)s.rej[2];
        r.rejected_insufficient_reduction = (uint32_t)s.rej[3];
        r.rejected_steps = r.rejected_linearization + r.rejected_step_solve + r.rejected_candidate +
                           r.rejected_insufficient_reduction;
        r.initial_cost = s.have_initial ? s.initial_cost : INFINITY;
        r.final_cost = std::isfinite(fin[w].cost) ? fin[w].cost : INFINITY;
        r.initial_rmse = s.have_initial ? std::sqrt(s.initial_sq / (2.0 * no)) * std::sqrt(2.0) : INFINITY;
        r.final_rmse = std::sqrt(fin[w].sq / (2.0 * no)) * std::sqrt(2.0);
        r.gradient_inf_norm = std::fmax(fin[w].sgin, fin[w].sgin_prepared);
        r.step_norm = s.step_norm;
        r.initial_damping = 1e-4;
        r.final_damping = lam[w];
        r.last_gain_ratio = s.has_gain ? s.gain : NAN;
        r.reduced_dimension = (uint32_t)P.D;
        r.weak_point_count = (uint32_t)fin[w].weak;
        r.minimum_point_rcond = fin[w].min_rcond;
        r.maximum_pcg_relative_residual = NAN;
        r.normalization_scale = P.scene_scale;
        r.observation_count = P.n_obs();
        r.point_count = P.n_points();
        r.pose_count = P.n_cameras();
        r.quarantined_points = (uint32_t)P.n_quarantined;
        r.seconds_setup = t1 - t0;
        r.seconds_iterations = t2 - t1;
        r.refinement_residual = rel[w];
        r.precision_fallbacks = fallbacks[w];
    }
    if (trace) {
        trace->tile = (uint32_t)tb;
        trace->autotune_seconds[0] = tune[0];
        trace->autotune_seconds[1] = tune[1];
        trace->seconds_loop = t2 - t1;
        if (trace->capacity) {
            const size_t n = (size_t)std::min(trace->capacity, cap) * W;
            if (trace->cost) HIP_OK(hipMemcpy(trace->cost, tr_cost.p, n * sizeof(double), hipMemcpyDeviceToHost));
            if (trace->damping) HIP_OK(hipMemcpy(trace->damping, tr_lam.p, n * sizeof(double), hipMemcpyDeviceToHost));
            if (trace->candidate_cost)
                HIP_OK(hipMemcpy(trace->candidate_cost, tr_cand.p, n * sizeof(double), hipMemcpyDeviceToHost));
            if (trace->accepted) HIP_OK(hipMemcpy(trace->accepted, tr_acc.p, n, hipMemcpyDeviceToHost));
        }
    }
    return Status{};
}

int finish_batch(const Status& st, char* msg, size_t len) {
    if (msg && len) std::snprintf(msg, len, "%s", st.message.c_str());
    return st.code;
}

}  // namespace
}  // namespace ba

using namespace ba;

extern "C" int ba_hip_solve_batched(ba_hip_handle* h, size_t n_windows, const ba_problem* problems,
                                    const ba_solve_config* cfg, ba_solution* solutions, ba_solve_report* reports,
                                    ba_batch_trace* trace, int32_t mode, char* msg, size_t len) {
    if (!h || !problems || !cfg || !solutions) return finish_batch(error(BA_ERR_INVALID_CONFIG, "null argument"), msg, len);
    if (n_windows == 0) return BA_OK;
    Status st = validate_config(cfg);
    if (!st.ok()) return finish_batch(st, msg, len);
    if (cfg->linear_solver == BA_SOLVER_ITERATIVE)
        return finish_batch(error(BA_ERR_UNSUPPORTED, "batched windows use the direct camera step"), msg, len);
    if (mode != BA_BATCH_GRAPH && mode != BA_BATCH_STREAM)
        return finish_batch(error(BA_ERR_INVALID_CONFIG, "mode"), msg, len);
    try {
        if (hipSetDevice(h->device) != hipSuccess) return finish_batch(error(BA_ERR_INTERNAL, "hipSetDevice"), msg, len);
        st = cfg->precision == BA_PRECISION_MIXED
                 ? run_batched<float>(h, problems, n_windows, cfg, solutions, reports, trace, mode)
                 : run_batched<double>(h, problems, n_windows, cfg, solutions, reports, trace, mode);
        return finish_batch(st, msg, len);
    } catch (const HipError& e) {
        return finish_batch(
            error(BA_ERR_INTERNAL, std::string("HIP error: ") + hipGetErrorString(e.code) + " in " + e.what), msg, len);
    } catch (const std::bad_alloc&) {
        return finish_batch(error(BA_ERR_RESOURCE_EXHAUSTED, "host allocation failed"), msg, len);
    }
}
