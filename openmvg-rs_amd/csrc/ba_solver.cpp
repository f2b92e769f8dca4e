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
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/ba_hip.h"
#include "ba_handle.hpp"
#include "ba_host.hpp"
#include "ba_kernels.hpp"

namespace ba {
namespace {

constexpr int REDUCE_BLOCKS = 512;

struct HipError {
    hipError_t code;
    const char* what;
};
#define HIP_OK(expr)                                 \
    do {                                             \
        hipError_t _e = (expr);                      \
        if (_e != hipSuccess) throw HipError{_e, #expr}; \
    } while (0)

template <class T>
struct Buf {
    T* p = nullptr;
    size_t n = 0;
    Buf() = default;
    Buf(const Buf&) = delete;
    Buf& operator=(const Buf&) = delete;
    ~Buf() {
        if (p) (void)hipFree(p);
    }
    void alloc(size_t count) {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = count;
        if (count) HIP_OK(hipMalloc(&p, count * sizeof(T)));
    }
    void upload(const T* h, size_t count, hipStream_t s) {
        alloc(count);
        if (count) HIP_OK(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void upload(const std::vector<T>& v, hipStream_t s) { upload(v.data(), v.size(), s); }
};

template <class T>
struct Pinned {
    T* p = nullptr;
    Pinned() { HIP_OK(hipHostMalloc(&p, sizeof(T))); }
    ~Pinned() {
        if (p) (void)hipHostFree(p);
    }
};

double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Nielsen {  // numerics/refine/nonlinear/controller.rs:15-91
    double value, nu = 2.0, gain = NAN;
    bool has_gain = false;
    void increase() {
        value = std::fmin(value * nu, 1e12);
        nu = std::fmin(nu * 2.0, 1e12);
    }
    void reject() {
        has_gain = false;
        increase();
    }
    bool accept(double actual, double predicted) {
        double g = (std::isfinite(actual) && std::isfinite(predicted) && predicted > 0.0) ? actual / predicted : NAN;
        has_gain = std::isfinite(g);
        gain = g;
        if (!(std::isfinite(g) && g > 0.0)) {
            increase();
            return false;
        }
        double c = 2.0 * g - 1.0;
        value = std::fmax(value * std::fmax(1.0 - c * c * c, 1.0 / 3.0), 1e-12);
        nu = 2.0;
        return true;
    }
};

// Device copy of a lowered problem and every buffer of the loop.  T is the storage and
// arithmetic type of the per-observation Jacobians and per-point blocks: double for the plain
// form, float for the mixed form (state, sums, reduced-system vectors and the LM control stay
// FP64 in both).
template <class T>
class Engine {
   public:
    Engine(hipStream_t s, const Lowered& L, bool iterative, int profile = BA_PROFILE_NONE)
        : s_(s), L_(L), iterative_(iterative), profile_(profile) {}
    ~Engine() {
        for (auto& e : events_) {
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
    }

    // Run `launch`, bracketed by HIP events on this stream when `which` is the profiled kernel.
    template <class F>
    void timed(int which, F&& launch) {
        if (which != profile_) {
            launch();
            return;
        }
        if (used_ == events_.size()) {
            hipEvent_t a, b;
            HIP_OK(hipEventCreate(&a));
            HIP_OK(hipEventCreate(&b));
            events_.emplace_back(a, b);
        }
        auto& e = events_[used_++];
        HIP_OK(hipEventRecord(e.first, s_));
        launch();
        HIP_OK(hipEventRecord(e.second, s_));
    }
    // Launch count and summed duration of the profiled kernel so far.
    void profile_result(uint32_t* launches, double* seconds) {
        HIP_OK(hipStreamSynchronize(s_));
        double total = 0.0;
        for (size_t k = 0; k < used_; ++k) {
            float ms = 0.0f;
            HIP_OK(hipEventElapsedTime(&ms, events_[k].first, events_[k].second));
            total += ms * 1e-3;
        }
        *launches = (uint32_t)used_;
        *seconds = total;
    }

    void setup() {
        B_ = build_block_lists(L_);
        PL_ = build_pair_lists(L_, B_, !iterative_);
        const size_t nc = L_.n_cameras(), ng = L_.n_groups(), np = L_.n_points(), no = L_.n_obs(), D = L_.D;
        obs_cam_.upload(L_.obs_cam, s_);
        obs_pt_.upload(L_.obs_point, s_);
        pt_off_.upload(L_.point_off, s_);
        pt_fixed_.upload(L_.point_fixed, s_);
        cam_group_.upload(L_.cam_group, s_);
        cam_off_.upload(L_.cam_off, s_);
        cam_dim_.upload(L_.cam_dim, s_);
        grp_off_.upload(L_.grp_off, s_);
        grp_dim_.upload(L_.grp_dim, s_);
        blk_off_.upload(B_.block_off, s_);
        blk_dim_.upload(B_.block_dim, s_);
        blk_obs_off_.upload(B_.obs_off, s_);
        blk_obs_.upload(B_.obs, s_);
        {  // cameras of each group, ascending
            std::vector<uint32_t> off(ng + 1, 0), cams(nc);
            for (size_t c = 0; c < nc; ++c) off[L_.cam_group[c] + 1]++;
            for (size_t g = 0; g < ng; ++g) off[g + 1] += off[g];
            std::vector<uint32_t> cur(off.begin(), off.end() - 1);
            for (size_t c = 0; c < nc; ++c) cams[cur[L_.cam_group[c]]++] = (uint32_t)c;
            grp_cam_off_.upload(off, s_);
            grp_cams_.upload(cams, s_);
            HIP_OK(hipStreamSynchronize(s_));  // the host vectors end with this scope
        }
        uv_.upload(L_.uv, s_);
        for (int k = 0; k < 2; ++k) {
            R_[k].alloc(9 * nc);
            t_[k].alloc(3 * nc);
            gp_[k].alloc(3 * ng);
            X_[k].alloc(3 * np);
        }
        HIP_OK(hipMemcpyAsync(R_[0].p, L_.R.data(), 9 * nc * 8, hipMemcpyHostToDevice, s_));
        HIP_OK(hipMemcpyAsync(t_[0].p, L_.t.data(), 3 * nc * 8, hipMemcpyHostToDevice, s_));
        HIP_OK(hipMemcpyAsync(gp_[0].p, L_.gp.data(), 3 * ng * 8, hipMemcpyHostToDevice, s_));
        HIP_OK(hipMemcpyAsync(X_[0].p, L_.X.data(), 3 * np * 8, hipMemcpyHostToDevice, s_));
        fast_cam_ = iterative_ && simple_topology();
        build_matvec_order(iterative_);
        jpm_.alloc(SOA_ROWS * no);
        if (!fast_cam_) jc_.alloc(18 * no);  // block-pair path
        ocost_.alloc(no);
        osq_.alloc(no);
        valid_.alloc(no);
        vinv_.alloc(6 * np);
        vdamp_.alloc(6 * np);
        y_.alloc(3 * np);
        ddiag_.alloc(3 * np);
        if (!fast_cam_) rrows_.alloc(27 * no);
        pscal_.alloc(6 * np);
        cscal_.alloc(4 * np);
        grad_.alloc(D);
        fgrad_.alloc(D);
        diag_.alloc(D);
        step_.alloc(D);
        rhs_.alloc(D);
        partials_.alloc(8 * REDUCE_BLOCKS);
        asm_.alloc(1);
        cand_.alloc(1);
        gauge_.alloc(1);
        fail_.alloc(1);
        dots_.alloc(4);
        pair_a_.upload(PL_.block_a, s_);
        pair_b_.upload(PL_.block_b, s_);
        term_off_.upload(PL_.term_off, s_);
        term_i_.upload(PL_.term_i, s_);
        term_j_.upload(PL_.term_j, s_);
        value_off_.upload(PL_.value_off, s_);
        const size_t nvals = PL_.value_off.back();
        u_vals_.alloc(nvals);
        s_vals_.alloc(nvals);
        if (!iterative_) {
            S_.alloc(D * D);
            work_.alloc(cholesky_work_doubles((uint32_t)D));
            if (std::is_same<T, float>::value) {
                S32_.alloc(D * D);
                done_.alloc(1);
            }
        } else {
            std::vector<int32_t> dp(B_.n_blocks, -1), ep(nc, -1);
            for (size_t k = 0; k < PL_.block_a.size(); ++k) {
                uint32_t a = PL_.block_a[k], b = PL_.block_b[k];
                if (a == b) dp[a] = (int32_t)k;
                else if (a >= nc && b < nc) ep[b] = (int32_t)k;
            }
            diag_pair_.upload(dp, s_);
            edge_pair_.upload(ep, s_);
            factors_.alloc(54 * B_.n_blocks);
            jcm_.alloc(SOA_ROWS * no);
            pP_.alloc(9 * nc);
            pvcam_.alloc(9 * nc);
            ppart_.alloc(std::max<size_t>(1, (B_.n_blocks + 255) / 256));
            px_.alloc(D);
            pr_.alloc(D);
            pz_.alloc(D);
            pd_.alloc(D);
            pq_.alloc(D);
            pw_.alloc(3 * np);
            pcg_.alloc(1);
        }
        tp_.nc = (uint32_t)nc;
        tp_.ng = (uint32_t)ng;
        tp_.np = (uint32_t)np;
        tp_.no = (uint32_t)no;
        tp_.D = (uint32_t)D;
        tp_.nblocks = (uint32_t)B_.n_blocks;
        tp_.obs_cam = obs_cam_.p;
        tp_.obs_pt = obs_pt_.p;
        tp_.pt_off = pt_off_.p;
        tp_.pt_fixed = pt_fixed_.p;
        tp_.cam_group = cam_group_.p;
        tp_.cam_off = cam_off_.p;
        tp_.cam_dim = cam_dim_.p;
        tp_.grp_off = grp_off_.p;
        tp_.grp_dim = grp_dim_.p;
        tp_.blk_off = blk_off_.p;
        tp_.blk_dim = blk_dim_.p;
        tp_.blk_obs_off = blk_obs_off_.p;
        tp_.blk_obs = blk_obs_.p;
        tp_.grp_cam_off = grp_cam_off_.p;
        tp_.grp_cams = grp_cams_.p;
        tp_.nchunks = (uint32_t)(chunk_pt_.n ? chunk_pt_.n - 1 : 0);
        tp_.chunk_pt = chunk_pt_.p;
        tp_.chunk_obs = chunk_obs_.p;
        tp_.cm_off = cm_off_.p;
        tp_.cm_obs = cm_obs_.p;
        tp_.cm_pt = cm_pt_.p;
        tp_.loss = L_.loss;
        tp_.loss_scale = L_.loss_scale;
        HIP_OK(hipStreamSynchronize(s_));
    }

    // Point chunks (at most NT observations of whole points, or one larger point; at most NT
    // points) and the camera-major observation order of the matrix-free Schur product.
    // Every intrinsic group has one camera and no camera observes a point twice: each diagonal
    // block term of the reduced system is then one observation's (the camera-side assembly).
    bool simple_topology() const {
        const size_t nc = L_.n_cameras(), ng = L_.n_groups(), np = L_.n_points();
        std::vector<uint32_t> per_group(ng, 0);
        for (size_t c = 0; c < nc; ++c) per_group[L_.cam_group[c]]++;
        for (uint32_t k : per_group)
            if (k != 1) return false;
        std::vector<uint32_t> seen(nc, UINT32_MAX);
        for (size_t p = 0; p < np; ++p)
            for (uint32_t o = L_.point_off[p]; o < L_.point_off[p + 1]; ++o) {
                const uint32_t c = L_.obs_cam[o];
                if (seen[c] == (uint32_t)p) return false;
                seen[c] = (uint32_t)p;
            }
        return true;
    }

    void build_matvec_order(bool camera_major) {
        const size_t np = L_.n_points(), no = L_.n_obs(), nc = L_.n_cameras();
        const uint32_t cap = 256;
        std::vector<uint32_t> cpt{0}, cob{0};
        uint32_t obs_in = 0, pts_in = 0;
        for (size_t p = 0; p < np; ++p) {
            const uint32_t k = L_.point_off[p + 1] - L_.point_off[p];
            if (pts_in > 0 && (obs_in + k > cap || pts_in == cap)) {
                cpt.push_back((uint32_t)p);
                cob.push_back(L_.point_off[p]);
                obs_in = pts_in = 0;
            }
            obs_in += k;
            pts_in += 1;
        }
        if (pts_in > 0) {
            cpt.push_back((uint32_t)np);
            cob.push_back((uint32_t)no);
        }
        chunk_pt_.upload(cpt, s_);
        chunk_obs_.upload(cob, s_);
        if (!camera_major) {
            HIP_OK(hipStreamSynchronize(s_));  // the host vectors end with this scope
            return;
        }
        std::vector<uint32_t> off(nc + 1, 0), obs(no), pts(no);
        for (size_t o = 0; o < no; ++o) off[L_.obs_cam[o] + 1]++;
        for (size_t c = 0; c < nc; ++c) off[c + 1] += off[c];
        std::vector<uint32_t> cur(off.begin(), off.end() - 1);
        for (size_t o = 0; o < no; ++o) {
            const uint32_t j = cur[L_.obs_cam[o]]++;
            obs[j] = (uint32_t)o;
            pts[j] = L_.obs_point[o];
        }
        cm_off_.upload(off, s_);
        cm_obs_.upload(obs, s_);
        cm_pt_.upload(pts, s_);
        HIP_OK(hipStreamSynchronize(s_));  // the host vectors end with this scope
    }

    PcgVecs pcg_vecs() {
        PcgVecs V;
        V.grad = grad_.p;
        V.factors = factors_.p;
        V.x = px_.p;
        V.r = pr_.p;
        V.z = pz_.p;
        V.d = pd_.p;
        V.q = pq_.p;
        V.w = pw_.p;
        V.P = pP_.p;
        V.vcam = pvcam_.p;
        V.part = ppart_.p;
        V.st = pcg_.p;
        return V;
    }

    Lin<T> lin() { return Lin<T>{nullptr, jc_.p, ocost_.p, osq_.p, valid_.p, jpm_.p, jcm_.p}; }
    PointElim<T> pe() { return PointElim<T>{vinv_.p, vdamp_.p, y_.p, ddiag_.p, rrows_.p, pscal_.p}; }

    // Use a caller-given gauge state instead of the one derived from the translations.
    void set_gauge(const GaugeDev& g) {
        fixed_gauge_ = true;
        HIP_OK(hipMemcpyAsync(gauge_.p, &g, sizeof(GaugeDev), hipMemcpyHostToDevice, s_));
        HIP_OK(hipStreamSynchronize(s_));
    }

    // One assembly at damping lambda with loss `loss` (schur/assembly.rs:832-1069).
    AsmScalars assemble(int state, double lambda, int loss, double loss_scale, bool reduced_system) {
        Topo tp = tp_;
        tp.loss = loss;
        tp.loss_scale = loss_scale;
        if (!fixed_gauge_) launch_gauge<double>(s_, t_[state].p, L_.gauge_camera, L_.scene_scale, gauge_.p);
        timed(BA_PROFILE_LINEARIZE, [&] {
            launch_linearize<T>(s_, tp, R_[state].p, t_[state].p, gp_[state].p, X_[state].p, uv_.p, gauge_.p,
                                     lin());
        });
        if (iterative_ && (reduced_system || fast_cam_))
            launch_linearize_cm<T>(s_, tp, R_[state].p, t_[state].p, gp_[state].p, X_[state].p, uv_.p, gauge_.p,
                                        lin());
        lambda_ = lambda;
        timed(BA_PROFILE_POINT_ELIM, [&] { launch_point_elim<T>(s_, tp, lin(), (T)lambda, lambda == 0.0, pe()); });
        launch_reduce_points(s_, pscal_.p, tp.np, partials_.p, asm_.p);
        if (fast_cam_) {
            timed(BA_PROFILE_PAIRS, [&] {
                launch_camera_assembly<T>(s_, tp, lin(), pe(), diag_pair_.p, value_off_.p, grad_.p, fgrad_.p,
                                               diag_.p, u_vals_.p, s_vals_.p);
            });
        } else {
            launch_block_gradient<T>(s_, tp, lin(), pe(), grad_.p, fgrad_.p, diag_.p);
        }
        if (reduced_system && !fast_cam_) {
            timed(BA_PROFILE_PAIRS, [&] {
                launch_pairs<T>(s_, tp, lin(), pe(), (uint32_t)PL_.block_a.size(), pair_a_.p, pair_b_.p,
                                     term_off_.p, term_i_.p, term_j_.p, value_off_.p, u_vals_.p, s_vals_.p);
            });
        }
        if (reduced_system && !iterative_) {  // dense S with lambda diag on its diagonal
            HIP_OK(hipMemsetAsync(S_.p, 0, S_.n * sizeof(double), s_));
            launch_scatter_dense(s_, tp, (uint32_t)PL_.block_a.size(), pair_a_.p, pair_b_.p, value_off_.p, u_vals_.p,
                                 s_vals_.p, S_.p, tp.D);
            launch_finalize(s_, tp, lambda, diag_.p, fgrad_.p, grad_.p, S_.p, tp.D, nullptr, value_off_.p, u_vals_.p,
                            s_vals_.p, asm_.p);
        } else if (reduced_system) {  // lambda diag on the diagonal blocks of the preconditioner
            launch_finalize(s_, tp, lambda, diag_.p, fgrad_.p, grad_.p, nullptr, tp.D, diag_pair_.p, value_off_.p,
                            u_vals_.p, s_vals_.p, asm_.p);
        } else {
            launch_finalize(s_, tp, lambda, diag_.p, fgrad_.p, grad_.p, nullptr, tp.D, nullptr, value_off_.p, u_vals_.p,
                            s_vals_.p, asm_.p);
        }
        HIP_OK(hipMemcpyAsync(h_asm_.p, asm_.p, sizeof(AsmScalars), hipMemcpyDeviceToHost, s_));
        HIP_OK(hipStreamSynchronize(s_));
        return *h_asm_.p;
    }

    // Direct camera step: S step = -g by dense Cholesky (camera_step.rs:122-128).  The mixed form
    // factors S in FP32 and refines against the FP64 S; when the FP32 factor breaks down or the
    // refinement stalls above 1e-10 relative residual, the step is solved with the FP64 factor.
    bool solve_direct() {
        if (tp_.D == 0) return true;
        launch_negate(s_, tp_.D, grad_.p, rhs_.p);
        if (std::is_same<T, float>::value) {
            timed(BA_PROFILE_CHOLESKY, [&] {
                launch_cholesky_solve_mixed(s_, S_.p, S32_.p, tp_.D, tp_.D, rhs_.p, step_.p, work_.p, fail_.p, done_.p,
                                            8);
            });
            double st[3];
            HIP_OK(hipMemcpyAsync(h_int_.p, fail_.p, sizeof(int), hipMemcpyDeviceToHost, s_));
            HIP_OK(hipMemcpyAsync(st, work_.p + 8, 3 * sizeof(double), hipMemcpyDeviceToHost, s_));
            HIP_OK(hipStreamSynchronize(s_));
            refine_rel_ = std::sqrt(st[1] / st[2]);
            if (*h_int_.p == 0 && std::isfinite(refine_rel_) && refine_rel_ <= 1e-10) return true;
            mixed_fallbacks_ += 1;
        }
        timed(BA_PROFILE_CHOLESKY,
              [&] { launch_cholesky_solve(s_, S_.p, tp_.D, tp_.D, rhs_.p, step_.p, work_.p, fail_.p); });
        HIP_OK(hipMemcpyAsync(h_int_.p, fail_.p, sizeof(int), hipMemcpyDeviceToHost, s_));
        HIP_OK(hipStreamSynchronize(s_));
        return *h_int_.p == 0;
    }
    uint32_t mixed_fallbacks() const { return mixed_fallbacks_; }
    double refinement_residual() const { return refine_rel_; }

    // Retained PCG (self_calibration_retained.rs:512-579).  Returns accepted; terms/rel out.
    bool solve_pcg(double tol, int limit, uint32_t* terms, double* rel) {
        *terms = 0;
        *rel = NAN;
        if (tp_.D == 0) {
            HIP_OK(hipMemsetAsync(step_.p, 0, 0, s_));
            return true;
        }
        HIP_OK(hipMemsetAsync(fail_.p, 0, sizeof(int), s_));
        launch_precond_factor(s_, tp_, diag_pair_.p, value_off_.p, u_vals_.p, s_vals_.p, factors_.p, fail_.p);
        HIP_OK(hipMemcpyAsync(h_int_.p, fail_.p, sizeof(int), hipMemcpyDeviceToHost, s_));
        const PcgVecs V = pcg_vecs();
        launch_pcg_init(s_, tp_, V, tol, limit);
        HIP_OK(hipStreamSynchronize(s_));
        if (*h_int_.p) return false;
        const int batch = 4;
        const bool profiling_pcg = profile_ == BA_PROFILE_PCG_POINTS || profile_ == BA_PROFILE_PCG_CAMERAS;
        size_t events_before = used_;
        int products_before = 0;
        for (;;) {
            HIP_OK(hipMemcpyAsync(h_pcg_.p, pcg_.p, sizeof(PcgState), hipMemcpyDeviceToHost, s_));
            HIP_OK(hipStreamSynchronize(s_));
            // launches after the state turned done returned at once: drop their timings
            if (profiling_pcg) used_ = events_before + (size_t)(h_pcg_.p->products - products_before);
            if (h_pcg_.p->done) break;
            events_before = used_;
            products_before = h_pcg_.p->products;
            for (int k = 0; k < batch; ++k) {
                timed(BA_PROFILE_PCG_POINTS, [&] { launch_pcg_points<T>(s_, tp_, lin(), pe(), V); });
                timed(BA_PROFILE_PCG_CAMERAS, [&] { launch_pcg_cameras<T>(s_, tp_, lin(), V); });
                launch_pcg_blocks(s_, tp_, lambda_, diag_.p, px_.p, pd_.p, pP_.p, pq_.p, pcg_.p);
                launch_pcg_update(s_, tp_, V);
            }
        }
        *terms = (uint32_t)h_pcg_.p->iter;
        *rel = h_pcg_.p->rel;
        if (h_pcg_.p->accepted)
            HIP_OK(hipMemcpyAsync(step_.p, px_.p, tp_.D * sizeof(double), hipMemcpyDeviceToDevice, s_));
        return h_pcg_.p->accepted != 0;
    }

    struct Candidate {
        CandScalars c;
        double dots[4];
    };
    Candidate candidate(int state, int loss, double loss_scale, bool with_product) {
        const int cs = 1 - state;
        Topo tp = tp_;
        tp.loss = loss;
        tp.loss_scale = loss_scale;
        launch_retract(s_, tp, R_[state].p, t_[state].p, gp_[state].p, step_.p, gauge_.p, R_[cs].p, t_[cs].p, gp_[cs].p);
        timed(BA_PROFILE_CANDIDATE, [&] {
            launch_candidate<T>(s_, tp, R_[state].p, t_[state].p, R_[cs].p, t_[cs].p, gp_[cs].p, X_[state].p,
                                     uv_.p, lin(), pe(), step_.p, gauge_.p, X_[cs].p, cscal_.p);
        });
        launch_reduce_candidates(s_, cscal_.p, tp.np, partials_.p, cand_.p);
        launch_step_dots(s_, tp.D, grad_.p, diag_.p, step_.p, with_product ? pq_.p : nullptr, dots_.p);
        HIP_OK(hipMemcpyAsync(h_cand_.p, cand_.p, sizeof(CandScalars), hipMemcpyDeviceToHost, s_));
        HIP_OK(hipMemcpyAsync(h_dots_, dots_.p, 4 * sizeof(double), hipMemcpyDeviceToHost, s_));
        HIP_OK(hipStreamSynchronize(s_));
        Candidate out;
        out.c = *h_cand_.p;
        std::memcpy(out.dots, h_dots_, sizeof(out.dots));
        return out;
    }

    void download(int state, std::vector<double>& R, std::vector<double>& t, std::vector<double>& gp,
                  std::vector<double>& X) {
        R.resize(9 * tp_.nc);
        t.resize(3 * tp_.nc);
        gp.resize(3 * tp_.ng);
        X.resize(3 * tp_.np);
        HIP_OK(hipMemcpyAsync(R.data(), R_[state].p, R.size() * 8, hipMemcpyDeviceToHost, s_));
        HIP_OK(hipMemcpyAsync(t.data(), t_[state].p, t.size() * 8, hipMemcpyDeviceToHost, s_));
        HIP_OK(hipMemcpyAsync(gp.data(), gp_[state].p, gp.size() * 8, hipMemcpyDeviceToHost, s_));
        HIP_OK(hipMemcpyAsync(X.data(), X_[state].p, X.size() * 8, hipMemcpyDeviceToHost, s_));
        HIP_OK(hipStreamSynchronize(s_));
    }

    const Topo& topo() const { return tp_; }
    bool iterative() const { return iterative_; }
    const T* jpm() const { return jpm_.p; }
    const double* obs_cost() const { return ocost_.p; }
    const uint8_t* valid() const { return valid_.p; }
    const double* S() const { return S_.p; }
    const double* grad() const { return grad_.p; }
    const double* full_grad() const { return fgrad_.p; }
    const double* diag() const { return diag_.p; }
    const double* step() const { return step_.p; }
    const double* pcg_x() const { return px_.p; }
    hipStream_t stream() const { return s_; }

   private:
    hipStream_t s_;
    const Lowered& L_;
    bool iterative_;
    BlockLists B_;
    PairLists PL_;
    Topo tp_{};
    Buf<uint32_t> obs_cam_, obs_pt_, pt_off_, cam_group_, cam_off_, cam_dim_, grp_off_, grp_dim_, blk_off_, blk_dim_,
        blk_obs_off_, blk_obs_, grp_cam_off_, grp_cams_, pair_a_, pair_b_, term_off_, term_i_, term_j_, chunk_pt_,
        chunk_obs_, cm_off_, cm_obs_, cm_pt_;
    Buf<uint8_t> pt_fixed_, valid_;
    Buf<uint64_t> value_off_;
    Buf<int32_t> diag_pair_, edge_pair_;
    Buf<double> uv_, R_[2], t_[2], gp_[2], X_[2], ocost_, osq_, pscal_, cscal_, grad_, fgrad_, diag_, step_, rhs_,
        partials_, u_vals_, s_vals_, S_, work_, factors_, px_, pr_, pz_, pd_, pq_, pw_, dots_, pP_, pvcam_, ppart_;
    Buf<T> jc_, vinv_, vdamp_, y_, ddiag_, rrows_, jpm_, jcm_;
    Buf<float> S32_;  // FP32 factor of the mixed direct form
    Buf<int> done_;
    Buf<AsmScalars> asm_;
    Buf<CandScalars> cand_;
    Buf<GaugeDev> gauge_;
    Buf<int> fail_;
    Buf<PcgState> pcg_;
    Pinned<AsmScalars> h_asm_;
    Pinned<CandScalars> h_cand_;
    Pinned<int> h_int_;
    Pinned<PcgState> h_pcg_;
    double h_dots_[4];
    bool fixed_gauge_ = false;
    double lambda_ = 0.0;
    double refine_rel_ = NAN;
    uint32_t mixed_fallbacks_ = 0;
    bool fast_cam_ = false;
    int profile_ = BA_PROFILE_NONE;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events_;
    size_t used_ = 0;
};

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
template <class T>
Status refine(Engine<T>& E, const Lowered& L, const ba_solve_config* cfg, ba_solve_report& rep,
              ba_iteration_trace* trace_rows, int* final_state, double* loop_seconds) {
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
    for (int si = 0; si < 3 && !stop; ++si) {
        bool confirm = false;
        for (uint32_t it = 0; it < sched[si].iters; ++it) {
            iterations += 1;
            const double lambda = damping.value;
            AsmScalars a = E.assemble(state, lambda, sched[si].loss, sched[si].scale, true);
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
                if (std::isfinite(rel)) pcg_max_rel = std::isfinite(pcg_max_rel) ? std::fmax(pcg_max_rel, rel) : rel;
            }
            if (!solved) {
                rej[1]++;
                damping.reject();
                trace_push(trace, current, lambda, NAN, NAN, NAN, false, terms);
                continue;
            }
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
            if (!iterative) {  // exact_model_reduction (self_calibration/optimizer.rs:30-52)
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
template <class T>
Status run_solve(ba_hip_handle* h, const ba_problem* p, const ba_solve_config* cfg, const Lowered& L,
                 ba_solution* sol, ba_solve_report& r, ba_iteration_trace* trace, double t0) {
    const bool iterative = cfg->linear_solver == BA_SOLVER_ITERATIVE;
    Engine<T> E(h->stream, L, iterative, cfg->profile_kernel);
    E.setup();
    double t1 = now();
    int state = 0;
    double loop_seconds = 0.0;
    Status st = refine(E, L, cfg, r, trace, &state, &loop_seconds);
    if (!st.ok()) return st;
    std::vector<double> R, t, gp, X;
    E.download(state, R, t, gp, X);
    raise_solution(p, L, R.data(), t.data(), gp.data(), X.data(), sol);
    double t3 = now();
    r.seconds_setup = t1 - t0;
    r.seconds_iterations = loop_seconds;
    r.seconds_final = (t3 - t1) - loop_seconds;
    E.profile_result(&r.profile_launches, &r.profile_seconds);
    r.precision_fallbacks = E.mixed_fallbacks();
    r.refinement_residual = E.refinement_residual();
    return st;
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
    *out = h;
    return BA_OK;
}

void ba_hip_destroy(ba_hip_handle* h) {
    if (!h) return;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

int ba_hip_prepare(const ba_problem* p, ba_native_sizes* sizes, ba_native_problem* n, char* msg, size_t len) {
    Status st = validate_problem(p);
    if (!st.ok()) return finish(st, msg, len);
    Lowered L;
    st = lower_problem(p, L, true);
    if (!st.ok()) return finish(st, msg, len);
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


int ba_hip_solve(ba_hip_handle* h, const ba_problem* p, const ba_solve_config* cfg, ba_solution* sol,
                 ba_solve_report* rep, ba_iteration_trace* trace, char* msg, size_t len) {
    if (!h || !cfg || !sol) return finish(error(BA_ERR_INVALID_CONFIG, "null argument"), msg, len);
    Status st = validate_config(cfg);
    if (st.ok()) st = validate_problem(p);
    if (!st.ok()) return finish(st, msg, len);
    try {
        if (hipSetDevice(h->device) != hipSuccess) return finish(error(BA_ERR_INTERNAL, "hipSetDevice"), msg, len);
        double t0 = now();
        Lowered L;
        st = lower_problem(p, L, false);
        if (!st.ok()) return finish(st, msg, len);
        if (trace) trace->count = 0;
        ba_solve_report r;
        st = cfg->precision == BA_PRECISION_MIXED ? run_solve<float>(h, p, cfg, L, sol, r, trace, t0)
                                                  : run_solve<double>(h, p, cfg, L, sol, r, trace, t0);
        if (!st.ok()) return finish(st, msg, len);
        if (rep) *rep = r;
        return BA_OK;
    } catch (const HipError& e) {
        return finish(error(BA_ERR_INTERNAL, std::string("HIP error: ") + hipGetErrorString(e.code) + " in " + e.what),
                      msg, len);
    } catch (const std::bad_alloc&) {
        return finish(error(BA_ERR_RESOURCE_EXHAUSTED, "host allocation failed"), msg, len);
    }
}

#define HOOK_BEGIN                                                       \
    if (!h) return BA_ERR_INVALID_CONFIG;                                \
    try {                                                                \
        if (hipSetDevice(h->device) != hipSuccess) return BA_ERR_INTERNAL;
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
    E.setup();
    E.set_gauge(gauge_from(dp));
    E.assemble(0, 0.0, L.loss, L.loss_scale, false);
    const size_t no = L.n_obs();
    hipStream_t s = h->stream;
    if (no) {
        // structure-of-arrays rows -> per-observation records
        auto rows = [&](double* dst, int first, int count) {
            for (int k = 0; k < count; ++k)
                HIP_OK(hipMemcpy2DAsync(dst + k, count * sizeof(double), E.jpm() + (size_t)(first + k) * no,
                                        sizeof(double), sizeof(double), no, hipMemcpyDeviceToDevice, s));
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
    E.setup();
    E.set_gauge(gauge_from(dp));
    AsmScalars a = E.assemble(0, lambda, L.loss, L.loss_scale, true);
    uint32_t terms = 0;
    double rel = NAN;
    bool ok = (a.invalid == 0.0 && a.nonfinite == 0.0) && E.solve_pcg(tol, (int)max_iterations, &terms, &rel);
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
