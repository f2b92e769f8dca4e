// ba_engine.hpp — device state and kernel sequencing of one LM solve (internal to the library).
//
// Engine<T> holds a lowered problem resident in HBM (topology, two copies of the state, the
// per-observation linearisation, per-point blocks, the reduced system) and runs the passes of
// one LM iteration: assemble (linearise, eliminate points, assemble the camera side), the camera
// step (dense Cholesky or PCG), and the candidate pass.  T is the storage type of Jacobians and
// point blocks (double: plain form; float: mixed form).
#pragma once

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
#include "ba_ledger.hpp"
#include "ba_policy.hpp"

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

// While a solve sets its engine up, the device bytes that Buf allocates are added here (the
// workspace size that an explicit budget limits).
inline thread_local uint64_t* g_device_tally = nullptr;

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
        if (count) {
            HIP_OK(hipMalloc(&p, count * sizeof(T)));
            if (g_device_tally) *g_device_tally += count * sizeof(T);
        }
    }
    void upload(const T* h, size_t count, hipStream_t s) {
        alloc(count);
        if (count) HIP_OK(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void upload(const std::vector<T>& v, hipStream_t s) { upload(v.data(), v.size(), s); }
    // Upload v into a buffer of count >= v.size() elements, the tail zeroed.
    void upload_padded(const std::vector<T>& v, size_t count, hipStream_t s) {
        alloc(count);
        if (v.size()) HIP_OK(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
        if (count > v.size()) HIP_OK(hipMemsetAsync(p + v.size(), 0, (count - v.size()) * sizeof(T), s));
    }
};

struct HipEventOps {
    void create(hipEvent_t& e) { HIP_OK(hipEventCreate(&e)); }
    void destroy(hipEvent_t e) {
        if (e) (void)hipEventDestroy(e);
    }
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
        for (hipEvent_t e : pcg_ev_)
            if (e) (void)hipEventDestroy(e);
    }

    // Run `launch`, bracketed by HIP events on this stream when `which` is the profiled kernel.
    template <class F>
    void timed(int which, F&& launch) {
        if (which != profile_) {
            launch();
            return;
        }
        auto& e = ledger_.next();
        HIP_OK(hipEventRecord(e.first, s_));
        launch();
        HIP_OK(hipEventRecord(e.second, s_));
    }
    // Launch count and summed duration of the profiled kernel so far.
    void profile_result(uint32_t* launches, double* seconds) {
        HIP_OK(hipStreamSynchronize(s_));
        double total = 0.0;
        for (size_t k = 0; k < ledger_.count(); ++k) {
            float ms = 0.0f;
            HIP_OK(hipEventElapsedTime(&ms, ledger_[k].first, ledger_[k].second));
            total += ms * 1e-3;
        }
        *launches = (uint32_t)ledger_.count();
        *seconds = total;
    }

    // Device memory allocated by setup().
    uint64_t device_bytes() const { return device_bytes_; }
    // An explicit workspace budget: the optional explicit reduced system is set up only within it.
    void set_budget(uint64_t bytes) { budget_ = bytes; }

    void setup() {
        struct Tally {
            explicit Tally(uint64_t* t) { g_device_tally = t; }
            ~Tally() { g_device_tally = nullptr; }
        } tally(&device_bytes_);
        B_ = build_block_lists(L_);
        // With one camera per intrinsic group and no camera observing a point twice, the FP64
        // direct step assembles its dense S from the explicit system's camera rows (camera-side
        // assembly, camera-pair matrix-core pass, scatter) instead of the block pairs' terms.
        const bool simple = (iterative_ || (!batched_ && std::is_same<T, double>::value)) && simple_topology();
        dense_rows_ = !iterative_ && !pair_terms_ && simple && count_schur_terms(L_) < (uint64_t)UINT32_MAX;
        fast_cam_ = simple && (iterative_ || dense_rows_);
        // the camera-side assembly reads only the pairs' value offsets, not their terms
        PL_ = build_pair_lists(L_, B_, !iterative_ && !dense_rows_, !fast_cam_);
        const size_t nc = L_.n_cameras(), ng = L_.n_groups(), np = L_.n_points(), no = L_.n_obs(), D = L_.D;
        npad_ = soa_pad(no);
        obs_cam_.upload_padded(L_.obs_cam, npad_, s_);
        obs_pt_.upload_padded(L_.obs_point, npad_, s_);
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
        pose_rows_ = keep_pose_rows_ || !fast_cam_;  // the block-gradient pass reads them
        if (fast_cam_) prec_.alloc(PE_REC * np);          // read by the camera assembly
        srec_.alloc(PCG_CAM_REC * std::max<size_t>(1, nc));
        build_matvec_order(iterative_ || dense_rows_);
        jpm_.alloc(SOA_ROWS * npad_);
        HIP_OK(hipMemsetAsync(jpm_.p, 0, jpm_.n * sizeof(T), s_));  // finite padding
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
        dots_.alloc(5);
        pair_a_.upload(PL_.block_a, s_);
        pair_b_.upload(PL_.block_b, s_);
        term_off_.upload(PL_.term_off, s_);
        term_i_.upload(PL_.term_i, s_);
        term_j_.upload(PL_.term_j, s_);
        value_off_.upload(PL_.value_off, s_);
        const size_t nvals = PL_.value_off.back();
        u_vals_.alloc(nvals);
        s_vals_.alloc(nvals);
        if (!iterative_ && !batched_) {
            S_.alloc(D * D);
            work_.alloc(cholesky_work_doubles((uint32_t)D));
            if (std::is_same<T, float>::value) {
                S32_.alloc(D * D);
                done_.alloc(1);
            }
            if (dense_rows_) {
                upload_local_pairs();
                jcm_.alloc(SOA_ROWS * npad_);
                HIP_OK(hipMemsetAsync(jcm_.p, 0, jcm_.n * sizeof(T), s_));  // finite padding
                casm_.alloc((size_t)CAM_ASM_VALUES * std::max<size_t>(1, nseg_));
                build_explicit();
            }
        } else {
            upload_local_pairs();
            factors_.alloc(54 * B_.n_blocks);
            jcm_.alloc(SOA_ROWS * npad_);
            HIP_OK(hipMemsetAsync(jcm_.p, 0, jcm_.n * sizeof(T), s_));  // finite padding
            pseg_.alloc(9 * std::max<size_t>(1, nseg_));
            casm_.alloc((size_t)CAM_ASM_VALUES * std::max<size_t>(1, nseg_));
            pvcam_.alloc(PCG_CAM_REC * nc);
            pwcam_.alloc(PCG_WREC * std::max<size_t>(1, nc));
            {
                int dev = 0, cu = 0;
                HIP_OK(hipGetDevice(&dev));
                HIP_OK(hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev));
                ncu_ = (uint32_t)std::max(cu, 1);
            }
            HIP_OK(hipMemsetAsync(pvcam_.p, 0, pvcam_.n * sizeof(double), s_));
            const size_t qparts = std::max<size_t>((B_.n_blocks + PCG_Q_WG / 64 - 1) / (PCG_Q_WG / 64),
                                                   nc);  // implicit or explicit product
            ppart_.alloc(std::max<size_t>(1, qparts) +
                         2 * std::max<size_t>(1, (B_.n_blocks + PCG_BLOCK_WG - 1) / PCG_BLOCK_WG));
            px_.alloc(D);
            pr_.alloc(D);
            pz_.alloc(D);
            pd_[0].alloc(D);
            pd_[1].alloc(D);
            pq_.alloc(D);
            pxw_.alloc(PCG_XW * np);
            HIP_OK(hipMemsetAsync(pxw_.p, 0, pxw_.n * sizeof(double), s_));
            pcg_.alloc(1);
            pcg1_.alloc(1);
            setup_explicit();  // last: it is optional and fits in what the budget leaves
        }
        tp_.nc = (uint32_t)nc;
        tp_.ng = (uint32_t)ng;
        tp_.np = (uint32_t)np;
        tp_.no = (uint32_t)no;
        tp_.D = (uint32_t)D;
        tp_.nblocks = (uint32_t)B_.n_blocks;
        gturn_.alloc(std::max<size_t>(1, ng));
        tp_.gturn = gturn_.p;
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
        tp_.cm_cam = cm_cam_.p;
        tp_.cm_uv = cm_uv_.p;
        tp_.loss = L_.loss;
        tp_.loss_scale = L_.loss_scale;
        HIP_OK(hipStreamSynchronize(s_));
    }

    // The explicit reduced camera system: chosen (see ba_schur_product) after the camera-major
    // order exists; builds the neighbour rows and the camera pairs' observation pairs on the host.
    void setup_explicit() {
        explicit_ = false;
        if (!fast_cam_ || schur_mode_ == BA_SCHUR_IMPLICIT) return;
        const size_t nc = L_.n_cameras(), no = L_.n_obs();
        const uint64_t terms = count_schur_terms(L_);
        const uint64_t pairs = std::min<uint64_t>(terms, (uint64_t)nc * (nc > 0 ? nc - 1 : 0) / 2);
        const double bytes = (double)sizeof(T) * ((double)CREC * no + 81.0 * ((double)nc + (double)pairs)) +
                             8.0 * (double)terms + 88.0 * (double)pairs;
        if (schur_mode_ != BA_SCHUR_EXPLICIT) {
            size_t fr = 0, tot = 0;
            HIP_OK(hipMemGetInfo(&fr, &tot));
            if (bytes > 0.4 * (double)fr) return;
        }
        // the rest of the workspace is allocated before this point
        if (budget_ && (double)device_bytes_ + bytes > (double)budget_) return;
        if (terms >= (uint64_t)UINT32_MAX) return;
        build_explicit();
    }
    // The diagonal pair of each block and the (group, camera) pair of each camera.
    void upload_local_pairs() {
        const size_t nc = L_.n_cameras();
        std::vector<int32_t> dp(B_.n_blocks, -1), ep(nc, -1);
        for (size_t k = 0; k < PL_.block_a.size(); ++k) {
            uint32_t a = PL_.block_a[k], b = PL_.block_b[k];
            if (a == b) dp[a] = (int32_t)k;
            else if (a >= nc && b < nc) ep[b] = (int32_t)k;
        }
        diag_pair_.upload(dp, s_);
        edge_pair_.upload(ep, s_);
        HIP_OK(hipStreamSynchronize(s_));  // the host vectors end with this scope
    }
    void build_explicit() {
        const size_t nc = L_.n_cameras(), no = L_.n_obs();
        std::vector<uint32_t> pos(no), cur(nc + 1, 0);
        for (size_t o = 0; o < no; ++o) cur[L_.obs_cam[o] + 1]++;
        for (size_t c = 0; c < nc; ++c) cur[c + 1] += cur[c];
        for (size_t o = 0; o < no; ++o) pos[o] = cur[L_.obs_cam[o]]++;
        const SchurLists SL = build_schur_lists(L_, pos);
        snbr_off_.upload(SL.nbr_off, s_);
        snbr_.upload(SL.nbr, s_);
        srow_base_.upload(SL.row_base, s_);
        spair_off_.upload(SL.pair_off, s_);
        srev_off_.upload(SL.rev_off, s_);
        srev_pair_.upload(SL.rev_pair, s_);
        spair_cam_.upload(SL.pair_cam, s_);
        sterm_off_.upload(SL.term_off, s_);
        sterm_.upload(SL.term.data(), SL.term.size(), s_);
        npairs_ = (uint32_t)(SL.pair_cam.size() / 2);
        max_row_slots_ = 0;
        for (size_t c = 0; c < nc; ++c) max_row_slots_ = std::max(max_row_slots_, SL.nbr_off[c + 1] - SL.nbr_off[c]);
        {
            std::vector<uint32_t> grp;
            std::vector<uint8_t> grn;
            build_pair_groups(SL.pair_cam, (uint32_t)nc, grp, grn);
            sgroup_.upload(grp, s_);
            sgroup_n_.upload(grn, s_);
        }
        crec_.alloc((size_t)CREC * std::max<size_t>(1, no));
        svals_.alloc(SL.n_values);
        pvexp_.alloc(9 * std::max<size_t>(1, nc));
        stpair_.alloc(9 * std::max<size_t>(1, npairs_));
        sqrow_.alloc(9 * std::max<size_t>(1, nc));
        HIP_OK(hipMemsetAsync(pvexp_.p, 0, pvexp_.n * sizeof(double), s_));
        HIP_OK(hipStreamSynchronize(s_));  // the host lists end with this scope
        explicit_ = true;
        use_explicit_ = true;
    }
    SchurRows<T> schur_rows() {
        return use_explicit_ ? SchurRows<T>{svals_.p, srow_base_.p, snbr_off_.p, snbr_.p, spair_off_.p,
                                        srev_off_.p, srev_pair_.p, stpair_.p, sqrow_.p}
                         : SchurRows<T>{};
    }
    SchurPairs schur_pairs() {
        return SchurPairs{npairs_, spair_cam_.p, sterm_off_.p, sterm_.p, (uint32_t)sgroup_.n, sgroup_.p, sgroup_n_.p};
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
        if (camera_major) {  // chunks of the PCG point pass: whole points in one wavefront's lanes
            std::vector<uint32_t> ppt{0}, pob{0};
            uint32_t first = 0, pts = 0;
            for (size_t p = 0; p < np; ++p) {
                const uint32_t end = L_.point_off[p + 1];
                if (pts > 0 && end - first > PCG_WAVE_OBS) {
                    ppt.push_back((uint32_t)p);
                    pob.push_back(L_.point_off[p]);
                    first = L_.point_off[p];
                    pts = 0;
                }
                pts += 1;
            }
            if (pts > 0) {
                ppt.push_back((uint32_t)np);
                pob.push_back((uint32_t)no);
            }
            npchunks_ = (uint32_t)(ppt.size() - 1);
            pchunk_pt_.upload(ppt, s_);
            pchunk_obs_.upload(pob, s_);
        }
        if (!camera_major) {
            HIP_OK(hipStreamSynchronize(s_));  // the host vectors end with this scope
            return;
        }
        std::vector<uint32_t> off(nc + 1, 0), obs(no), pts(no), cams(no);
        std::vector<double> uvj(2 * no);
        for (size_t o = 0; o < no; ++o) off[L_.obs_cam[o] + 1]++;
        for (size_t c = 0; c < nc; ++c) off[c + 1] += off[c];
        std::vector<uint32_t> cur(off.begin(), off.end() - 1);
        for (size_t o = 0; o < no; ++o) {
            const uint32_t j = cur[L_.obs_cam[o]]++;
            obs[j] = (uint32_t)o;
            pts[j] = L_.obs_point[o];
            cams[j] = L_.obs_cam[o];
            uvj[2 * j] = L_.uv[2 * o];
            uvj[2 * j + 1] = L_.uv[2 * o + 1];
        }
        cm_off_.upload(off, s_);
        cm_obs_.upload(obs, s_);
        cm_pt_.upload_padded(pts, npad_, s_);
        cm_cam_.upload_padded(cams, npad_, s_);
        cm_uv_.upload(uvj, s_);
        // camera segments of the PCG camera pass: windows from the aligned first position
        std::vector<uint32_t> cseg(nc + 1, 0), segc;
        for (size_t c = 0; c < nc; ++c) {
            const uint32_t vn = pcg_vec<T>(), seg = pcg_seg<T>();
            const uint32_t n = off[c + 1] > off[c] ? off[c + 1] - off[c] / vn * vn : 0;
            const uint32_t k = (n + seg - 1) / seg;
            cseg[c + 1] = cseg[c] + k;
            for (uint32_t i = 0; i < k; ++i) segc.push_back((uint32_t)c);
        }
        nseg_ = (uint32_t)segc.size();
        cseg_off_.upload(cseg, s_);
        seg_cam_.upload(segc, s_);
        HIP_OK(hipStreamSynchronize(s_));  // the host vectors end with this scope
    }

    PcgVecs pcg_vecs() {
        PcgVecs V;
        V.grad = grad_.p;
        V.factors = factors_.p;
        V.diag = diag_.p;
        V.R = R_[lin_state_].p;
        V.X = X_[lin_state_].p;
        V.gp = gp_[lin_state_].p;
        V.gauge = gauge_.p;
        V.lambda = lambda_;
        V.x = px_.p;
        V.r = pr_.p;
        V.z = pz_.p;
        V.q = pq_.p;
        V.d[0] = pd_[0].p;
        V.d[1] = pd_[1].p;
        V.xw = pxw_.p;
        V.seg = pseg_.p;
        V.vcam = pvcam_.p;
        V.part = ppart_.p;
        V.nparts = pcg_partials(tp_);
        V.nqparts = use_explicit_ ? pcg_product_partials(tp_) : pcg_q_partials(tp_);
        V.vexp = use_explicit_ ? pvexp_.p : nullptr;
        V.nseg = nseg_;
        V.npchunks = npchunks_;
        V.pchunk_pt = pchunk_pt_.p;
        V.pchunk_obs = pchunk_obs_.p;
        V.ncu = ncu_;
        V.wcam = pwcam_.p;
        V.seg_cam = seg_cam_.p;
        V.cseg_off = cseg_off_.p;
        V.st = pcg_.p;
        V.st1 = pcg1_.p;
        return V;
    }

    Lin<T> lin() { return Lin<T>{nullptr, jc_.p, ocost_.p, osq_.p, valid_.p, jpm_.p, jcm_.p, pose_rows_ ? 1 : 0}; }
    void set_schur_product(int mode) { schur_mode_ = mode; }
    // Whether a PCG solve of this engine has used the explicit system.
    bool explicit_product() const { return used_explicit_; }
    // The product of the next LM iteration (before its assembly): use_explicit_product.
    void plan_product(uint32_t expected_pcg) { use_explicit_ = use_explicit_product(explicit_, schur_mode_, expected_pcg); }
    // Assemble a direct step's S from the block pairs' terms, U - sum W V^-1 W^T in the oracle's
    // order of operations, also for a simple topology (before setup()).
    void keep_pair_terms() { pair_terms_ = true; }
    // Keep the pose columns of J~_c in the point-major copy (for the per-observation hook).
    void keep_pose_rows() { keep_pose_rows_ = true; }
    PointElim<T> pe() { return PointElim<T>{vinv_.p, vdamp_.p, y_.p, ddiag_.p, rrows_.p, prec_.p, pscal_.p}; }

    // Use a caller-given gauge state instead of the one derived from the translations.
    void set_gauge(const GaugeDev& g) {
        fixed_gauge_ = true;
        HIP_OK(hipMemcpyAsync(gauge_.p, &g, sizeof(GaugeDev), hipMemcpyHostToDevice, s_));
        HIP_OK(hipStreamSynchronize(s_));
    }

    // One assembly at damping lambda with loss `loss` (schur/assembly.rs:832-1069).
    // same_state: the caller has not written state `state` since the last assembly (an LM step
    // was rejected); the linearisation of that assembly is then kept when it was taken of this
    // state with this loss (it is a function of them alone).
    AsmScalars assemble(int state, double lambda, int loss, double loss_scale, bool reduced_system,
                        bool same_state = false) {
        Topo tp = tp_;
        tp.loss = loss;
        tp.loss_scale = loss_scale;
        lin_state_ = state;
        const bool cm = (iterative_ && reduced_system) || fast_cam_;
        const bool keep = same_state && lin_kept_.valid && lin_kept_.state == state && lin_kept_.loss == loss &&
                          lin_kept_.scale == loss_scale && (lin_kept_.cm || !cm);
        if (!keep) {
            if (!fixed_gauge_) launch_gauge<double>(s_, t_[state].p, L_.gauge_camera, L_.scene_scale, gauge_.p);
            timed(BA_PROFILE_LINEARIZE, [&] {
                launch_linearize<T>(s_, tp, R_[state].p, t_[state].p, gp_[state].p, X_[state].p, uv_.p, gauge_.p,
                                    lin());
            });
            if (cm)
                launch_linearize_cm<T>(s_, tp, R_[state].p, t_[state].p, gp_[state].p, X_[state].p, uv_.p, gauge_.p,
                                       lin());
            lin_kept_ = {true, state, loss, loss_scale, cm};
        }
        lambda_ = lambda;
        timed(BA_PROFILE_POINT_ELIM, [&] { launch_point_elim<T>(s_, tp, lin(), (T)lambda, lambda == 0.0, pe()); });
        const bool ex = use_explicit_ && reduced_system;  // the explicit system only for a step
        launch_reduce_points(s_, pscal_.p, tp.np, partials_.p, asm_.p);
        if (fast_cam_) {
            timed(BA_PROFILE_PAIRS, [&] {
                launch_camera_assembly<T>(s_, tp, lin(), pe(), R_[state].p, gp_[state].p, gauge_.p, nseg_,
                                          seg_cam_.p, cseg_off_.p, casm_.p, diag_pair_.p, value_off_.p, grad_.p,
                                          fgrad_.p, diag_.p, u_vals_.p, s_vals_.p, ex ? crec_.p : nullptr,
                                          ex ? schur_rows() : SchurRows<T>{});
            });
            if (ex)
                timed(BA_PROFILE_SCHUR_PAIRS, [&] {
                    launch_schur_pairs<T>(s_, schur_pairs(), crec_.p, R_[state].p, gp_[state].p, cam_group_.p,
                                          cam_dim_.p, gauge_.p, schur_rows());
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
            if (dense_rows_)
                launch_scatter_rows<T>(s_, tp, schur_rows(), max_row_slots_, S_.p, tp.D);
            else
                launch_scatter_dense(s_, tp, (uint32_t)PL_.block_a.size(), pair_a_.p, pair_b_.p, value_off_.p,
                                     u_vals_.p, s_vals_.p, S_.p, tp.D);
            launch_finalize(s_, tp, lambda, diag_.p, fgrad_.p, grad_.p, S_.p, tp.D, nullptr, value_off_.p, u_vals_.p,
                            partials_.p, asm_.p);
        } else if (reduced_system) {  // lambda diag on the diagonal blocks of the preconditioner
            launch_finalize(s_, tp, lambda, diag_.p, fgrad_.p, grad_.p, nullptr, tp.D, diag_pair_.p, value_off_.p,
                            u_vals_.p, partials_.p, asm_.p);
        } else {
            launch_finalize(s_, tp, lambda, diag_.p, fgrad_.p, grad_.p, nullptr, tp.D, nullptr, value_off_.p, u_vals_.p,
                            partials_.p, asm_.p);
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
        launch_pcg_init<T>(s_, tp_, V, tol, limit);
        HIP_OK(hipStreamSynchronize(s_));
        if (*h_int_.p) return false;
        used_explicit_ = used_explicit_ || use_explicit_;
        const int batch = 4;
        const SchurRows<T> SR = schur_rows();
        auto enqueue_batch = [&] {
            for (int k = 0; k < batch; ++k) {
                if (use_explicit_) {
                    timed(BA_PROFILE_PCG_PRODUCT, [&] { launch_pcg_product_upper<T>(s_, tp_, SR, V); });
                    launch_pcg_product_lower<T>(s_, tp_, SR, V);
                    launch_pcg_step_explicit<T>(s_, tp_, V);
                } else {
                    timed(BA_PROFILE_PCG_POINTS, [&] { launch_pcg_points<T>(s_, tp_, lin(), pe(), V); });
                    timed(BA_PROFILE_PCG_CAMERAS, [&] { launch_pcg_cameras<T>(s_, tp_, lin(), V); });
                    launch_pcg_step<T>(s_, tp_, V);
                }
            }
        };
        const PcgState* fin = nullptr;
        const bool profiled_product = use_explicit_ ? profile_ == BA_PROFILE_PCG_PRODUCT
                                                : (profile_ == BA_PROFILE_PCG_POINTS || profile_ == BA_PROFILE_PCG_CAMERAS);
        if (profiled_product) {
            // one batch at a time, so that launches after the state turned done (which return at
            // once) can be dropped from the timings by the state's product count
            size_t events_before = ledger_.count();
            int products_before = 0;
            for (;;) {
                HIP_OK(hipMemcpyAsync(h_pcg_[0].p, pcg_.p, sizeof(PcgState), hipMemcpyDeviceToHost, s_));
                HIP_OK(hipStreamSynchronize(s_));
                const int applied = h_pcg_[0].p->products - products_before;
                ledger_.keep(events_before + (size_t)std::max(applied, 0));
                if (h_pcg_[0].p->done) break;
                events_before = ledger_.count();
                products_before = h_pcg_[0].p->products;
                enqueue_batch();
            }
            fin = h_pcg_[0].p;
        } else {
            // Two batches in flight: the next batch is queued before the host reads the state
            // after the previous one, so the device never waits for the host.  Every kernel of
            // the batch queued after the state turned done returns at once.
            for (int i = 0; i < 2; ++i)
                if (!pcg_ev_[i]) HIP_OK(hipEventCreateWithFlags(&pcg_ev_[i], hipEventDisableTiming));
            enqueue_batch();
            HIP_OK(hipMemcpyAsync(h_pcg_[0].p, pcg_.p, sizeof(PcgState), hipMemcpyDeviceToHost, s_));
            HIP_OK(hipEventRecord(pcg_ev_[0], s_));
            for (int k = 0;; ++k) {
                const int cur = k & 1, nxt = cur ^ 1;
                enqueue_batch();
                HIP_OK(hipMemcpyAsync(h_pcg_[nxt].p, pcg_.p, sizeof(PcgState), hipMemcpyDeviceToHost, s_));
                HIP_OK(hipEventRecord(pcg_ev_[nxt], s_));
                HIP_OK(hipEventSynchronize(pcg_ev_[cur]));
                if (h_pcg_[cur].p->done) {
                    fin = h_pcg_[cur].p;
                    break;
                }
            }
        }
        *terms = (uint32_t)fin->iter;
        *rel = fin->rel;
        if (fin->accepted)
            HIP_OK(hipMemcpyAsync(step_.p, px_.p, tp_.D * sizeof(double), hipMemcpyDeviceToDevice, s_));
        return fin->accepted != 0;
    }

    // The calibrated path's step limit (optimizer.rs:677-702), between the camera solve and the
    // candidate pass: scales the step, and the PCG product of the step when with_product, so
    // that the candidate, its dot products and the model reduction all use the limited step.
    void limit_step(double maximum, bool with_product) {
        launch_limit_step(s_, tp_, maximum, step_.p, with_product ? pq_.p : nullptr, dots_.p + 4);
        limited_ = true;
    }

    struct Candidate {
        CandScalars c;
        double dots[4];
        double alpha;  // the step limit's factor (1 when not limited or no limit ran)
    };
    Candidate candidate(int state, int loss, double loss_scale, bool with_product) {
        const int cs = 1 - state;
        Topo tp = tp_;
        tp.loss = loss;
        tp.loss_scale = loss_scale;
        launch_retract(s_, tp, R_[state].p, t_[state].p, gp_[state].p, step_.p, gauge_.p, R_[cs].p, t_[cs].p, gp_[cs].p);
        timed(BA_PROFILE_CANDIDATE, [&] {
            launch_candidate<T>(s_, tp, R_[state].p, t_[state].p, gp_[state].p, R_[cs].p, t_[cs].p, gp_[cs].p, X_[state].p,
                                uv_.p, lin(), pe(), step_.p, gauge_.p, srec_.p, X_[cs].p, cscal_.p);
        });
        launch_reduce_candidates(s_, cscal_.p, tp.np, partials_.p, cand_.p);
        launch_step_dots(s_, tp.D, grad_.p, diag_.p, step_.p, with_product ? pq_.p : nullptr, dots_.p);
        HIP_OK(hipMemcpyAsync(h_cand_.p, cand_.p, sizeof(CandScalars), hipMemcpyDeviceToHost, s_));
        HIP_OK(hipMemcpyAsync(h_dots_, dots_.p, (limited_ ? 5 : 4) * sizeof(double), hipMemcpyDeviceToHost, s_));
        HIP_OK(hipStreamSynchronize(s_));
        Candidate out;
        out.c = *h_cand_.p;
        std::memcpy(out.dots, h_dots_, sizeof(out.dots));
        out.alpha = limited_ ? h_dots_[4] : 1.0;
        limited_ = false;
        return out;
    }

    // Inlier count and RMSE of the residual norms within `limit` at the last linearisation
    // (solver.rs:1353-1371); the RMSE is NaN when none is within it.
    void inliers(double limit, uint64_t* count, double* rmse) {
        double v[2] = {0.0, 0.0};
        if (tp_.no) {
            launch_inliers(s_, osq_.p, valid_.p, tp_.no, limit, partials_.p, dots_.p);
            HIP_OK(hipMemcpyAsync(v, dots_.p, sizeof(v), hipMemcpyDeviceToHost, s_));
            HIP_OK(hipStreamSynchronize(s_));
        }
        *count = (uint64_t)v[0];
        *rmse = v[0] > 0.0 ? std::sqrt(v[1] / v[0]) : NAN;
        if (!std::isfinite(*rmse)) *rmse = NAN;
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
    size_t jpm_size() const { return jpm_.n; }
    const double* obs_cost() const { return ocost_.p; }
    const uint8_t* valid() const { return valid_.p; }
    const double* S() const { return S_.p; }
    double* S_mut() { return S_.p; }
    const double* grad() const { return grad_.p; }
    const double* full_grad() const { return fgrad_.p; }
    const double* diag() const { return diag_.p; }
    const double* step() const { return step_.p; }
    const double* pcg_x() const { return px_.p; }
    hipStream_t stream() const { return s_; }

   protected:
    hipStream_t s_;
    const Lowered& L_;
    bool iterative_;
    BlockLists B_;
    PairLists PL_;
    Topo tp_{};
    Buf<uint32_t> obs_cam_, obs_pt_, pt_off_, cam_group_, cam_off_, cam_dim_, grp_off_, grp_dim_, blk_off_, blk_dim_,
        blk_obs_off_, blk_obs_, grp_cam_off_, grp_cams_, pair_a_, pair_b_, term_off_, term_i_, term_j_, chunk_pt_,
        chunk_obs_, cm_off_, cm_obs_, cm_pt_, cm_cam_, cseg_off_, seg_cam_, pchunk_pt_, pchunk_obs_;
    uint32_t nseg_ = 0, npchunks_ = 0, ncu_ = 1;
    size_t npad_ = 0;  // padded observation count of the tiled copies and index arrays
    Buf<uint8_t> pt_fixed_, valid_;
    Buf<uint64_t> value_off_;
    Buf<int32_t> diag_pair_, edge_pair_;
    Buf<double> uv_, R_[2], t_[2], gp_[2], X_[2], ocost_, osq_, pscal_, cscal_, grad_, fgrad_, diag_, step_, rhs_,
        partials_, u_vals_, s_vals_, S_, work_, factors_, px_, pr_, pz_, pd_[2], pq_, dots_, pseg_, ppart_;
    Buf<double> pxw_, pvcam_, pwcam_;  // PCG point, camera and world camera records
    Buf<double> srec_;         // candidate step records
    Buf<double> gturn_;        // first radial turn per group
    Buf<double> cm_uv_;        // pixels in camera-major order
    Buf<double> casm_;         // camera assembly: per-segment sums
    Buf<T> jc_, vinv_, vdamp_, y_, ddiag_, rrows_, prec_, jpm_, jcm_;
    Buf<float> S32_;  // FP32 factor of the mixed direct form
    Buf<int> done_;
    Buf<AsmScalars> asm_;
    Buf<CandScalars> cand_;
    Buf<GaugeDev> gauge_;
    Buf<int> fail_;
    Buf<PcgState> pcg_, pcg1_;
    int lin_state_ = 0;  // state of the last assembly (the PCG's linearisation point)
    Pinned<AsmScalars> h_asm_;
    Pinned<CandScalars> h_cand_;
    Pinned<int> h_int_;
    Pinned<PcgState> h_pcg_[2];
    hipEvent_t pcg_ev_[2] = {nullptr, nullptr};
    double h_dots_[5];
    bool limited_ = false;
    bool fixed_gauge_ = false;
    struct {
        bool valid;
        int state, loss;
        double scale;
        bool cm;
    } lin_kept_ = {false, 0, 0, 0.0, false};  // the linearisation held in lin() (see assemble)
    double lambda_ = 0.0;
    bool batched_ = false;  // independent windows: per-window dense systems instead of one S
    double refine_rel_ = NAN;
    uint32_t mixed_fallbacks_ = 0;
    bool fast_cam_ = false;
    bool dense_rows_ = false;  // the direct step's dense S from the explicit system's rows
    uint32_t max_row_slots_ = 0;  // the longest camera row of the explicit system (blocks)
    bool pair_terms_ = false;  // ... never (keep_pair_terms)
    int schur_mode_ = BA_SCHUR_AUTO;
    bool explicit_ = false;      // the explicit system is set up
    bool use_explicit_ = false;  // ... and the current LM iteration's PCG uses it
    bool used_explicit_ = false;
    uint32_t npairs_ = 0;
    Buf<T> crec_, svals_;
    Buf<uint64_t> srow_base_;
    Buf<uint32_t> snbr_off_, snbr_, spair_off_, srev_off_, srev_pair_, spair_cam_, sterm_off_, sterm_, sgroup_;
    Buf<uint8_t> sgroup_n_;
    Buf<double> pvexp_, stpair_, sqrow_;
    bool pose_rows_ = true, keep_pose_rows_ = false;
    int profile_ = BA_PROFILE_NONE;
    uint64_t device_bytes_ = 0, budget_ = 0;
    LaunchLedger<hipEvent_t, HipEventOps> ledger_;  // event pairs of the profiled kernel's launches
};

}  // namespace
}  // namespace ba
