// ba_kernels.hpp — launchers of the HIP kernels of the LM inner loop.
//
// Every launcher enqueues on `stream` and returns immediately.  Template parameter T is the
// arithmetic of the per-observation and per-point work: double (plain form) or float (mixed
// form).  Reductions run in a fixed order with no floating-point atomics, so the results do not
// depend on scheduling.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace ba {

// Gauge state of one linearisation (schur/layout.rs:79-125), computed on the device.
struct GaugeDev {
    double baseline;
    double scene_scale;
    double direction[3];
    double basis[6];
    int camera;  // -1 when free
    int _pad;
};

// Topology shared by all kernels (device pointers).
struct Topo {
    uint32_t nc, ng, np, no, D, nblocks;
    const uint32_t* obs_cam;
    const uint32_t* obs_pt;
    const uint32_t* pt_off;
    const uint8_t* pt_fixed;
    const uint32_t* cam_group;
    const uint32_t* cam_off;
    const uint32_t* cam_dim;
    const uint32_t* grp_off;
    const uint32_t* grp_dim;
    const uint32_t* blk_off;      // parameter offset per block
    const uint32_t* blk_dim;
    const uint32_t* blk_obs_off;  // observations per block
    const uint32_t* blk_obs;
    const uint32_t* grp_cam_off;  // cameras per group (for intrinsic-block edge terms)
    const uint32_t* grp_cams;
    // Point chunks of the point-major passes: chunk k holds points chunk_pt[k]..chunk_pt[k+1]
    // and their observations chunk_obs[k]..chunk_obs[k+1] (at most NT, or one point with more).
    uint32_t nchunks;
    const uint32_t* chunk_pt;
    const uint32_t* chunk_obs;
    // Camera-major order: positions cm_off[c]..cm_off[c+1] hold camera c's observations in
    // ascending point-major index; cm_obs gives that index and cm_pt its point.  The camera-major
    // arrays are padded to soa_pad(no) positions.
    const uint32_t* cm_off;
    const uint32_t* cm_obs;
    const uint32_t* cm_pt;
    const uint32_t* cm_cam;  // camera of position j
    const double* cm_uv;     // pixel of position j (principal point removed), 2 per position
    // Batched windows (null for one problem): window of each camera and point, and the damping of
    // each window; the gauge argument of the kernels then holds one GaugeDev per window.
    const uint32_t* cam_win;
    const uint32_t* pt_win;
    const double* win_lambda;
    int loss;
    double loss_scale;
    double* gturn;  // ng scratch: first radial turn per group of the intrinsics being projected with
};

// Rows of the structure-of-arrays linearisation copies (each row holds one value per observation):
// r~ (2); the compact state n = (x/z, y/z), sqrt(w)/z, sqrt(w) (4); the rotated point R X (3,
// camera-major copy only); J~_X (2x3 row-major, point-major copy); J~_c (2x9 row-major, point-major
// copy, written only for the block-gradient path and the per-observation hook).
enum : int { SOA_R = 0, SOA_Q = 2, SOA_RX = 6, SOA_JX = 9, SOA_JC = 15, SOA_ROWS = 33 };
// The copies are tiled: each run of SOA_TILE consecutive observations keeps its SOA_ROWS rows
// together (SOA_ROWS x SOA_TILE values), so one wavefront's load of a row is one contiguous
// segment and the rows it reads share DRAM pages.  soa_pad(no) positions are allocated (and the
// per-observation index arrays padded to that length): SOA_SLACK spare positions after the last
// observation let passes read whole fixed-size windows and aligned groups with vector loads.
constexpr size_t SOA_TILE = 64, SOA_SLACK = 256;
__host__ __device__ inline size_t soa_at(int row, size_t o) {
    return (o / SOA_TILE) * (SOA_TILE * SOA_ROWS) + (size_t)row * SOA_TILE + o % SOA_TILE;
}
inline size_t soa_pad(size_t no) { return (no + SOA_SLACK + SOA_TILE - 1) / SOA_TILE * SOA_TILE; }

// Per-observation linearisation (kernel (a)) outputs; T-typed.
template <class T>
struct Lin {
    T* rj;      // 8 per observation: r~ (2), J~_X (2x3 row-major); or null (block-pair path only)
    T* jc;      // 18 per observation: J~_c (2x9), rows [pose 0..5 | intrinsic 6..8]; or null
    double* cost;  // robust cost per observation (FP64 in both forms)
    double* sq;    // squared residual norm per observation
    uint8_t* valid;
    T* jpm;     // tiled rows (soa_at), point-major
    T* jcm;     // tiled rows (soa_at), camera-major (position j holds observation cm_obs[j]); or null
    int pose_rows;  // k_linearize writes the pose columns of J~_c into jpm (else only jcm has them)
};

// Per-point elimination (kernel (b)) outputs.  V^-1 and y are zero for points that are fixed or
// not eliminated.  PE_REC values per point make the packed record one 128-byte line in FP64.
constexpr int PE_REC = 16;
template <class T>
struct PointElim {
    T* vinv;     // 6 per point: packed inverse of the damped block V^
    T* vdamp;    // 6 per point: packed damped block V^
    T* y;        // 3 per point: V^-1 g_p
    T* ddiag;    // 3 per point: damping diagonal max(V_ii, floor)
    T* rrows;    // 27 per observation: R = V^-1 W rows (9 rows x 3); or null (block-pair path only)
    T* prec;     // PE_REC per point: [V^-1 (6), y (3)] packed for camera-side gathers; or null
    double* pscal;  // 6 per point: cost, squared residual, point model reduction, scaled gradient
                    // norm, rcond, flags (bit0 linearised, bit1 eliminated, bit2 weak)
};

// Scalars of one assembly (double) gathered by reduce_assembly.
struct AsmScalars {
    double cost, sq, pmr, sgin;
    double min_rcond;
    double weak;
    double invalid;     // number of points that failed linearisation or elimination
    double unprojectable;  // number of points with an observation outside the projection domain
    double sgin_prepared, min_diag, max_diag, clamped, nonfinite;
};

struct CandScalars {
    double cost, step2, damping_norm, rejected;
};

template <class T>
void launch_gauge(hipStream_t s, const T* cam_t, int camera, double scene_scale, GaugeDev* out);
template <class T>
void launch_linearize(hipStream_t s, const Topo& tp, const double* R, const double* t, const double* gp,
                      const double* X, const double* uv, const GaugeDev* gauge, Lin<T> out);
template <class T>
void launch_point_elim(hipStream_t s, const Topo& tp, const Lin<T> lin, T lambda, bool want_rcond,
                       PointElim<T> out);
// Sum of per-point scalars in a fixed order -> AsmScalars fields cost..invalid.
// Inlier count and squared-norm sum of the residuals within `limit` (out: 2 doubles); partials
// holds 2 * 512 doubles.
void launch_inliers(hipStream_t s, const double* sq, const uint8_t* valid, uint32_t no, double limit, double* partials,
                    double* out);
void launch_reduce_points(hipStream_t s, const double* pscal, uint32_t np, double* partials,
                          AsmScalars* out);
// ---- explicit reduced camera system (same topologies as the camera-side assembly) ----
// S = U - W V^-1 W^T in camera blocks of 9 rows (pose 0..5, the camera's intrinsics 6..8; rows
// and columns beyond the free dimensions are zero).  S is symmetric; its diagonal and upper blocks
// are stored as row slices: row a of camera c holds 9 values per slot s of its upper row (slot 0
// the camera itself, then its partners j > c ascending, nbr[nbr_off[c] + s]):
// vals[row_base[c] + a * schur_row_len(nn_c) + 9 * s + b], nn_c = nbr_off[c + 1] - nbr_off[c];
// each row slice is padded to a multiple of 4 values (16-byte groups in both precisions; the
// padding is never read as data).  Pair number pair_off[c] + s - 1 is the block in slot s >= 1.
__host__ __device__ inline uint32_t schur_row_len(uint32_t nn) { return (9 * nn + 3) & ~3u; }
template <class T>
struct SchurRows {
    T* vals;
    const uint64_t* row_base;   // nc
    const uint32_t* nbr_off;    // nc + 1
    const uint32_t* nbr;        // upper neighbour cameras (the camera first)
    const uint32_t* pair_off;   // nc + 1: pairs (c, j > c) of row c
    const uint32_t* rev_off;    // nc + 1: per camera j, the range of its pairs (c < j, j) in rev_pair
    const uint32_t* rev_pair;   // the pair numbers, ascending c per camera j
    double* tpair;              // 9 per pair: S_cj^T v_c (the product's first pass)
    double* qrow;               // 9 per camera: sum over the camera's upper row of S v
};
// Camera pairs c < j that see common points, with the observation pairs (camera-major positions
// of the observations by c and by j) of those points in point order.
struct SchurPairs {
    uint32_t npairs;
    const uint32_t* cam;       // 2 per pair: c, j
    const uint32_t* term_off;  // npairs + 1
    const uint32_t* term;      // 2 per term
    uint32_t ngroups;
    const uint32_t* group;     // per workgroup: first pair (SP_GROUP pairs of one camera row at most)
    const uint8_t* group_n;    // per workgroup: its pair count (0: an empty slot)
};
// Per camera-major observation, the compact record that the explicit system's pair assembly
// rebuilds the observation's Jacobians from: n = (x/z, y/z), sqrt(w)/z, sqrt(w), the rotated point
// R X, a zero, then the packed V^-1 of its point and two zeros (values 0..7 are the first half
// line, which is all a partner's record needs).
constexpr int CREC = 16;
constexpr int CREC_HALF = 8;
// Off-diagonal blocks S_cj = -sum_t J~_c,o^T (J~_X,o V^-1 J~_X,o'^T) J~_c,o' (= -W_o V^-1 W_o'^T)
// over each pair's terms, the Jacobians rebuilt from the records, the cameras' R, f, k1, k2 and
// the gauge; stored in row c.
template <class T>
void launch_schur_pairs(hipStream_t s, const SchurPairs& P, const T* crec, const double* R, const double* gp,
                        const uint32_t* cam_group, const uint32_t* cam_dim, const GaugeDev* gauge, SchurRows<T> S);
// Pairs per workgroup of the pair assembly.
constexpr uint32_t SP_GROUP = 4;
// Workgroup slots of the pair assembly: camera row c's pairs go, in groups of SP_GROUP, to slots
// b = 8 i + (c mod 8), rows in ascending order within each residue, so that with workgroups dealt
// round-robin to the eight XCDs a row stays on one XCD's L2 while all XCDs sweep the same band of
// rows.
void build_pair_groups(const std::vector<uint32_t>& pair_cam, uint32_t nc, std::vector<uint32_t>& group,
                       std::vector<uint8_t>& group_n);

// Camera-side assembly of the PCG form when every intrinsic group has one camera and no camera
// sees a point twice: per camera, the pose and intrinsic blocks' full and reduced gradients,
// unclamped diagonal, and the diagonal blocks of U - W V^-1 W^T (into the diagonal pairs' U
// values; their Schur values are set to zero).  Runs over the PCG camera segments; part is
// scratch of CAM_ASM_VALUES doubles per segment.  With z set it also writes each camera-major
// observation's Z record and, with S.vals set, each camera's full diagonal block of S.  The Jacobians are rebuilt from the camera-major compact state
// and rotated points with the rotations R and group intrinsics gp.
constexpr int CAM_ASM_VALUES = 72;
template <class T>
void launch_camera_assembly(hipStream_t s, const Topo& tp, const Lin<T> lin, const PointElim<T> pe, const double* R,
                            const double* gp, const GaugeDev* gauge, uint32_t nseg, const uint32_t* seg_cam,
                            const uint32_t* cseg_off, double* part, const int32_t* diag_pair,
                            const uint64_t* value_off, double* grad, double* full_grad, double* diag, double* u_vals,
                            double* s_vals, T* crec = nullptr, SchurRows<T> S = SchurRows<T>{});
// Per block: full gradient J~^T r~, reduced gradient (full minus W V^-1 g), unclamped diagonal.
template <class T>
void launch_block_gradient(hipStream_t s, const Topo& tp, const Lin<T> lin, const PointElim<T> pe,
                           double* grad, double* full_grad, double* diag);
// Block-pair values: U (Gram) and Schur (W V^-1 W^T) parts, in pair value space.
template <class T>
void launch_pairs(hipStream_t s, const Topo& tp, const Lin<T> lin, const PointElim<T> pe, uint32_t npairs,
                  const uint32_t* pair_a, const uint32_t* pair_b, const uint32_t* term_off,
                  const uint32_t* term_i, const uint32_t* term_j, const uint64_t* value_off, double* u_vals,
                  double* s_vals);
// Dense S (row-major, both triangles) = U - Schur from pair value space.
void launch_scatter_dense(hipStream_t s, const Topo& tp, uint32_t npairs, const uint32_t* pair_a,
                          const uint32_t* pair_b, const uint64_t* value_off, const double* u_vals,
                          const double* s_vals, double* S, uint32_t ld);
// Dense S (row-major, both triangles) from the explicit system's camera rows (one camera per
// intrinsic group, no camera observing a point twice); max_slots: the longest row's blocks.
template <class T>
void launch_scatter_rows(hipStream_t s, const Topo& tp, const SchurRows<T> S, uint32_t max_slots, double* Sd,
                         uint32_t ld);
// Clamp the damping diagonal per block, add lambda * diag to the reduced diagonal (dense S when
// S != null, else the diagonal pairs' U and Schur-block values), and finish the scaled gradient.
void launch_finalize(hipStream_t s, const Topo& tp, double lambda, double* diag, const double* full_grad,
                     const double* grad, double* S, uint32_t ld, const int32_t* diag_pair,
                     const uint64_t* value_off, double* u_vals, double* partials, AsmScalars* sc);

// Candidate pass (kernel (a')): back-substitution and candidate rule per point; R, t, gp the
// current state, Rc, tc, gpc the candidate cameras; srec is scratch of PCG_CAM_REC doubles per
// camera for the step records.
void launch_retract(hipStream_t s, const Topo& tp, const double* R, const double* t, const double* gp,
                    const double* step, const GaugeDev* gauge, double* Rc, double* tc, double* gpc);
template <class T>
void launch_candidate(hipStream_t s, const Topo& tp, const double* R, const double* t, const double* gp,
                      const double* Rc, const double* tc, const double* gpc, const double* X, const double* uv,
                      const Lin<T> lin, const PointElim<T> pe, const double* step, const GaugeDev* gauge, double* srec,
                      double* Xc, double* cscal);
void launch_reduce_candidates(hipStream_t s, const double* cscal, uint32_t np, double* partials,
                              CandScalars* out);
// dots[0] = g.x, dots[1] = diag.x^2, dots[2] = x.x, dots[3] = x.product
void launch_step_dots(hipStream_t s, uint32_t D, const double* g, const double* diag, const double* x,
                      const double* product, double* dots);
// Calibrated step limit (optimizer.rs:677-702): when the largest camera rotation or translation
// block norm of `step` exceeds `maximum`, scale step (and product, when not null) by
// maximum / largest; *alpha receives the factor (1 when not limited).
void launch_limit_step(hipStream_t s, const Topo& tp, double maximum, double* step, double* product, double* alpha);

// ---- implicit-Schur PCG (kernel (c)) ----
// The camera pass reads the structure-of-arrays rows in 16-byte groups, pcg_vec<T>() consecutive
// observations per load, and runs one wavefront per camera segment: camera c's camera-major
// positions from cm_off[c] rounded down to a multiple of pcg_vec<T>() are cut into windows of
// pcg_seg<T>() (one group per lane), segments cseg_off[c]..cseg_off[c+1].
template <class T>
constexpr uint32_t pcg_vec() { return 16 / sizeof(T); }
template <class T>
constexpr uint32_t pcg_seg() { return 64 * pcg_vec<T>(); }
static_assert(64 * 16 / sizeof(float) <= SOA_SLACK, "camera windows may end past the spare positions");
static_assert(SOA_TILE % (16 / sizeof(float)) == 0, "16-byte groups stay inside one tile");
constexpr uint32_t PCG_CAM_REC = 24;  // doubles per camera record: R, f, k1, k2, motion (21 used)
constexpr uint32_t PCG_XW = 4;        // doubles per point record: the point pass output w (3 used)
constexpr uint32_t PCG_WREC = 9;      // doubles per camera world record
constexpr uint32_t PCG_WAVE_OBS = 64; // observations per point-pass chunk (whole points)
constexpr uint32_t PCG_BLOCK_WG = 64;  // workgroup of the block-parallel PCG kernels
constexpr uint32_t PCG_Q_WG = 1024;    // workgroup of the block product pass (a wavefront per block)
struct PcgState {
    double rho, n0, rel, tol;
    double tr;     // true residual ratio of a failed confirmation (restart)
    int mode, iter, max_iter, done, accepted;
    int products;  // Schur products applied (each update that ran follows one product)
    int par;       // direction buffer holding d
    int _pad;
};
// Vectors of one PCG solve (device).
struct PcgVecs {
    const double* grad;     // D, reduced gradient (the right-hand side is -grad)
    const double* factors;  // 54 per block, block-Jacobi factors
    const double* diag;     // D, clamped damping diagonal
    const double* R;        // 9 per camera: rotations of the linearisation state
    const double* X;        // 3 per point: points of the linearisation state
    const double* gp;       // 3 per group: intrinsics (f, k1, k2) of the linearisation state
    const GaugeDev* gauge;  // gauge of the linearisation
    double lambda;
    double *x, *r, *z, *q;  // D each
    double* d[2];           // D each: the direction alternates between the two
    double* xw;    // PCG_XW per point: the point pass output w
    double* seg;   // 9 per camera segment: the camera pass sums (E, F, I)
    double* vcam;  // PCG_CAM_REC per camera: rotation, intrinsics and the product input
    double* wcam;  // PCG_WREC per camera: the product input in the world frame (point pass)
    double* vexp;  // 9 per camera: the product input in camera coordinates (explicit product), or null
    double* part;  // per-workgroup partial dot products: v.q (nqparts), r.r and r.z (nparts each)
    uint32_t nqparts, nparts;
    uint32_t nseg;
    const uint32_t* seg_cam;   // camera of each segment
    const uint32_t* cseg_off;  // nc + 1 segment offsets
    uint32_t npchunks;           // point-pass chunks: whole points in at most PCG_WAVE_OBS lanes
    const uint32_t* pchunk_pt;   // npchunks + 1 point offsets
    const uint32_t* pchunk_obs;  // npchunks + 1 observation offsets
    uint32_t ncu;                // compute units (persistent point-pass grid)
    PcgState* st;   // read by the product passes; written by the direction update
    PcgState* st1;  // written by the step update; read by the direction update
};
// Per-workgroup partial counts of the block-parallel PCG kernels and of the block product pass.
uint32_t pcg_partials(const Topo& tp);
uint32_t pcg_q_partials(const Topo& tp);
// Factor the block-Jacobi preconditioner: U_AA(damped) - Schur_AA per block.
void launch_precond_factor(hipStream_t s, const Topo& tp, const int32_t* diag_pair, const uint64_t* value_off,
                           const double* u_vals, const double* s_vals, double* factors, int* fail);
// x = 0, r = -g, z = M^-1 r, the first direction and its per-camera input.
template <class T>
void launch_pcg_init(hipStream_t s, const Topo& tp, const PcgVecs& V, double tol, int max_iter);
// Implicit Schur product q = S v (v = d or x by the PCG mode), S = U + lambda diag - W V^-1 W^T:
//   point pass   w_p = V^-1 sum_o J~_X,o^T J~_c,o v                   (point chunks, point-major J)
//   camera pass  per segment sums of J~_c,o^T (J~_c,o v - J~_X,o w_p)  (camera-major J)
// Both read only J~_X and the intrinsic columns of J~_c per observation: the pose columns follow
// from J~_X, the point and the camera rotation.
template <class T>
void launch_pcg_points(hipStream_t s, const Topo& tp, const Lin<T> lin, const PointElim<T> pe, const PcgVecs& V);
template <class T>
void launch_pcg_cameras(hipStream_t s, const Topo& tp, const Lin<T> lin, const PcgVecs& V);
// q = S v per block, then the step (x, r, z) and direction (d, per-camera input) updates.
template <class T>
void launch_pcg_step(hipStream_t s, const Topo& tp, const PcgVecs& V);
// Explicit product q = S v + lambda diag v (v = d or x by the PCG mode, from V.vexp) with the
// partials of v.q, in two launches: the stored upper rows (row sums into S.qrow, column sums of
// the off-diagonal blocks into S.tpair), then per camera the row part, its blocks' transposed
// parts and the damping.  Then the step and direction updates.
template <class T>
void launch_pcg_product_upper(hipStream_t s, const Topo& tp, const SchurRows<T> S, const PcgVecs& V);
template <class T>
void launch_pcg_product_lower(hipStream_t s, const Topo& tp, const SchurRows<T> S, const PcgVecs& V);
template <class T>
void launch_pcg_step_explicit(hipStream_t s, const Topo& tp, const PcgVecs& V);
// Partials of the explicit product (one per camera).
uint32_t pcg_product_partials(const Topo& tp);
// Camera-major linearisation: the SoA camera-major Jacobian copy (lin.jcm).
template <class T>
void launch_linearize_cm(hipStream_t s, const Topo& tp, const double* R, const double* t, const double* gp,
                         const double* X, const double* uv, const GaugeDev* gauge, Lin<T> out);

// ---- batched windows: independent problems concatenated, one LM control per window ----
struct WinTopo {
    uint32_t W;
    const uint32_t* cam_off;    // W+1 camera ranges
    const uint32_t* grp_off;    // W+1 group ranges
    const uint32_t* pt_off;     // W+1 point ranges
    const uint32_t* param_off;  // W+1 parameter ranges (D_w = difference)
    const uint64_t* S_off;      // W+1 dense per-window matrix offsets (D_w^2 each)
    const uint64_t* linv_off;   // W+1 offsets of the per-window diagonal-tile inverses
    const uint32_t* grp_win;    // window of each group
    const uint32_t* blk_win;    // window of each block
    const int32_t* gauge_cam;   // W, global baseline camera or -1
    const double* scene;        // W, scene scale
};
// LM control of one window (sparse/self_calibration/mod.rs:212-639 for a single-stage loss and
// the direct camera step).
struct LmWin {
    double nu, gain, current, pmr, initial_cost, initial_sq, step_norm;
    int iterations, accepted, target_accepted, status;
    int done, go, accepted_now, has_gain, have_initial;
    int rej[4];
};
struct LmParams {
    uint32_t max_iterations;
    double ftol, gtol, ptol;
};
struct WinTrace {  // [iteration * W + window], optional
    double* cost;
    double* lambda;
    double* candidate_cost;
    uint8_t* accepted;
};
void launch_win_gauge(hipStream_t s, const WinTopo& wt, const double* t, GaugeDev* out);
void launch_win_reduce_points(hipStream_t s, const WinTopo& wt, const double* pscal, AsmScalars* out);
void launch_win_zero(hipStream_t s, double* p, size_t n);
void launch_win_scatter(hipStream_t s, const Topo& tp, const WinTopo& wt, uint32_t npairs, const uint32_t* pair_a,
                        const uint32_t* pair_b, const uint64_t* value_off, const double* u_vals,
                        const double* s_vals, double* S);
// lambda: damping per window (Topo::win_lambda carries the same array to the point elimination).
void launch_win_finalize(hipStream_t s, const Topo& tp, const WinTopo& wt, const double* lambda, double* diag,
                         const double* full_grad, const double* grad, double* S, AsmScalars* out, bool damp);
// Per window: dense Cholesky of S_w (FP64, T = double) or of a Jacobi-scaled FP32 copy with FP64
// iterative refinement (T = float), solving S_w x = -g_w; one workgroup per window, tile edge tb
// (32 or 64, chosen by the autotuner).  Windows with go == 0 are skipped.
// The mixed form (T = float) then re-solves, with the FP64 factor of S_w, every window whose
// refinement stayed above 1e-10 relative residual (linv64: FP64 tile inverses; fallbacks counted
// per window).
template <class T>
void launch_win_cholesky(hipStream_t s, const WinTopo& wt, int tb, double* S, T* F, T* linv, const double* grad,
                         double* x, double* vecs, const LmWin* lm, int* fail, double* refine_rel,
                         double* linv64, uint32_t* fallbacks);
void launch_win_step_dots(hipStream_t s, const WinTopo& wt, const double* g, const double* diag, const double* x,
                          double* dots);
void launch_win_reduce_cand(hipStream_t s, const WinTopo& wt, const double* cscal, CandScalars* out);
void launch_lm_begin(hipStream_t s, const WinTopo& wt, LmWin* lm, double* lambda, const AsmScalars* a, LmParams P,
                     WinTrace tr);
void launch_lm_solved(hipStream_t s, const WinTopo& wt, LmWin* lm, double* lambda, const int* fail, LmParams P);
void launch_lm_end(hipStream_t s, const WinTopo& wt, LmWin* lm, double* lambda, const CandScalars* c,
                   const double* dots, LmParams P, WinTrace tr);
void launch_win_accept(hipStream_t s, const Topo& tp, const WinTopo& wt, const LmWin* lm, double* R0, double* t0,
                       double* gp0, double* X0, const double* R1, const double* t1, const double* gp1,
                       const double* X1);
// Gauge state of one camera's translation (schur/layout.rs:101-113).
__host__ __device__ inline GaugeDev make_gauge(const double* t3, int camera, double scene_scale) {
    GaugeDev g;
    g.camera = camera;
    g._pad = 0;
    g.scene_scale = scene_scale;
    g.baseline = 0.0;
    for (int k = 0; k < 3; ++k) g.direction[k] = 0.0;
    for (int k = 0; k < 6; ++k) g.basis[k] = 0.0;
    if (camera < 0) return g;
    const double t[3] = {t3[0], t3[1], t3[2]};
    const double b = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
    g.baseline = b;
    const double dir[3] = {t[0] / b, t[1] / b, t[2] / b};
    for (int k = 0; k < 3; ++k) g.direction[k] = dir[k];
    // s2_tangent_basis (numerics/manifold.rs:129-141)
    const double n = sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
    const double tv[3] = {dir[0] / n, dir[1] / n, dir[2] / n};
    double ref[3] = {0, 0, 0};
    const double a0 = fabs(tv[0]), a1 = fabs(tv[1]), a2 = fabs(tv[2]);
    if (a0 <= a1 && a0 <= a2)
        ref[0] = 1.0;
    else if (a1 <= a2)
        ref[1] = 1.0;
    else
        ref[2] = 1.0;
    const double c0[3] = {tv[1] * ref[2] - tv[2] * ref[1], tv[2] * ref[0] - tv[0] * ref[2], tv[0] * ref[1] - tv[1] * ref[0]};
    const double n0 = sqrt(c0[0] * c0[0] + c0[1] * c0[1] + c0[2] * c0[2]);
    const double b0[3] = {c0[0] / n0, c0[1] / n0, c0[2] / n0};
    const double c1[3] = {tv[1] * b0[2] - tv[2] * b0[1], tv[2] * b0[0] - tv[0] * b0[2], tv[0] * b0[1] - tv[1] * b0[0]};
    const double n1 = sqrt(c1[0] * c1[0] + c1[1] * c1[1] + c1[2] * c1[2]);
    for (int k = 0; k < 3; ++k) {
        g.basis[k] = b0[k];
        g.basis[3 + k] = c1[k] / n1;
    }
    return g;
}

// ---- dense Cholesky (kernel (d)) ----
// Factor the lower triangle of the n x n row-major matrix A in place (MFMA FP64 trailing
// updates) and solve A x = b.  *fail != 0 after a pivot at or below 64 eps max|diag(A)|.
void launch_cholesky_solve(hipStream_t s, double* A, uint32_t n, uint32_t ld, const double* b, double* x,
                           double* work, int* fail);
size_t cholesky_work_doubles(uint32_t n);
// Mixed form: factor an FP32 copy F of the lower triangle of S (FP32 matrix cores), solve, then
// refine x against the FP64 S for at most max_steps steps (r = b - S x in FP64).  work[8..10]
// receive the refinement state (done flag in *done, last ||r||^2, ||b||^2); *fail != 0 after an
// FP32 pivot at or below 64 eps_32 max|diag|.
void launch_cholesky_solve_mixed(hipStream_t s, const double* S, float* F, uint32_t n, uint32_t ld, const double* b,
                                 double* x, double* work, int* fail, int* done, int max_steps);
void launch_negate(hipStream_t s, uint32_t n, const double* in, double* out);
// Factor only (no right-hand side), in place, with the stored L_kk^-1 tiles in `work`; *fail != 0
// after a pivot that is not finite and above rel_floor * max|diag(A)|, or, with tol_row, above
// tol_row[k] for pivot k (the rank certificate's tolerances of launch_rank_tolerances).
void launch_cholesky_factor(hipStream_t s, double* A, uint32_t n, uint32_t ld, double* work, int* fail,
                            double rel_floor, const double* tol_row = nullptr);
// The rank certificate's per-pivot tolerances of the lower triangle A (n x n, leading dimension
// ld): scratch holds 3 n doubles, the tolerances at scratch + 2 n.  Pivot k of A fails when it is
// at most 64 eps n g max(A_kk, 64 eps max_j A_jj), g the Gershgorin bound of the diagonally
// equilibrated matrix: the scale-aware test of the reference's spectral rank certificate
// (kernel/multi_view/sparse/direct.rs:342-457) applied to the pivots.
void launch_rank_tolerances(hipStream_t s, const double* A, uint32_t n, uint32_t ld, double* scratch);
// Marginal covariances of the trailing coordinates [j0, n) of a factored S (launch_cholesky_factor):
// for each group g with offset goff[g] >= j0 and dimension gdim[g] <= 3, cov[9g..] = w (S^-1)_gg
// (row-major 3x3, zero beyond gdim[g]); nothing when *fail.  R: trailing_inverse_doubles(n, j0)
// doubles of scratch.
size_t trailing_inverse_doubles(uint32_t n, uint32_t j0);
void launch_group_marginals(hipStream_t s, const double* L, uint32_t n, uint32_t ld, const double* work,
                            const int* fail, uint32_t j0, uint32_t ngroups, const uint32_t* goff,
                            const uint32_t* gdim, double w, double* R, double* cov);
}  // namespace ba
