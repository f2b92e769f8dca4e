// ba_host.hpp — host side of the MI355X bundle adjuster: lowering of a BundleProblem to the
// native self-calibrating problem, its index maps, and the lists that drive the
// deterministic reductions on the GPU.
#pragma once

#include <cstddef>
#include <cstdint>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "../../include/ba_hip.h"

namespace ba {

struct Status {
    int code = BA_OK;
    std::string message;
    bool ok() const { return code == BA_OK; }
};
inline Status error(int code, std::string message) { return Status{code, std::move(message)}; }

// IntrinsicVariables::is_fixed (problem.rs:136-185): no calibration coordinate is variable.
inline bool is_fixed_intrinsic(const ba_intrinsic& in) {
    return !in.var_focal && !in.var_principal_point && !in.var_radial && !in.var_tangential;
}

// The lowered problem: PreparedSelfCalibratingProblem (native/self_calibration.rs:30-44) plus the
// workspace layout and index maps (workspace/topology.rs:351-397, workspace/pattern.rs:306-468).
struct Lowered {
    // input bookkeeping
    size_t n_input_points = 0, n_quarantined = 0;
    std::vector<uint8_t> quarantined;      // per input point
    std::vector<uint32_t> point_source;    // native point -> input point
    std::vector<uint32_t> camera_source;   // native camera -> input pose
    std::vector<uint32_t> obs_source;      // native observation -> input observation
    double anchor_rotation[9];             // WorldRebase (native/world_rebase.rs:12-87)
    double anchor_center[3];
    double rebase_scale = 1.0;
    // native state, anchor-first cameras
    std::vector<double> R, t, gp, X, uv;   // 9/cam, 3/cam, 3/group, 3/point, 2/obs
    std::vector<uint32_t> cam_group, obs_cam, obs_point, point_off;
    std::vector<uint8_t> cam_fixed, point_fixed, grp_fixed;
    std::vector<uint32_t> grp_free;
    int gauge_camera = -1;
    // the calibrated (fixed-intrinsic) path of native/calibrated.rs: every intrinsic group fixed,
    // the camera step limited (optimizer.rs:677-702) and the direct step's model reduction from
    // the damped quadratic form (calibrated/camera_step.rs:40-62); its anchor stays free when
    // survey controls fix the frame without a fixed observed pose (native/gauge.rs:153-162)
    bool calibrated = false;
    bool anchor_free = false;
    int loss = 0;
    double loss_scale = 0.0;               // kernel scale (Cauchy: scale / sqrt 2)
    double scene_scale = 1.0;
    // layout
    size_t D = 0;
    std::vector<uint32_t> cam_off, cam_dim, grp_off, grp_dim;
    uint64_t n_cross = 0;  // point parameters (the size of params), always counted
    // reference index maps (built only for build_pattern)
    std::vector<uint32_t> cross_off, params, col_ptr, row_idx;
    size_t n_cameras() const { return cam_group.size(); }
    size_t n_groups() const { return gp.size() / 3; }
    size_t n_points() const { return point_fixed.size(); }
    size_t n_obs() const { return obs_cam.size(); }
};

Status validate_problem(const ba_problem* p);
// TrackGeometry verdict of every point (quarantine.rs:57-96) at the given camera centres
// (3 per pose) and points (3 per point): rays without parallax, or a camera on the point.
std::vector<uint8_t> degenerate_points(const ba_problem* p, const double* centers, const double* xyz);
Status validate_config(const ba_solve_config* c);
// quarantine = false keeps every point (the problems of the re-solves, whose input-degenerate
// points were held out before the first solve).
Status lower_problem(const ba_problem* p, Lowered& L, bool build_pattern, bool quarantine = true);
// Blocks and point-parameter count from the fixed flags (topology.rs:351-397).
void build_layout(Lowered& L);
// Hold the given intrinsic groups at their current values: fixed blocks, layout rebuilt
// (native/self_calibration.rs:240-247).
void hold_groups(Lowered& L, const std::vector<uint32_t>& groups);

// LinearSolver::Auto's camera-step choice for a lowered problem (execution_policy.rs:190-227).
void plan_execution(const Lowered& L, const ba_solve_config* c, ba_execution_plan& out);
// ExecutionPlan::rank_certificate_fits (execution_policy.rs:48-66) of an Auto plan that chose the
// retained PCG step: its estimate plus the certificate's reduced values within 4/5 of the budget.
bool rank_certificate_fits(const Lowered& L, const ba_execution_plan& plan);

// Block b < n_cameras is the pose block of camera b; block n_cameras + g the intrinsic block
// of group g.  Rows of a per-observation 9-row camera Jacobian: pose rows 0..5, intrinsic 6..8.
struct BlockLists {
    size_t n_blocks = 0;
    std::vector<uint32_t> block_off, block_dim;  // parameter offset and dimension
    std::vector<uint32_t> obs_off;               // n_blocks+1: observations touching a block
    std::vector<uint32_t> obs;                   // ascending observation index per block
};
BlockLists build_block_lists(const Lowered& L);

// Lower-triangle block pairs (A >= B) of the reduced camera system and their terms.  A term
// (i, j) is an ordered observation pair of one free point whose camera blocks hold A (from i)
// and B (from j); FLAG_U marks i == j, where the pair also carries the U Gram term.
constexpr uint32_t TERM_U = 1u << 30;       // add J~_i[A]^T J~_i[B]
constexpr uint32_t TERM_SCHUR = 1u << 31;   // subtract W_i[A] V^-1 W_j[B]^T
constexpr uint32_t TERM_INDEX = (1u << 30) - 1;
struct PairLists {
    std::vector<uint32_t> block_a, block_b;     // per pair
    std::vector<uint32_t> term_off;             // n_pairs+1
    std::vector<uint32_t> term_i, term_j;       // term_j carries the flags in its top bits
    std::vector<uint64_t> value_off;            // n_pairs+1: offset of the pair's dA*dB values
};
// all = every block pair of the pattern (direct solve); otherwise the diagonal blocks and the
// pose-intrinsic edges of each camera (the retained operator and its preconditioner).  Without
// with_terms (local pairs only) term_i and term_j stay empty; term_off still counts the terms.
PairLists build_pair_lists(const Lowered& L, const BlockLists& B, bool all, bool with_terms = true);

// An allocator that leaves value-initialised elements uninitialised: a large host array is then
// first written, page by page, by the threads that fill it rather than zeroed by one thread.
template <class T>
struct NoInitAlloc : std::allocator<T> {
    template <class U>
    struct rebind {
        using other = NoInitAlloc<U>;
    };
    NoInitAlloc() = default;
    template <class U>
    NoInitAlloc(const NoInitAlloc<U>&) noexcept {}
    template <class U>
    void construct(U*) noexcept {}
    template <class U, class... A>
    void construct(U* p, A&&... a) {
        ::new ((void*)p) U(std::forward<A>(a)...);
    }
};

// Lists of the explicit reduced camera system, for topologies with one camera per intrinsic
// group and at most one observation per point and camera.  S is symmetric and only its diagonal
// and upper blocks are stored: row c holds the camera itself (slot 0) and then its partners j > c
// (cameras sharing a free point with c) ascending.  The camera pairs c < j, numbered in (c, j)
// order, carry the observation pairs (camera-major positions pos[o] of the observations by c and
// by j) of their common free points, in point order.  The pairs (c < j, j) of each camera j,
// in ascending c, form its reverse range rev_off[j]..rev_off[j + 1] of rev_pair.
struct SchurLists {
    std::vector<uint32_t> nbr_off, nbr;        // nc + 1, upper neighbours (the camera first)
    std::vector<uint64_t> row_base;            // nc: first value of row c
    std::vector<uint32_t> pair_off, pair_cam;  // nc + 1 (pairs of row c), 2 per pair: (c, j)
    std::vector<uint32_t> rev_off, rev_pair;   // nc + 1, pair numbers
    std::vector<uint32_t> term_off;            // n_pairs + 1
    std::vector<uint32_t, NoInitAlloc<uint32_t>> term;  // 2 per term
    uint64_t n_values = 0;
};
// Number of observation pairs (sum of k(k-1)/2 over free points with k observations).
uint64_t count_schur_terms(const Lowered& L);
SchurLists build_schur_lists(const Lowered& L, const std::vector<uint32_t>& pos);

// Raise the native state back into the caller's frame (native/self_calibration.rs:375-436,
// native/world_rebase.rs:61-86).
void raise_solution(const ba_problem* p, const Lowered& L, const double* R, const double* t,
                    const double* gp, const double* X, ba_solution* out);

}  // namespace ba
