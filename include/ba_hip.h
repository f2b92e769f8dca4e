/*
 * ba_hip.h — C ABI of the MI355X (gfx950) Levenberg–Marquardt bundle-adjustment path.
 *
 * This library is a drop-in for one entry point of openmvg-ba:
 *
 *     impl BundleAdjuster for X {
 *         fn name(&self) -> &'static str;
 *         fn solve(&mut self, &BundleProblem, SolveConfig) -> Result<Solution, SolveError>;
 *     }                                   (crates/openmvg-ba/src/solver.rs:1562-1573)
 *
 * restricted to two of the paths the native adjuster takes: problems with free intrinsics
 * (crates/openmvg-ba/src/native/mod.rs:352-367 -> native/self_calibration.rs:223-249):
 * pixel residuals, perspective cameras with one focal length, free (f[, k1[, k2]]),
 * a fixed principal point, full-pose or fixed poses; and problems whose observed intrinsics
 * are all fixed (native/mod.rs:336-455 -> native/calibrated.rs): pixel residuals, pinhole or
 * Brown-Conrady cameras with one focal length and radial k1/k2, full-pose or fixed poses,
 * not a two-view problem (native/two_view.rs:14-50).  Anything else returns
 * BA_ERR_UNSUPPORTED, which callers treat like SolveError::Unsupported
 * (crates/openmvg-sfm/src/ba/native.rs:79-85: the caller falls back to another adjuster).
 *
 * Conventions
 *   - Plain C types only; no allocation crosses the boundary.  Every output buffer is
 *     allocated by the caller with the sizes of the problem it passes.
 *   - Problem arrays are host pointers for ba_hip_solve (the library copies them to HBM)
 *     and device pointers for the per-kernel hooks (ba_hip_linearize ... ba_hip_cost_grad).
 *   - Rotations are row-major 3x3 world-to-camera matrices; a pose is (R, C) with
 *     camera point = R (X - C), as openmvg-core's WorldToCameraPose.
 *   - Return codes mirror SolveError (crates/openmvg-ba/src/solver.rs:1474-1493).
 *   - One solve at a time per handle (the reference trait takes &mut self; SfM serialises
 *     calls through a Mutex, crates/openmvg-sfm/src/ba/native.rs:30-33).  Results do not
 *     depend on scheduling: every reduction runs in a fixed order, with no float atomics.
 */
#ifndef OPENMVG_BA_HIP_H
#define OPENMVG_BA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes: SolveError (solver.rs:1474-1493) plus ProblemError (problem.rs build) ---- */
enum ba_status {
    BA_OK = 0,
    BA_ERR_INVALID_CONFIG = 1,      /* SolveError::InvalidConfig */
    BA_ERR_UNSUPPORTED = 2,         /* SolveError::Unsupported: the sanctioned decline */
    BA_ERR_RESOURCE_EXHAUSTED = 3,  /* SolveError::ResourceExhausted */
    BA_ERR_NUMERICAL = 4,           /* SolveError::Numerical */
    BA_ERR_INTERNAL = 5,            /* SolveError::Internal (also: HIP runtime failures) */
    BA_ERR_INVALID_PROBLEM = 6      /* ProblemError from BundleProblemBuilder::build (problem.rs:492-576) */
};

/* ---- camera model: openmvg-core Camera, the variants self-calibration reads ---- */
enum ba_camera_model {
    BA_CAMERA_PINHOLE = 0,          /* Camera::Pinhole */
    BA_CAMERA_BROWN_NONE = 1,       /* Camera::BrownConrady with RadialDistortion::None */
    BA_CAMERA_BROWN_K1 = 2,         /* RadialDistortion::K1 */
    BA_CAMERA_BROWN_K1K2 = 3,       /* RadialDistortion::K1K2 */
    BA_CAMERA_BROWN_K1K2K3 = 4,     /* RadialDistortion::K1K2K3 */
    BA_CAMERA_FISHEYE = 5,          /* unsupported for self-calibration */
    BA_CAMERA_SPHERICAL = 6         /* unsupported for self-calibration */
};

/* BundleIntrinsic (problem.rs:136-185): calibration, variable mask, certification. */
typedef struct ba_intrinsic {
    int32_t model;                  /* enum ba_camera_model */
    double fx, fy, cx, cy;
    double k1, k2, k3, p1, p2;
    uint32_t width, height;
    uint8_t fixed_calibration;      /* CameraIntrinsics::is_fixed */
    uint8_t var_focal, var_principal_point, var_radial, var_tangential; /* IntrinsicVariables */
    uint8_t certification_waived;   /* IntrinsicCertification::Waived (else Required) */
    uint8_t _pad[2];
} ba_intrinsic;

enum ba_residual_space { BA_RESIDUAL_PIXELS = 0, BA_RESIDUAL_ANGULAR = 1 };
enum ba_robust_loss { BA_LOSS_SQUARED = 0, BA_LOSS_HUBER = 1, BA_LOSS_CAUCHY = 2 };
enum ba_gauge { BA_GAUGE_FREE = 0, BA_GAUGE_FIXED_BASELINE = 1 };

/* BundleProblem (problem.rs:447-455) as structure-of-arrays views. */
typedef struct ba_problem {
    size_t n_poses;
    const double* pose_rotation;          /* n_poses*9, row-major world-to-camera */
    const double* pose_center;            /* n_poses*3, world camera centre */
    const uint8_t* pose_var_rotation;     /* n_poses, PoseVariables::rotation */
    const uint8_t* pose_var_translation;  /* n_poses, PoseVariables::translation */
    size_t n_intrinsics;
    const ba_intrinsic* intrinsics;       /* n_intrinsics */
    size_t n_points;
    const double* point_xyz;              /* n_points*3 */
    const uint8_t* point_fixed;           /* n_points, or NULL when no point is fixed */
    size_t n_observations;
    const uint32_t* obs_pose;             /* n_observations */
    const uint32_t* obs_intrinsic;        /* n_observations */
    const uint32_t* obs_point;            /* n_observations */
    const double* obs_pixel;              /* n_observations*2, pixels */
    int32_t residual_space;               /* enum ba_residual_space */
    int32_t robust_loss;                  /* enum ba_robust_loss */
    double loss_scale;                    /* Huber/Cauchy ResidualScale value, pixels */
    int32_t gauge;                        /* enum ba_gauge */
    uint32_t gauge_first, gauge_second;   /* Gauge::FixedBaseline */
    double gauge_distance;
} ba_problem;

enum ba_linear_solver { BA_SOLVER_AUTO = 0, BA_SOLVER_DIRECT = 1, BA_SOLVER_ITERATIVE = 2 };

/* Arithmetic of the inner loop.  PLAIN runs kernels (a)-(d) in FP64; MIXED runs (a)-(c)
 * in FP32 with compensated cost and gradient sums, factors (d) in FP32 and recovers the
 * Newton step by FP64 iterative refinement.  The LM gain ratio, damping update and
 * acceptance test are FP64 in both forms.  (No counterpart in the reference, which is FP64.) */
enum ba_precision { BA_PRECISION_FP64 = 0, BA_PRECISION_MIXED = 1 };

/* SolveConfig (solver.rs:74-155) plus the two knobs of this implementation. */
typedef struct ba_solve_config {
    uint32_t max_iterations;              /* default 50 */
    double function_tolerance;            /* default 1e-6 */
    double gradient_tolerance;            /* default 1e-10 */
    double parameter_tolerance;           /* default 1e-8 */
    int32_t linear_solver;                /* enum ba_linear_solver */
    uint64_t workspace_budget_bytes;      /* 0 = automatic: a 1 GiB planning budget for AUTO and no
                                             limit; else the planning budget of AUTO and a limit on
                                             the solve's device memory (BA_ERR_RESOURCE_EXHAUSTED) */
    int32_t precision;                    /* enum ba_precision */
    uint32_t pcg_max_iterations;          /* 0 = the native cap min(600, D) */
    double pcg_relative_tolerance;        /* 0 = the native forcing tolerance 1e-1 */
    int32_t profile_kernel;               /* enum ba_profile_kernel: time this kernel's launches */
    int32_t certification;                /* enum ba_certification; 0 (the reference's) by default */
} ba_solve_config;

/* What follows the LM loop.  FULL is the reference's native solve: the final rank certificate
 * (sparse/self_calibration/mod.rs:498-560, direct.rs:288-334), the intrinsic marginal covariances
 * with the re-solve that holds unconstrained groups at their input values
 * (native/self_calibration.rs:223-303), the weak-point re-solves (native/mod.rs:260-350) and the
 * round-off canonicalisation (solver.rs:1030-1101).  LOOP_ONLY returns the LM loop's result with
 * the certificate not assessed (no counterpart in the reference). */
enum ba_certification { BA_CERTIFY_FULL = 0, BA_CERTIFY_LOOP_ONLY = 1 };

/* Observability (solver.rs:187-208) of the final undamped reduced camera system. */
enum ba_observability {
    BA_OBSERVABILITY_FULL_RANK = 0,
    BA_OBSERVABILITY_ONLY_EXPECTED_GAUGE = 1,
    BA_OBSERVABILITY_RANK_DEFICIENT = 2,
    BA_OBSERVABILITY_NOT_ASSESSED = 3
};

/* Limits of the certificate this library forms (a dense factor of the reduced system on the
 * device, assembled from the block pairs of every point's observations): at most this reduced
 * dimension and this many block-pair terms (the sum over free points of (2 x observations)^2).
 * Beyond them the solve reports BA_OBSERVABILITY_NOT_ASSESSED, forms no marginals and holds no
 * group, where the reference would factor its sparse system. */
#define BA_CERTIFY_MAX_DIMENSION 8192
#define BA_CERTIFY_MAX_TERMS 67108864

/* Kernels that ba_solve_config.profile_kernel can time with HIP events on the solve's stream. */
enum ba_profile_kernel {
    BA_PROFILE_NONE = 0,
    BA_PROFILE_LINEARIZE = 1,      /* kernel (a) */
    BA_PROFILE_POINT_ELIM = 2,     /* kernel (b), per-point elimination */
    BA_PROFILE_PCG_POINTS = 3,     /* kernel (c), point pass of the implicit Schur product */
    BA_PROFILE_PCG_CAMERAS = 4,    /* kernel (c), camera pass of the implicit Schur product */
    BA_PROFILE_PAIRS = 5,          /* kernel (b), block-pair segmented reduction */
    BA_PROFILE_CHOLESKY = 6,       /* kernel (d), factor and solve */
    BA_PROFILE_CANDIDATE = 7,      /* candidate pass */
    BA_PROFILE_SCHUR_PAIRS = 8,    /* kernel (b), off-diagonal blocks of the explicit S */
    BA_PROFILE_PCG_PRODUCT = 9     /* kernel (c), explicit block-sparse product with S */
};

/* SolveStatus (solver.rs:172-185) with ConvergenceReason folded in. */
enum ba_solve_status {
    BA_STATUS_CONVERGED_GRADIENT = 0,
    BA_STATUS_CONVERGED_STEP = 1,
    BA_STATUS_CONVERGED_COST = 2,
    BA_STATUS_MAX_ITERATIONS = 3,
    BA_STATUS_STALLED = 4,
    BA_STATUS_DEGENERATE = 5,
    BA_STATUS_SKIPPED = 6             /* the input state returned (round-off non-descent) */
};

/* SolveReport (solver.rs:461-473) and the kernel report fields it is built from
 * (kernel/multi_view/sparse/self_calibration/mod.rs:569-638, native/reporting.rs:24-76). */
typedef struct ba_solve_report {
    int32_t status;                       /* enum ba_solve_status */
    int32_t linear_solver_used;           /* BA_SOLVER_DIRECT or BA_SOLVER_ITERATIVE */
    uint32_t iterations;
    uint32_t accepted_steps;
    uint32_t rejected_steps;
    uint32_t rejected_linearization, rejected_step_solve, rejected_candidate,
             rejected_insufficient_reduction;
    double initial_cost, final_cost;
    double initial_rmse, final_rmse;      /* per-axis pixel RMSE x sqrt(2), as reporting.rs:72-73 */
    double gradient_inf_norm;             /* scaled, at the final zero-damping assembly */
    double step_norm;
    double initial_damping, final_damping;
    double last_gain_ratio;               /* NaN when none */
    uint32_t reduced_dimension;           /* D of ReducedLayout::try_radial */
    uint32_t weak_point_count;
    double minimum_point_rcond;
    uint32_t pcg_iterations;              /* retained_inner_iterations */
    uint32_t quarantined_points;
    double maximum_pcg_relative_residual; /* NaN when no PCG step ran */
    double normalization_scale;
    uint64_t observation_count, point_count, pose_count;
    /* wall-clock split of the call (no counterpart in the reference report) */
    double seconds_setup;                 /* lowering, index maps, uploads */
    double seconds_iterations;            /* the LM loop: first iteration's start to the last one's end */
    double seconds_final;                 /* projection check, final zero-damping assembly, download */
    uint32_t profile_launches;            /* launches of the profiled kernel */
    double profile_seconds;               /* their summed HIP-event durations */
    /* mixed form: direct steps solved with the FP64 factor because the FP32 factor broke down or
     * its refinement stalled, and the relative residual of the last refined step */
    uint32_t precision_fallbacks;
    double refinement_residual;           /* NaN when no refined step ran */
    /* camera-step choice (SparseExecutionDecision, kernel/multi_view/sparse/api.rs) */
    int32_t plan_reason;                  /* enum ba_plan_reason */
    uint64_t workspace_bytes;             /* device memory the solve allocated (checked against an
                                             explicit workspace_budget_bytes, native/mod.rs:557-570) */
    /* SolveReport::inlier_count / final_inlier_rmse (solver.rs:585-608, 1353-1371): final residual
     * norms within the robust-loss scale (every retained observation under squared loss) */
    uint64_t inlier_count;
    double final_inlier_rmse;             /* NaN when no observation is within the scale */
    /* SolveDiagnostics (solver.rs:297-308, native/reporting.rs:24-100) */
    int32_t observability;                /* enum ba_observability */
    uint32_t pose_rank, pose_dimension;   /* ObservabilityDiagnostics; 0, 0 when not assessed */
    uint32_t unconstrained_intrinsics;    /* groups held at their input values (SolveReport) */
    uint32_t weak_point_resolves;         /* weak-point re-solves run (at most 4) */
    uint32_t _pad_certify;
} ba_solve_report;

/* SparseExecutionReason (kernel/multi_view/sparse/api.rs) of the camera-step choice. */
enum ba_plan_reason {
    BA_PLAN_EXPLICIT_SOLVER = 0,          /* ExplicitBackend: linear_solver DIRECT or ITERATIVE */
    BA_PLAN_DIRECT_FITS_BUDGET = 1,       /* DirectFitsBudget: AUTO picks the direct step */
    BA_PLAN_DIRECT_EXCEEDS_BUDGET = 2,    /* DirectExceedsBudget: AUTO picks PCG, which fits */
    BA_PLAN_LOWEST_MEMORY = 3,            /* LowestMemoryBackend: neither fits; AUTO picks PCG */
    BA_PLAN_WORKLOAD_PREFERS_RETAINED = 4 /* WorkloadPrefersCachedRetained (calibrated path only):
                                             at least 250,000 observations and D >= 768 */
};

/* The camera-step choice of LinearSolver::Auto (execution_policy.rs:190-227, estimates of
 * workspace/self_calibration.rs:466-535): the direct step when the reference's direct-workspace
 * estimate is at most 4/5 of the planning budget (workspace_budget_bytes, or 1 GiB when it is 0:
 * SparseWorkspaceMemoryBudget::AUTO_BYTES, api.rs:43), else PCG.  The estimates restate
 * estimated_direct_bytes (execution_policy.rs:98-120) with serial direct assembly and
 * estimated_self_calibrating_retained_bytes (self_calibration_retained.rs:233-290) without its
 * thread-count-dependent worker term; they size the reference's CPU workspace, which is what the
 * policy compares, not this library's device memory (ba_solve_report.workspace_bytes). */
typedef struct ba_execution_plan {
    int32_t linear_solver;                /* BA_SOLVER_DIRECT or BA_SOLVER_ITERATIVE */
    int32_t reason;                       /* enum ba_plan_reason */
    uint64_t budget_bytes;                /* planning budget (0 for an explicit solver) */
    uint64_t estimated_direct_bytes;      /* 0 for an explicit solver */
    uint64_t estimated_retained_bytes;    /* 0 for an explicit solver */
    uint64_t reduced_nonzeros;            /* lower-triangle nonzeros of S; 0 when the bounds
                                             D(D+1)/2 and the diagonal blocks decided alone */
    uint32_t reduced_dimension;
    uint32_t _pad;
} ba_execution_plan;

/* The camera step ba_hip_solve would take for `problem` under `config` (host only, no device
 * work).  Same errors as ba_hip_solve's validation and lowering. */
int ba_hip_plan(const ba_problem* problem, const ba_solve_config* config, ba_execution_plan* plan,
                char* message, size_t message_len);

/* Optional per-iteration record: the LM trajectory a parity test compares. */
typedef struct ba_iteration_trace {
    uint32_t capacity;                    /* entries each array can hold */
    uint32_t count;                       /* written by the library */
    double* cost;                         /* cost at the linearisation */
    double* damping;                      /* lambda used by the iteration */
    double* candidate_cost;               /* NaN when no candidate was formed */
    double* predicted_reduction;          /* NaN when none */
    double* gain_ratio;                   /* NaN when none */
    uint8_t* accepted;
    uint32_t* pcg_iterations;
    double* seconds;                      /* wall time from the start of the LM loop to the end
                                             of the iteration */
} ba_iteration_trace;

/* Solution (solver.rs:803-876): caller-allocated buffers with the problem's sizes. */
typedef struct ba_solution {
    double* pose_rotation;                /* n_poses*9 */
    double* pose_center;                  /* n_poses*3 */
    double* intrinsic_focal;              /* n_intrinsics (fx = fy) */
    double* intrinsic_k1;                 /* n_intrinsics */
    double* intrinsic_k2;                 /* n_intrinsics */
    double* point_xyz;                    /* n_points*3 */
    uint8_t* point_quarantined;           /* n_points, or NULL */
    uint8_t* intrinsic_unconstrained;     /* n_intrinsics, or NULL */
    /* n_intrinsics*9, or NULL: each free group's marginal covariance in its local coordinates
     * (log f, k1, k2; row-major 3x3, zero beyond its free dimension), weight-normalised, at the
     * first solve's final state (sparse/self_calibration/mod.rs:110-209); NaN where none was
     * formed.  No counterpart in the reference's Solution, which keeps only the verdict. */
    double* intrinsic_covariance;
} ba_solution;

typedef struct ba_hip_handle ba_hip_handle;

/* BundleAdjuster::name (solver.rs:1563). */
const char* ba_hip_name(void);

/* Create a handle bound to HIP device `device` (its streams and buffers).  The handle is
 * the adjuster object of the reference (NativeBundleAdjuster, native/mod.rs:117-118). */
int ba_hip_create(int device, ba_hip_handle** out);
void ba_hip_destroy(ba_hip_handle* handle);

/* Form of the PCG product with S = U - W V^-1 W^T (no counterpart in the reference, whose
 * retained operator is matrix-free: self_calibration_retained/operator.rs:26-76).  IMPLICIT
 * multiplies matrix-free from the per-observation Jacobians (two passes over the observations
 * per product); EXPLICIT assembles S once per LM iteration in 9 x 9 camera blocks (its diagonal
 * and upper blocks) and multiplies block-sparse.  The explicit system is set up when every
 * intrinsic group has one camera, no camera observes a point twice and (under AUTO) its buffers
 * fit in 40 % of the free device memory; EXPLICIT then uses it in every LM iteration, AUTO (the
 * default) in the iterations that follow one whose PCG took at least 10 iterations (the
 * assembly costs about as much as ten matrix-free products save), and IMPLICIT never.  On
 * another topology every mode is IMPLICIT.  The forms agree to rounding. */
enum ba_schur_product { BA_SCHUR_AUTO = 0, BA_SCHUR_IMPLICIT = 1, BA_SCHUR_EXPLICIT = 2 };
int ba_hip_set_schur_product(ba_hip_handle* handle, int32_t mode);
/* The form of the handle's last iterative solve: EXPLICIT when any of its PCG solves used the
 * explicit system, else IMPLICIT (AUTO before any). */
int32_t ba_hip_last_schur_product(const ba_hip_handle* handle);

/* BundleAdjuster::solve (solver.rs:1572) for the self-calibrating and the calibrated pixel paths
 * (native/mod.rs:260-455, native/self_calibration.rs:223-249, native/calibrated.rs).  Host buffers.
 * `trace` and `message` may be NULL.  On error, `solution` is left unspecified and the
 * problem is untouched (the reference's atomic-error contract, sparse/self_calibration/mod.rs:44-48). */
int ba_hip_solve(ba_hip_handle* handle, const ba_problem* problem, const ba_solve_config* config,
                 ba_solution* solution, ba_solve_report* report, ba_iteration_trace* trace,
                 char* message, size_t message_len);

/* ----------------------------------------------------------------------------------------
 * Lowered ("native") problem: what PreparedSelfCalibratingProblem::try_new builds
 * (native/self_calibration.rs:47-197) and the workspace index maps
 * (kernel/multi_view/sparse/workspace/topology.rs:351-397, workspace/pattern.rs:270-468).
 * ---------------------------------------------------------------------------------------- */
typedef struct ba_native_sizes {
    size_t n_cameras, n_groups, n_points, n_observations;
    size_t dimension;                     /* reduced camera dimension D */
    size_t n_point_parameters;            /* length of the per-point active-parameter list */
    size_t n_pattern_entries;             /* lower-triangle nonzeros of the reduced pattern */
    size_t n_quarantined;                 /* points held out before the solve (quarantine.rs:33-53) */
} ba_native_sizes;

typedef struct ba_native_problem {
    /* state, anchor-first camera order, anchor-frame coordinates */
    double* camera_rotation;              /* n_cameras*9 */
    double* camera_translation;           /* n_cameras*3 */
    double* group_params;                 /* n_groups*3: focal, k1, k2 */
    uint32_t* camera_group;               /* n_cameras */
    uint32_t* camera_source;              /* n_cameras: input pose index of each camera */
    double* point_xyz;                    /* n_points*3 (input points minus the quarantined) */
    uint32_t* point_source;               /* n_points: input point index */
    uint32_t* obs_camera;                 /* n_observations, sorted by (point, input index) */
    uint32_t* obs_source;                 /* n_observations: input observation index */
    double* obs_uv;                       /* n_observations*2, principal point subtracted */
    uint32_t* point_offsets;              /* n_points+1 */
    uint8_t* camera_fixed;                /* n_cameras */
    uint8_t* point_fixed;                 /* n_points */
    /* layout */
    uint32_t* camera_offset;              /* n_cameras */
    uint32_t* camera_dim;                 /* n_cameras: 0, 5 (baseline camera) or 6 */
    uint32_t* group_offset;               /* n_groups */
    uint32_t* group_dim;                  /* n_groups: 0..3 */
    uint32_t* cross_offsets;              /* n_points+1 (workspace/pattern.rs:306-379) */
    uint32_t* point_parameters;           /* n_point_parameters, first-encounter order */
    uint32_t* pattern_col_ptr;            /* dimension+1, lower CSC (pattern.rs:441-468) */
    uint32_t* pattern_row_idx;            /* n_pattern_entries */
    /* gauge: FixedBaseline camera or -1 for Free */
    int32_t gauge_camera;
    int32_t _pad;
    double scene_scale;                   /* optimizer.rs:620-627 at the initial state */
} ba_native_problem;

/* Lower `problem` (host buffers).  Call once with native == NULL to get the sizes, then with
 * buffers of those sizes. */
int ba_hip_prepare(const ba_problem* problem, ba_native_sizes* sizes, ba_native_problem* native,
                   char* message, size_t message_len);

/* ---- per-kernel hooks on device buffers (the tests' parity seams) ----------------------- */

/* Device view of a lowered problem at one state. */
typedef struct ba_device_problem {
    size_t n_cameras, n_groups, n_points, n_observations, dimension;
    const double* camera_rotation;        /* n_cameras*9 */
    const double* camera_translation;     /* n_cameras*3 */
    const double* group_params;           /* n_groups*3 */
    const uint32_t* camera_group;
    const double* point_xyz;              /* n_points*3 */
    const uint32_t* obs_camera;
    const uint32_t* obs_point;            /* n_observations: owning point */
    const double* obs_uv;                 /* n_observations*2 */
    const uint32_t* point_offsets;        /* n_points+1 */
    const uint8_t* point_fixed;           /* n_points */
    const uint32_t* camera_offset;
    const uint32_t* camera_dim;
    const uint32_t* group_offset;
    const uint32_t* group_dim;
    int32_t loss;                         /* enum ba_robust_loss, kernel scale below */
    double loss_scale;                    /* Huber delta / Cauchy kernel scale (scale/sqrt 2) */
    int32_t gauge_camera;                 /* -1 = Free */
    double scene_scale;
    double baseline;                      /* |t| of the gauge camera */
    double basis[6];                      /* S^2 tangent basis rows of the gauge camera */
} ba_device_problem;

/* Kernel (a): per observation, robust-weighted residual r~ (2), point Jacobian J~_X (2x3) and
 * camera-block Jacobian J~_c (2x9 row-major: pose columns 0..5, intrinsic columns 6..8, zero
 * beyond the free dimension of each run), the robust
 * cost and the validity flag.  Reference: schur/point.rs:239-355 and
 * self_calibration/model.rs:26-129,308-324.  Outputs are device arrays of n_observations rows. */
int ba_hip_linearize(ba_hip_handle* handle, const ba_device_problem* problem,
                     double* residual /*2*n*/, double* point_jacobian /*6*n*/,
                     double* camera_jacobian /*18*n*/, double* cost /*n*/, uint8_t* valid /*n*/);

/* Kernel (b): point elimination and the reduced camera system at damping `lambda`:
 * S (dense D x D, lower triangle filled, row-major), reduced gradient g^ (D), unreduced
 * gradient (D), clamped damping diagonal (D), per-point inverse damped block (6), inverse
 * gradient (3), and the scalar outputs (cost, point model reduction, scaled gradient norm,
 * valid flag).  Reference: schur/assembly.rs:893-1069, point_block.rs:85-168. */
typedef struct ba_schur_scalars {
    double cost, squared_residual, point_model_reduction, scaled_gradient_inf_norm;
    double minimum_point_rcond;
    uint32_t weak_point_count, valid;
} ba_schur_scalars;
int ba_hip_schur(ba_hip_handle* handle, const ba_device_problem* problem, double lambda,
                 double* schur /*D*D*/, double* gradient /*D*/, double* full_gradient /*D*/,
                 double* diagonal /*D*/, ba_schur_scalars* scalars /*host*/);

/* Kernel (c): block-Jacobi PCG on the implicit Schur system at damping `lambda`
 * (self_calibration_retained.rs:512-579, pcg.rs:23-119).  Writes the step (D, device) and
 * returns iterations and the true relative residual. */
int ba_hip_pcg(ba_hip_handle* handle, const ba_device_problem* problem, double lambda,
               double relative_tolerance, uint32_t max_iterations, double* step /*D*/,
               uint32_t* iterations, double* relative_residual, int32_t* accepted);

/* Kernel (d): dense Cholesky factor of the n x n SPD matrix `a` (device, row-major) followed by
 * the solve a x = b; `x` (device, n) receives the solution.
 * Fails (BA_ERR_NUMERICAL) on a pivot <= 64 eps max|diag| (dense/linear_solver.rs:9-61).
 * precision: BA_PRECISION_FP64 (lower triangle read), or BA_PRECISION_MIXED: FP32 factor on the
 * matrix cores and FP64 iterative refinement against the full FP64 matrix (both triangles read),
 * failing when the refined relative residual stays above 1e-10. */
int ba_hip_cholesky_solve(ba_hip_handle* handle, size_t n, const double* a, const double* b,
                          double* x, int32_t precision);

/* Allocation-free form of ba_hip_cholesky_solve (same kernels, same results, same errors) for
 * callers that cannot allocate device memory: code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc.  No counterpart in the reference.  Its workspace (the matrix
 * copy or FP32 factor, the factor's scratch and two flags) lives in a fixed device array that each
 * device holds for the life of the process (about 33 MiB), resolved per device by ba_hip_create;
 * the call makes no hipMalloc, hipFree or other allocating call.  n above
 * BA_NOALLOC_MAX_DIMENSION returns BA_ERR_RESOURCE_EXHAUSTED.  Calls on one device share its
 * array: calls that overlap on one device (from several handles or threads) race. */
#define BA_NOALLOC_MAX_DIMENSION 2048
int ba_hip_cholesky_solve_noalloc(ba_hip_handle* handle, size_t n, const double* a, const double* b,
                                  double* x, int32_t precision);

/* ---- Batched form -------------------------------------------------------------------------
 * Many independent windows (local bundle adjustments) solved together with the direct camera
 * step: the result of each window equals ba_hip_solve on that window alone (same per-window LM
 * control: sparse/self_calibration/mod.rs:212-639 for squared or Huber loss).  All windows run in
 * every launch; each window keeps its damping, acceptance and stop tests on the device, so one LM
 * iteration is a fixed launch sequence, captured once in a HIP graph and replayed
 * (BA_BATCH_GRAPH), or issued on the stream each iteration (BA_BATCH_STREAM).  The tile edge of
 * the per-window dense Cholesky (32 or 64) is chosen by timing both on the batch's window sizes.
 * precision MIXED factors Jacobi-scaled FP32 copies with FP64 iterative refinement.  No
 * counterpart entry point in the reference, whose callers solve windows one by one through
 * BundleAdjuster::solve (solver.rs:1562-1573). */
enum ba_batch_mode { BA_BATCH_GRAPH = 0, BA_BATCH_STREAM = 1 };

typedef struct ba_batch_trace {
    uint32_t capacity;                    /* iterations per window the arrays hold (0: none) */
    double* cost;                         /* [iteration * n_windows + window], cost at the linearisation */
    double* damping;                      /* lambda used */
    double* candidate_cost;               /* NaN when no candidate was formed */
    uint8_t* accepted;
    uint32_t tile;                        /* out: tile edge the autotuner chose */
    double seconds_loop;                  /* out: wall time of the LM iterations of all windows */
    double autotune_seconds[2];           /* out: measured per-window Cholesky time, tiles 32 and 64 */
} ba_batch_trace;

/* problems, solutions and reports are arrays of n_windows (reports may be NULL); the config
 * (max_iterations, tolerances, precision) applies to every window; linear_solver ITERATIVE is
 * unsupported.  seconds_iterations of each report is the loop time of the whole batch. */
int ba_hip_solve_batched(ba_hip_handle* handle, size_t n_windows, const ba_problem* problems,
                         const ba_solve_config* config, ba_solution* solutions,
                         ba_solve_report* reports, ba_batch_trace* trace, int32_t mode,
                         char* message, size_t message_len);

/* ---- Differentiation form ------------------------------------------------------------------
 * Robust reprojection cost  C = sum_o rho(|r_o|)  over observations o = (camera c, point p, pixel m)
 * with the radial model of self_calibration/model.rs:26-129 (X_c = exp(w) X + t, n = X_c.xy / X_c.z,
 * r_o = f (1 + k1 |n|^2 + k2 |n|^4) n - m, m already relative to the principal point) and rho of
 * numerics/robust.rs:8-44.  Cameras are 9 values [w(3) angle-axis, t(3), f, k1, k2]; points are
 * [X, Y, Z].  Observations outside the model's projection domain (model.rs:26-75) contribute
 * nothing and are counted.  The backward pass recomputes every Jacobian; its camera rotation
 * gradient is J_l(w)^T times the tangent gradient of the reference's pose motion
 * (schur/point.rs:730-762).  No counterpart entry points in the reference: they expose its
 * objective to automatic differentiation.  Sums run over fixed partitions in a fixed order, so
 * results do not depend on scheduling. */
typedef struct ba_reprojection_topology ba_reprojection_topology;

/* Index the observations once: obs_camera / obs_point are HOST arrays of n_observations entries,
 * checked against n_cameras / n_points.  The topology keeps device copies and the workspace of
 * every later ba_hip_cost_grad call on it. */
int ba_hip_reprojection_topology_create(ba_hip_handle* handle, size_t n_cameras, size_t n_points,
                                        size_t n_observations, const uint32_t* obs_camera,
                                        const uint32_t* obs_point, ba_reprojection_topology** out,
                                        char* message, size_t message_len);
void ba_hip_reprojection_topology_destroy(ba_reprojection_topology* topology);

/* Forward (grad_camera and grad_point NULL) or forward + backward on `stream` (a hipStream_t,
 * NULL for the handle's stream; no host synchronisation when a stream is given).  All pointers
 * are device memory of the element type is_fp32 ? float : double: cameras (9 n_cameras), points
 * (3 n_points), obs_uv (2 n_observations), cost_out (1).  grad_scale (1 element, NULL = 1)
 * multiplies both gradients: dC/dcameras (9 n_cameras) and dC/dpoints (3 n_points).
 * invalid_out (uint32, NULL to skip) receives the number of out-of-domain observations. */
int ba_hip_cost_grad(ba_hip_handle* handle, const ba_reprojection_topology* topology, int32_t is_fp32,
                     const void* cameras, const void* points, const void* obs_uv, int32_t loss,
                     double loss_scale, const void* grad_scale, void* cost_out, uint32_t* invalid_out,
                     void* grad_camera, void* grad_point, void* stream, char* message,
                     size_t message_len);

#ifdef __cplusplus
}
#endif
#endif /* OPENMVG_BA_HIP_H */
