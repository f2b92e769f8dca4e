// ba_cholesky.hip — kernel (d): dense blocked Cholesky of the reduced camera system on gfx950.
//
// Right-looking blocked LL^T on a row-major lower triangle with 64x64 tiles, per diagonal tile k:
//   k_potrf_tile   one wave factors the diagonal tile in LDS, checking every pivot against
//                  64 eps max|diag(A)| as the reference's dense camera solve does
//                  (kernel/multi_view/dense/linear_solver.rs:9-61), and forms L_kk^-1;
//   k_trsm_gemm    the panel below, X = A_ik L_kk^-T, as a 64x64x64 product per tile on the
//                  matrix cores;
//   k_syrk_update  the trailing update A_ij -= L_ik L_jk^T on the matrix cores;
// (v_mfma_f64_16x16x4f64, or v_mfma_f32_16x16x4f32 for the FP32 factor; four waves per 64x64
// tile).  The solves L y = b and L^T x = y run tile by tile in one workgroup: each step is a
// product with the already solved part (coalesced row reads) and one with L_kk^-1.
// The mixed form factors a Jacobi-scaled FP32 copy and recovers the FP64 solution by iterative
// refinement against the FP64 matrix (residuals r = b - S x in FP64).
#include <hip/hip_runtime.h>

#include <cfloat>

#include "ba_kernels.hpp"

#define HIP_OK_OR_IGNORE(expr) (void)(expr)

namespace ba {
namespace {

constexpr int TB = 64;  // tile edge
constexpr int TT = TB * TB;

// Matrix-core step C += A B on 16x16x4 tiles: lane l holds A[l%16][l/16] and B[l/16][l%16];
// element v of the accumulator holds D[row(l/16, v)][l%16] (layouts measured on gfx950).
template <class T>
struct Mfma;
template <>
struct Mfma<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 mac(double a, double b, v4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lk, int v) { return lk + 4 * v; }
    static constexpr double eps = DBL_EPSILON;
};
template <>
struct Mfma<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 mac(float a, float b, v4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lk, int v) { return 4 * lk + v; }
    static constexpr double eps = FLT_EPSILON;
};

template <class T>
__global__ void __launch_bounds__(1024) k_diag_tol(const T* A, uint32_t n, uint32_t ld, double* work, int* fail) {
    __shared__ double sh[1024];
    double m = 0.0;
    for (uint32_t i = threadIdx.x; i < n; i += 1024) m = fmax(m, fabs((double)A[(size_t)i * ld + i]));
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        work[0] = 64.0 * Mfma<T>::eps * fmax(sh[0], DBL_MIN);
        *fail = 0;
    }
}

// Factor the diagonal tile at k0 (size nb <= 64) in LDS with four waves: thread (i, q) owns the
// entries of row i in columns c = q (mod 4).  Right-looking, one column per step; every pivot is
// checked against 64 eps max|diag(A)|.  L_kk is written back, then inverted in place column by
// column from the last (X L = I: X_ij = -(sum_{j<k<=i} X_ik L_kj) / L_jj, the four partial sums
// of a row combined by lane shuffles in a fixed order) and L_kk^-1 written (identity beyond nb).
// Once a pivot fails, *fail is set and later tiles return at once.
template <class T>
__global__ void __launch_bounds__(256) k_potrf_tile(T* A, uint32_t ld, uint32_t k0, uint32_t nb, const double* work,
                                                   int* fail, T* Linv) {
    __shared__ T a[TB][TB + 1];
    if (*fail) return;
    const int t = threadIdx.x, i = t / 4, q = t % 4;
    for (int c = q; c < TB; c += 4)
        a[i][c] = (i < (int)nb && c <= i) ? A[(size_t)(k0 + i) * ld + k0 + c] : (i == c ? T(1) : T(0));
    __syncthreads();
    const T tol = (T)work[0];
    const bool row_in = i < (int)nb;
    for (int j = 0; j < (int)nb; ++j) {
        const T piv = a[j][j];
        if (!(isfinite(piv) && piv > tol)) {
            if (t == 0) *fail = 1;
            return;  // uniform: every thread reads the same pivot
        }
        const T l = sqrt(piv);
        __syncthreads();
        if (q == 0 && row_in && i > j) a[i][j] /= l;
        if (t == 0) a[j][j] = l;
        __syncthreads();
        if (row_in && i > j) {
            const T lij = a[i][j];
            // columns c in (j, i] with c = q (mod 4)
            int c = j + 1 + ((q - (j + 1)) % 4 + 4) % 4;
            for (; c <= i; c += 4) a[i][c] -= lij * a[c][j];
        }
        __syncthreads();
    }
    if (row_in)
        for (int c = q; c <= i; c += 4) A[(size_t)(k0 + i) * ld + k0 + c] = a[i][c];
    for (int j = (int)nb - 1; j >= 0; --j) {
        const T xjj = T(1) / a[j][j];
        T s = T(0);
        if (row_in && i > j) {
            int k = j + 1 + ((q - (j + 1)) % 4 + 4) % 4;
            for (; k <= i; k += 4) s += a[i][k] * a[k][j];
        }
        // lanes 4i .. 4i+3 are consecutive in one wave: fixed-order sum of the four partials
        s += __shfl_xor(s, 1);
        s += __shfl_xor(s, 2);
        __syncthreads();
        if (q == 0 && row_in && i > j) a[i][j] = -s * xjj;
        if (t == 0) a[j][j] = xjj;
        __syncthreads();
    }
    for (int c = q; c < TB; c += 4)
        Linv[i * TB + c] = (row_in && c < (int)nb) ? (c <= i ? a[i][c] : T(0)) : (i == c ? T(1) : T(0));
}

// Panel tile ti (row tile below k0): X = A_ik L_kk^-T on the matrix cores (K = 64), in place.
template <class T>
__global__ void __launch_bounds__(256) k_trsm_gemm(T* A, uint32_t n, uint32_t ld, uint32_t k0, const T* Linv,
                                                   const int* fail) {
    typedef typename Mfma<T>::v4 v4;
    __shared__ T X[TB][TB + 1];
    __shared__ T Li[TB][TB + 1];
    if (*fail) return;
    const uint32_t r0 = k0 + TB * (1 + blockIdx.x);
    const int t = threadIdx.x;
    for (int e = t; e < TT; e += 256) {
        const int r = e / TB, c = e % TB;
        X[r][c] = (r0 + r < n && k0 + c < n) ? A[(size_t)(r0 + r) * ld + k0 + c] : T(0);
        Li[r][c] = Linv[e];
    }
    __syncthreads();
    const int wave = t / 64, lane = t % 64;
    const int qr = (wave / 2) * 32, qc = (wave % 2) * 32;
    const int li = lane % 16, lk = lane / 16;
    v4 acc[2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) acc[a][b] = v4{T(0), T(0), T(0), T(0)};
#pragma unroll
    for (int k = 0; k < TB; k += 4) {
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            const T av = X[qr + 16 * a + li][k + lk];
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                const T bv = Li[qc + 16 * b + li][k + lk];  // B[k][j] = Linv[j][k]
                acc[a][b] = Mfma<T>::mac(av, bv, acc[a][b]);
            }
        }
    }
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int v = 0; v < 4; ++v) {
                const uint32_t r = r0 + qr + 16 * a + Mfma<T>::row(lk, v);
                const uint32_t c = k0 + qc + 16 * b + li;
                if (r < n && c < n) A[(size_t)r * ld + c] = acc[a][b][v];
            }
}

// Trailing update of tile pair (ti, tj), tj <= ti, both after the panel at k0:
// A_ij -= L_ik L_jk^T with K = nb, staged through LDS in K chunks of 32.  Four waves, each a
// 32x32 quadrant of 2x2 16x16x4 matrix-core tiles.
constexpr int KC = 32;
template <class T>
__global__ void __launch_bounds__(256) k_syrk_update(T* A, uint32_t n, uint32_t ld, uint32_t k0, uint32_t nb,
                                                     uint32_t first_tile, const int* fail) {
    typedef typename Mfma<T>::v4 v4;
    __shared__ T Li[TB][KC + 1];
    __shared__ T Lj[TB][KC + 1];
    if (*fail) return;
    // blockIdx.x enumerates the lower-triangular tile pairs (ti, tj) of the trailing matrix
    const uint32_t idx = blockIdx.x;
    uint32_t ti = (uint32_t)((sqrt(8.0 * idx + 1.0) - 1.0) * 0.5);
    while ((ti + 1) * (ti + 2) / 2 <= idx) ++ti;
    while (ti * (ti + 1) / 2 > idx) --ti;
    const uint32_t tj = idx - ti * (ti + 1) / 2;
    const uint32_t r0 = (first_tile + ti) * TB, c0 = (first_tile + tj) * TB;
    const int t = threadIdx.x;
    const int wave = t / 64, lane = t % 64;
    const int qr = (wave / 2) * 32, qc = (wave % 2) * 32;
    const int li = lane % 16, lk = lane / 16;
    v4 acc[2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) acc[a][b] = v4{T(0), T(0), T(0), T(0)};
    for (int kc = 0; kc < (int)nb; kc += KC) {
        __syncthreads();
        for (int e = t; e < TB * KC; e += 256) {
            int r = e / KC, c = e % KC;
            uint32_t ri = r0 + r, rj = c0 + r;
            bool kin = kc + c < (int)nb;
            Li[r][c] = (ri < n && kin) ? A[(size_t)ri * ld + k0 + kc + c] : T(0);
            Lj[r][c] = (rj < n && kin) ? A[(size_t)rj * ld + k0 + kc + c] : T(0);
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < KC; k += 4) {
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                T av = Li[qr + 16 * a + li][k + lk];
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    T bv = Lj[qc + 16 * b + li][k + lk];  // B[k][j] = L_jk[j][k]
                    acc[a][b] = Mfma<T>::mac(av, bv, acc[a][b]);
                }
            }
        }
    }
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int v = 0; v < 4; ++v) {
                uint32_t r = r0 + qr + 16 * a + Mfma<T>::row(lk, v);
                uint32_t c = c0 + qc + 16 * b + li;
                if (r < n && c < n && c <= r) A[(size_t)r * ld + c] -= acc[a][b][v];
            }
}

// Forward solve L y = b (one workgroup; FP64 vectors), left-looking by tiles:
//   y_k = L_kk^-1 (b_k - L_k,<k y_<k).
// Thread (r, g) of 64 x 16 sums row r of tile k against the solved entries j = g, g + 16, ...
// (16 consecutive lanes read one row); the 16 partials are summed in g order.  `skip` (optional)
// turns the solve into a no-op.
template <class T>
__global__ void __launch_bounds__(1024) k_trsv_lower(const T* L, const T* Linv, uint32_t n, uint32_t ld, const double* b,
                                                     double* y, const int* skip) {
    __shared__ double part[16][TB];
    __shared__ double rhs[TB];
    if (skip && *skip) return;
    const int t = threadIdx.x;
    const int sr = t / 16, sg = t % 16;  // summation: 16 consecutive lanes share a row
    const uint32_t ntile = (n + TB - 1) / TB;
    for (uint32_t kt = 0; kt < ntile; ++kt) {
        const uint32_t k0 = kt * TB;
        double s = 0.0;
        if (k0 + sr < n) {  // four independent chains, combined in a fixed order
            const T* Lr = L + (size_t)(k0 + sr) * ld;
            double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
            uint32_t j = sg;
            for (; j + 48 < k0; j += 64) {
                s0 += (double)Lr[j] * y[j];
                s1 += (double)Lr[j + 16] * y[j + 16];
                s2 += (double)Lr[j + 32] * y[j + 32];
                s3 += (double)Lr[j + 48] * y[j + 48];
            }
            for (; j < k0; j += 16) s0 += (double)Lr[j] * y[j];
            s = (s0 + s1) + (s2 + s3);
        }
        part[sg][sr] = s;
        __syncthreads();
        const uint32_t row = k0 + t;
        if (t < TB) {
            double v = row < n ? b[row] : 0.0;
            for (int q = 0; q < 16; ++q) v -= part[q][t];
            rhs[t] = v;
        }
        __syncthreads();
        if (t < TB && row < n) {
            const T* Li = Linv + (size_t)kt * TT + t * TB;
            double v = 0.0;
            for (int c = 0; c <= t; ++c) v += (double)Li[c] * rhs[c];
            y[row] = v;
        }
        __syncthreads();
    }
}

// Backward solve L^T x = y (one workgroup; FP64 vectors), by tiles from the last:
//   x_k = L_kk^-T (y_k - L_>k,k^T x_>k).
// Thread (g, c) of 16 x 64 sums column c of tile k over the solved rows i = k0 + 64 + g + 16 m.
template <class T>
__global__ void __launch_bounds__(1024) k_trsv_upper(const T* L, const T* Linv, uint32_t n, uint32_t ld, const double* y,
                                                     double* x, const int* skip) {
    __shared__ double part[16][TB];
    __shared__ double rhs[TB];
    if (skip && *skip) return;
    const int t = threadIdx.x, c = t % TB, g = t / TB;
    const uint32_t ntile = (n + TB - 1) / TB;
    for (int kt = (int)ntile - 1; kt >= 0; --kt) {
        const uint32_t k0 = (uint32_t)kt * TB;
        const uint32_t col = k0 + c;
        double s = 0.0;
        if (col < n) {  // four independent chains, combined in a fixed order
            double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
            uint32_t i = k0 + TB + g;
            for (; i + 48 < n; i += 64) {
                s0 += (double)L[(size_t)i * ld + col] * x[i];
                s1 += (double)L[(size_t)(i + 16) * ld + col] * x[i + 16];
                s2 += (double)L[(size_t)(i + 32) * ld + col] * x[i + 32];
                s3 += (double)L[(size_t)(i + 48) * ld + col] * x[i + 48];
            }
            for (; i < n; i += 16) s0 += (double)L[(size_t)i * ld + col] * x[i];
            s = (s0 + s1) + (s2 + s3);
        }
        part[g][c] = s;
        __syncthreads();
        if (t < TB) {
            double v = col < n ? y[col] : 0.0;
            for (int q = 0; q < 16; ++q) v -= part[q][c];
            rhs[c] = v;
        }
        __syncthreads();
        if (t < TB && col < n) {  // x_c = sum_r Linv[r][c] rhs[r], r >= c
            const T* Li = Linv + (size_t)kt * TT;
            double v = 0.0;
            for (int rr = c; rr < TB; ++rr) v += (double)Li[rr * TB + c] * rhs[rr];
            x[col] = v;
        }
        __syncthreads();
    }
}

// Iterative refinement: r = b - S x from the FP64 matrix (one wave per row, fixed lane order),
// then the convergence test (one workgroup): stop when ||r|| <= 1e-11 ||b|| or when ||r|| no
// longer halves; state[0] = done flag, state[1] = ||r||^2 of the last step, state[2] = ||b||^2.
__global__ void __launch_bounds__(256) k_refine_residual(const double* S, uint32_t n, uint32_t ld, const double* b,
                                                         const double* x, double* r, const int* done) {
    if (*done) return;
    const uint32_t row = blockIdx.x * 4 + threadIdx.x / 64;
    const int lane = threadIdx.x % 64;
    if (row >= n) return;
    const double* Sr = S + (size_t)row * ld;
    double s = 0.0;
    for (uint32_t c = lane; c < n; c += 64) s += Sr[c] * x[c];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if (lane == 0) r[row] = b[row] - s;
}
__global__ void __launch_bounds__(1024) k_refine_check(uint32_t n, const double* b, const double* r, double* state,
                                                       int* done, int first) {
    __shared__ double sh[2][1024];
    if (*done) return;
    double rr = 0.0, bb = 0.0;
    for (uint32_t i = threadIdx.x; i < n; i += 1024) {
        rr += r[i] * r[i];
        bb += b[i] * b[i];
    }
    sh[0][threadIdx.x] = rr;
    sh[1][threadIdx.x] = bb;
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + s];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    const double r2 = sh[0][0], b2 = sh[1][0];
    const bool small = r2 <= 1e-22 * b2;  // ||r|| <= 1e-11 ||b||
    const bool stalled = !first && !(r2 < 0.25 * state[1]);
    if (small || stalled || !isfinite(r2)) *done = 1;
    if (!stalled || first) state[1] = r2;
    state[2] = b2;
}
__global__ void __launch_bounds__(256) k_refine_axpy(uint32_t n, const double* dx, double* x, const int* done) {
    if (*done) return;
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) x[i] += dx[i];
}
// Jacobi scaling of the FP32 copy: F = D^-1/2 S D^-1/2 (unit diagonal), scale[i] = S_ii^-1/2 (1
// when S_ii is not positive; the factor then reports the breakdown).  The FP32 factor of the
// scaled matrix keeps the units of pose, intrinsic and gauge columns apart from its rounding.
__global__ void __launch_bounds__(256) k_jacobi_scale(const double* S, uint32_t n, uint32_t ld, double* scale) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const double d = S[(size_t)i * ld + i];
    scale[i] = (isfinite(d) && d > 0.0) ? 1.0 / sqrt(d) : 1.0;
}
__global__ void __launch_bounds__(256) k_lower_to_f32(const double* S, uint32_t n, uint32_t ld, const double* scale,
                                                      float* F) {
    for (uint32_t row = blockIdx.y; row < n; row += gridDim.y)
        for (uint32_t c = blockIdx.x * 256 + threadIdx.x; c <= row && c < n; c += 256 * gridDim.x)
            F[(size_t)row * ld + c] = (float)(S[(size_t)row * ld + c] * scale[row] * scale[c]);
}
// y = scale * v (elementwise), skipped once `done` is set
__global__ void __launch_bounds__(256) k_scale_vec(uint32_t n, const double* scale, const double* v, double* y,
                                                   const int* done) {
    if (done && *done) return;
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) y[i] = scale[i] * v[i];
}

// Workspace (doubles): [0, 16) scalars | tmp n | r n | dx n | scale n | sv n | L_kk^-1 tiles.
inline size_t linv_offset(uint32_t n) { return 16 + 5 * (size_t)n; }

template <class T>
void factor(hipStream_t s, T* A, uint32_t n, uint32_t ld, double* work, int* fail) {
    hipLaunchKernelGGL(k_diag_tol<T>, dim3(1), dim3(1024), 0, s, A, n, ld, work, fail);
    T* Linv = reinterpret_cast<T*>(work + linv_offset(n));
    const uint32_t ntile = (n + TB - 1) / TB;
    for (uint32_t kt = 0; kt < ntile; ++kt) {
        const uint32_t k0 = kt * TB;
        const uint32_t nb = n - k0 < TB ? n - k0 : TB;
        hipLaunchKernelGGL(k_potrf_tile<T>, dim3(1), dim3(256), 0, s, A, ld, k0, nb, (const double*)work, fail,
                           Linv + (size_t)kt * TT);
        const uint32_t rest = ntile - kt - 1;
        if (rest == 0) break;
        hipLaunchKernelGGL(k_trsm_gemm<T>, dim3(rest), dim3(256), 0, s, A, n, ld, k0, (const T*)(Linv + (size_t)kt * TT),
                           (const int*)fail);
        const uint32_t pairs = rest * (rest + 1) / 2;
        hipLaunchKernelGGL(k_syrk_update<T>, dim3(pairs), dim3(256), 0, s, A, n, ld, k0, nb, kt + 1, (const int*)fail);
    }
}
template <class T>
void apply(hipStream_t s, const T* L, uint32_t n, uint32_t ld, const double* b, double* x, double* work, const int* skip) {
    const T* Linv = reinterpret_cast<const T*>(work + linv_offset(n));
    double* tmp = work + 16;
    hipLaunchKernelGGL(k_trsv_lower<T>, dim3(1), dim3(1024), 0, s, L, Linv, n, ld, b, tmp, skip);
    hipLaunchKernelGGL(k_trsv_upper<T>, dim3(1), dim3(1024), 0, s, L, Linv, n, ld, (const double*)tmp, x, skip);
}

}  // namespace

size_t cholesky_work_doubles(uint32_t n) {
    const size_t ntile = (n + TB - 1) / TB;
    return linv_offset(n) + ntile * TT;
}

void launch_cholesky_solve(hipStream_t s, double* A, uint32_t n, uint32_t ld, const double* b, double* x, double* work,
                           int* fail) {
    if (n == 0) return;
    factor<double>(s, A, n, ld, work, fail);
    apply<double>(s, A, n, ld, b, x, work, (const int*)fail);
}

void launch_cholesky_solve_mixed(hipStream_t s, const double* S, float* F, uint32_t n, uint32_t ld, const double* b,
                                 double* x, double* work, int* fail, int* done, int max_steps) {
    if (n == 0) return;
    double* r = work + 16 + n;
    double* dx = r + n;
    double* scale = dx + n;
    double* sv = scale + n;
    const unsigned nb = (n + 255) / 256;
    hipLaunchKernelGGL(k_jacobi_scale, dim3(nb), dim3(256), 0, s, S, n, ld, scale);
    hipLaunchKernelGGL(k_lower_to_f32, dim3(nb, n < 65535u ? n : 65535u), dim3(256), 0, s, S, n, ld,
                       (const double*)scale, F);
    factor<float>(s, F, n, ld, work, fail);
    HIP_OK_OR_IGNORE(hipMemsetAsync(done, 0, sizeof(int), s));
    // x = D^-1/2 F^-1 D^-1/2 b
    hipLaunchKernelGGL(k_scale_vec, dim3(nb), dim3(256), 0, s, n, (const double*)scale, b, sv, (const int*)nullptr);
    apply<float>(s, F, n, ld, sv, dx, work, (const int*)fail);
    hipLaunchKernelGGL(k_scale_vec, dim3(nb), dim3(256), 0, s, n, (const double*)scale, (const double*)dx, x,
                       (const int*)nullptr);
    const unsigned rows = (n + 3) / 4;
    for (int k = 0; k < max_steps; ++k) {
        hipLaunchKernelGGL(k_refine_residual, dim3(rows), dim3(256), 0, s, S, n, ld, b, (const double*)x, r,
                           (const int*)done);
        hipLaunchKernelGGL(k_refine_check, dim3(1), dim3(1024), 0, s, n, b, (const double*)r, work + 8, done, (int)(k == 0));
        hipLaunchKernelGGL(k_scale_vec, dim3(nb), dim3(256), 0, s, n, (const double*)scale, (const double*)r, sv,
                           (const int*)done);
        apply<float>(s, F, n, ld, sv, dx, work, done);
        hipLaunchKernelGGL(k_scale_vec, dim3(nb), dim3(256), 0, s, n, (const double*)scale, (const double*)dx, dx,
                           (const int*)done);
        hipLaunchKernelGGL(k_refine_axpy, dim3(nb), dim3(256), 0, s, n, (const double*)dx, x, (const int*)done);
    }
}

}  // namespace ba
