// ba_cholesky.hip — kernel (d): dense blocked Cholesky of the reduced camera system on gfx950.
//
// Right-looking blocked LL^T on a row-major lower triangle with 64x64 tiles:
//   k_potrf_tile   one workgroup factors the diagonal tile in LDS, checking every pivot
//                  against 64 eps max|diag(A)| as the reference's dense camera solve does
//                  (kernel/multi_view/dense/linear_solver.rs:9-61);
//   k_trsm_panel   one wave per row tile solves X L^T = A against the factored diagonal tile;
//   k_syrk_update  the trailing update A_ij -= L_ik L_jk^T on the FP64 matrix cores
//                  (v_mfma_f64_16x16x4f64), four waves per 64x64 tile;
// then blocked forward and backward substitution in one workgroup each.
#include <hip/hip_runtime.h>

#include <cfloat>

#include "ba_kernels.hpp"

namespace ba {
namespace {

constexpr int TB = 64;  // tile edge
typedef double d4 __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(1024) k_diag_tol(const double* A, uint32_t n, uint32_t ld, double* work, int* fail) {
    __shared__ double sh[1024];
    double m = 0.0;
    for (uint32_t i = threadIdx.x; i < n; i += 1024) m = fmax(m, fabs(A[(size_t)i * ld + i]));
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        work[0] = 64.0 * DBL_EPSILON * fmax(sh[0], DBL_MIN);
        *fail = 0;
    }
}

// Factor the diagonal tile at k0 (size nb <= 64) in LDS, right-looking, column by column.
__global__ void __launch_bounds__(256) k_potrf_tile(double* A, uint32_t ld, uint32_t k0, uint32_t nb, const double* work,
                                                   int* fail) {
    __shared__ double a[TB][TB + 1];
    const int t = threadIdx.x;
    for (int idx = t; idx < TB * TB; idx += 256) {
        int r = idx / TB, c = idx % TB;
        a[r][c] = (r < (int)nb && c <= r) ? A[(size_t)(k0 + r) * ld + k0 + c] : 0.0;
    }
    __syncthreads();
    const double tol = work[0];
    __shared__ int bad;
    if (t == 0) bad = 0;
    __syncthreads();
    for (int j = 0; j < (int)nb; ++j) {
        double piv = a[j][j];
        if (!(isfinite(piv) && piv > tol)) {
            if (t == 0) bad = 1;
            break;  // uniform: every thread reads the same pivot
        }
        double l = sqrt(piv);
        __syncthreads();
        if (t == 0) a[j][j] = l;
        for (int i = j + 1 + t; i < (int)nb; i += 256) a[i][j] /= l;
        __syncthreads();
        // trailing update of the lower triangle below column j
        const int m = (int)nb - j - 1;
        const int tot = m * (m + 1) / 2;
        for (int idx = t; idx < tot; idx += 256) {
            // idx -> (i, c) with j < c <= i < nb
            int ii = (int)((sqrt(8.0 * idx + 1.0) - 1.0) * 0.5);
            while ((ii + 1) * (ii + 2) / 2 <= idx) ++ii;
            while (ii * (ii + 1) / 2 > idx) --ii;
            int cc = idx - ii * (ii + 1) / 2;
            int i = j + 1 + ii, c = j + 1 + cc;
            a[i][c] -= a[i][j] * a[c][j];
        }
        __syncthreads();
    }
    __syncthreads();
    if (bad) {
        if (t == 0) *fail = 1;
        return;
    }
    for (int idx = t; idx < TB * TB; idx += 256) {
        int r = idx / TB, c = idx % TB;
        if (r < (int)nb && c <= r) A[(size_t)(k0 + r) * ld + k0 + c] = a[r][c];
    }
}

// Rows of tile `ti` (ti > k tile) below the factored diagonal tile: X L_kk^T = A_ik.  One lane
// per row; the substitution is unrolled over the 64 columns so the row stays in registers.
__global__ void __launch_bounds__(64) k_trsm_panel(double* A, uint32_t n, uint32_t ld, uint32_t k0, uint32_t nb) {
    __shared__ double L[TB][TB + 1];
    const int t = threadIdx.x;
    for (int idx = t; idx < TB * TB; idx += 64) {
        int r = idx / TB, c = idx % TB;
        L[r][c] = (r < (int)nb && c <= r) ? A[(size_t)(k0 + r) * ld + k0 + c] : (r == c ? 1.0 : 0.0);
    }
    __syncthreads();
    const uint32_t row = k0 + nb + blockIdx.x * TB + t;
    if (row >= n) return;
    double x[TB];
    double* Ar = A + (size_t)row * ld + k0;
#pragma unroll
    for (int c = 0; c < TB; ++c) x[c] = c < (int)nb ? Ar[c] : 0.0;
#pragma unroll
    for (int c = 0; c < TB; ++c) {
        double v = x[c];
#pragma unroll
        for (int m = 0; m < c; ++m) v -= x[m] * L[c][m];
        x[c] = v / L[c][c];
    }
#pragma unroll
    for (int c = 0; c < TB; ++c)
        if (c < (int)nb) Ar[c] = x[c];
}

// Trailing update of tile pair (ti, tj), tj <= ti, both after the panel at k0:
// A_ij -= L_ik L_jk^T with K = nb, staged through LDS in K chunks of 32.  Four waves, each a
// 32x32 quadrant of 2x2 MFMA 16x16x4 FP64 tiles.
constexpr int KC = 32;
__global__ void __launch_bounds__(256) k_syrk_update(double* A, uint32_t n, uint32_t ld, uint32_t k0, uint32_t nb,
                                                     uint32_t first_tile) {
    __shared__ double Li[TB][KC + 1];
    __shared__ double Lj[TB][KC + 1];
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
    d4 acc[2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) acc[a][b] = d4{0.0, 0.0, 0.0, 0.0};
    for (int kc = 0; kc < (int)nb; kc += KC) {
        __syncthreads();
        for (int e = t; e < TB * KC; e += 256) {
            int r = e / KC, c = e % KC;
            uint32_t ri = r0 + r, rj = c0 + r;
            bool kin = kc + c < (int)nb;
            Li[r][c] = (ri < n && kin) ? A[(size_t)ri * ld + k0 + kc + c] : 0.0;
            Lj[r][c] = (rj < n && kin) ? A[(size_t)rj * ld + k0 + kc + c] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < KC; k += 4) {
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                double av = Li[qr + 16 * a + li][k + lk];
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    double bv = Lj[qc + 16 * b + li][k + lk];  // B[k][j] = L_jk[j][k]
                    acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc[a][b], 0, 0, 0);
                }
            }
        }
    }
    // v_mfma_f64_16x16x4f64 leaves D[lane/16 + 4v][lane%16] in accumulator element v
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int v = 0; v < 4; ++v) {
                uint32_t r = r0 + qr + 16 * a + lk + 4 * v;
                uint32_t c = c0 + qc + 16 * b + li;
                if (r < n && c < n && c <= r) A[(size_t)r * ld + c] -= acc[a][b][v];
            }
}

// Forward substitution L y = b (one workgroup, tile by tile).
__global__ void __launch_bounds__(1024) k_trsv_lower(const double* L, uint32_t n, uint32_t ld, const double* b, double* y) {
    __shared__ double yt[TB];
    const int t = threadIdx.x;
    for (uint32_t i = t; i < n; i += 1024) y[i] = b[i];
    __syncthreads();
    for (uint32_t k0 = 0; k0 < n; k0 += TB) {
        const uint32_t nb = n - k0 < TB ? n - k0 : TB;
        if (t < 64) {  // one wave solves the diagonal tile column by column
            double v = t < (int)nb ? y[k0 + t] : 0.0;
            for (int c = 0; c < (int)nb; ++c) {
                double yc = __shfl(v, c) / L[(size_t)(k0 + c) * ld + k0 + c];
                if (t == c) v = yc;
                if (t > c && t < (int)nb) v -= L[(size_t)(k0 + t) * ld + k0 + c] * yc;
            }
            if (t < (int)nb) {
                yt[t] = v;
                y[k0 + t] = v;
            }
        }
        __syncthreads();
        for (uint32_t r = k0 + nb + t; r < n; r += 1024) {
            const double* Lr = L + (size_t)r * ld + k0;
            double s = 0.0;
            for (int m = 0; m < (int)nb; ++m) s += Lr[m] * yt[m];
            y[r] -= s;
        }
        __syncthreads();
    }
}

// Backward substitution L^T x = y (one workgroup).
__global__ void __launch_bounds__(1024) k_trsv_upper(const double* L, uint32_t n, uint32_t ld, const double* y, double* x) {
    __shared__ double xt[TB];
    const int t = threadIdx.x;
    for (uint32_t i = t; i < n; i += 1024) x[i] = y[i];
    __syncthreads();
    if (n == 0) return;
    uint32_t ntile = (n + TB - 1) / TB;
    for (int kt = (int)ntile - 1; kt >= 0; --kt) {
        const uint32_t k0 = (uint32_t)kt * TB;
        const uint32_t nb = n - k0 < TB ? n - k0 : TB;
        if (t < 64) {
            double v = t < (int)nb ? x[k0 + t] : 0.0;
            for (int c = (int)nb - 1; c >= 0; --c) {
                double xc = __shfl(v, c) / L[(size_t)(k0 + c) * ld + k0 + c];
                if (t == c) v = xc;
                if (t < c) v -= L[(size_t)(k0 + c) * ld + k0 + t] * xc;
            }
            if (t < (int)nb) {
                xt[t] = v;
                x[k0 + t] = v;
            }
        }
        __syncthreads();
        for (uint32_t r = t; r < k0; r += 1024) {
            double s = 0.0;
            for (int m = 0; m < (int)nb; ++m) s += L[(size_t)(k0 + m) * ld + r] * xt[m];
            x[r] -= s;
        }
        __syncthreads();
    }
}

}  // namespace

size_t cholesky_work_doubles(uint32_t n) { return 8 + (size_t)n; }

void launch_cholesky_solve(hipStream_t s, double* A, uint32_t n, uint32_t ld, const double* b, double* x, double* work,
                           int* fail) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_diag_tol, dim3(1), dim3(1024), 0, s, A, n, ld, work, fail);
    const uint32_t ntile = (n + TB - 1) / TB;
    for (uint32_t kt = 0; kt < ntile; ++kt) {
        const uint32_t k0 = kt * TB;
        const uint32_t nb = n - k0 < TB ? n - k0 : TB;
        hipLaunchKernelGGL(k_potrf_tile, dim3(1), dim3(256), 0, s, A, ld, k0, nb, work, fail);
        const uint32_t rest = ntile - kt - 1;
        if (rest == 0) break;
        hipLaunchKernelGGL(k_trsm_panel, dim3(rest), dim3(64), 0, s, A, n, ld, k0, nb);
        const uint32_t pairs = rest * (rest + 1) / 2;
        hipLaunchKernelGGL(k_syrk_update, dim3(pairs), dim3(256), 0, s, A, n, ld, k0, nb, kt + 1);
    }
    double* y = work + 8;
    hipLaunchKernelGGL(k_trsv_lower, dim3(1), dim3(1024), 0, s, A, n, ld, b, y);
    hipLaunchKernelGGL(k_trsv_upper, dim3(1), dim3(1024), 0, s, A, n, ld, y, x);
}

}  // namespace ba
