// ba_autodiff.hip — robust reprojection cost and its hand-written gradient (gfx950), FP32 and FP64.
//
//   k_cg_prep     one lane per camera: R = exp(w) (numerics/manifold/so3.rs:254-265) and the first
//                 radial turn (self_calibration/model.rs:266-292), packed as 16 values
//   k_cg_obs      one lane per observation, input order: residual and robust cost
//                 (model.rs:26-129, numerics/robust.rs:8-44); with gradients, dC/dX_c = J_p^T w r is
//                 split into the point part R^T e and the camera part [e, RX x e, df, dk1, dk2]
//                 (pose motion of schur/point.rs:730-762) and written per observation
//   k_cg_points   one lane per point: fixed-order sum of its observations' point parts
//   k_cg_cameras  one workgroup per camera: fixed-order tree sum of its observations' camera parts,
//                 then the rotation gradient J_l(w)^T g_omega
//   k_cg_final    one workgroup: fixed-order sum of the per-block cost partials
// No floating-point atomics; every sum has a fixed partition and order for a given topology.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ba_hip.h"
#include "ba_handle.hpp"
#include "ba_math.hpp"

struct ba_reprojection_topology {
    int device = 0;
    size_t nc = 0, np = 0, no = 0;
    uint32_t* obs_cam = nullptr;  // input order
    uint32_t* obs_pt = nullptr;
    uint32_t* pt_off = nullptr;  // np + 1: observations of each point, ascending input index
    uint32_t* pt_obs = nullptr;
    uint32_t* cam_off = nullptr;  // nc + 1: observations of each camera, ascending input index
    uint32_t* cam_obs = nullptr;
    double* prep = nullptr;      // 16 per camera (holds floats in the FP32 form)
    double* pcontrib = nullptr;  // 3 per observation
    double* ccontrib = nullptr;  // 9 per observation
    double* cost_part = nullptr;
    uint32_t* invalid_part = nullptr;
    uint32_t blocks = 0;  // fixed partition of the observation pass
};

namespace ba {
namespace {

constexpr int NT = 256;
constexpr uint32_t OBS_BLOCKS = 2048;  // fixed partition of the per-observation pass

struct HipFail {
    hipError_t code;
    const char* what;
};
#define CG_OK(expr)                                  \
    do {                                             \
        hipError_t _e = (expr);                      \
        if (_e != hipSuccess) throw HipFail{_e, #expr}; \
    } while (0)

// Left Jacobian of SO(3), J_l(w) = I + A [w]x + B [w]x^2, A = (1 - cos t)/t^2, B = (t - sin t)/t^3,
// with series for small t (both forms are the textbook expansion; truncation below 1e-17).
template <class T>
__device__ void left_jacobian(const T w[3], T J[9]) {
    T t2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
    T th = sqrt(t2);
    const T small = sizeof(T) == 4 ? T(0.1) : T(1e-2);
    T A, B;
    if (th < small) {
        A = T(0.5) - t2 / T(24) + t2 * t2 / T(720);
        B = T(1) / T(6) - t2 / T(120) + t2 * t2 / T(5040);
    } else {
        T s = sin(T(0.5) * th);
        A = T(2) * s * s / t2;
        B = (th - sin(th)) / (t2 * th);
    }
    T K[9] = {T(0), -w[2], w[1], w[2], T(0), -w[0], -w[1], w[0], T(0)};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            T k2 = K[3 * i] * K[j] + K[3 * i + 1] * K[3 + j] + K[3 * i + 2] * K[6 + j];
            J[3 * i + j] = (i == j ? T(1) : T(0)) + A * K[3 * i + j] + B * k2;
        }
}

template <class T>
__global__ void __launch_bounds__(NT) k_cg_prep(uint32_t nc, const T* __restrict__ cams, T* __restrict__ prep) {
    uint32_t c = blockIdx.x * NT + threadIdx.x;
    if (c >= nc) return;
    const T* q = cams + 9 * (size_t)c;
    T w[3] = {q[0], q[1], q[2]};
    T R[9];
    exp_so3(w, R);
    T* o = prep + 16 * (size_t)c;
    for (int k = 0; k < 9; ++k) o[k] = R[k];
    o[9] = q[3];
    o[10] = q[4];
    o[11] = q[5];
    o[12] = q[6];
    o[13] = q[7];
    o[14] = q[8];
    T root;
    o[15] = first_turn(q[7], q[8], &root) ? root : T(INFINITY);
}

template <class T, bool GRAD>
__global__ void __launch_bounds__(NT) k_cg_obs(uint64_t no, const T* __restrict__ prep, const T* __restrict__ pts,
                                               const uint32_t* __restrict__ obs_cam,
                                               const uint32_t* __restrict__ obs_pt, const T* __restrict__ uv,
                                               int loss, T scale, double* __restrict__ cost_part,
                                               uint32_t* __restrict__ invalid_part, T* __restrict__ pc,
                                               T* __restrict__ cc) {
    __shared__ double sh[NT];
    __shared__ uint32_t shi[NT];
    double cost = 0.0;
    uint32_t invalid = 0;
    const uint64_t stride = (uint64_t)gridDim.x * NT;
    for (uint64_t i = (uint64_t)blockIdx.x * NT + threadIdx.x; i < no; i += stride) {
        const uint32_t c = obs_cam[i], p = obs_pt[i];
        const T* cp = prep + 16 * (size_t)c;
        T R[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) R[k] = cp[k];
        const T f = cp[12], k1 = cp[13], k2 = cp[14], turn = cp[15];
        const T X[3] = {pts[3 * (size_t)p], pts[3 * (size_t)p + 1], pts[3 * (size_t)p + 2]};
        T RX[3];
        rotate(R, X, RX);
        const T Xc[3] = {RX[0] + cp[9], RX[1] + cp[10], RX[2] + cp[11]};
        const T u = uv[2 * i], v = uv[2 * i + 1];
        T res[2];
        if constexpr (!GRAD) {
            if (!radial_residual(f, k1, k2, turn, true, Xc, u, v, res)) {
                ++invalid;
                continue;
            }
            cost += (double)loss_cost(loss, scale, sqrt(res[0] * res[0] + res[1] * res[1]));
        } else {
            T J[2][3], Ji[2][3];
            T* po = pc + 3 * (size_t)i;
            T* co = cc + 9 * (size_t)i;
            if (!radial_linearize(f, k1, k2, turn, true, Xc, u, v, res, J, Ji)) {
                ++invalid;
#pragma unroll
                for (int k = 0; k < 3; ++k) po[k] = T(0);
#pragma unroll
                for (int k = 0; k < 9; ++k) co[k] = T(0);
                continue;
            }
            const T nrm = sqrt(res[0] * res[0] + res[1] * res[1]);
            cost += (double)loss_cost(loss, scale, nrm);
            const T sw = loss_sqrt_weight(loss, scale, nrm);
            const T wgt = sw * sw;
            const T psi0 = wgt * res[0], psi1 = wgt * res[1];
            // e = dC/dX_c
            const T e[3] = {J[0][0] * psi0 + J[1][0] * psi1, J[0][1] * psi0 + J[1][1] * psi1,
                            J[0][2] * psi0 + J[1][2] * psi1};
            T gx[3];
            rotate_t(R, e, gx);
            po[0] = gx[0];
            po[1] = gx[1];
            po[2] = gx[2];
            // camera part: d/dt = e, tangent rotation d/domega = RX x e, intrinsics
            co[0] = RX[1] * e[2] - RX[2] * e[1];
            co[1] = RX[2] * e[0] - RX[0] * e[2];
            co[2] = RX[0] * e[1] - RX[1] * e[0];
            co[3] = e[0];
            co[4] = e[1];
            co[5] = e[2];
            co[6] = (Ji[0][0] * psi0 + Ji[1][0] * psi1) / f;  // Ji column 0 is d/d(log f)
            co[7] = Ji[0][1] * psi0 + Ji[1][1] * psi1;
            co[8] = Ji[0][2] * psi0 + Ji[1][2] * psi1;
        }
    }
    sh[threadIdx.x] = cost;
    shi[threadIdx.x] = invalid;
    __syncthreads();
    for (int s = NT / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sh[threadIdx.x] += sh[threadIdx.x + s];
            shi[threadIdx.x] += shi[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        cost_part[blockIdx.x] = sh[0];
        invalid_part[blockIdx.x] = shi[0];
    }
}

template <class T>
__global__ void __launch_bounds__(NT) k_cg_points(uint32_t np, const uint32_t* __restrict__ off,
                                                  const uint32_t* __restrict__ list, const T* __restrict__ pc,
                                                  const T* __restrict__ gscale, T* __restrict__ out) {
    uint32_t p = blockIdx.x * NT + threadIdx.x;
    if (p >= np) return;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (uint32_t k = off[p]; k < off[p + 1]; ++k) {
        const T* q = pc + 3 * (size_t)list[k];
        s0 += (double)q[0];
        s1 += (double)q[1];
        s2 += (double)q[2];
    }
    const double g = gscale ? (double)gscale[0] : 1.0;
    out[3 * (size_t)p] = (T)(g * s0);
    out[3 * (size_t)p + 1] = (T)(g * s1);
    out[3 * (size_t)p + 2] = (T)(g * s2);
}

template <class T>
__global__ void __launch_bounds__(NT) k_cg_cameras(const uint32_t* __restrict__ off, const uint32_t* __restrict__ list,
                                                   const T* __restrict__ cc, const T* __restrict__ cams,
                                                   const T* __restrict__ gscale, T* __restrict__ out) {
    __shared__ double sh[9][NT];
    const uint32_t c = blockIdx.x;
    double acc[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) acc[k] = 0.0;
    for (uint32_t k = off[c] + threadIdx.x; k < off[c + 1]; k += NT) {
        const T* q = cc + 9 * (size_t)list[k];
#pragma unroll
        for (int j = 0; j < 9; ++j) acc[j] += (double)q[j];
    }
#pragma unroll
    for (int j = 0; j < 9; ++j) sh[j][threadIdx.x] = acc[j];
    __syncthreads();
    for (int s = NT / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
#pragma unroll
            for (int j = 0; j < 9; ++j) sh[j][threadIdx.x] += sh[j][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    const double g = gscale ? (double)gscale[0] : 1.0;
    const T* q = cams + 9 * (size_t)c;
    T w[3] = {q[0], q[1], q[2]};
    T Jl[9];
    left_jacobian(w, Jl);
    double gw[3];
    for (int j = 0; j < 3; ++j)  // J_l^T g_omega
        gw[j] = (double)Jl[j] * sh[0][0] + (double)Jl[3 + j] * sh[1][0] + (double)Jl[6 + j] * sh[2][0];
    T* o = out + 9 * (size_t)c;
    for (int j = 0; j < 3; ++j) o[j] = (T)(g * gw[j]);
    for (int j = 3; j < 9; ++j) o[j] = (T)(g * sh[j][0]);
}

template <class T>
__global__ void __launch_bounds__(1024) k_cg_final(uint32_t nb, const double* __restrict__ part,
                                                   const uint32_t* __restrict__ ipart, T* __restrict__ cost,
                                                   uint32_t* __restrict__ invalid) {
    __shared__ double sh[1024];
    __shared__ uint32_t shi[1024];
    double s = 0.0;
    uint32_t n = 0;
    for (uint32_t b = threadIdx.x; b < nb; b += 1024) {
        s += part[b];
        n += ipart[b];
    }
    sh[threadIdx.x] = s;
    shi[threadIdx.x] = n;
    __syncthreads();
    for (int k = 512; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) {
            sh[threadIdx.x] += sh[threadIdx.x + k];
            shi[threadIdx.x] += shi[threadIdx.x + k];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        cost[0] = (T)sh[0];
        if (invalid) invalid[0] = shi[0];
    }
}

inline unsigned grid_for(size_t n, int nt) { return (unsigned)((n + nt - 1) / nt); }

template <class T>
void run(const ba_reprojection_topology* tp, hipStream_t s, const void* cams, const void* pts, const void* uv,
         int loss, double scale, const void* gscale, void* cost, uint32_t* invalid, void* gcam, void* gpt) {
    const T* C = static_cast<const T*>(cams);
    T* prep = reinterpret_cast<T*>(tp->prep);
    hipLaunchKernelGGL(k_cg_prep<T>, dim3(grid_for(tp->nc, NT)), dim3(NT), 0, s, (uint32_t)tp->nc, C, prep);
    const bool grad = gcam != nullptr;
    T* pc = reinterpret_cast<T*>(tp->pcontrib);
    T* cc = reinterpret_cast<T*>(tp->ccontrib);
    if (grad)
        hipLaunchKernelGGL((k_cg_obs<T, true>), dim3(tp->blocks), dim3(NT), 0, s, (uint64_t)tp->no, prep,
                           static_cast<const T*>(pts), tp->obs_cam, tp->obs_pt, static_cast<const T*>(uv), loss,
                           (T)scale, tp->cost_part, tp->invalid_part, pc, cc);
    else
        hipLaunchKernelGGL((k_cg_obs<T, false>), dim3(tp->blocks), dim3(NT), 0, s, (uint64_t)tp->no, prep,
                           static_cast<const T*>(pts), tp->obs_cam, tp->obs_pt, static_cast<const T*>(uv), loss,
                           (T)scale, tp->cost_part, tp->invalid_part, pc, cc);
    hipLaunchKernelGGL(k_cg_final<T>, dim3(1), dim3(1024), 0, s, tp->blocks, tp->cost_part, tp->invalid_part,
                       static_cast<T*>(cost), invalid);
    if (!grad) return;
    const T* g = static_cast<const T*>(gscale);
    hipLaunchKernelGGL(k_cg_points<T>, dim3(grid_for(tp->np, NT)), dim3(NT), 0, s, (uint32_t)tp->np, tp->pt_off,
                       tp->pt_obs, pc, g, static_cast<T*>(gpt));
    hipLaunchKernelGGL(k_cg_cameras<T>, dim3((unsigned)tp->nc), dim3(NT), 0, s, tp->cam_off, tp->cam_obs, cc, C, g,
                       static_cast<T*>(gcam));
}

int report(int code, const std::string& m, char* msg, size_t len) {
    if (msg && len) std::snprintf(msg, len, "%s", m.c_str());
    return code;
}

// Stable bucket order of observations by key: offsets (n_keys + 1) and indices ascending per key.
void bucket(const uint32_t* key, size_t n, size_t nk, std::vector<uint32_t>& off, std::vector<uint32_t>& list) {
    off.assign(nk + 1, 0);
    for (size_t i = 0; i < n; ++i) off[key[i] + 1]++;
    for (size_t k = 0; k < nk; ++k) off[k + 1] += off[k];
    list.assign(n, 0);
    std::vector<uint32_t> fill(off.begin(), off.end() - 1);
    for (size_t i = 0; i < n; ++i) list[fill[key[i]]++] = (uint32_t)i;
}

void free_topology(ba_reprojection_topology* t) {
    if (!t) return;
    (void)hipSetDevice(t->device);
    for (void* p : {(void*)t->obs_cam, (void*)t->obs_pt, (void*)t->pt_off, (void*)t->pt_obs, (void*)t->cam_off,
                    (void*)t->cam_obs, (void*)t->prep, (void*)t->pcontrib, (void*)t->ccontrib, (void*)t->cost_part,
                    (void*)t->invalid_part})
        if (p) (void)hipFree(p);
    delete t;
}

}  // namespace
}  // namespace ba

using namespace ba;

extern "C" {

int ba_hip_reprojection_topology_create(ba_hip_handle* h, size_t nc, size_t np, size_t no, const uint32_t* obs_camera,
                                        const uint32_t* obs_point, ba_reprojection_topology** out, char* msg,
                                        size_t len) {
    if (!h || !out) return report(BA_ERR_INVALID_CONFIG, "null argument", msg, len);
    *out = nullptr;
    if (nc == 0 || np == 0 || no == 0) return report(BA_ERR_INVALID_PROBLEM, "empty cameras, points or observations", msg, len);
    if (!obs_camera || !obs_point) return report(BA_ERR_INVALID_PROBLEM, "null observation index", msg, len);
    if (no >= (size_t)UINT32_MAX || nc >= (size_t)UINT32_MAX || np >= (size_t)UINT32_MAX)
        return report(BA_ERR_UNSUPPORTED, "more than 2^32 - 1 cameras, points or observations", msg, len);
    for (size_t i = 0; i < no; ++i) {
        if (obs_camera[i] >= nc) return report(BA_ERR_INVALID_PROBLEM, "observation camera out of bounds", msg, len);
        if (obs_point[i] >= np) return report(BA_ERR_INVALID_PROBLEM, "observation point out of bounds", msg, len);
    }
    std::vector<uint32_t> pt_off, pt_obs, cam_off, cam_obs;
    bucket(obs_point, no, np, pt_off, pt_obs);
    bucket(obs_camera, no, nc, cam_off, cam_obs);
    auto* t = new ba_reprojection_topology();
    t->device = h->device;
    t->nc = nc;
    t->np = np;
    t->no = no;
    t->blocks = (uint32_t)std::min<size_t>(OBS_BLOCKS, (no + NT - 1) / NT);
    try {
        CG_OK(hipSetDevice(h->device));
        auto up = [&](uint32_t** dst, const uint32_t* src, size_t n) {
            CG_OK(hipMalloc((void**)dst, n * sizeof(uint32_t)));
            CG_OK(hipMemcpy(*dst, src, n * sizeof(uint32_t), hipMemcpyHostToDevice));
        };
        up(&t->obs_cam, obs_camera, no);
        up(&t->obs_pt, obs_point, no);
        up(&t->pt_off, pt_off.data(), np + 1);
        up(&t->pt_obs, pt_obs.data(), no);
        up(&t->cam_off, cam_off.data(), nc + 1);
        up(&t->cam_obs, cam_obs.data(), no);
        CG_OK(hipMalloc((void**)&t->prep, 16 * nc * sizeof(double)));
        CG_OK(hipMalloc((void**)&t->pcontrib, 3 * no * sizeof(double)));
        CG_OK(hipMalloc((void**)&t->ccontrib, 9 * no * sizeof(double)));
        CG_OK(hipMalloc((void**)&t->cost_part, t->blocks * sizeof(double)));
        CG_OK(hipMalloc((void**)&t->invalid_part, t->blocks * sizeof(uint32_t)));
    } catch (const HipFail& e) {
        free_topology(t);
        if (e.code == hipErrorOutOfMemory)
            return report(BA_ERR_RESOURCE_EXHAUSTED, "device memory for the topology", msg, len);
        return report(BA_ERR_INTERNAL, std::string("HIP error: ") + hipGetErrorString(e.code) + " in " + e.what, msg,
                      len);
    }
    *out = t;
    return BA_OK;
}

void ba_hip_reprojection_topology_destroy(ba_reprojection_topology* t) { free_topology(t); }

int ba_hip_cost_grad(ba_hip_handle* h, const ba_reprojection_topology* t, int32_t is_fp32, const void* cams,
                     const void* pts, const void* uv, int32_t loss, double loss_scale, const void* gscale, void* cost,
                     uint32_t* invalid, void* gcam, void* gpt, void* stream, char* msg, size_t len) {
    if (!h || !t || !cams || !pts || !uv || !cost) return report(BA_ERR_INVALID_CONFIG, "null argument", msg, len);
    if ((gcam == nullptr) != (gpt == nullptr))
        return report(BA_ERR_INVALID_CONFIG, "grad_camera and grad_point must both be set or both be NULL", msg, len);
    if (loss != BA_LOSS_SQUARED && loss != BA_LOSS_HUBER && loss != BA_LOSS_CAUCHY)
        return report(BA_ERR_INVALID_CONFIG, "loss", msg, len);
    if (loss != BA_LOSS_SQUARED && !(std::isfinite(loss_scale) && loss_scale > 0.0))
        return report(BA_ERR_INVALID_CONFIG, "loss scale must be finite and positive", msg, len);
    if (t->device != h->device) return report(BA_ERR_INVALID_CONFIG, "topology belongs to another device", msg, len);
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : h->stream;
    try {
        CG_OK(hipSetDevice(h->device));
        if (is_fp32)
            run<float>(t, s, cams, pts, uv, loss, loss_scale, gscale, cost, invalid, gcam, gpt);
        else
            run<double>(t, s, cams, pts, uv, loss, loss_scale, gscale, cost, invalid, gcam, gpt);
        CG_OK(hipGetLastError());
        if (!stream) CG_OK(hipStreamSynchronize(s));
    } catch (const HipFail& e) {
        return report(BA_ERR_INTERNAL, std::string("HIP error: ") + hipGetErrorString(e.code) + " in " + e.what, msg,
                      len);
    }
    return BA_OK;
}

}  // extern "C"
