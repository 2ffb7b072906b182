// Closed-form segment integral of the trilinear feature field over one voxel
// (utils/integrator.py:62-113) and its adjoint (scatter-add into the grid).
//
// Layout: voxels [N+1][N+1][N+1][C] fp32, z-major, channel fastest; coords [K][3] (x, y, z) in grid
// units.  A row is handled by C/4 lanes, each owning 4 channels (one 16-byte load per corner),
// so the 8 corner reads of a row are whole 128-byte lines at C=32; the adjoint's atomic adds go
// one lane per channel, whole rows per instruction.
#include "common.h"
#include "kernels.h"

namespace diver {

struct CornerWeights {
    float w[8];
    int cx, cy, cz;
};

// Corner weights of the mean of the trilinear interpolant over [p_in, p_out] inside the voxel
// min(int(p_in+1e-4), int(p_out+1e-4)).  Same polynomial as integrator.py:92-111, arranged as
// in utils/mlp_evaluation/src/mlp_eval.cu:83-101 (corner j = (j&1, (j>>1)&1, j>>2)).
__device__ __forceinline__ CornerWeights corner_weights(float x0, float y0, float z0, float x1,
                                                        float y1, float z1) {
    CornerWeights r;
    {
#pragma clang fp contract(off)
        int ax = (int)(x0 + 1e-4f), bx = (int)(x1 + 1e-4f);
        int ay = (int)(y0 + 1e-4f), by = (int)(y1 + 1e-4f);
        int az = (int)(z0 + 1e-4f), bz = (int)(z1 + 1e-4f);
        r.cx = ax < bx ? ax : bx;
        r.cy = ay < by ? ay : by;
        r.cz = az < bz ? az : bz;
    }
    x0 -= (float)r.cx;
    x1 -= (float)r.cx;
    y0 -= (float)r.cy;
    y1 -= (float)r.cy;
    z0 -= (float)r.cz;
    z1 -= (float)r.cz;
    float z01 = z0 + z1, y01 = y0 + y1, x01 = x0 + x1;
    float w7 = (2.0f * x0 * y0 * z0 + 2.0f * x1 * y1 * z1 + x01 * y01 * z01) * (1.0f / 12.0f);
    float w2 = (y01 * z01 + y0 * z0 + y1 * z1) * (1.0f / 6.0f);
    float w5 = (x01 * z01 + x0 * z0 + x1 * z1) * (1.0f / 6.0f) - w7;
    float w4 = (x01 * y01 + x0 * y0 + x1 * y1) * (1.0f / 6.0f);
    x01 *= 0.5f;
    y01 *= 0.5f;
    z01 *= 0.5f;
    r.w[7] = w7;
    r.w[5] = w5;
    r.w[4] = z01 - w5 - w2;
    r.w[0] = 1.0f - r.w[4] - y01 - x01 + w4;
    r.w[6] = w2 - w7;
    r.w[2] = y01 - r.w[6] - w4;
    r.w[1] = x01 - w5 - w4;
    r.w[3] = w4 - w7;
    return r;
}

template <int LANES>  // lanes per row (lanes_for(C))
__global__ void __launch_bounds__(256) integrate_kernel(const float* __restrict__ voxels, int n1,
                                                        int C, int64_t K,
                                                        const float* __restrict__ ci,
                                                        const float* __restrict__ co,
                                                        float* __restrict__ out) {
    int64_t gt = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t row = gt / LANES;
    int sub = (int)(gt % LANES);
    if (row >= K) return;
    CornerWeights cw = corner_weights(ci[3 * row], ci[3 * row + 1], ci[3 * row + 2], co[3 * row],
                                      co[3 * row + 1], co[3 * row + 2]);
    const int64_t sz = (int64_t)n1 * n1 * C, sy = (int64_t)n1 * C, sx = C;
    // the upper corner is clamped to the grid as in integrator.py:79; the lower one only guards
    // malformed coordinates (valid ones lie in [0, N])
    int xs[2] = {min(max(cw.cx, 0), n1 - 1), min(max(cw.cx + 1, 0), n1 - 1)};
    int ys[2] = {min(max(cw.cy, 0), n1 - 1), min(max(cw.cy + 1, 0), n1 - 1)};
    int zs[2] = {min(max(cw.cz, 0), n1 - 1), min(max(cw.cz + 1, 0), n1 - 1)};
    for (int c4 = sub; c4 < C / 4; c4 += LANES) {
        float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            int64_t base = zs[j >> 2] * sz + ys[(j >> 1) & 1] * sy + (int64_t)xs[j & 1] * sx;
            float4 f = *reinterpret_cast<const float4*>(voxels + base + 4 * c4);
            float w = cw.w[j];
            acc.x += w * f.x;
            acc.y += w * f.y;
            acc.z += w * f.z;
            acc.w += w * f.w;
        }
        *reinterpret_cast<float4*>(out + row * C + 4 * c4) = acc;
    }
}

// The adjoint scatters with one lane per channel (LPR lanes per row: C rounded up to a power of
// two, at most 64; channels past 64 loop), so each atomic instruction of a wave adds whole corner
// rows -- 64 / LPR rows of C contiguous floats -- where a float4-per-lane layout would touch a
// 16-byte piece of eight rows per instruction (four times the cache-line requests for the same
// adds; the atomics bound this kernel).
template <int LPR>
__global__ void __launch_bounds__(256) integrate_backward_kernel(int n1, int C, int64_t K,
                                                                 const float* __restrict__ ci,
                                                                 const float* __restrict__ co,
                                                                 const float* __restrict__ gout,
                                                                 float* __restrict__ gvox) {
    const int64_t gt = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t row = gt / LPR;
    const int sub = (int)(gt % LPR);
    if (row >= K) return;
    const CornerWeights cw = corner_weights(ci[3 * row], ci[3 * row + 1], ci[3 * row + 2],
                                            co[3 * row], co[3 * row + 1], co[3 * row + 2]);
    const int64_t sz = (int64_t)n1 * n1 * C, sy = (int64_t)n1 * C, sx = C;
    // the upper corner is clamped to the grid as in integrator.py:79; the lower one only guards
    // malformed coordinates (valid ones lie in [0, N])
    const int xs[2] = {min(max(cw.cx, 0), n1 - 1), min(max(cw.cx + 1, 0), n1 - 1)};
    const int ys[2] = {min(max(cw.cy, 0), n1 - 1), min(max(cw.cy + 1, 0), n1 - 1)};
    const int zs[2] = {min(max(cw.cz, 0), n1 - 1), min(max(cw.cz + 1, 0), n1 - 1)};
    for (int c = sub; c < C; c += LPR) {
        const float g = gout[row * C + c];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int64_t base = zs[j >> 2] * sz + ys[(j >> 1) & 1] * sy + (int64_t)xs[j & 1] * sx;
            unsafeAtomicAdd(gvox + base + c, cw.w[j] * g);
        }
    }
}

// ---- implicit grid (integrator.py:3-60): features come from a table of distinct corners -----

// The 8 corners of each row's integral as keys z n1^2 + y n1 + x (corner j = (j&1, (j>>1)&1,
// j>>2), clamped as integrate_kernel does).  The implicit stage evaluates its MLP once per
// distinct key and integrates from that table.
__global__ void __launch_bounds__(256) corner_keys_kernel(int n1, int64_t K,
                                                          const float* __restrict__ ci,
                                                          const float* __restrict__ co,
                                                          int64_t* __restrict__ keys) {
    const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= K) return;
    const CornerWeights cw = corner_weights(ci[3 * row], ci[3 * row + 1], ci[3 * row + 2],
                                            co[3 * row], co[3 * row + 1], co[3 * row + 2]);
    const int xs[2] = {min(max(cw.cx, 0), n1 - 1), min(max(cw.cx + 1, 0), n1 - 1)};
    const int ys[2] = {min(max(cw.cy, 0), n1 - 1), min(max(cw.cy + 1, 0), n1 - 1)};
    const int zs[2] = {min(max(cw.cz, 0), n1 - 1), min(max(cw.cz + 1, 0), n1 - 1)};
#pragma unroll
    for (int j = 0; j < 8; ++j)
        keys[8 * row + j] = ((int64_t)zs[j >> 2] * n1 + ys[(j >> 1) & 1]) * n1 + xs[j & 1];
}

// integrate_kernel with corner j of row r read from table row crow[8 r + j] ([U][C]).
template <int LANES>
__global__ void __launch_bounds__(256) integrate_table_kernel(const float* __restrict__ table,
                                                              int C, int64_t K,
                                                              const int64_t* __restrict__ crow,
                                                              const float* __restrict__ ci,
                                                              const float* __restrict__ co,
                                                              float* __restrict__ out) {
    const int64_t gt = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t row = gt / LANES;
    const int sub = (int)(gt % LANES);
    if (row >= K) return;
    const CornerWeights cw = corner_weights(ci[3 * row], ci[3 * row + 1], ci[3 * row + 2],
                                            co[3 * row], co[3 * row + 1], co[3 * row + 2]);
    for (int c4 = sub; c4 < C / 4; c4 += LANES) {
        float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const float4 f = *reinterpret_cast<const float4*>(table + crow[8 * row + j] * C + 4 * c4);
            const float w = cw.w[j];
            acc.x += w * f.x;
            acc.y += w * f.y;
            acc.z += w * f.z;
            acc.w += w * f.w;
        }
        *reinterpret_cast<float4*>(out + row * C + 4 * c4) = acc;
    }
}

// integrate_backward_kernel with corner j of row r scattered into table row crow[8 r + j] (one
// lane per channel, as there).
template <int LPR>
__global__ void __launch_bounds__(256) integrate_table_backward_kernel(
    int C, int64_t K, const int64_t* __restrict__ crow, const float* __restrict__ ci,
    const float* __restrict__ co, const float* __restrict__ gout, float* __restrict__ gtable) {
    const int64_t gt = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t row = gt / LPR;
    const int sub = (int)(gt % LPR);
    if (row >= K) return;
    const CornerWeights cw = corner_weights(ci[3 * row], ci[3 * row + 1], ci[3 * row + 2],
                                            co[3 * row], co[3 * row + 1], co[3 * row + 2]);
    for (int c = sub; c < C; c += LPR) {
        const float g = gout[row * C + c];
#pragma unroll
        for (int j = 0; j < 8; ++j)
            unsafeAtomicAdd(gtable + crow[8 * row + j] * C + c, cw.w[j] * g);
    }
}

hipError_t launch_corner_keys(int n1, int64_t K, const float* ci, const float* co, int64_t* keys,
                              hipStream_t s) {
    if (K == 0) return hipSuccess;
    hipLaunchKernelGGL(corner_keys_kernel, dim3((unsigned)((K + 255) / 256)), dim3(256), 0, s, n1,
                       K, ci, co, keys);
    return hipGetLastError();
}

// Lanes per row: C/4 float4 groups rounded up to a power of two (at most 16), so the grid of
// K*lanes threads matches the kernel's row = thread / LANES for every C % 4 == 0 (C = 12 gives 4
// lanes, one of them idle; C = 48 gives 16 lanes, each taking 3 groups).
static int lanes_for(int C) {
    int lanes = 1;
    while (lanes < C / 4 && lanes < 16) lanes *= 2;
    return lanes;
}

#define DIVER_LANES_LAUNCH(KERNEL, ...)                                                        \
    do {                                                                                       \
        const int lanes = lanes_for(C);                                                        \
        const int64_t threads = K * lanes;                                                     \
        const dim3 grid((unsigned)((threads + 255) / 256));                                    \
        switch (lanes) {                                                                       \
            case 1: hipLaunchKernelGGL(KERNEL<1>, grid, dim3(256), 0, s, __VA_ARGS__); break;  \
            case 2: hipLaunchKernelGGL(KERNEL<2>, grid, dim3(256), 0, s, __VA_ARGS__); break;  \
            case 4: hipLaunchKernelGGL(KERNEL<4>, grid, dim3(256), 0, s, __VA_ARGS__); break;  \
            case 8: hipLaunchKernelGGL(KERNEL<8>, grid, dim3(256), 0, s, __VA_ARGS__); break;  \
            default: hipLaunchKernelGGL(KERNEL<16>, grid, dim3(256), 0, s, __VA_ARGS__); break; \
        }                                                                                      \
    } while (0)

hipError_t launch_integrate_table(const float* table, int C, int64_t K, const int64_t* crow,
                                  const float* ci, const float* co, float* out, hipStream_t s) {
    if (K == 0) return hipSuccess;
    DIVER_LANES_LAUNCH(integrate_table_kernel, table, C, K, crow, ci, co, out);
    return hipGetLastError();
}

// Lanes per row of the scatters: C rounded up to a power of two, at most 64 (C = 12 gives 16
// lanes, 4 of them idle; C = 96 gives 64 lanes, some taking two channels).
static int scatter_lanes_for(int C) {
    int lanes = 1;
    while (lanes < C && lanes < 64) lanes *= 2;
    return lanes;
}

#define DIVER_SCATTER_LAUNCH(KERNEL, ...)                                                      \
    do {                                                                                       \
        const int lanes = scatter_lanes_for(C);                                                \
        const int64_t threads = K * lanes;                                                     \
        const dim3 grid((unsigned)((threads + 255) / 256));                                    \
        switch (lanes) {                                                                       \
            case 1: hipLaunchKernelGGL(KERNEL<1>, grid, dim3(256), 0, s, __VA_ARGS__); break;  \
            case 2: hipLaunchKernelGGL(KERNEL<2>, grid, dim3(256), 0, s, __VA_ARGS__); break;  \
            case 4: hipLaunchKernelGGL(KERNEL<4>, grid, dim3(256), 0, s, __VA_ARGS__); break;  \
            case 8: hipLaunchKernelGGL(KERNEL<8>, grid, dim3(256), 0, s, __VA_ARGS__); break;  \
            case 16: hipLaunchKernelGGL(KERNEL<16>, grid, dim3(256), 0, s, __VA_ARGS__); break; \
            case 32: hipLaunchKernelGGL(KERNEL<32>, grid, dim3(256), 0, s, __VA_ARGS__); break; \
            default: hipLaunchKernelGGL(KERNEL<64>, grid, dim3(256), 0, s, __VA_ARGS__); break; \
        }                                                                                      \
    } while (0)

hipError_t launch_integrate_table_backward(int C, int64_t K, const int64_t* crow, const float* ci,
                                           const float* co, const float* gout, float* gtable,
                                           hipStream_t s) {
    if (K == 0) return hipSuccess;
    DIVER_SCATTER_LAUNCH(integrate_table_backward_kernel, C, K, crow, ci, co, gout, gtable);
    return hipGetLastError();
}

hipError_t launch_integrate(const float* voxels, int n1, int C, int64_t K, const float* ci,
                            const float* co, float* out, hipStream_t s) {
    if (K == 0) return hipSuccess;
    DIVER_LANES_LAUNCH(integrate_kernel, voxels, n1, C, K, ci, co, out);
    return hipGetLastError();
}

hipError_t launch_integrate_backward(int n1, int C, int64_t K, const float* ci, const float* co,
                                     const float* gout, float* gvox, hipStream_t s) {
    if (K == 0) return hipSuccess;
    DIVER_SCATTER_LAUNCH(integrate_backward_kernel, n1, C, K, ci, co, gout, gvox);
    return hipGetLastError();
}

#undef DIVER_LANES_LAUNCH
#undef DIVER_SCATTER_LAUNCH

}  // namespace diver
