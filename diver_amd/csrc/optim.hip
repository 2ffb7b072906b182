// One-pass Adam update for the trainable grid (and any dense FP32 parameter).
//
// The reference trains with torch.optim.Adam(params, lr, weight_decay) (train.py:39-47): L2 weight
// decay folded into the gradient, no amsgrad.  Per element, for step t:
//   g' = g + wd * p
//   m  = b1 m + (1 - b1) g'
//   v  = b2 v + (1 - b2) g'^2
//   p  = p - step_size * m / (sqrt(v) / bc2_sqrt + eps),  step_size = lr / (1 - b1^t),
//                                                          bc2_sqrt = sqrt(1 - b2^t)
// The caller computes step_size and bc2_sqrt for t.  The pass reads p, g, m, v once and writes p, m,
// v once (28 B per parameter), and can clear g in the same pass (4 B more, no separate memset), so it
// runs at the HBM bound: 15.2 GB per step for a 257^3 x 32 grid.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.h"

namespace diver {

// omb1, omb2 = 1 - beta, rounded to FP32 once from the double betas (as torch's update does)
struct AdamScalars {
    float b1, b2, omb1, omb2, eps, wd, step_size, bc2_sqrt;
};

// Every product-sum is an explicit fmaf, so no kernel's code layout can change how the compiler
// contracts them: the dense and row kernels compute the same bits for the same element.
__device__ __forceinline__ void adam_one(float& p, float& g, float& m, float& v, const AdamScalars& s) {
    const float gw = s.wd != 0.0f ? fmaf(s.wd, p, g) : g;
    m = fmaf(s.b1, m, s.omb1 * gw);
    v = fmaf(s.b2, v, (s.omb2 * gw) * gw);
    const float denom = sqrtf(v) / s.bc2_sqrt + s.eps;
    p = fmaf(-s.step_size, m / denom, p);
}

typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void adam4(f4& p, const f4& g, f4& m, f4& v, const AdamScalars& s) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        float pp = p[k], gg = g[k], mm = m[k], vv = v[k];
        adam_one(pp, gg, mm, vv, s);
        p[k] = pp;
        m[k] = mm;
        v[k] = vv;
    }
}

// Grid-stride over float4 groups (VEC: every array 16-byte aligned) or single elements; with VEC,
// elements past the last full group are done one by one by the first threads.
template <bool CLEAR, bool VEC>
__global__ void __launch_bounds__(256) adam_kernel(float* __restrict__ p, float* __restrict__ g,
                                                   float* __restrict__ m, float* __restrict__ v,
                                                   int64_t n, AdamScalars s) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if constexpr (!VEC) {
        for (int64_t i = t0; i < n; i += stride) {
            float pp = p[i], gg = g[i], mm = m[i], vv = v[i];
            adam_one(pp, gg, mm, vv, s);
            p[i] = pp;
            m[i] = mm;
            v[i] = vv;
            if (CLEAR) g[i] = 0.0f;
        }
        return;
    }
    const int64_t n4 = n >> 2;
    f4* P = reinterpret_cast<f4*>(p);
    f4* G = reinterpret_cast<f4*>(g);
    f4* M = reinterpret_cast<f4*>(m);
    f4* V = reinterpret_cast<f4*>(v);
    const f4 zero = {0.f, 0.f, 0.f, 0.f};
    // two groups per iteration (8 loads in flight per thread); every value is used once, so loads
    // and stores are non-temporal
    for (int64_t i = t0; i < n4; i += 2 * stride) {
        const int64_t j = i + stride;
        const bool two = j < n4;
        f4 pp0 = __builtin_nontemporal_load(P + i), gg0 = __builtin_nontemporal_load(G + i);
        f4 mm0 = __builtin_nontemporal_load(M + i), vv0 = __builtin_nontemporal_load(V + i);
        f4 pp1 = pp0, gg1 = gg0, mm1 = mm0, vv1 = vv0;
        if (two) {
            pp1 = __builtin_nontemporal_load(P + j);
            gg1 = __builtin_nontemporal_load(G + j);
            mm1 = __builtin_nontemporal_load(M + j);
            vv1 = __builtin_nontemporal_load(V + j);
        }
        adam4(pp0, gg0, mm0, vv0, s);
        __builtin_nontemporal_store(pp0, P + i);
        __builtin_nontemporal_store(mm0, M + i);
        __builtin_nontemporal_store(vv0, V + i);
        if (CLEAR) __builtin_nontemporal_store(zero, G + i);
        if (two) {
            adam4(pp1, gg1, mm1, vv1, s);
            __builtin_nontemporal_store(pp1, P + j);
            __builtin_nontemporal_store(mm1, M + j);
            __builtin_nontemporal_store(vv1, V + j);
            if (CLEAR) __builtin_nontemporal_store(zero, G + j);
        }
    }
    const int64_t tail = n4 * 4 + t0;
    if (tail < n) {
        float pp = p[tail], gg = g[tail], mm = m[tail], vv = v[tail];
        adam_one(pp, gg, mm, vv, s);
        p[tail] = pp;
        m[tail] = mm;
        v[tail] = vv;
        if (CLEAR) g[tail] = 0.0f;
    }
}

// The same update over listed rows only: row r covers elements [r * row_len, (r + 1) * row_len)
// (row_len % 4 == 0, arrays 16-byte aligned), one float4 group per thread.  Rows outside
// [0, n / row_len) are skipped.
template <bool CLEAR>
__global__ void __launch_bounds__(256) adam_rows_kernel(float* __restrict__ p, float* __restrict__ g,
                                                        float* __restrict__ m, float* __restrict__ v,
                                                        int64_t n_rows_total,
                                                        const int64_t* __restrict__ rows,
                                                        int64_t n_rows, int q, AdamScalars s) {
    const int64_t units = n_rows * q;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    f4* P = reinterpret_cast<f4*>(p);
    f4* G = reinterpret_cast<f4*>(g);
    f4* M = reinterpret_cast<f4*>(m);
    f4* V = reinterpret_cast<f4*>(v);
    const f4 zero = {0.f, 0.f, 0.f, 0.f};
    for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < units; u += stride) {
        const int64_t k = u / q;
        const int c = (int)(u - k * q);
        const int64_t r = rows[k];
        if (r < 0 || r >= n_rows_total) continue;
        const int64_t i = r * q + c;
        f4 pp = __builtin_nontemporal_load(P + i), gg = __builtin_nontemporal_load(G + i);
        f4 mm = __builtin_nontemporal_load(M + i), vv = __builtin_nontemporal_load(V + i);
        adam4(pp, gg, mm, vv, s);
        __builtin_nontemporal_store(pp, P + i);
        __builtin_nontemporal_store(mm, M + i);
        __builtin_nontemporal_store(vv, V + i);
        if (CLEAR) __builtin_nontemporal_store(zero, G + i);
    }
}

hipError_t launch_adam_rows(float* p, float* g, float* m, float* v, int64_t n, const int64_t* rows,
                            int64_t n_rows, int row_len, double b1, double b2, float eps,
                            float step_size, float bc2_sqrt, bool clear, hipStream_t st) {
    if (n_rows <= 0) return hipSuccess;
    const AdamScalars s{(float)b1, (float)b2, (float)(1.0 - b1), (float)(1.0 - b2), eps, 0.0f,
                        step_size, bc2_sqrt};
    const int q = row_len / 4;
    const int64_t units = n_rows * q;
    int64_t blocks = (units + 255) / 256;
    if (blocks > 256 * 32) blocks = 256 * 32;
    const dim3 grid((unsigned)blocks), block(256);
    if (clear)
        hipLaunchKernelGGL((adam_rows_kernel<true>), grid, block, 0, st, p, g, m, v, n / row_len,
                           rows, n_rows, q, s);
    else
        hipLaunchKernelGGL((adam_rows_kernel<false>), grid, block, 0, st, p, g, m, v, n / row_len,
                           rows, n_rows, q, s);
    return hipGetLastError();
}

hipError_t launch_adam(float* p, float* g, float* m, float* v, int64_t n, double b1, double b2,
                       float eps, float wd, float step_size, float bc2_sqrt, bool clear,
                       hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const AdamScalars s{(float)b1, (float)b2, (float)(1.0 - b1), (float)(1.0 - b2), eps, wd,
                        step_size, bc2_sqrt};
    const bool vec = ((reinterpret_cast<uintptr_t>(p) | reinterpret_cast<uintptr_t>(g) |
                       reinterpret_cast<uintptr_t>(m) | reinterpret_cast<uintptr_t>(v)) & 15) == 0;
    const int64_t units = vec ? (n >> 2) : n;
    // up to 32 workgroups per CU; the grid-stride loop gives each thread several units
    int64_t blocks = (units + 255) / 256;
    if (blocks > 256 * 32) blocks = 256 * 32;
    if (blocks < 1) blocks = 1;
    const dim3 grid((unsigned)blocks), block(256);
    if (clear && vec) hipLaunchKernelGGL((adam_kernel<true, true>), grid, block, 0, st, p, g, m, v, n, s);
    if (clear && !vec) hipLaunchKernelGGL((adam_kernel<true, false>), grid, block, 0, st, p, g, m, v, n, s);
    if (!clear && vec) hipLaunchKernelGGL((adam_kernel<false, true>), grid, block, 0, st, p, g, m, v, n, s);
    if (!clear && !vec) hipLaunchKernelGGL((adam_kernel<false, false>), grid, block, 0, st, p, g, m, v, n, s);
    return hipGetLastError();
}

}  // namespace diver
