// Recompute pass of the render adjoint.
//
// The forward fused kernel keeps no per-interval samples: only (rgb, T) and the number of
// intervals composited per ray.  The backward re-walks each ray (same walk, same FP32 integral,
// same compensated transmittance, so the same values) and writes one row per composited
// interval at the ray's row offset:
//   [p_in (3), p_out (3), s = w_s.F + b_s, T_before, alpha, F (C)]
// The decoder and compositing adjoints then run as batched GEMMs over the rows, and the grid
// gradient is scattered by the segment-integral adjoint.
#include "common.h"
#include "decoder.h"
#include "kernels.h"
#include "raystate.h"

namespace diver {

// One wave per workgroup; lanes [0, rpw) take consecutive rays (rays_per_wave: small batches
// spread over more waves).
template <class D, bool RT>
__global__ void __launch_bounds__(64) rows_kernel(RenderArgs a, const int32_t* __restrict__ count,
                                                  const int64_t* __restrict__ row_off,
                                                  float* __restrict__ rows, int rpw) {
    static_assert(D::NPRE == 0, "rows are written for decoders without pre-layers");
    constexpr int C = D::C;
    constexpr int RW = 9 + C;
    const int64_t r = (int64_t)blockIdx.x * rpw + threadIdx.x;
    if ((int)threadIdx.x >= rpw || r >= a.n_rays) return;
    const int n = a.n;
    RayWalk ray;
    if constexpr (RT)
        ray.init_rt(a.o[3 * r], a.o[3 * r + 1], a.o[3 * r + 2], a.d[3 * r], a.d[3 * r + 1],
                    a.d[3 * r + 2], a.xyzmin, a.voxel_size, n, a.octrees, true);
    else
        ray.init(a.o[3 * r], a.o[3 * r + 1], a.o[3 * r + 2], a.d[3 * r], a.d[3 * r + 1],
                 a.d[3 * r + 2], a.xyzmin, a.voxel_size, (float)n, true);
    const int todo = count[r];
    float* out = rows + row_off[r] * RW;
    Kahan S;
    float T = 1.0f;
    // occupancy through the block/brick bits when the grid has them (the forward's lookup)
    const bool cached = a.blocks != nullptr && a.bricks != nullptr;
    OccCache occ;
    for (int k = 0; k < todo; ++k) {
        float px, py, pz, qx, qy, qz;
        const bool hit = cached ? next_cached<RT>(ray, occ, a.mask, a.blocks, a.bricks, n, px, py,
                                                  pz, qx, qy, qz)
                                : ray.next<RT>(a.mask, n, px, py, pz, qx, qy, qz);
        if (!hit) break;
        float w[8];
        int64_t off[8];
        segment_corners<C>(px, py, pz, qx, qy, qz, n, w, off);
        float f[C];
        gather_integral<C>(a.voxels, w, off, f);
        float s = a.W[D::bs_off];
#pragma unroll
        for (int i = 0; i < C; ++i) s += a.W[D::ws_off + i] * f[i];
        float sigma, alpha;
        density_alpha<D>(s, sigma, alpha);
        float* o = out + (int64_t)k * RW;
        o[0] = px; o[1] = py; o[2] = pz;
        o[3] = qx; o[4] = qy; o[5] = qz;
        o[6] = s;
        o[7] = T;
        o[8] = alpha;
#pragma unroll
        for (int i = 0; i < C; ++i) o[9 + i] = f[i];
        S.add(sigma);
        T = __expf(-S.s);
    }
}

// Per-ray suffix sums of the compositing adjoint: for row k of ray r (rows [off_r, off_r + cnt_r)),
// suffix[k] = sum over the ray's rows j > k of w_j c_j, plus bg_r.  One lane per ray runs backwards
// over its rows with FP64 sums (the rows of a ray are a few to a few hundred).
__global__ void __launch_bounds__(256) ray_suffix_kernel(int64_t R, const int64_t* __restrict__ off,
                                                         const int32_t* __restrict__ cnt,
                                                         const float* __restrict__ w,
                                                         const float* __restrict__ c,
                                                         const float* __restrict__ bg,
                                                         float* __restrict__ out) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R) return;
    const int64_t o = off[r];
    const int n = cnt[r];
    const double b = bg ? (double)bg[r] : 0.0;
    double s0 = b, s1 = b, s2 = b;
#pragma unroll 4
    for (int k = n - 1; k >= 0; --k) {
        const int64_t i = o + k;
        out[3 * i + 0] = (float)s0;
        out[3 * i + 1] = (float)s1;
        out[3 * i + 2] = (float)s2;
        const double wk = (double)w[i];
        s0 += wk * (double)c[3 * i + 0];
        s1 += wk * (double)c[3 * i + 1];
        s2 += wk * (double)c[3 * i + 2];
    }
}

hipError_t launch_ray_suffix(int64_t R, const int64_t* off, const int32_t* cnt, const float* w,
                             const float* c, const float* bg, float* out, hipStream_t s) {
    if (R <= 0) return hipSuccess;
    hipLaunchKernelGGL(ray_suffix_kernel, dim3((unsigned)((R + 255) / 256)), dim3(256), 0, s, R, off,
                       cnt, w, c, bg, out);
    return hipGetLastError();
}

template <class D>
static hipError_t launch_rows_d(const RenderArgs& a, const int32_t* count, const int64_t* off,
                                float* rows, hipStream_t s) {
    const int rpw = rays_per_wave(a.n_rays);
    dim3 grid((unsigned)((a.n_rays + rpw - 1) / rpw)), block(64);
    if (a.traversal == 1)
        hipLaunchKernelGGL((rows_kernel<D, true>), grid, block, 0, s, a, count, off, rows, rpw);
    else
        hipLaunchKernelGGL((rows_kernel<D, false>), grid, block, 0, s, a, count, off, rows, rpw);
    return hipGetLastError();
}

hipError_t launch_rows(int variant, const RenderArgs& a, const int32_t* count, const int64_t* off,
                       float* rows, hipStream_t s) {
    if (a.n_rays == 0) return hipSuccess;
    switch (variant) {
        case 0: return launch_rows_d<DecB32>(a, count, off, rows, s);
        case 1: return launch_rows_d<DecB64>(a, count, off, rows, s);
        case 2: return launch_rows_d<DecB64C64>(a, count, off, rows, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace diver
