// The fused render kernel's launches (render_tile.h holds the kernel): grids whose features fit
// 32-bit byte offsets, the FP32 matrix-core operand packing and the launch dispatch.
#include "render_tile.h"

namespace diver {

// FP32 matrix-core A operands (shade_tile_f32's order) from the canonical parameters of a decoder
// without pre-layers: one float per (k step, lane).
__global__ void __launch_bounds__(256) pack_f32_kernel(int C, int H, int npost,
                                                       const float* __restrict__ W,
                                                       float* __restrict__ out, int64_t total) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    const int r = (int)(idx & 31), h = (int)((idx >> 5) & 1);
    int64_t step = idx >> 6;
    const int64_t wc = C + 1;  // w_s[C], b_s, then W_c[H][C]
    const int64_t post0 = wc + (int64_t)H * C + (int64_t)H * kPE + H;
    const int64_t outo = post0 + (int64_t)npost * (H * H + H);
    const int64_t head_steps = (int64_t)(H / 32) * (C / 2);
    const int64_t post_steps = (int64_t)(H / 32) * (H / 2);
    float v;
    if (step < head_steps) {
        const int mt = (int)(step / (C / 2)), kc = (int)(step % (C / 2));
        v = W[wc + (int64_t)(32 * mt + r) * C + 2 * kc + h];
    } else if ((step -= head_steps) < npost * post_steps) {
        const int j = (int)(step / post_steps);
        const int64_t rem = step % post_steps;
        const int mt = (int)(rem / (H / 2)), q = (int)(rem % (H / 2));
        v = W[post0 + (int64_t)j * (H * H + H) + (int64_t)(32 * mt + r) * H + d_order_feature(q, h)];
    } else {
        const int q = (int)(step - npost * post_steps);
        v = r < 3 ? W[outo + (int64_t)r * H + d_order_feature(q, h)] : 0.0f;
    }
    out[idx] = v;
}

int64_t pack_f32_size(int C, int H, int npost) {
    return (int64_t)H * C + (int64_t)npost * H * H + 32 * (int64_t)H;
}

hipError_t launch_pack_f32(int C, int H, int npost, const float* W, float* out, hipStream_t s) {
    if (C % 2 || H % 32 || C <= 0) return hipErrorInvalidValue;
    const int64_t total = pack_f32_size(C, H, npost);
    hipLaunchKernelGGL(pack_f32_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, C, H,
                       npost, W, out, total);
    return hipGetLastError();
}

size_t render_tile_lds_bytes(int n, int H, int prec) {
    int nb = n >> 3;
    int64_t nwords = kBlocksInLds ? ((int64_t)nb * nb * nb + 31) >> 5 : 0;
    return (size_t)((nwords + 3) & ~3) * 4 + kWavesPerGroup * (size_t)wave_lds_floats(H, prec) * sizeof(float);
}

bool tile_fits(int n) {
    const int nb = n >> 3;
    const int64_t nwords = ((int64_t)nb * nb * nb + 31) >> 5;
    return n % 8 == 0 && (!kBlocksInLds || nwords <= kMaxLdsBlockWords);
}

hipError_t launch_render_tiles(int variant, const BatchArgs& args, int prec, bool dch0,
                               hipStream_t s) {
    if (args.total == 0) return hipSuccess;
    if (prec != kPrecFP32 && !args.g.pack16) return hipErrorInvalidValue;  // no 16x16x32 operands
    BatchArgs p = args;
    p.rays_per_wave = rays_per_wave(p.total);
    const int64_t n1 = (int64_t)p.g.n + 1;
    const bool b32 = n1 * n1 * n1 * p.g.channels * (int64_t)sizeof(float) <= (int64_t)UINT32_MAX;
    return b32 ? launch_tiles<true>(variant, p, prec, dch0, s)
               : launch_tiles_wide(variant, p, prec, dch0, s);
}

}  // namespace diver
