// Real-time fused decode + composite, one ray per lane (the plain FP32 path):
//   mlp_eval <- utils/mlp_evaluation/src/mlp_eval.cu:32-231 (DIVeR32 blob) and
//               utils/mlp_evaluation64/src/mlp_eval.cu:32-231 (DIVeR64 blob),
// generalised to any canonical decoder.  Per unfinished ray it walks the <= 8 intervals a march
// round left in `coord` ([8][B][6], hit-major), integrates the features of each interval's voxel
// (sparse rows addressed through the (N+1)^3 corner map, clamped like the reference's texture),
// decodes density, skips the colour when alpha < alpha_skip (transmittance unchanged, as
// mlp_eval.cu:153-155), composites front to back and finishes the ray when transmittance drops
// below t_stop (mlp_eval.cu:218-222) or the interval list ends (x0 < 0, mlp_eval.cu:67-70).
// The reference maps rays through an 800-wide 16x24 tiling (mlp_eval.cu:44-55) that enumerates
// every index below the batch exactly once; here rays are a flat 1-D grid, so any image size works.
#include "common.h"
#include "decoder.h"
#include "kernels.h"

namespace diver {

__device__ __forceinline__ int clampi(int x, int lo, int hi) { return x < lo ? lo : (x > hi ? hi : x); }

template <class D>
__global__ void __launch_bounds__(256) mlp_eval_kernel(
    int64_t batch, const float* __restrict__ coord, const float* __restrict__ feats,
    const int32_t* __restrict__ vmap, int map_size, const float* __restrict__ v,
    float* __restrict__ rgba, uint8_t* __restrict__ finish, const float* __restrict__ W,
    float alpha_skip, float t_stop) {
    int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= batch) return;
    if (finish[b]) return;
    float hv[D::H];
    head_view<D>(W, v[3 * b], v[3 * b + 1], v[3 * b + 2], hv);
    float acc_r = rgba[4 * b + 0], acc_g = rgba[4 * b + 1], acc_b = rgba[4 * b + 2];
    float T = rgba[4 * b + 3];
    bool done = false;
    const int ms1 = map_size - 1;
    for (int hit = 0; hit < kMaxHit; ++hit) {
        const float* c = coord + (int64_t)hit * batch * 6 + b * 6;
        float x0 = c[0], y0 = c[1], z0 = c[2], x1 = c[3], y1 = c[4], z1 = c[5];
        if (x0 < 0) {
            done = true;
            break;
        }
        int cx, cy, cz;
        {
#pragma clang fp contract(off)
            int ax = (int)(x0 + 1e-4f), bx = (int)(x1 + 1e-4f);
            int ay = (int)(y0 + 1e-4f), by = (int)(y1 + 1e-4f);
            int az = (int)(z0 + 1e-4f), bz = (int)(z1 + 1e-4f);
            cx = ax < bx ? ax : bx;
            cy = ay < by ? ay : by;
            cz = az < bz ? az : bz;
        }
        x0 -= cx; x1 -= cx; y0 -= cy; y1 -= cy; z0 -= cz; z1 -= cz;
        float z01 = z0 + z1, y01 = y0 + y1, x01 = x0 + x1;
        float w[8];
        w[7] = (2.0f * x0 * y0 * z0 + 2.0f * x1 * y1 * z1 + x01 * y01 * z01) * (1.0f / 12.0f);
        float w2 = (y01 * z01 + y0 * z0 + y1 * z1) * (1.0f / 6.0f);
        w[5] = (x01 * z01 + x0 * z0 + x1 * z1) * (1.0f / 6.0f) - w[7];
        float w4 = (x01 * y01 + x0 * y0 + x1 * y1) * (1.0f / 6.0f);
        x01 *= 0.5f; y01 *= 0.5f; z01 *= 0.5f;
        w[4] = z01 - w[5] - w2;
        w[0] = 1.0f - w[4] - y01 - x01 + w4;
        w[6] = w2 - w[7];
        w[2] = y01 - w[6] - w4;
        w[1] = x01 - w[5] - w4;
        w[3] = w4 - w[7];
        float f[D::C];
#pragma unroll
        for (int i = 0; i < D::C; ++i) f[i] = 0.0f;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            int gx = clampi(cx + (j & 1), 0, ms1);
            int gy = clampi(cy + ((j >> 1) & 1), 0, ms1);
            int gz = clampi(cz + (j >> 2), 0, ms1);
            int64_t row = vmap[((int64_t)gz * map_size + gy) * map_size + gx];
            const float4* src = reinterpret_cast<const float4*>(feats + row * D::C);
#pragma unroll
            for (int i = 0; i < D::C / 4; ++i) {
                float4 q = src[i];
                f[4 * i + 0] += w[j] * q.x;
                f[4 * i + 1] += w[j] * q.y;
                f[4 * i + 2] += w[j] * q.z;
                f[4 * i + 3] += w[j] * q.w;
            }
        }
        float z[D::Z];
        float s = pre_and_sigma<D>(W, f, z);
        float alpha = 1.0f - __expf(-density_act<D>(s));
        if (alpha < alpha_skip) continue;
        float rgb[3];
        colour<D>(W, z, hv, rgb);
        float wgt = T * alpha;
        acc_r += wgt * rgb[0];
        acc_g += wgt * rgb[1];
        acc_b += wgt * rgb[2];
        T = T * (1.0f - alpha);
        if (T < t_stop) {
            done = true;
            break;
        }
    }
    rgba[4 * b + 0] = acc_r;
    rgba[4 * b + 1] = acc_g;
    rgba[4 * b + 2] = acc_b;
    rgba[4 * b + 3] = T;
    if (done) finish[b] = 1;
}

template <class D>
static hipError_t launch_mlp_eval_t(int64_t batch, const float* coord, const float* feats,
                                    const int32_t* vmap, int map_size, const float* v, float* rgba,
                                    uint8_t* finish, const float* W, float alpha_skip, float t_stop,
                                    hipStream_t s) {
    if (batch == 0) return hipSuccess;
    hipLaunchKernelGGL(mlp_eval_kernel<D>, dim3((unsigned)((batch + 255) / 256)), dim3(256), 0, s,
                       batch, coord, feats, vmap, map_size, v, rgba, finish, W, alpha_skip, t_stop);
    return hipGetLastError();
}

hipError_t launch_mlp_eval(int variant, int64_t batch, const float* coord, const float* feats,
                           const int32_t* vmap, int map_size, const float* v, float* rgba,
                           uint8_t* finish, const float* W, float alpha_skip, float t_stop,
                           hipStream_t s) {
    switch (variant) {
        case 0: return launch_mlp_eval_t<DecB32>(batch, coord, feats, vmap, map_size, v, rgba, finish, W, alpha_skip, t_stop, s);
        case 1: return launch_mlp_eval_t<DecB64>(batch, coord, feats, vmap, map_size, v, rgba, finish, W, alpha_skip, t_stop, s);
        case 2: return launch_mlp_eval_t<DecB64C64>(batch, coord, feats, vmap, map_size, v, rgba, finish, W, alpha_skip, t_stop, s);
        case 3: return launch_mlp_eval_t<DecR32>(batch, coord, feats, vmap, map_size, v, rgba, finish, W, alpha_skip, t_stop, s);
        case 4: return launch_mlp_eval_t<DecR64>(batch, coord, feats, vmap, map_size, v, rgba, finish, W, alpha_skip, t_stop, s);
        case 5: return launch_mlp_eval_t<DecR32Blob>(batch, coord, feats, vmap, map_size, v, rgba, finish, W, alpha_skip, t_stop, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace diver
