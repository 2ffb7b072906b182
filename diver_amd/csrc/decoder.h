// Canonical decoder form shared by every kernel that decodes intervals.
//
// All decoders the reference path uses reduce to one shape (the layer fusion of convert.py:70-77,
// applied generally):
//   z   = F                                  (integrated feature, C wide)
//   z   = relu(W_i z + b_i)                  for each of NPRE pre-layers (H wide)
//   s   = w_s . z + b_s ;  sigma = relu(s) or softplus(s)
//   y   = W_c z + W_v PE(d) + b_c            (H wide; PE: model/mlps.py:12-18, 4 bands, 27 wide)
//   y   = W_j relu(y) + b_j                  for each of NPOST post-layers
//   rgb = sigmoid(W_o relu(y) + b_o)
// Reference DIVeR32 (mlp1 32-32-32-33, mlp2 59-32-3): NPRE=2, relu density.  DIVeR64 (mlp1
// 32-64-65, mlp2 91-64-3): NPRE=1.  The real-time DIVeR32 blob stores features pre-multiplied by
// layer 0 (convert.py:149-150): NPRE=2 with PRE_ID (first pre-layer weight = identity, bias only).
// The benchmark decoder (softplus(f[0]); colour MLP on [f[1:], PE(d)]): NPRE=0, NPOST=1.
//
// Packed parameter layout (float32, row-major [out][in]):
//   pre 0: PRE_ID ? b[H] : W[H][C], b[H];   pre i>0: W[H][H], b[H]
//   head : w_s[Z], b_s, W_c[H][Z], W_v[H][27], b_c[H]        (Z = NPRE ? H : C)
//   post : W[H][H], b[H] (each)
//   out  : W_o[3][H], b_o[3]
#pragma once
#include <hip/hip_runtime.h>

namespace diver {

constexpr int kPE = 27;  // 3 * (2*4 + 1)

template <int C_, int H_, int NPRE_, bool PREID_, int DACT_, int NPOST_>
struct Dec {
    static constexpr int C = C_, H = H_, NPRE = NPRE_, DACT = DACT_, NPOST = NPOST_;
    static constexpr bool PREID = PREID_;
    static constexpr int Z = NPRE ? H : C;
    static constexpr int pre_size(int i) {
        return (i == 0) ? (PREID ? H : H * C + H) : (H * H + H);
    }
    static constexpr int pre_off(int i) {
        int o = 0;
        for (int k = 0; k < i; ++k) o += pre_size(k);
        return o;
    }
    static constexpr int head_off = pre_off(NPRE);
    static constexpr int ws_off = head_off;
    static constexpr int bs_off = ws_off + Z;
    static constexpr int wc_off = bs_off + 1;
    static constexpr int wv_off = wc_off + H * Z;
    static constexpr int bc_off = wv_off + H * kPE;
    static constexpr int post_off0 = bc_off + H;
    static constexpr int post_off(int j) { return post_off0 + j * (H * H + H); }
    static constexpr int out_off = post_off(NPOST);
    static constexpr int total = out_off + 3 * H + 3;
};

// The decoder variants compiled into the library.
//   id 0: benchmark decoder, C=32 H=32      id 1: benchmark decoder, C=32 H=64
//   id 2: benchmark decoder, C=64 H=64      id 3: reference DIVeR32 (offline)
//   id 4: reference DIVeR64 (offline)       id 5: reference DIVeR32 real-time blob (PRE_ID)
using DecB32 = Dec<32, 32, 0, false, 1, 1>;
using DecB64 = Dec<32, 64, 0, false, 1, 1>;
using DecB64C64 = Dec<64, 64, 0, false, 1, 1>;
using DecR32 = Dec<32, 32, 2, false, 0, 0>;
using DecR64 = Dec<32, 64, 1, false, 0, 0>;
using DecR32Blob = Dec<32, 32, 2, true, 0, 0>;

__device__ __forceinline__ float softplusf(float x) {
    // log(1 + e^x) = max(x, 0) + log1p(e^-|x|)
    return fmaxf(x, 0.0f) + log1pf(__expf(-fabsf(x)));
}

__device__ __forceinline__ float sigmoidf(float x) { return 1.0f / (1.0f + __expf(-x)); }

// PE(d) in the reference's torch order: [d, sin(pi d), cos(pi d), sin(2 pi d), ... cos(8 pi d)].
// sin(2^l pi d) is evaluated as sinpi(2^l d): 2^l d is exact, so only the sinpi rounding remains
// (the reference rounds pi * 2^l * d first, model/mlps.py:15-16).
__device__ __forceinline__ void positional_encoding(float dx, float dy, float dz, float* pe) {
    pe[0] = dx;
    pe[1] = dy;
    pe[2] = dz;
    const float d[3] = {dx, dy, dz};
#pragma unroll
    for (int l = 0; l < 4; ++l) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            float s, c;
            sincospif((float)(1 << l) * d[a], &s, &c);
            pe[3 + 6 * l + a] = s;
            pe[6 + 6 * l + a] = c;
        }
    }
}

// Per-ray hoisted part of the head: hv = W_v PE(d) + b_c   (FP32).
template <class D>
__device__ __forceinline__ void head_view(const float* __restrict__ W, float dx, float dy, float dz,
                                          float (&hv)[D::H]) {
    float pe[kPE];
    positional_encoding(dx, dy, dz, pe);
#pragma unroll
    for (int j = 0; j < D::H; ++j) {
        float a = W[D::bc_off + j];
#pragma unroll
        for (int i = 0; i < kPE; ++i) a += W[D::wv_off + j * kPE + i] * pe[i];
        hv[j] = a;
    }
}

// Pre-layers + density of one row (FP32, one lane).  z ends as the head input (Z wide).
template <class D>
__device__ __forceinline__ float pre_and_sigma(const float* __restrict__ W, const float* f,
                                               float (&z)[D::Z]) {
    if constexpr (D::NPRE == 0) {
#pragma unroll
        for (int i = 0; i < D::Z; ++i) z[i] = f[i];
    } else {
        float h[D::H];
        if constexpr (D::PREID) {
#pragma unroll
            for (int j = 0; j < D::H; ++j) h[j] = fmaxf(f[j] + W[D::pre_off(0) + j], 0.0f);
        } else {
#pragma unroll
            for (int j = 0; j < D::H; ++j) {
                float a = W[D::pre_off(0) + D::H * D::C + j];
#pragma unroll
                for (int i = 0; i < D::C; ++i) a += W[D::pre_off(0) + j * D::C + i] * f[i];
                h[j] = fmaxf(a, 0.0f);
            }
        }
#pragma unroll
        for (int l = 1; l < D::NPRE; ++l) {
            float g[D::H];
#pragma unroll
            for (int j = 0; j < D::H; ++j) {
                float a = W[D::pre_off(l) + D::H * D::H + j];
#pragma unroll
                for (int i = 0; i < D::H; ++i) a += W[D::pre_off(l) + j * D::H + i] * h[i];
                g[j] = fmaxf(a, 0.0f);
            }
#pragma unroll
            for (int j = 0; j < D::H; ++j) h[j] = g[j];
        }
#pragma unroll
        for (int i = 0; i < D::Z; ++i) z[i] = h[i];
    }
    float s = W[D::bs_off];
#pragma unroll
    for (int i = 0; i < D::Z; ++i) s += W[D::ws_off + i] * z[i];
    return s;
}

template <class D>
__device__ __forceinline__ float density_act(float s) {
    if constexpr (D::DACT == 0) return fmaxf(s, 0.0f);
    else return softplusf(s);
}

// Density and opacity of one interval from the pre-activation s: sigma = act(s) and
// alpha = 1 - exp(-sigma).  For softplus, exp(-softplus(s)) = 1 / (1 + e^s), so with
// log(1 + e) for e in [0, 1], as 2 atanh(z) with z = e / (2 + e) in [0, 1/3]: z is corrected
// for the rounding of 2 + e and of the reciprocal, then the odd series runs to z^13 (truncation
// below 2^-26 relative).  Within 1.25 ulp of log1p over the whole interval, in 16 operations.
__device__ __forceinline__ float log1p_unit(float e) {
    const float d = 2.0f + e;
    const float derr = e - (d - 2.0f);  // 2 + e = d + derr exactly
    const float r = __builtin_amdgcn_rcpf(d);
    float z = e * r;
    float res = fmaf(-z, d, e);
    res = fmaf(-z, derr, res);
    z = fmaf(res, r, z);
    const float z2 = z * z;
    float p = 1.0f / 13.0f;
    p = fmaf(p, z2, 1.0f / 11.0f);
    p = fmaf(p, z2, 1.0f / 9.0f);
    p = fmaf(p, z2, 1.0f / 7.0f);
    p = fmaf(p, z2, 1.0f / 5.0f);
    p = fmaf(p, z2, 1.0f / 3.0f);
    const float t = 2.0f * z;
    return fmaf(t * z2, p, t);
}

// The sparsity term log(1 + s^2 / 0.5) of one interval (train.py:84-85, where torch takes it in
// FP32 as log(1 + x)): log1p_unit for x = 2 s^2 <= 1, and log(1 + x) above, where rounding 1 + x
// costs at most 2^-24 / log 2 relative.  About 20 operations against ~110 for the library log1pf.
__device__ __forceinline__ float sparsity_term(float s) {
    const float x = 2.0f * s * s;
    return x <= 1.0f ? log1p_unit(x) : __logf(1.0f + x);
}

// e = exp(-|s|): alpha = 1 / (1 + e) for s > 0 and e / (1 + e) otherwise, which shares the
// exponential with softplus.
template <class D>
__device__ __forceinline__ void density_alpha(float s, float& sigma, float& alpha) {
    if constexpr (D::DACT == 0) {
        sigma = fmaxf(s, 0.0f);
        alpha = -expm1f(-sigma);
    } else {
        const float e = __expf(-fabsf(s));
        sigma = fmaxf(s, 0.0f) + log1p_unit(e);
        alpha = (s > 0.0f ? 1.0f : e) * __builtin_amdgcn_rcpf(1.0f + e);
    }
}

// Colour of one row from the head input z and the hoisted hv (FP32, one lane).
template <class D>
__device__ __forceinline__ void colour(const float* __restrict__ W, const float (&z)[D::Z],
                                       const float (&hv)[D::H], float (&rgb)[3]) {
    float y[D::H];
#pragma unroll
    for (int j = 0; j < D::H; ++j) {
        float a = hv[j];
#pragma unroll
        for (int i = 0; i < D::Z; ++i) a += W[D::wc_off + j * D::Z + i] * z[i];
        y[j] = a;
    }
#pragma unroll
    for (int l = 0; l < D::NPOST; ++l) {
        float g[D::H];
#pragma unroll
        for (int j = 0; j < D::H; ++j) {
            float a = W[D::post_off(l) + D::H * D::H + j];
#pragma unroll
            for (int i = 0; i < D::H; ++i) a += W[D::post_off(l) + j * D::H + i] * fmaxf(y[i], 0.0f);
            g[j] = a;
        }
#pragma unroll
        for (int j = 0; j < D::H; ++j) y[j] = g[j];
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float a = W[D::out_off + 3 * D::H + k];
#pragma unroll
        for (int i = 0; i < D::H; ++i) a += W[D::out_off + k * D::H + i] * fmaxf(y[i], 0.0f);
        rgb[k] = sigmoidf(a);
    }
}

}  // namespace diver
