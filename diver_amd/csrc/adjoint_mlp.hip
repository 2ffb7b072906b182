// Decoder and compositing adjoint over the recompute pass's rows, on the matrix cores.
//
// The rows of diver_render_rows ([p_in, p_out, s, T_before, alpha, F], one per composited interval,
// grouped by ray in walk order) feed two kernels, 32 rows per tile and one wave per tile:
//
//   adj_forward_kernel: the colour decoder of each row, as the fused renderer evaluates it (FP32
//     operands on v_mfma_f32_32x32x2_f32, weights as A, rows as B):
//        y1 = W_c F + (W_v PE(d) + b_c);  r1 = relu(y1);  y2 = W_2 r1 + b_2;  r2 = relu(y2);
//        c = sigmoid(W_o r2 + b_o);  w = T_before * alpha
//     and writes r1, r2, c and w.  Per-ray suffix sums of w c (diver_ray_suffix) follow.
//   adj_backward_kernel: the compositing adjoint of C = sum_k T_k a_k c_k + bg T_n and the
//     decoder's backward pass,
//        dsigma = g . (T_after c - suffix);  ds = dsigma act'(s) (+ the sparsity term);
//        do = (g w) c (1 - c);  dr2 = (W_o^T do) [r2 > 0];  dr1 = (W_2^T dr2) [r1 > 0];
//        dF = W_c^T dr1 + ds w_s
//     with the transposed weights as A operands (packed by adj_pack_kernel, in the D order of the
//     previous product, so each D tile is the next product's B operand in place), and writes dF
//     (the grid adjoint's input), do, ds, dr1 and dr2 (the weight gradients' operands).
// The weight gradients are reductions over all rows (dW_o = do^T r2, dW_2 = dr2^T r1, ...):
// wgrad_partial_kernel and wgrad_reduce_kernel at the end of this file.
#include <mutex>

#include "common.h"
#include "decoder.h"
#include "kernels.h"
#include "mfma.h"

namespace diver {

// A-operand steps of the adjoint, 64 floats each: W_o^T (natural k pairs over the 3 outputs,
// padded to 4), W_2^T and W_c^T (k steps in the D order of the product they follow).
template <class D>
struct AdjPack {
    static constexpr int MT = D::H / 32, KH = D::H / 2;
    static constexpr int wo = 0;                              // MT x 2 steps
    static constexpr int w2 = wo + MT * 2;                     // MT x KH steps
    static constexpr int wc = w2 + D::NPOST * MT * KH;         // (C / 32) x KH steps
    static constexpr int steps = wc + (D::C / 32) * KH;
};

int64_t pack_adj_size(int C, int H, int npost) {
    return 64 * ((int64_t)(H / 32) * 2 + (int64_t)npost * (H / 32) * (H / 2) + (int64_t)(C / 32) * (H / 2));
}

__global__ void __launch_bounds__(256) adj_pack_kernel(int C, int H, int npost,
                                                       const float* __restrict__ W,
                                                       float* __restrict__ out, int64_t total) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    const int r = (int)(idx & 31), h = (int)((idx >> 5) & 1);
    int64_t step = idx >> 6;
    const int MT = H / 32, KH = H / 2;
    const int64_t wc_off = C + 1;                                   // w_s[C], b_s, then W_c[H][C]
    const int64_t post0 = wc_off + (int64_t)H * C + (int64_t)H * kPE + H;
    const int64_t out0 = post0 + (int64_t)npost * (H * H + H);
    float v;
    if (step < 2 * MT) {
        const int mt = (int)(step / 2), kc = (int)(step % 2);
        const int k = 2 * kc + h;
        v = k < 3 ? W[out0 + (int64_t)k * H + 32 * mt + r] : 0.0f;
    } else if ((step -= 2 * MT) < (int64_t)npost * MT * KH) {
        const int mt = (int)(step / KH), q = (int)(step % KH);
        v = W[post0 + (int64_t)d_order_feature(q, h) * H + 32 * mt + r];
    } else {
        step -= (int64_t)npost * MT * KH;
        const int mt = (int)(step / KH), q = (int)(step % KH);
        v = W[wc_off + (int64_t)d_order_feature(q, h) * C + 32 * mt + r];
    }
    out[idx] = v;
}

hipError_t launch_pack_adj(int C, int H, int npost, const float* W, float* out, hipStream_t s) {
    if (C % 32 || H % 32 || npost != 1) return hipErrorInvalidValue;
    const int64_t total = pack_adj_size(C, H, npost);
    hipLaunchKernelGGL(adj_pack_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, C, H,
                       npost, W, out, total);
    return hipGetLastError();
}

// D-layout unit of register i of m tile mt on this lane half: 32 mt + 8 (i / 4) + 4 h + i % 4, so
// registers 4 g .. 4 g + 3 are the units 32 mt + 8 g + 4 h + (0..3), one float4 of a row.
__device__ __forceinline__ int d_unit(int mt, int i, int h) { return 32 * mt + 8 * (i >> 2) + 4 * h + (i & 3); }

template <int MT>
__device__ __forceinline__ void store_d_rows(float* __restrict__ dst, const f32x16_t (&v)[MT], int h,
                                             bool relu) {
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            float4 x = make_float4(v[mt][4 * g], v[mt][4 * g + 1], v[mt][4 * g + 2], v[mt][4 * g + 3]);
            if (relu) {
                x.x = fmaxf(x.x, 0.0f);
                x.y = fmaxf(x.y, 0.0f);
                x.z = fmaxf(x.z, 0.0f);
                x.w = fmaxf(x.w, 0.0f);
            }
            *reinterpret_cast<float4*>(dst + 32 * mt + 8 * g + 4 * h) = x;
        }
}

template <class D>
__global__ void __launch_bounds__(64) adj_forward_kernel(AdjArgs p) {
    static_assert(D::NPRE == 0 && D::NPOST == 1 && D::C % 32 == 0 && D::H % 32 == 0, "decoder shape");
    constexpr int C = D::C, H = D::H, MT = H / 32, KC = C / 2, KH = H / 2, HC = C / 2;
    const int lane = threadIdx.x, c = lane & 31, h = lane >> 5;
    const float* W = p.W;
    const float* pk = p.pack;
    for (int64_t tile = blockIdx.x; tile * 32 < p.n_rows; tile += gridDim.x) {
        // per-tile opaque pointers keep the operand and weight loads inside the tile loop
        asm volatile("" : "+s"(pk));
        asm volatile("" : "+s"(W));
        const int64_t row = tile * 32 + c;
        const bool live = row < p.n_rows;
        const int64_t rr = live ? row : 0;
        const float* rw = p.rows + rr * p.row_stride;
        const int64_t ray = p.ray[rr];
        float pe[kPE];
        positional_encoding(p.d[3 * ray], p.d[3 * ray + 1], p.d[3 * ray + 2], pe);
        // y1 starts at this lane's units of W_v PE(d) + b_c
        f32x16_t y[MT];
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                const int u = d_unit(mt, i, h);
                float a = W[D::bc_off + u];
#pragma unroll
                for (int j = 0; j < kPE; ++j) a += W[D::wv_off + u * kPE + j] * pe[j];
                y[mt][i] = a;
            }
        // B operands of the head: lane c + 32 h holds channels [h C/2, (h + 1) C/2) of row c;
        // one swap per pair of k steps puts (2 kc, 2 kc + 1) on the two lane halves
        float g[HC];
#pragma unroll
        for (int i = 0; i < HC; ++i) g[i] = rw[9 + h * HC + i];
        float b[KC];
#pragma unroll
        for (int kc = 0; kc < HC / 2; ++kc) {
            float x = g[2 * kc], z = g[2 * kc + 1];
            swap32f(x, z);
            b[kc] = x;
            b[kc + HC / 2] = z;
        }
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) y[mt] = mfma_chain_f32<KC>(pk + mt * KC * 64, lane, b, y[mt]);
        if (live) store_d_rows<MT>(p.r1 + rr * H, y, h, true);
        float e[KH];
#pragma unroll
        for (int q = 0; q < KH; ++q) e[q] = fmaxf(y[q / 16][q % 16], 0.0f);
        const int pofs = MT * KC * 64;
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
            f32x16_t s;
#pragma unroll
            for (int i = 0; i < 16; ++i) s[i] = W[D::post_off(0) + H * H + d_unit(mt, i, h)];
            y[mt] = mfma_chain_f32<KH>(pk + pofs + mt * KH * 64, lane, e, s);
        }
        if (live) store_d_rows<MT>(p.r2 + rr * H, y, h, true);
#pragma unroll
        for (int q = 0; q < KH; ++q) e[q] = fmaxf(y[q / 16][q % 16], 0.0f);
        f32x16_t o = {};
        o = mfma_chain_f32<KH>(pk + pofs + MT * KH * 64, lane, e, o);
        // outputs 0..2 of row c sit in registers 0..2 of lanes 0-31
        if (h == 0 && live) {
            const float* bo = W + D::out_off + 3 * H;
            p.color[3 * rr + 0] = sigmoidf(o[0] + bo[0]);
            p.color[3 * rr + 1] = sigmoidf(o[1] + bo[1]);
            p.color[3 * rr + 2] = sigmoidf(o[2] + bo[2]);
            p.weight[rr] = rw[7] * rw[8];
        }
    }
}

template <class D>
__global__ void __launch_bounds__(64) adj_backward_kernel(AdjArgs p) {
    static_assert(D::NPRE == 0 && D::NPOST == 1 && D::C % 32 == 0 && D::H % 32 == 0, "decoder shape");
    constexpr int C = D::C, H = D::H, MT = H / 32, KH = H / 2, CT = C / 32;
    using P = AdjPack<D>;
    const int lane = threadIdx.x, c = lane & 31, h = lane >> 5;
    const float* W = p.W;
    const float* pa = p.pack_adj;
    const float sscale = p.sparsity_scale ? *p.sparsity_scale : 0.0f;
    for (int64_t tile = blockIdx.x; tile * 32 < p.n_rows; tile += gridDim.x) {
        asm volatile("" : "+s"(pa));
        asm volatile("" : "+s"(W));
        const int64_t row = tile * 32 + c;
        const bool live = row < p.n_rows;
        const bool valid = row < p.n_valid;
        const int64_t rr = live ? row : 0;
        const float* rw = p.rows + rr * p.row_stride;
        const float s = rw[6], T = rw[7], alpha = rw[8];
        const float c0 = p.color[3 * rr], c1 = p.color[3 * rr + 1], c2 = p.color[3 * rr + 2];
        const float wgt = p.weight[rr];
        const int64_t ray = p.ray[rr];
        const float g0 = p.g[3 * ray], g1 = p.g[3 * ray + 1], g2 = p.g[3 * ray + 2];
        float u0 = 0.0f, u1 = 0.0f, u2 = 0.0f;
        if (valid) {
            u0 = p.suffix[3 * rr];
            u1 = p.suffix[3 * rr + 1];
            u2 = p.suffix[3 * rr + 2];
        }
        // compositing adjoint (every lane for its row c; both lane halves alike)
        const float t_after = T * (1.0f - alpha);
        const float dsigma = g0 * (t_after * c0 - u0) + g1 * (t_after * c1 - u1) + g2 * (t_after * c2 - u2);
        float ds = dsigma * (D::DACT == 1 ? sigmoidf(s) : (s > 0.0f ? 1.0f : 0.0f));
        // d/ds log(1 + s^2 / 0.5) = 4 s / (1 + 2 s^2); padding rows have s = 0
        ds += sscale * (4.0f * s / (1.0f + 2.0f * s * s));
        const float do0 = g0 * wgt * c0 * (1.0f - c0);
        const float do1 = g1 * wgt * c1 * (1.0f - c1);
        const float do2 = g2 * wgt * c2 * (1.0f - c2);
        // dr2 = W_o^T do: k pairs (do0, do1), (do2, 0) over the lane halves
        const float bA = h ? do1 : do0, bB = h ? 0.0f : do2;
        f32x16_t t[MT];
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
            f32x16_t acc = {};
            acc = mfma32_f32(pa[(P::wo + 2 * mt) * 64 + lane], bA, acc);
            acc = mfma32_f32(pa[(P::wo + 2 * mt + 1) * 64 + lane], bB, acc);
            // mask by r2 > 0 (r2 = relu(y2) written by the forward, in the same units)
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const float4 r = *reinterpret_cast<const float4*>(p.r2 + rr * H + 32 * mt + 8 * g + 4 * h);
                acc[4 * g + 0] = r.x > 0.0f ? acc[4 * g + 0] : 0.0f;
                acc[4 * g + 1] = r.y > 0.0f ? acc[4 * g + 1] : 0.0f;
                acc[4 * g + 2] = r.z > 0.0f ? acc[4 * g + 2] : 0.0f;
                acc[4 * g + 3] = r.w > 0.0f ? acc[4 * g + 3] : 0.0f;
            }
            t[mt] = acc;
        }
        if (live) store_d_rows<MT>(p.dr2 + rr * H, t, h, false);
        float e[KH];
#pragma unroll
        for (int q = 0; q < KH; ++q) e[q] = t[q / 16][q % 16];
        // dr1 = W_2^T dr2, masked by r1 > 0
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
            f32x16_t acc = {};
            acc = mfma_chain_f32<KH>(pa + (P::w2 + mt * KH) * 64, lane, e, acc);
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const float4 r = *reinterpret_cast<const float4*>(p.r1 + rr * H + 32 * mt + 8 * g + 4 * h);
                acc[4 * g + 0] = r.x > 0.0f ? acc[4 * g + 0] : 0.0f;
                acc[4 * g + 1] = r.y > 0.0f ? acc[4 * g + 1] : 0.0f;
                acc[4 * g + 2] = r.z > 0.0f ? acc[4 * g + 2] : 0.0f;
                acc[4 * g + 3] = r.w > 0.0f ? acc[4 * g + 3] : 0.0f;
            }
            t[mt] = acc;
        }
        if (live) store_d_rows<MT>(p.dr1 + rr * H, t, h, false);
#pragma unroll
        for (int q = 0; q < KH; ++q) e[q] = t[q / 16][q % 16];
        // dF = W_c^T dr1 + ds w_s
        f32x16_t f[CT];
#pragma unroll
        for (int mt = 0; mt < CT; ++mt) {
            f32x16_t acc;
#pragma unroll
            for (int i = 0; i < 16; ++i) acc[i] = ds * W[D::ws_off + d_unit(mt, i, h)];
            f[mt] = mfma_chain_f32<KH>(pa + (P::wc + mt * KH) * 64, lane, e, acc);
        }
        if (valid) store_d_rows<CT>(p.dF + rr * C, f, h, false);
        if (h == 0 && live) {
            p.dout[3 * rr + 0] = do0;
            p.dout[3 * rr + 1] = do1;
            p.dout[3 * rr + 2] = do2;
            p.ds[rr] = ds;
        }
    }
}

// PE(d) per ray (the fused kernels' convention), [R][27].
__global__ void __launch_bounds__(256) pe_kernel(int64_t R, const float* __restrict__ d,
                                                 float* __restrict__ out) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R) return;
    float pe[kPE];
    positional_encoding(d[3 * r], d[3 * r + 1], d[3 * r + 2], pe);
#pragma unroll
    for (int j = 0; j < kPE; ++j) out[r * kPE + j] = pe[j];
}

hipError_t launch_positional_encoding(int64_t R, const float* d, float* pe, hipStream_t s) {
    if (R <= 0) return hipSuccess;
    hipLaunchKernelGGL(pe_kernel, dim3((unsigned)((R + 255) / 256)), dim3(256), 0, s, R, d, pe);
    return hipGetLastError();
}

// Tiles per launch: one wave per 32-row tile up to 8 waves per CU's worth, then each wave loops.
static dim3 adj_grid(int64_t n_rows) {
    const int64_t tiles = (n_rows + 31) / 32;
    return dim3((unsigned)(tiles < 256 * 8 ? tiles : 256 * 8));
}

template <class D>
static hipError_t launch_adj_d(bool backward, const AdjArgs& p, hipStream_t s) {
    if (backward)
        hipLaunchKernelGGL((adj_backward_kernel<D>), adj_grid(p.n_rows), dim3(64), 0, s, p);
    else
        hipLaunchKernelGGL((adj_forward_kernel<D>), adj_grid(p.n_rows), dim3(64), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_adjoint(int variant, bool backward, const AdjArgs& p, hipStream_t s) {
    if (p.n_rows == 0) return hipSuccess;
    switch (variant) {
        case 0: return launch_adj_d<DecB32>(backward, p, s);
        case 1: return launch_adj_d<DecB64>(backward, p, s);
        case 2: return launch_adj_d<DecB64C64>(backward, p, s);
        default: return hipErrorInvalidValue;
    }
}

// ---------------------------------------------------------------------------------------------
// Weight gradients of the decoder adjoint: reductions over every row of outer products,
//   dw_s = ds^T F, db_s = sum ds, dW_c = dr1^T F, dW_v = dr1^T PE(d), db_c = sum dr1,
//   dW_2 = dr2^T r1, db_2 = sum dr2, dW_o = do^T r2, db_o = sum do,
// written straight into the canonical gradient vector.  Two passes, both in a fixed order (the
// result does not depend on timing): each workgroup sums a contiguous slab of rows into a partial
// of its own (wgrad_partial_kernel), then one lane per gradient element adds the workgroups'
// partials in workgroup order (wgrad_reduce_kernel).
//
// Rows are staged through LDS, TK at a time, as two vectors per row:
//   a = [dr1 (H) | dr2 (H) | do (3) | ds]
//   b = [F (C) | PE (27) | 1 | 0 ... to CP | r1 (H) | r2 (H) | 1 0 0 0 0 0 0 0]
// and each lane accumulates one 4 x 8 block ("tile") of sum a b^T in registers.  The tiles cover
// dr1 x [F | PE | 1] (dW_c, dW_v, db_c), dr2 x r1 (dW_2), dr2 x 1 (db_2) and [do | ds] x (r2, F, 1)
// (dW_o, dw_s, db_o, db_s); the few products the blocks compute beyond those are never read.
template <int C, int H>
struct WgradLayout {
    static constexpr int CP = (C + kPE + 1 + 7) / 8 * 8;
    static constexpr int LA = 2 * H + 4;
    static constexpr int LB = CP + 2 * H + 8;
    static constexpr int B_R1 = CP, B_R2 = CP + H, B_ONE = CP + 2 * H;
    static constexpr int N1 = (H / 4) * (CP / 8);
    static constexpr int N2 = (H / 4) * (H / 8);
    static constexpr int N3 = H / 4;
    static constexpr int N4 = H / 8 + C / 8 + 1;
    static constexpr int NT = N1 + N2 + N3 + N4;
    static constexpr int THREADS = (NT + 63) / 64 * 64;
    static constexpr int NP = C + 1 + H * C + H * kPE + H + H * H + H + 3 * H + 3;

    __device__ static void tile(int t, int& a0, int& b0) {
        if (t < N1) { a0 = 4 * (t / (CP / 8)); b0 = 8 * (t % (CP / 8)); return; }
        t -= N1;
        if (t < N2) { a0 = H + 4 * (t / (H / 8)); b0 = B_R1 + 8 * (t % (H / 8)); return; }
        t -= N2;
        if (t < N3) { a0 = H + 4 * t; b0 = B_ONE; return; }
        t -= N3;
        a0 = 2 * H;
        b0 = t < H / 8 ? B_R2 + 8 * t : (t < H / 8 + C / 8 ? 8 * (t - H / 8) : B_ONE);
    }
    // the tile holding element (ia, ib) of a b^T, for the (ia, ib) the gradients read
    __device__ static int tile_of(int ia, int ib) {
        if (ia < H) return (ia / 4) * (CP / 8) + ib / 8;
        if (ia < 2 * H)
            return ib == B_ONE ? N1 + N2 + (ia - H) / 4 : N1 + ((ia - H) / 4) * (H / 8) + (ib - B_R1) / 8;
        if (ib == B_ONE) return N1 + N2 + N3 + H / 8 + C / 8;
        return ib >= B_R2 ? N1 + N2 + N3 + (ib - B_R2) / 8 : N1 + N2 + N3 + H / 8 + ib / 8;
    }
    // canonical gradient element p -> (ia, ib)
    __device__ static void element(int p, int& ia, int& ib) {
        if (p < C) { ia = 2 * H + 3; ib = p; return; }                 // w_s
        if (p == C) { ia = 2 * H + 3; ib = B_ONE; return; }            // b_s
        int q = p - (C + 1);
        if (q < H * C) { ia = q / C; ib = q % C; return; }             // W_c
        q -= H * C;
        if (q < H * kPE) { ia = q / kPE; ib = C + q % kPE; return; }   // W_v
        q -= H * kPE;
        if (q < H) { ia = q; ib = C + kPE; return; }                   // b_c
        q -= H;
        if (q < H * H) { ia = H + q / H; ib = B_R1 + q % H; return; }  // W_2
        q -= H * H;
        if (q < H) { ia = H + q; ib = B_ONE; return; }                 // b_2
        q -= H;
        if (q < 3 * H) { ia = 2 * H + q / H; ib = B_R2 + q % H; return; }  // W_o
        q -= 3 * H;
        ia = 2 * H + q; ib = B_ONE;                                    // b_o
    }
};

constexpr int kWgradTK = 32;

// The weight gradients' first pass.  One LDS stage (TK rows) is staged in registers: every global
// load of the stage is issued before any is used (fully unrolled loops of compile-time trip count),
// so a stage costs one memory round trip, and the next stage's loads are in flight while the
// current one is summed.  The H-wide row arrays (dr1, dr2, r1, r2) move as float4 chunks; F (rows
// 9 + C floats apart), do, ds and the directions as single floats.
template <int C, int H>
__global__ void __launch_bounds__((WgradLayout<C, H>::THREADS))
wgrad_partial_kernel(WgradArgs p) {
    using L = WgradLayout<C, H>;
    constexpr int TK = kWgradTK;
    constexpr int QH = H / 4;
    constexpr int NQ = (TK * QH + L::THREADS - 1) / L::THREADS;
    constexpr int NF = (TK * C + L::THREADS - 1) / L::THREADS;
    constexpr int NS = (TK * 4 + L::THREADS - 1) / L::THREADS;
    constexpr int ND = (TK * 3 + L::THREADS - 1) / L::THREADS;
    __shared__ __align__(16) float sa[TK][L::LA];
    __shared__ __align__(16) float sb[TK][L::LB];
    const int tid = threadIdx.x;
    const int64_t k_begin = (int64_t)blockIdx.x * p.rows_per_group;
    const int64_t k_end = k_begin + p.rows_per_group < p.n_rows ? k_begin + p.rows_per_group : p.n_rows;
    // constant columns: the 1 after PE and its zero pad, and the unit block
    constexpr int NPAD = L::CP - (C + kPE);
    for (int i = tid; i < TK * NPAD; i += L::THREADS)
        sb[i / NPAD][C + kPE + i % NPAD] = (i % NPAD) == 0 ? 1.0f : 0.0f;
    for (int i = tid; i < TK * 8; i += L::THREADS) sb[i / 8][L::B_ONE + i % 8] = (i % 8) == 0 ? 1.0f : 0.0f;
    int a0 = 0, b0 = 0;
    const bool active = tid < L::NT;
    if (active) L::tile(tid, a0, b0);
    float acc[4][8];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 8; ++j) acc[i][j] = 0.0f;

    float4 q1[NQ], q2[NQ], q3[NQ], q4[NQ];  // dr1, dr2, r1, r2
    float vf[NF], vs[NS], vd[ND];           // F, [do | ds], d
    // rows past the slab's end are staged as zeros (a zero a with a finite b adds nothing)
    auto load = [&](int64_t k0) {
#pragma unroll
        for (int u = 0; u < NQ; ++u) {
            const int i = tid + u * L::THREADS;
            const int64_t k = k0 + i / QH;
            const int j = i % QH;
            // values, not a select between a global address and a local zero (which would put
            // the zero in scratch memory and load through the selected address)
            q1[u] = q2[u] = q3[u] = q4[u] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (i < TK * QH && k < k_end) {
                q1[u] = reinterpret_cast<const float4*>(p.dr1 + k * H)[j];
                q2[u] = reinterpret_cast<const float4*>(p.dr2 + k * H)[j];
                q3[u] = reinterpret_cast<const float4*>(p.r1 + k * H)[j];
                q4[u] = reinterpret_cast<const float4*>(p.r2 + k * H)[j];
            }
        }
#pragma unroll
        for (int u = 0; u < NF; ++u) {
            const int i = tid + u * L::THREADS;
            const int64_t k = k0 + i / C;
            vf[u] = (i < TK * C && k < k_end) ? p.rows[k * (9 + C) + 9 + i % C] : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < NS; ++u) {
            const int i = tid + u * L::THREADS;
            const int64_t k = k0 + i / 4;
            const int c = i % 4;
            vs[u] = (i < TK * 4 && k < k_end) ? (c < 3 ? p.dout[k * 3 + c] : p.ds[k]) : 0.0f;
        }
        // the row's direction (a dependent load through ray[k])
#pragma unroll
        for (int u = 0; u < ND; ++u) {
            const int i = tid + u * L::THREADS;
            const int64_t k = k0 + i / 3;
            vd[u] = (i < TK * 3 && k < k_end) ? p.d[p.ray[k] * 3 + i % 3] : 0.0f;
        }
    };
    auto store = [&]() {
#pragma unroll
        for (int u = 0; u < NQ; ++u) {
            const int i = tid + u * L::THREADS;
            if (i < TK * QH) {
                const int r = i / QH, c = 4 * (i % QH);
                *reinterpret_cast<float4*>(&sa[r][c]) = q1[u];
                *reinterpret_cast<float4*>(&sa[r][H + c]) = q2[u];
                *reinterpret_cast<float4*>(&sb[r][L::B_R1 + c]) = q3[u];
                *reinterpret_cast<float4*>(&sb[r][L::B_R2 + c]) = q4[u];
            }
        }
#pragma unroll
        for (int u = 0; u < NF; ++u) {
            const int i = tid + u * L::THREADS;
            if (i < TK * C) sb[i / C][i % C] = vf[u];
        }
#pragma unroll
        for (int u = 0; u < NS; ++u) {
            const int i = tid + u * L::THREADS;
            if (i < TK * 4) sa[i / 4][2 * H + i % 4] = vs[u];
        }
        // PE(d) of the row's ray, element by element as positional_encoding forms it
#pragma unroll
        for (int u = 0; u < ND; ++u) {
            const int i = tid + u * L::THREADS;
            if (i < TK * 3) {
                const int r = i / 3, ax = i % 3;
                const float dv = vd[u];
                sb[r][C + ax] = dv;
#pragma unroll 1
                for (int l = 0; l < 4; ++l) {
                    float sv, cv;
                    sincospif((float)(1 << l) * dv, &sv, &cv);
                    sb[r][C + 3 + 6 * l + ax] = sv;
                    sb[r][C + 6 + 6 * l + ax] = cv;
                }
            }
        }
    };

    load(k_begin);
    for (int64_t k0 = k_begin; k0 < k_end; k0 += TK) {
        __syncthreads();  // the previous stage's sums have read the LDS
        store();
        __syncthreads();
        if (k0 + TK < k_end) load(k0 + TK);
        if (active) {
#pragma unroll 4
            for (int r = 0; r < TK; ++r) {
                const float4 av = *reinterpret_cast<const float4*>(&sa[r][a0]);
                const float4 bl = *reinterpret_cast<const float4*>(&sb[r][b0]);
                const float4 bh = *reinterpret_cast<const float4*>(&sb[r][b0 + 4]);
                const float a[4] = {av.x, av.y, av.z, av.w};
                const float b[8] = {bl.x, bl.y, bl.z, bl.w, bh.x, bh.y, bh.z, bh.w};
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 8; ++j) acc[i][j] = fmaf(a[i], b[j], acc[i][j]);
            }
        }
    }
    if (active) {
        float* out = p.partial + ((int64_t)blockIdx.x * L::NT + tid) * 32;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            *reinterpret_cast<float4*>(out + i * 8) = make_float4(acc[i][0], acc[i][1], acc[i][2], acc[i][3]);
            *reinterpret_cast<float4*>(out + i * 8 + 4) = make_float4(acc[i][4], acc[i][5], acc[i][6], acc[i][7]);
        }
    }
}

// Second pass: 16 gradient elements per workgroup, the workgroups' partials split over 16 slices
// (slice s adds groups s, s + 16, ...), then the 16 slice sums added in slice order: a fixed order.
constexpr int kWgradSlices = 16;

template <int C, int H>
__global__ void __launch_bounds__(256) wgrad_reduce_kernel(const float* __restrict__ partial,
                                                           int groups, float* __restrict__ grad) {
    using L = WgradLayout<C, H>;
    __shared__ double red[kWgradSlices][16];
    const int el = threadIdx.x & 15, sl = threadIdx.x >> 4;
    const int e = blockIdx.x * 16 + el;
    double s0 = 0.0, s1 = 0.0;
    if (e < L::NP) {
        int ia, ib;
        L::element(e, ia, ib);
        const int64_t at = (int64_t)L::tile_of(ia, ib) * 32 + (ia & 3) * 8 + (ib & 7);
        const int64_t stride = (int64_t)L::NT * 32;
        int g = sl;
        for (; g + kWgradSlices < groups; g += 2 * kWgradSlices) {
            s0 += (double)partial[g * stride + at];
            s1 += (double)partial[(g + kWgradSlices) * stride + at];
        }
        if (g < groups) s0 += (double)partial[g * stride + at];
    }
    red[sl][el] = s0 + s1;
    __syncthreads();
    if (sl == 0 && e < L::NP) {
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < kWgradSlices; ++j) s += red[j][el];
        grad[e] = (float)s;
    }
}

int wgrad_tiles(int C, int H) {
    if (C == 32 && H == 32) return WgradLayout<32, 32>::NT;
    if (C == 32 && H == 64) return WgradLayout<32, 64>::NT;
    if (C == 64 && H == 64) return WgradLayout<64, 64>::NT;
    return 0;
}

// LDS stages of TK rows per workgroup: the fewest that keep the workgroups within the maximum
static int64_t wgrad_stages_per_group(int64_t n_rows) {
    const int64_t tiles = (n_rows + kWgradTK - 1) / kWgradTK;
    return (tiles + kWgradMaxGroups - 1) / kWgradMaxGroups;
}

int64_t wgrad_groups(int64_t n_rows) {
    const int64_t tiles = (n_rows + kWgradTK - 1) / kWgradTK;
    if (tiles == 0) return 1;
    const int64_t per = wgrad_stages_per_group(n_rows);
    return (tiles + per - 1) / per;
}

template <int C, int H>
static hipError_t launch_wgrad_ch(WgradArgs p, float* grad, hipStream_t s) {
    using L = WgradLayout<C, H>;
    const int64_t groups = wgrad_groups(p.n_rows);
    p.rows_per_group = wgrad_stages_per_group(p.n_rows) * kWgradTK;
    hipLaunchKernelGGL((wgrad_partial_kernel<C, H>), dim3((unsigned)groups), dim3(L::THREADS), 0, s, p);
    hipLaunchKernelGGL((wgrad_reduce_kernel<C, H>), dim3((L::NP + 15) / 16), dim3(256), 0, s,
                       (const float*)p.partial, (int)groups, grad);
    return hipGetLastError();
}

// Allocation-free form of the weight gradients: the per-workgroup partials live in this file-scope
// device array instead of a caller's workspace.  Its size covers the largest partial any row count
// needs (the first pass never uses more than kWgradMaxGroups workgroups), 23.1 MB per device, held
// until the process ends.  Host code never passes `scratch` itself: the array's device address is
// looked up once per device (hipGetSymbolAddress, with that device current) and kept here.
constexpr int64_t kWgradScratchFloats = (int64_t)kWgradMaxGroups * WgradLayout<64, 64>::NT * 32;
static_assert(WgradLayout<64, 64>::NT >= WgradLayout<32, 64>::NT &&
              WgradLayout<64, 64>::NT >= WgradLayout<32, 32>::NT, "largest partial");
static __device__ __align__(256) float scratch[kWgradScratchFloats];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
float* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

hipError_t wgrad_scratch(float** ptr, int64_t* bytes) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* addr = nullptr;
        e = hipGetSymbolAddress(&addr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[dev] = static_cast<float*>(addr);
    }
    *ptr = g_scratch_ptr[dev];
    *bytes = kWgradScratchFloats * (int64_t)sizeof(float);
    return hipSuccess;
}

hipError_t launch_weight_grads(int variant, const WgradArgs& p, float* grad, hipStream_t s) {
    switch (variant) {
        case 0: return launch_wgrad_ch<32, 32>(p, grad, s);
        case 1: return launch_wgrad_ch<32, 64>(p, grad, s);
        case 2: return launch_wgrad_ch<64, 64>(p, grad, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace diver
