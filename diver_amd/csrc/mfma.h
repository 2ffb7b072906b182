// CDNA4 (gfx950) wavefront matrix helpers for the decoder: v_mfma_f32_32x32x16_{bf16,f16} with the
// "weights as A, rows as B" formulation, and v_permlane32_swap to move row vectors between the
// one-row-per-lane layout and the MFMA operand layouts without LDS.
//
// Layouts (wave64, 32x32x16):
//   A (32 x 16): lane l holds A[l % 32][8*(l/32) + 0..7]            (8 x 16-bit)
//   B (16 x 32): lane l holds B[8*(l/32) + 0..7][l % 32]
//   D (32 x 32): lane l, register i holds D[8*(i/4) + 4*(l/32) + (i%4)][l % 32]   (fp32)
// A data tile of 64 rows (one per lane) is processed as two 32-column tiles: tile 0 = rows of
// lanes 0..31, tile 1 = rows of lanes 32..63.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace diver {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8_t;
typedef __attribute__((ext_vector_type(8))) _Float16 f16x8_t;
typedef __attribute__((ext_vector_type(2))) __bf16 bf16x2_t;
typedef __attribute__((ext_vector_type(2))) _Float16 f16x2_t;
typedef __attribute__((ext_vector_type(16))) float f32x16_t;
typedef __attribute__((ext_vector_type(4))) uint32_t u32x4_t;

enum { kPrecFP32 = 0, kPrecBF16 = 1, kPrecFP16 = 2 };

template <int PREC>
__device__ __forceinline__ uint32_t pack2(float a, float b) {
    if constexpr (PREC == kPrecBF16) {
        bf16x2_t v = {(__bf16)a, (__bf16)b};
        return __builtin_bit_cast(uint32_t, v);
    } else {
        f16x2_t v = {(_Float16)a, (_Float16)b};
        return __builtin_bit_cast(uint32_t, v);
    }
}

// After the call: x holds {lanes 0-31: x, lanes 32-63: y from lanes 0-31},
//                 y holds {lanes 0-31: x from lanes 32-63, lanes 32-63: y}.
__device__ __forceinline__ void swap32(uint32_t& x, uint32_t& y) {
    auto r = __builtin_amdgcn_permlane32_swap(x, y, false, false);
    x = r[0];
    y = r[1];
}

__device__ __forceinline__ void swap32f(float& x, float& y) {
    uint32_t a = __builtin_bit_cast(uint32_t, x), b = __builtin_bit_cast(uint32_t, y);
    swap32(a, b);
    x = __builtin_bit_cast(float, a);
    y = __builtin_bit_cast(float, b);
}

template <int PREC>
__device__ __forceinline__ f32x16_t mfma32(u32x4_t a, u32x4_t b, f32x16_t c) {
    if constexpr (PREC == kPrecBF16) {
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8_t, a),
                                                       __builtin_bit_cast(bf16x8_t, b), c, 0, 0, 0);
    } else {
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(f16x8_t, a),
                                                      __builtin_bit_cast(f16x8_t, b), c, 0, 0, 0);
    }
}

// v_mfma_f32_16x16x32_{bf16,f16} (gfx950), lane layouts (checked against a host matmul):
//   A (16 x 32): lane l holds A[l % 16][8*(l/16) + 0..7]
//   B (32 x 16): lane l holds B[8*(l/16) + 0..7][l % 16]
//   D (16 x 16): lane l, register i holds D[4*(l/16) + i][l % 16]                (fp32)
typedef __attribute__((ext_vector_type(4))) float f32x4_t;

template <int PREC>
__device__ __forceinline__ f32x4_t mfma16(u32x4_t a, u32x4_t b, f32x4_t c) {
    if constexpr (PREC == kPrecBF16) {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_t, a),
                                                       __builtin_bit_cast(bf16x8_t, b), c, 0, 0, 0);
    } else {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8_t, a),
                                                      __builtin_bit_cast(f16x8_t, b), c, 0, 0, 0);
    }
}

// FP32 products on the matrix cores: v_mfma_f32_32x32x2_f32 (FP32 inputs, FP32 accumulation).
//   A (32 x 2): lane l holds A[l % 32][l / 32];   B (2 x 32): lane l holds B[l / 32][l % 32]
//   D: the 32x32 layout above.
// A hidden layer's D tiles feed the next layer's B operands in place when that layer's k steps
// follow the D layout: step q = 16*mt + v pairs feature 32*mt + 8*(v/4) + (v%4) (k = 0, lanes
// 0-31) with feature 32*mt + 8*(v/4) + 4 + (v%4) (k = 1, lanes 32-63), i.e. register v of D
// tile mt is the B operand of step q.  The packed A operands follow the same order.
__device__ __forceinline__ f32x16_t mfma32_f32(float a, float b, f32x16_t c) {
    return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
}

// acc += sum over k steps q < K of A(step q) B(bv[q]) with FP32 operands; the A operands (64
// floats per step at pk) are loaded 8 steps at a time, the next 8 in flight during the current
// products, which bounds the registers they hold.
template <int K>
__device__ __forceinline__ f32x16_t mfma_chain_f32(const float* __restrict__ pk, int lane,
                                                   const float (&bv)[K], f32x16_t acc) {
    static_assert(K % 8 == 0, "k steps in chunks of 8");
    float w[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) w[i] = pk[i * 64 + lane];
#pragma unroll
    for (int k0 = 0; k0 < K; k0 += 8) {
        float wn[8];
        if (k0 + 8 < K) {
#pragma unroll
            for (int i = 0; i < 8; ++i) wn[i] = pk[(k0 + 8 + i) * 64 + lane];
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) acc = mfma32_f32(w[i], bv[k0 + i], acc);
        __builtin_amdgcn_sched_barrier(0);
        if (k0 + 8 < K) {
#pragma unroll
            for (int i = 0; i < 8; ++i) w[i] = wn[i];
        }
    }
    return acc;
}

// Feature index of k step q, half h (0: lanes 0-31, 1: lanes 32-63) in the D-ordered k steps.
__host__ __device__ constexpr int d_order_feature(int q, int h) {
    return 32 * (q / 16) + 8 * ((q % 16) / 4) + 4 * h + (q % 4);
}

// Row vectors (K wide, one row per lane, fp32) -> FP32 B operands of both column tiles, one per
// k pair (features 2kc, 2kc + 1): b0[kc] feeds tile 0 (rows of lanes 0-31), b1[kc] tile 1.
template <int K>
__device__ __forceinline__ void rows_to_b_f32(const float (&z)[K], float (&b0)[K / 2],
                                              float (&b1)[K / 2]) {
#pragma unroll
    for (int kc = 0; kc < K / 2; ++kc) {
        float x = z[2 * kc], y = z[2 * kc + 1];
        swap32f(x, y);
        b0[kc] = x;
        b1[kc] = y;
    }
}

// Row vectors (K wide, one row per lane, fp32) -> B operands of both column tiles, one per
// 16-wide k chunk.  b0[kc] feeds tile 0, b1[kc] tile 1.
template <int PREC, int K>
__device__ __forceinline__ void rows_to_b(const float (&z)[K], u32x4_t (&b0)[K / 16],
                                          u32x4_t (&b1)[K / 16]) {
#pragma unroll
    for (int kc = 0; kc < K / 16; ++kc) {
        uint32_t x[4], y[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            x[q] = pack2<PREC>(z[16 * kc + 2 * q], z[16 * kc + 2 * q + 1]);
            y[q] = pack2<PREC>(z[16 * kc + 8 + 2 * q], z[16 * kc + 8 + 2 * q + 1]);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) swap32(x[q], y[q]);
        b0[kc] = u32x4_t{x[0], x[1], x[2], x[3]};
        b1[kc] = u32x4_t{y[0], y[1], y[2], y[3]};
    }
}

// One D tile (32 features x 32 rows, fp32) -> relu -> B operands of the two 16-wide k chunks it
// covers (same column tile).
template <int PREC>
__device__ __forceinline__ void d_to_b_relu(const f32x16_t& d, u32x4_t& c0, u32x4_t& c1) {
    float r[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) r[i] = fmaxf(d[i], 0.0f);
    // chunk kk covers register groups g = 2kk (Va) and 2kk+1 (Vb)
    uint32_t a0 = pack2<PREC>(r[0], r[1]), a1 = pack2<PREC>(r[2], r[3]);
    uint32_t b0 = pack2<PREC>(r[4], r[5]), b1 = pack2<PREC>(r[6], r[7]);
    swap32(a0, b0);
    swap32(a1, b1);
    c0 = u32x4_t{a0, a1, b0, b1};
    uint32_t e0 = pack2<PREC>(r[8], r[9]), e1 = pack2<PREC>(r[10], r[11]);
    uint32_t f0 = pack2<PREC>(r[12], r[13]), f1 = pack2<PREC>(r[14], r[15]);
    swap32(e0, f0);
    swap32(e1, f1);
    c1 = u32x4_t{e0, e1, f0, f1};
}

// Per-row fp32 vector (H wide, own row per lane) -> D-layout initial accumulators of both column
// tiles, one f32x16 per 32-feature m tile.
template <int H>
__device__ __forceinline__ void rows_to_d(const float (&hv)[H], f32x16_t (&d0)[H / 32],
                                          f32x16_t (&d1)[H / 32]) {
#pragma unroll
    for (int mt = 0; mt < H / 32; ++mt) {
#pragma unroll
        for (int g = 0; g < 4; ++g) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                float p = hv[32 * mt + 8 * g + r];
                float q = hv[32 * mt + 8 * g + 4 + r];
                swap32f(p, q);
                d0[mt][4 * g + r] = p;
                d1[mt][4 * g + r] = q;
            }
        }
    }
}

}  // namespace diver
