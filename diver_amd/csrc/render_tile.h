// Fused render kernel for decoders whose density comes from the integrated feature alone
// (canonical form with no pre-layers): traversal and density run per lane, colour decoding runs
// in dense 64-row matrix-core batches.
//
// One wavefront owns a tile of 64 consecutive ray ids of the launch (one ray per lane); the ids
// run over a table of requests (camera views in 8x8 pixel tiles, or explicit ray batches), so
// rays of several requests of different sizes share one launch.  A ray's result depends only on
// its own intervals (its colour sum adds them front to back), never on the rays it shares a
// wave or a launch with.  The wave alternates two phases:
//   A. every live lane advances the exact DDA one voxel per iteration (lanes never wait on each
//      other's hits); at an occupied voxel it integrates only what density needs (channel 0 for
//      softplus(f[0]) decoders), composites transmittance (compensated log-sum, early stop at
//      T < t_stop) and pushes the row (entry, exit, weight = T*alpha, owner lane) into the
//      wave's LDS queue.  The phase ends when >= 32 rows are queued or every lane is done.
//   B. the wave takes up to 64 queued rows: full FP32 segment integral of each row's voxel, then
//      the colour decoder on the matrix cores, and an LDS atomic add of weight * colour into the
//      owner ray's accumulator.  BF16/FP16 decoders: 16-row tiles with v_mfma_f32_16x16x32, the
//      integral gathered by 4 lanes per row so that each quad of lanes reads one 64-byte
//      half-line per load (shade_batch16).  FP32 decoder: one row per lane, two 32-row column
//      tiles with v_mfma_f32_32x32x2_f32.  The per-ray view term W_v PE(d) + b_c is computed
//      once per ray in FP32 and kept in LDS (as BF16 pairs for the BF16 decoder).
// Empty-space skipping: when a lane enters a new 4^3 brick it reads that brick's bit of the 8^3-block
// occupancy bitmap (4 KB at 256^3, cached in L1/L2) and the brick's 64-bit mask word together,
// and caches both; between full iterations it takes cheap steps while it stays in an empty block
// or on empty voxels of its brick (skip_empty).  Skipping changes no arithmetic of the walk, so
// the intervals are bit-identical to the drop-in traversal kernels.
// Feature addressing: when the grid's bytes fit 32 bits (B32; 256^3 x 32 channels is 2.2 GB) the
// corner offsets are formed in 32 bits and every load takes the grid pointer as its scalar base;
// larger grids (256^3 x 64 channels) use the 64-bit form (render_tile_wide.hip).
// Occupancy: a wave needs 7.75 KB of LDS and at most 128 VGPRs with the BF16 decoder at hidden
// width 32, so 16 waves fit a CU (10 with the bitmap copied to LDS and FP32 view terms).
#pragma once
#include "common.h"
#include "decoder.h"
#include "kernels.h"
#include "mfma.h"
#include "raystate.h"
#include "requests.h"

namespace diver {

constexpr int kQueue = 96;       // rows per wave queue (phase A stops at >= 32, adds <= 64)
constexpr int kFill = 32;        // phase A hands over to phase B at this many queued rows
constexpr int kRowFloats = 8;    // p_in (3), p_out (3), weight, owner (as int bits)
constexpr int kMaxLdsBlockWords = 8192;
// One wave per workgroup: a workgroup's LDS is released only when its last wave ends, so with
// several tiles per group a long tile would hold the LDS of its finished neighbours; single-wave
// groups also let the dispatcher spread consecutive tiles over the XCDs one by one.
constexpr int kWavesPerGroup = 1;
constexpr int kBlocksInLds = 0;    // 8^3-block bitmap copied to LDS (1) or read through L1 (0)
// A wave still walking after this many phase-A iterations gets the highest issue priority: the
// long central tiles finish sooner while short waves fill the gaps (2 % on batched frames).
constexpr int kLongWalk = 256;
constexpr int kCornerRound = 4;   // 32-channel corner rows whose loads are in flight together (16-bit phase B)

// The per-ray view terms of the BF16 decoder are kept as packed BF16 pairs, the decoder's operand
// precision (half the LDS of FP32 terms); the FP32 and FP16 decoders keep them in FP32.
__host__ __device__ constexpr bool view_packed(int prec) { return prec == kPrecBF16; }
__host__ __device__ constexpr int wave_lds_floats(int H, int prec) {
    return kQueue * kRowFloats + 64 * 3 + (view_packed(prec) ? 32 * H : 64 * H);
}

__device__ __forceinline__ float bf16_lo(uint32_t w) { return __builtin_bit_cast(float, w << 16); }
__device__ __forceinline__ float bf16_hi(uint32_t w) { return __builtin_bit_cast(float, w & 0xffff0000u); }

// relu of two 16x16 D tiles (units 4g + 0..3 of each) packed as one 16x16x32 B operand: k slots
// 0..3 from the first tile, 4..7 from the second.
template <int PREC>
__device__ __forceinline__ u32x4_t relu_b16(const f32x4_t& d0, const f32x4_t& d1) {
    return u32x4_t{pack2<PREC>(fmaxf(d0[0], 0.0f), fmaxf(d0[1], 0.0f)),
                   pack2<PREC>(fmaxf(d0[2], 0.0f), fmaxf(d0[3], 0.0f)),
                   pack2<PREC>(fmaxf(d1[0], 0.0f), fmaxf(d1[1], 0.0f)),
                   pack2<PREC>(fmaxf(d1[2], 0.0f), fmaxf(d1[3], 0.0f))};
}

// shade_tile with FP32 operands (v_mfma_f32_32x32x2_f32).  Packed A operands (floats, 64 per k
// step): head W_c in natural k pairs (2kc, 2kc + 1) for (mt, kc); post and output layers in the
// D order of d_order_feature for (mt, q), so relu(D register v of tile mt) is the B operand of
// step 16 mt + v.  The output layer is padded to 32 rows.
template <class D>
__device__ __forceinline__ void shade_tile_f32(const RenderArgs& a, const float* queue,
                                               const float* hvs, float* acc, int t, int nbatch,
                                               const float (&b)[D::Z / 2], int lane) {
    constexpr int H = D::H, Z = D::Z;
    constexpr int MT = H / 32, KH = Z / 2, KP = H / 2;
    const float* pk = reinterpret_cast<const float*>(a.pack);
    // The operands are the same for every batch; an opaque pointer keeps their loads inside the
    // batch (held across the walk they would take 128 registers at hidden width 64).
    asm volatile("" : "+s"(pk));
    const int c = lane & 31;
    const int half = lane & 32;
    const int row = 32 * t + c;
    const bool live = row < nbatch;
    const int owner = live ? __builtin_bit_cast(int, queue[row * kRowFloats + 7]) : 0;
    f32x16_t y[MT];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            float4 h4 = *reinterpret_cast<const float4*>(hvs + owner * H + 32 * mt + 8 * g + (half >> 3));
            y[mt][4 * g + 0] = h4.x;
            y[mt][4 * g + 1] = h4.y;
            y[mt][4 * g + 2] = h4.z;
            y[mt][4 * g + 3] = h4.w;
        }
        y[mt] = mfma_chain_f32<KH>(pk + mt * KH * 64, lane, b, y[mt]);
    }
    int pofs = MT * KH * 64;
    static_assert(D::NPOST <= 1, "one post layer at most");
    if constexpr (D::NPOST == 1) {
        float e[KP];
#pragma unroll
        for (int q = 0; q < KP; ++q) e[q] = fmaxf(y[q / 16][q % 16], 0.0f);
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
            f32x16_t s;
#pragma unroll
            for (int i = 0; i < 16; ++i)
                s[i] = a.W[D::post_off(0) + H * H + 32 * mt + 8 * (i >> 2) + (half >> 3) + (i & 3)];
            y[mt] = mfma_chain_f32<KP>(pk + pofs + mt * KP * 64, lane, e, s);
        }
        pofs += MT * KP * 64;
    }
    float e[KP];
#pragma unroll
    for (int q = 0; q < KP; ++q) e[q] = fmaxf(y[q / 16][q % 16], 0.0f);
    f32x16_t o = {};
    o = mfma_chain_f32<KP>(pk + pofs, lane, e, o);
    if (lane < 32 && live) {
        const float* bo = a.W + D::out_off + 3 * H;
        const float wgt = queue[row * kRowFloats + 6];
        atomicAdd(&acc[owner * 3 + 0], wgt * sigmoidf(o[0] + bo[0]));
        atomicAdd(&acc[owner * 3 + 1], wgt * sigmoidf(o[1] + bo[1]));
        atomicAdd(&acc[owner * 3 + 2], wgt * sigmoidf(o[2] + bo[2]));
    }
}

// Phase B of the BF16/FP16 decoders with v_mfma_f32_16x16x32: the batch's rows are shaded 16 at a
// time (one tile t).  The gather runs with 4 lanes per row: lane l takes row 16t + l/4 and the
// 16-byte piece p = l%4 of each 64-byte half of the corners' feature rows, i.e. channels
// 32kk + 4p + (0..3) and 32kk + 16 + 4p + (0..3) of each 32-wide k step kk.  The 4 lanes of a
// quad then read one 64-byte half-line per load instruction (one cache access per quad, against
// four when each lane gathers a row of its own: the texture path is what limits batched frames).
// One ds_bpermute per dword moves the sums to the MFMA layout (lane 16g + c: row 16t + c, lane
// group g), and the head's A operands take their k values in the same channel order.  The hidden
// layers keep their D tiles in registers: relu(D tiles 2kk, 2kk + 1) is the next layer's B
// operand for k step kk, its A operands packed in that unit order.
template <class D, int PREC, bool B32>
__device__ __forceinline__ void shade_batch16(const RenderArgs& a, const float* queue,
                                              const float* hvs, float* acc, int nbatch, int lane,
                                              int n) {
    constexpr int C = D::C, H = D::H;
    constexpr int KS0 = C / 32, MT = H / 16, KSH = H / 32;
    static_assert(D::Z == C && C % 32 == 0 && H % 32 == 0, "16x16x32 tiles need C, H in 32s");
    static_assert(D::NPOST <= 1, "one post layer at most");
    constexpr int post_pk = MT * KS0;                  // A-operand steps of the head layer
    constexpr int out_pk = post_pk + D::NPOST * MT * KSH;
    const u32x4_t* pk = reinterpret_cast<const u32x4_t*>(a.pack16);
    const int col = lane & 15, g = lane >> 4;          // MFMA layout
    const int gr = lane >> 2, gp = lane & 3;           // gather layout
    const int from = (4 * col + g) * 4;                // byte index of the gather lane of (col, g)
    for (int t = 0; 16 * t < nbatch; ++t) {
        const int grow = 16 * t + gr;
        u32x4_t b[KS0];
        if (grow < nbatch) {
            const float4 r0 = *reinterpret_cast<const float4*>(queue + grow * kRowFloats);
            const float4 r1 = *reinterpret_cast<const float4*>(queue + grow * kRowFloats + 4);
            float w[8];
            uint32_t off[8];
            segment_corners<C, uint32_t>(r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, n, w, off);
            // each channel is the FP32 sum over corners j = 0..7 in that order; the loads go in
            // rounds of kCornerRound / KS0 corners (every k step of a corner together), eight
            // 16-byte loads in flight per round whatever the width
            constexpr int kR = kCornerRound / KS0;
            static_assert(kR >= 1 && kCornerRound % KS0 == 0, "corner rounds");
            float s[KS0][8];
#pragma unroll
            for (int kk = 0; kk < KS0; ++kk)
#pragma unroll
                for (int c = 0; c < 8; ++c) s[kk][c] = 0.f;
#pragma unroll
            for (int j0 = 0; j0 < 8; j0 += kR) {
                float4 q0[kR][KS0], q1[kR][KS0];
#pragma unroll
                for (int j = 0; j < kR; ++j)
#pragma unroll
                    for (int kk = 0; kk < KS0; ++kk) {
                        const float* src = feature_at<B32>(a.voxels, off[j0 + j] + 32u * kk + 4u * gp);
                        q0[j][kk] = *reinterpret_cast<const float4*>(src);
                        q1[j][kk] = *reinterpret_cast<const float4*>(src + 16);
                    }
#pragma unroll
                for (int j = 0; j < kR; ++j)
#pragma unroll
                    for (int kk = 0; kk < KS0; ++kk) {
                        const float wj = w[j0 + j];
                        s[kk][0] += wj * q0[j][kk].x;
                        s[kk][1] += wj * q0[j][kk].y;
                        s[kk][2] += wj * q0[j][kk].z;
                        s[kk][3] += wj * q0[j][kk].w;
                        s[kk][4] += wj * q1[j][kk].x;
                        s[kk][5] += wj * q1[j][kk].y;
                        s[kk][6] += wj * q1[j][kk].z;
                        s[kk][7] += wj * q1[j][kk].w;
                    }
                __builtin_amdgcn_sched_barrier(0);
            }
#pragma unroll
            for (int kk = 0; kk < KS0; ++kk)
                b[kk] = u32x4_t{pack2<PREC>(s[kk][0], s[kk][1]), pack2<PREC>(s[kk][2], s[kk][3]),
                                pack2<PREC>(s[kk][4], s[kk][5]), pack2<PREC>(s[kk][6], s[kk][7])};
        } else {
#pragma unroll
            for (int kk = 0; kk < KS0; ++kk) b[kk] = u32x4_t{0u, 0u, 0u, 0u};  // +0 in both halves
        }
#pragma unroll
        for (int kk = 0; kk < KS0; ++kk)
#pragma unroll
            for (int q = 0; q < 4; ++q)
                b[kk][q] = (uint32_t)__builtin_amdgcn_ds_bpermute(from, (int)b[kk][q]);
        const int row = 16 * t + col;
        const bool live = row < nbatch;
        int owner = 0;
        float wgt = 0.0f;
        if (live) {
            const float2 ow = *reinterpret_cast<const float2*>(queue + row * kRowFloats + 6);
            wgt = ow.x;
            owner = __builtin_bit_cast(int, ow.y);
        }
        // head: W_c z + (W_v PE(d) + b_c) of the row's ray; lane holds units 16 mt + 4 g + i
        f32x4_t y[MT];
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
            if constexpr (view_packed(PREC)) {
                const uint32_t* hvp = reinterpret_cast<const uint32_t*>(hvs);
                const uint2 h = *reinterpret_cast<const uint2*>(hvp + owner * (H / 2) + 8 * mt + 2 * g);
                y[mt] = f32x4_t{bf16_lo(h.x), bf16_hi(h.x), bf16_lo(h.y), bf16_hi(h.y)};
            } else {
                const float4 h = *reinterpret_cast<const float4*>(hvs + owner * H + 16 * mt + 4 * g);
                y[mt] = f32x4_t{h.x, h.y, h.z, h.w};
            }
#pragma unroll
            for (int kk = 0; kk < KS0; ++kk)
                y[mt] = mfma16<PREC>(pk[(mt * KS0 + kk) * 64 + lane], b[kk], y[mt]);
        }
        if constexpr (D::NPOST == 1) {
            u32x4_t e[KSH];
#pragma unroll
            for (int kk = 0; kk < KSH; ++kk) e[kk] = relu_b16<PREC>(y[2 * kk], y[2 * kk + 1]);
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) {
                const float* bp = a.W + D::post_off(0) + H * H + 16 * mt + 4 * g;
                f32x4_t s = f32x4_t{bp[0], bp[1], bp[2], bp[3]};
#pragma unroll
                for (int kk = 0; kk < KSH; ++kk)
                    s = mfma16<PREC>(pk[(post_pk + mt * KSH + kk) * 64 + lane], e[kk], s);
                y[mt] = s;
            }
        }
        u32x4_t e[KSH];
#pragma unroll
        for (int kk = 0; kk < KSH; ++kk) e[kk] = relu_b16<PREC>(y[2 * kk], y[2 * kk + 1]);
        f32x4_t o = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kk = 0; kk < KSH; ++kk) o = mfma16<PREC>(pk[(out_pk + kk) * 64 + lane], e[kk], o);
        // outputs 0..2 of the row sit in registers 0..2 of lane group 0
        if (g == 0 && live) {
            const float* bo = a.W + D::out_off + 3 * H;
            atomicAdd(&acc[owner * 3 + 0], wgt * sigmoidf(o[0] + bo[0]));
            atomicAdd(&acc[owner * 3 + 1], wgt * sigmoidf(o[1] + bo[1]));
            atomicAdd(&acc[owner * 3 + 2], wgt * sigmoidf(o[2] + bo[2]));
        }
    }
}

// Per-lane state of phase A (one ray per lane) and the wave's queue count.
struct TileLane {
    RayWalk ray;
    OccCache occ;
    Kahan S;
    float T = 1.0f;
    int used = 0;       // intervals evaluated
    float reg = 0.0f;   // sparsity term: sum of log(1 + s^2 / 0.5) over the evaluated intervals
    int qn = 0;         // queued rows (wave-uniform)
    int iters = 0;      // phase-A iterations (wave-uniform)
    float tail = 0.0f;  // the walk ends once p.d passes this (tail_limit)
};

// Phase B's density step for rows queued without it (SEG).  Lane r holds row r's channel-0 integral
// s (the same corners, order and roundings as the walk's channel-0 sum); every row's sigma, alpha and
// sparsity term are formed in parallel into its (already consumed) entry slots, then each owner lane
// takes its rows in queue order, which is its walk order, and does what the walk does per interval:
// weight = T_before * alpha, the sparsity sum, the compensated sum of sigma and T = exp(-S).  The
// results are the same bits as the walk's.
template <class D>
__device__ __forceinline__ void defer_density(TileLane& L, const BatchArgs& p, float* queue, float s,
                                              bool mine, int nbatch, int lane) {
    if (mine) {
        float sigma, alpha;
        density_alpha<D>(s, sigma, alpha);
        float* rw = queue + lane * kRowFloats;
        rw[0] = sigma;
        rw[1] = alpha;
        rw[2] = p.sparsity ? sparsity_term(s) : 0.0f;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    for (int r = 0; r < nbatch; ++r) {
        float* rw = queue + r * kRowFloats;
        if (__builtin_bit_cast(int, rw[7]) == lane) {
            rw[6] = L.T * rw[1];
            if (p.sparsity) L.reg += rw[2];
            L.S.add(rw[0]);
            L.T = __expf(-L.S.s);
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Phase A: every live lane walks until the queue holds kFill rows or every lane is done.  AF: every
// live ray of the wave takes the reciprocal-division step (the usual case); such a walk needs no
// step bound, since each step moves one axis to its next integer plane ahead and the walk leaves
// the grid after at most 3(N+1) steps.
// SEG (the training forward: t_stop = 0, alpha_skip = 0, channel-0 density): record each evaluated
// interval's entry and exit at seg_base + its index along the ray, and queue it without its density
// (no early termination, so nothing on the walk depends on it); phase B adds the densities.
template <class D, bool RT, bool DCH0, bool AF, bool B32, bool SEG>
__device__ __forceinline__ void phase_a(TileLane& L, const BatchArgs& p, const uint32_t* blocks,
                                        float* queue, int lane, int64_t seg_base) {
    constexpr int C = D::C;
    const RenderArgs& a = p.g;
    const int n = a.n;
    const float nf = (float)n;
    RayWalk& ray = L.ray;
    while (L.qn < kFill) {
        if (!wave_any(ray.alive)) break;
        ++L.iters;
        if constexpr (kSkip > 0) skip_empty<RT, AF>(ray, L.occ, n);
        bool push = false;
        float px, py, pz, qx, qy, qz, wgt;  // read only where push is set
        if (ray.alive) {
            if (!AF && ray.steps >= ray.max_steps) {
                ray.alive = false;
            } else {
                if constexpr (!AF) ++ray.steps;
                float nx, ny, nz;
                if constexpr (AF)
                    dda_step_r(ray.lx, ray.ly, ray.lz, ray.dx, ray.dy, ray.dz, ray.rx, ray.ry, ray.rz,
                               ray.sx, ray.sy, ray.sz, nx, ny, nz);
                else
                    ray.step(ray.lx, ray.ly, ray.lz, nx, ny, nz);
                if ((RT ? outside_open(nx, ny, nz, nf) : outside(nx, ny, nz, nf)) ||
                    past_tail(ray, nx, ny, nz, L.tail)) {
                    ray.alive = false;
                } else {
                    int vx, vy, vz;
                    ray.voxel(nx, ny, nz, vx, vy, vz);
                    px = ray.lx; py = ray.ly; pz = ray.lz;
                    ray.lx = nx; ray.ly = ny; ray.lz = nz;
                    if (SEG && L.occ.lookup(a.mask, blocks, a.bricks, n, vx, vy, vz)) {
                        // training forward (no early termination): the row is queued without its
                        // density, which phase B works out for the whole batch at once
                        qx = nx; qy = ny; qz = nz;
                        L.used += 1;
                        if (L.used <= p.seg_cap) {
                            float2* sp = reinterpret_cast<float2*>(p.segments + (seg_base + L.used - 1) * 6);
                            sp[0] = make_float2(px, py);
                            sp[1] = make_float2(pz, qx);
                            sp[2] = make_float2(qy, qz);
                        }
                        wgt = 0.0f;
                        push = true;
                    } else if (!SEG && L.occ.lookup(a.mask, blocks, a.bricks, n, vx, vy, vz)) {
                        qx = nx; qy = ny; qz = nz;
                        float w[8];
                        uint32_t off[8];
                        segment_corners_cell<C, uint32_t>(vx, vy, vz, px, py, pz, qx, qy, qz, n, w, off);
                        float s;
                        if constexpr (DCH0) {
                            s = 0.0f;
#pragma unroll
                            for (int j = 0; j < 8; ++j) s += w[j] * *feature_at<B32>(a.voxels, off[j]);
                        } else {
                            float f[C];
                            gather_integral_at<C, B32>(a.voxels, w, off, f);
                            s = a.W[D::bs_off];
#pragma unroll
                            for (int i = 0; i < C; ++i) s += a.W[D::ws_off + i] * f[i];
                        }
                        float sigma, alpha;
                        density_alpha<D>(s, sigma, alpha);
                        L.used += 1;
                        if (p.sparsity) L.reg += sparsity_term(s);
                        if (alpha >= a.alpha_skip) {
                            wgt = L.T * alpha;
                            push = true;
                            L.S.add(sigma);
                            L.T = __expf(-L.S.s);
                            if (L.T < a.t_stop) ray.alive = false;
                        }
                    }
                }
            }
        }
        const uint64_t pm = __builtin_amdgcn_ballot_w64(push);
        if (push) {
            const int slot = L.qn + __builtin_popcountll(pm & ((1ull << lane) - 1ull));
            float* rw = queue + slot * kRowFloats;
            *reinterpret_cast<float4*>(rw) = make_float4(px, py, pz, qx);
            *reinterpret_cast<float4*>(rw + 4) = make_float4(qy, qz, wgt, __builtin_bit_cast(float, lane));
        }
        L.qn += __builtin_popcountll(pm);
    }
}

// Register budget: 3 waves per SIMD (168 VGPRs) for hidden width 32 (the BF16 kernel compiles to
// at most 128 VGPRs, which allows 4).  Hidden width 64 takes the 2-wave budget (the 32-channel
// kernels compile to about 130 VGPRs and run 3 waves anyway; the FP32 and linear-density forms
// need more than 168), except the 16-bit channel-0 kernels at 64 channels: held to 168 VGPRs they
// park the ray's output slot and request index (12 bytes) in scratch, stored once before the
// walk and loaded once after it, and shade a 256^3 x 64 scene 13 % faster than at 2 waves.
template <class D, int PREC, bool DCH0>
constexpr int tile_waves_per_simd() {
    if (D::H <= 32) return 3;
    return D::C > 32 && PREC != kPrecFP32 && DCH0 ? 3 : 2;
}

template <class D, int PREC, bool RT, bool DCH0, bool B32, bool SEG = false>
__global__ void __launch_bounds__(64 * kWavesPerGroup, (tile_waves_per_simd<D, PREC, DCH0>())) render_tile_kernel(BatchArgs p) {
    static_assert(D::NPRE == 0, "the tile kernel needs density from the integrated feature");
    constexpr int C = D::C, H = D::H, Z = D::Z;
    const RenderArgs& a = p.g;
    extern __shared__ uint32_t smem[];
    const int n = a.n;
    const int nb = n >> 3;
    const int nwords = kBlocksInLds ? (nb * nb * nb + 31) >> 5 : 0;
    const uint32_t* blocks = kBlocksInLds ? smem : a.blocks;
    float* wave_mem = reinterpret_cast<float*>(smem + ((nwords + 3) & ~3));

    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    float* queue = wave_mem + wid * wave_lds_floats(H, PREC);
    float* acc = queue + kQueue * kRowFloats;
    float* hvs = acc + 64 * 3;  // [64][H] per-ray view term (FP32, or [64][H / 2] packed pairs)
    acc[lane * 3 + 0] = 0.0f;
    acc[lane * 3 + 1] = 0.0f;
    acc[lane * 3 + 2] = 0.0f;

    // lanes [0, rays_per_wave) of each wave take consecutive ray ids; the others stay idle
    const int rpw = p.rays_per_wave;
    const int64_t g = ((int64_t)blockIdx.x * kWavesPerGroup + wid) * rpw + lane;
    int req = 0;
    int64_t out_idx = 0;
    float ox = 0.f, oy = 0.f, oz = 0.f, dx = 0.f, dy = 0.f, dz = 1.f;
    const bool valid = lane < rpw && g < p.total &&
                       ray_of(p, g, req, out_idx, ox, oy, oz, dx, dy, dz);
    if (!valid) {
        ox = oy = oz = 0.f;
        dx = dy = 0.f;
        dz = 1.f;
    }
    TileLane L;
    if constexpr (RT)
        L.ray.init_rt(ox, oy, oz, dx, dy, dz, a.xyzmin, a.voxel_size, n, a.octrees, valid);
    else
        L.ray.init(ox, oy, oz, dx, dy, dz, a.xyzmin, a.voxel_size, (float)n, valid);
    L.tail = tail_limit(a.tail, n, L.ray);

    // Most tiles of a view miss the occupied cells entirely: such a wave skips the bitmap copy and
    // the view term (nothing would read them) and goes straight to writing T = 1.
    const bool live = kWavesPerGroup > 1 || wave_any(L.ray.alive);
    if (kBlocksInLds && live) {
        if ((nwords & 3) == 0 && (reinterpret_cast<uintptr_t>(a.blocks) & 15) == 0) {
            const uint4* src = reinterpret_cast<const uint4*>(a.blocks);
            uint4* dst = reinterpret_cast<uint4*>(smem);
            for (int i = threadIdx.x; i < (nwords >> 2); i += blockDim.x) dst[i] = src[i];
        } else {
            for (int i = threadIdx.x; i < nwords; i += blockDim.x) smem[i] = a.blocks[i];
        }
        __syncthreads();
    }
    if (live) {
        float hv[H];
        head_view<D>(a.W, dx, dy, dz, hv);
        if constexpr (view_packed(PREC)) {
            uint32_t* hvp = reinterpret_cast<uint32_t*>(hvs);
#pragma unroll
            for (int j = 0; j < H; j += 4)
                *reinterpret_cast<uint2*>(hvp + lane * (H / 2) + j / 2) =
                    make_uint2(pack2<kPrecBF16>(hv[j], hv[j + 1]), pack2<kPrecBF16>(hv[j + 2], hv[j + 3]));
        } else {
#pragma unroll
            for (int j = 0; j < H; j += 4)
                *reinterpret_cast<float4*>(hvs + lane * H + j) = make_float4(hv[j], hv[j + 1], hv[j + 2], hv[j + 3]);
        }
    }
    // every live ray's direction takes the reciprocal-division step (wave-uniform)
    const bool all_fast = !wave_any(L.ray.alive && !L.ray.fast);
    const int64_t seg_base = SEG ? out_idx * (int64_t)p.seg_cap : 0;

    while (true) {
        // ---------------- phase A: walk, density, transmittance, enqueue ------------------------
        if (all_fast)
            phase_a<D, RT, DCH0, true, B32, SEG>(L, p, blocks, queue, lane, seg_base);
        else
            phase_a<D, RT, DCH0, false, B32, SEG>(L, p, blocks, queue, lane, seg_base);
        const int qn = L.qn;
        if (qn == 0) break;  // nothing queued and every lane done
        // a wave still walking after kLongWalk iterations shades and walks at the highest issue
        // priority: the long central tiles finish sooner while short waves fill the gaps
        if (L.iters >= kLongWalk) __builtin_amdgcn_s_setprio(3);

        // ---------------- phase B: shade up to 64 queued rows -----------------------------------
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        const int nbatch = qn < 64 ? qn : 64;
        if constexpr (PREC == kPrecFP32) {
            const bool mine = lane < nbatch;
            float w[8];
            uint32_t off[8];
            if (mine) {
                float4 r0 = *reinterpret_cast<const float4*>(queue + lane * kRowFloats);
                float4 r1 = *reinterpret_cast<const float4*>(queue + lane * kRowFloats + 4);
                segment_corners<C, uint32_t>(r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, n, w, off);
            }
            float z[Z];
            if (mine) {
                gather_integral_at<C, B32>(a.voxels, w, off, z);
            } else {
#pragma unroll
                for (int i = 0; i < Z; ++i) z[i] = 0.0f;
            }
            if constexpr (SEG) defer_density<D>(L, p, queue, z[0], mine, nbatch, lane);
            float b0[Z / 2], b1[Z / 2];
            rows_to_b_f32<Z>(z, b0, b1);
            shade_tile_f32<D>(a, queue, hvs, acc, 0, nbatch, b0, lane);
            if (nbatch > 32) shade_tile_f32<D>(a, queue, hvs, acc, 1, nbatch, b1, lane);
        } else {
            shade_batch16<D, PREC, B32>(a, queue, hvs, acc, nbatch, lane, n);
        }
        // drop the shaded rows: move rows [64, qn) down to [0, qn - 64)
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        const int rest = qn - nbatch;
        const bool move = lane < rest;
        float4 m0 = make_float4(0.f, 0.f, 0.f, 0.f), m1 = m0;
        if (move) {
            m0 = *reinterpret_cast<const float4*>(queue + (nbatch + lane) * kRowFloats);
            m1 = *reinterpret_cast<const float4*>(queue + (nbatch + lane) * kRowFloats + 4);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        if (move) {
            *reinterpret_cast<float4*>(queue + lane * kRowFloats) = m0;
            *reinterpret_cast<float4*>(queue + lane * kRowFloats + 4) = m1;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        L.qn = rest;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    if (valid) {
        const ReqDesc& rq = p.req[req];
        reinterpret_cast<float4*>(rq.rgba)[out_idx] =
            make_float4(acc[lane * 3], acc[lane * 3 + 1], acc[lane * 3 + 2], L.T);
        if (rq.count) rq.count[out_idx] = L.used;
        if (rq.sparsity) rq.sparsity[out_idx] = L.reg;
    }
}

size_t render_tile_lds_bytes(int n, int H, int prec);

template <class D, int PREC, bool RT, bool B32>
hipError_t launch_tiles3(const BatchArgs& p, bool dch0, hipStream_t s) {
    constexpr int threads = 64 * kWavesPerGroup;
    const int64_t per_group = (int64_t)p.rays_per_wave * kWavesPerGroup;
    dim3 grid((unsigned)((p.total + per_group - 1) / per_group)), block(threads);
    size_t lds = render_tile_lds_bytes(p.g.n, D::H, PREC);
    if (p.segments) {
        // segments are recorded by the training forward only: FP32, offline walk, channel-0 density
        if constexpr (PREC == kPrecFP32 && !RT) {
            if (!dch0 || p.n_req != 1) return hipErrorInvalidValue;
            hipLaunchKernelGGL((render_tile_kernel<D, PREC, RT, true, B32, true>), grid, block, lds, s, p);
            return hipGetLastError();
        } else {
            return hipErrorInvalidValue;
        }
    }
    if (dch0)
        hipLaunchKernelGGL((render_tile_kernel<D, PREC, RT, true, B32>), grid, block, lds, s, p);
    else
        hipLaunchKernelGGL((render_tile_kernel<D, PREC, RT, false, B32>), grid, block, lds, s, p);
    return hipGetLastError();
}

template <class D, bool RT, bool B32>
hipError_t launch_tiles2(const BatchArgs& p, int prec, bool dch0, hipStream_t s) {
    if (prec == kPrecBF16) return launch_tiles3<D, kPrecBF16, RT, B32>(p, dch0, s);
    if (prec == kPrecFP16) return launch_tiles3<D, kPrecFP16, RT, B32>(p, dch0, s);
    return launch_tiles3<D, kPrecFP32, RT, B32>(p, dch0, s);
}

// Every tile kernel of one feature addressing (B32: 32-bit byte offsets) for the decoder
// variant (0 = DIVeR32 shape, 1 = hidden 64, 2 = 64 channels / hidden 64).
template <bool B32>
hipError_t launch_tiles(int variant, const BatchArgs& p, int prec, bool dch0, hipStream_t s) {
    const bool rt = p.g.traversal == 1;
    switch (variant) {
        case 0: return rt ? launch_tiles2<DecB32, true, B32>(p, prec, dch0, s)
                          : launch_tiles2<DecB32, false, B32>(p, prec, dch0, s);
        case 1: return rt ? launch_tiles2<DecB64, true, B32>(p, prec, dch0, s)
                          : launch_tiles2<DecB64, false, B32>(p, prec, dch0, s);
        case 2: return rt ? launch_tiles2<DecB64C64, true, B32>(p, prec, dch0, s)
                          : launch_tiles2<DecB64C64, false, B32>(p, prec, dch0, s);
        default: return hipErrorInvalidValue;
    }
}

// the kernels with 64-bit feature addressing, instantiated in render_tile_wide.hip
hipError_t launch_tiles_wide(int variant, const BatchArgs& p, int prec, bool dch0, hipStream_t s);

}  // namespace diver
