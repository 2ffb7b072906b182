// Occupancy extraction: the alpha map of prune.py:83-143 in one launch.
//
// For every training ray the reference walks the masked_intersect recurrence (offline walk, exit
// > N) to each occupied voxel, integrates the voxel's features over the segment, evaluates the
// decoder's density (pre-layers included) and forms the rendering weight of model/diver.py:194-202,
// w_i = alpha_i * T_i with T_i the transmittance in front of the interval (no early termination).
// The weight goes to the voxel min(int(p_in + 1e-4), int(p_out + 1e-4)) at flat index
// x + y N + z N^2 (prune.py:118-129) unless that voxel lies within `pad` of the border, and the map
// keeps the maximum (prune.py:132-133).  Weights are non-negative, so the maximum is an integer
// atomicMax on their bit patterns (sign bit cleared): exact and independent of the order in which
// lanes arrive, so repeated runs and batch splits give the same bits.
//
// alpha_tile_kernel (grids with block and brick bits): one ray per lane, three phases per wave.
//   A. every live lane walks its ray (cached 4^3-brick occupancy, cheap steps over empty space, as
//      the fused renderer walks) and queues each occupied voxel's segment in the wave's LDS, the
//      lane's rows linked in walk order, until 64 rows are queued or every lane is done.
//   B. all queued rows: the FP32 segment integral, one row per lane, then the pre-layers on the
//      matrix cores (v_mfma_f32_32x32x2_f32: weights as A, 32 rows as B, relu(D) the next layer's B
//      in place) and the density pre-activation as a per-lane dot over the last layer's D tile
//      plus the partner half-wave's; decoders without pre-layers take the dot directly.
//   C. every lane follows its own rows in walk order: alpha, the weight T * alpha, the
//      compensated transmittance update and the map's maximum.
// A row's pre-activation depends only on its own integral (each MFMA output column is one row),
// so the map does not depend on which rays share a wave or a launch.
// alpha_map_kernel: the same per lane (walk over the byte mask, FP32 density head on the vector
// units), for grids without the bit structures.
#include "common.h"
#include "decoder.h"
#include "kernels.h"
#include "mfma.h"
#include "raystate.h"

namespace diver {

template <class D>
__global__ void __launch_bounds__(256) alpha_map_kernel(RenderArgs a, int pad,
                                                        unsigned int* __restrict__ map) {
    constexpr int C = D::C, Z = D::Z;
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = r < a.n_rays;
    const int64_t rr = valid ? r : 0;
    const int n = a.n;
    RayWalk ray;
    ray.init(a.o[3 * rr], a.o[3 * rr + 1], a.o[3 * rr + 2], a.d[3 * rr], a.d[3 * rr + 1],
             a.d[3 * rr + 2], a.xyzmin, a.voxel_size, (float)n, valid);
    Kahan S;
    float T = 1.0f;
    const int64_t n1 = n, n2 = (int64_t)n * n;
    while (true) {
        float px = 0.f, py = 0.f, pz = 0.f, qx = 0.f, qy = 0.f, qz = 0.f;
        const bool has = ray.next<false>(a.mask, n, px, py, pz, qx, qy, qz);
        if (!wave_any(has)) break;
        if (!has) continue;
        float w[8];
        int64_t off[8];
        segment_corners<C>(px, py, pz, qx, qy, qz, n, w, off);
        float f[C];
        gather_integral<C>(a.voxels, w, off, f);
        float z[Z];
        float sigma, alpha;
        density_alpha<D>(pre_and_sigma<D>(a.W, f, z), sigma, alpha);
        const float wgt = alpha * T;
        S.add(sigma);
        T = __expf(-S.s);
        int cx, cy, cz;
        {
#pragma clang fp contract(off)
            cx = min((int)(px + 1e-4f), (int)(qx + 1e-4f));
            cy = min((int)(py + 1e-4f), (int)(qy + 1e-4f));
            cz = min((int)(pz + 1e-4f), (int)(qz + 1e-4f));
        }
        const int hi = n - 1 - pad;
        // the weight is >= 0 but may be -0 (relu(-0) = -0 gives alpha = -0), whose bit pattern
        // would beat every positive weight: store its magnitude
        if (cx >= pad && cy >= pad && cz >= pad && cx <= hi && cy <= hi && cz <= hi)
            atomicMax(map + (cx + cy * n1 + cz * n2), __float_as_uint(wgt) & 0x7fffffffu);
    }
}

// ------------------------------- tile kernel ----------------------------------------------------

constexpr int kAlphaQueue = 128;   // rows per wave queue (phase A stops at >= 64, adds <= 63)
constexpr int kAlphaFill = 64;
constexpr int kAlphaRow = 8;       // p_in (3), p_out (3), next row of the lane (int bits), s

// Phase A: walk, queue the occupied voxels' segments.  head/last: the lane's first and latest row
// of this batch (-1: none); each row's link names the lane's next row.
template <bool AF>
__device__ __forceinline__ void alpha_walk(RayWalk& ray, OccCache& occ, const RenderArgs& a,
                                           float* queue, int lane, int& qn, int& head, int& last,
                                           float tail) {
    const int n = a.n;
    const float nf = (float)n;
    while (qn < kAlphaFill) {
        if (!wave_any(ray.alive)) break;
        skip_empty<false, AF>(ray, occ, n);
        bool push = false;
        float px, py, pz, qx, qy, qz;  // read only where push is set
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
                if (outside(nx, ny, nz, nf) || past_tail(ray, nx, ny, nz, tail)) {
                    ray.alive = false;
                } else {
                    int vx, vy, vz;
                    ray.voxel(nx, ny, nz, vx, vy, vz);
                    px = ray.lx; py = ray.ly; pz = ray.lz;
                    ray.lx = nx; ray.ly = ny; ray.lz = nz;
                    if (occ.lookup(a.mask, a.blocks, a.bricks, n, vx, vy, vz)) {
                        qx = nx; qy = ny; qz = nz;
                        push = true;
                    }
                }
            }
        }
        const uint64_t pm = __builtin_amdgcn_ballot_w64(push);
        if (push) {
            const int slot = qn + __builtin_popcountll(pm & ((1ull << lane) - 1ull));
            float* rw = queue + slot * kAlphaRow;
            *reinterpret_cast<float4*>(rw) = make_float4(px, py, pz, qx);
            *reinterpret_cast<float4*>(rw + 4) = make_float4(qy, qz, __builtin_bit_cast(float, -1), 0.0f);
            if (last >= 0)
                queue[last * kAlphaRow + 6] = __builtin_bit_cast(float, slot);
            else
                head = slot;
            last = slot;
        }
        qn += __builtin_popcountll(pm);
    }
}

// Pre-activation s = w_s . z + b_s of the 32 rows whose integrals are the FP32 B operands b (one
// per natural k pair), z the last pre-layer's output; written to the rows' s field.
template <class D>
__device__ __forceinline__ void alpha_density_tile(const float* W, const float* pk,
                                                   const float (&b)[D::C / 2],
                                                   int lane, float* queue, int row0, int nrows) {
    constexpr int C = D::C, H = D::H, MT = H / 32, KC = C / 2, KH = H / 2;
    static_assert(D::NPRE >= 1 && !D::PREID && H % 32 == 0 && C % 16 == 0, "pre-layer shape");
    // The operands, biases and head weights are the same for every batch; opaque pointers keep
    // their loads inside the batch (hoisted out of it they would stay live across the walk).
    asm volatile("" : "+s"(pk));
    asm volatile("" : "+s"(W));
    const int h = lane >> 5, c = lane & 31;
    // D-layout unit of register i of tile mt: 32 mt + 8 (i / 4) + 4 h + i % 4
    f32x16_t y[MT];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
        f32x16_t bias;
#pragma unroll
        for (int i = 0; i < 16; ++i) bias[i] = W[D::pre_off(0) + H * C + 32 * mt + 8 * (i >> 2) + 4 * h + (i & 3)];
        y[mt] = mfma_chain_f32<KC>(pk + mt * KC * 64, lane, b, bias);
    }
    int pofs = MT * KC * 64;
#pragma unroll
    for (int l = 1; l < D::NPRE; ++l) {
        float e[KH];
#pragma unroll
        for (int q = 0; q < KH; ++q) e[q] = fmaxf(y[q / 16][q % 16], 0.0f);
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
            f32x16_t bias;
#pragma unroll
            for (int i = 0; i < 16; ++i) bias[i] = W[D::pre_off(l) + H * H + 32 * mt + 8 * (i >> 2) + 4 * h + (i & 3)];
            y[mt] = mfma_chain_f32<KH>(pk + pofs + mt * KH * 64, lane, e, bias);
        }
        pofs += MT * KH * 64;
    }
    float s = 0.0f;
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int i = 0; i < 16; ++i)
            s += W[D::ws_off + 32 * mt + 8 * (i >> 2) + 4 * h + (i & 3)] * fmaxf(y[mt][i], 0.0f);
    // lanes c and c + 32 hold the two halves of row c's units: after the swap, t of lane c < 32
    // is lane c + 32's partial sum
    float o = s, t = s;
    swap32f(o, t);
    s = s + t + W[D::bs_off];
    if (h == 0 && c < nrows) queue[(row0 + c) * kAlphaRow + 7] = s;
}

// Phase B: the pre-activation of every queued row.  Decoders with pre-layers take 32 rows at a
// time, two lanes per row (lane c + 32 h gathers channels [h C/2, (h + 1) C/2) of row c), so one
// column tile's operands are live at once; the halves meet in the B layout through one
// permlane32 swap per pair of k steps.  Decoders without pre-layers take one row per lane.
template <class D, bool B32>
__device__ __forceinline__ void alpha_shade(const RenderArgs& a, const float* __restrict__ pk,
                                            float* queue, int qn, int lane) {
    constexpr int C = D::C;
    const int n = a.n;
    if constexpr (D::NPRE == 0) {
        for (int base = 0; base < qn; base += 64) {
            const int r = base + lane;
            if (r < qn) {
                const float4 r0 = *reinterpret_cast<const float4*>(queue + r * kAlphaRow);
                const float4 r1 = *reinterpret_cast<const float4*>(queue + r * kAlphaRow + 4);
                float w[8];
                uint32_t off[8];
                segment_corners<C, uint32_t>(r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, n, w, off);
                float f[C];
                gather_integral_at<C, B32>(a.voxels, w, off, f);
                float s = a.W[D::bs_off];
#pragma unroll
                for (int i = 0; i < C; ++i) s += a.W[D::ws_off + i] * f[i];
                queue[r * kAlphaRow + 7] = s;
            }
        }
    } else {
        constexpr int HC = C / 2;
        const int c = lane & 31, h = lane >> 5;
        for (int base = 0; base < qn; base += 32) {
            const int r = base + c;
            float g[HC];
            if (r < qn) {
                const float4 r0 = *reinterpret_cast<const float4*>(queue + r * kAlphaRow);
                const float4 r1 = *reinterpret_cast<const float4*>(queue + r * kAlphaRow + 4);
                float w[8];
                uint32_t off[8];
                segment_corners<C, uint32_t>(r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, n, w, off);
#pragma unroll
                for (int j = 0; j < 8; ++j) off[j] += (uint32_t)(h * HC);
                gather_integral_at<HC, B32>(a.voxels, w, off, g);
            } else {
#pragma unroll
                for (int i = 0; i < HC; ++i) g[i] = 0.0f;
            }
            // b[kc] = (feature 2 kc on lanes 0-31, 2 kc + 1 on lanes 32-63) of row c
            float b[C / 2];
#pragma unroll
            for (int kc = 0; kc < HC / 2; ++kc) {
                float x = g[2 * kc], y = g[2 * kc + 1];
                swap32f(x, y);
                b[kc] = x;
                b[kc + HC / 2] = y;
            }
            const int nb = qn - base < 32 ? qn - base : 32;
            alpha_density_tile<D>(a.W, pk, b, lane, queue, base, nb);
        }
    }
}

// Register budget: 4 waves per SIMD (3 with 64 channels) without spills.
template <class D>
constexpr int alpha_waves_per_simd() { return D::C > 32 ? 3 : 4; }

template <class D, bool B32>
__global__ void __launch_bounds__(64, alpha_waves_per_simd<D>()) alpha_tile_kernel(RenderArgs a, int pad, int rpw,
                                                        const float* __restrict__ pk,
                                                        unsigned int* __restrict__ map) {
    __shared__ float queue[kAlphaQueue * kAlphaRow];
    const int lane = threadIdx.x;
    const int64_t r = (int64_t)blockIdx.x * rpw + lane;
    const bool valid = lane < rpw && r < a.n_rays;
    const int64_t rr = valid ? r : 0;
    const int n = a.n;
    RayWalk ray;
    ray.init(a.o[3 * rr], a.o[3 * rr + 1], a.o[3 * rr + 2], a.d[3 * rr], a.d[3 * rr + 1],
             a.d[3 * rr + 2], a.xyzmin, a.voxel_size, (float)n, valid);
    OccCache occ;
    Kahan S;
    float T = 1.0f;
    int qn = 0, head = -1, last = -1;
    const float tail = tail_limit(a.tail, n, ray);
    const bool all_fast = !wave_any(ray.alive && !ray.fast);
    const int64_t n1 = n, n2 = (int64_t)n * n;
    const int hi = n - 1 - pad;
    while (true) {
        if (all_fast)
            alpha_walk<true>(ray, occ, a, queue, lane, qn, head, last, tail);
        else
            alpha_walk<false>(ray, occ, a, queue, lane, qn, head, last, tail);
        if (qn == 0) break;  // nothing queued and every lane done
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        alpha_shade<D, B32>(a, pk, queue, qn, lane);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        // Phase C: the lane's rows in walk order
        int k = head;
        while (wave_any(k >= 0)) {
            if (k >= 0) {
                const float4 r0 = *reinterpret_cast<const float4*>(queue + k * kAlphaRow);
                const float4 r1 = *reinterpret_cast<const float4*>(queue + k * kAlphaRow + 4);
                float sigma, alpha;
                density_alpha<D>(r1.w, sigma, alpha);
                const float wgt = alpha * T;
                S.add(sigma);
                T = __expf(-S.s);
                int cx, cy, cz;
                {
#pragma clang fp contract(off)
                    cx = min((int)(r0.x + 1e-4f), (int)(r0.w + 1e-4f));
                    cy = min((int)(r0.y + 1e-4f), (int)(r1.x + 1e-4f));
                    cz = min((int)(r0.z + 1e-4f), (int)(r1.y + 1e-4f));
                }
                // the weight is >= 0 but may be -0 (relu(-0) = -0 gives alpha = -0), whose bit
                // pattern would beat every positive weight: store its magnitude
                if (cx >= pad && cy >= pad && cz >= pad && cx <= hi && cy <= hi && cz <= hi)
                    atomicMax(map + (cx + cy * n1 + cz * n2), __float_as_uint(wgt) & 0x7fffffffu);
                k = __builtin_bit_cast(int, r1.z);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        qn = 0;
        head = -1;
        last = -1;
    }
}

// FP32 matrix-core A operands of the pre-layers (alpha_density_tile's order), one float per (k
// step, lane): pre-layer 0 in natural k pairs (2 kc, 2 kc + 1) for (mt, kc), later pre-layers in
// the D order of d_order_feature for (mt, q).
__global__ void __launch_bounds__(256) pack_pre_f32_kernel(int C, int H, int npre,
                                                           const float* __restrict__ W,
                                                           float* __restrict__ out, int64_t total) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    const int r = (int)(idx & 31), h = (int)((idx >> 5) & 1);
    int64_t step = idx >> 6;
    const int64_t first_steps = (int64_t)(H / 32) * (C / 2);
    const int64_t later_steps = (int64_t)(H / 32) * (H / 2);
    float v;
    if (step < first_steps) {
        const int mt = (int)(step / (C / 2)), kc = (int)(step % (C / 2));
        v = W[(int64_t)(32 * mt + r) * C + 2 * kc + h];
    } else {
        step -= first_steps;
        const int l = 1 + (int)(step / later_steps);
        const int64_t rem = step % later_steps;
        const int mt = (int)(rem / (H / 2)), q = (int)(rem % (H / 2));
        const int64_t off = (int64_t)H * C + H + (int64_t)(l - 1) * (H * H + H);
        v = W[off + (int64_t)(32 * mt + r) * H + d_order_feature(q, h)];
    }
    out[idx] = v;
}

int64_t pack_pre_f32_size(int C, int H, int npre) {
    return npre <= 0 ? 0 : (int64_t)H * C + (int64_t)(npre - 1) * H * H;
}

hipError_t launch_pack_pre_f32(int C, int H, int npre, const float* W, float* out, hipStream_t s) {
    if (npre <= 0 || C % 16 || H % 32) return hipErrorInvalidValue;
    const int64_t total = pack_pre_f32_size(C, H, npre);
    hipLaunchKernelGGL(pack_pre_f32_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s,
                       C, H, npre, W, out, total);
    return hipGetLastError();
}

template <class D>
static hipError_t launch_alpha_map_d(const RenderArgs& a, int pad, float* map, const float* pk,
                                     hipStream_t s) {
    unsigned int* m = reinterpret_cast<unsigned int*>(map);
    const int64_t n1 = (int64_t)a.n + 1;
    const bool tiles = a.blocks && a.bricks && a.n % 8 == 0 && a.n <= 4096 &&
                       n1 * n1 * n1 * D::C < ((int64_t)1 << 32) && (D::NPRE == 0 || pk);
    if (!tiles) {
        dim3 grid((unsigned)((a.n_rays + 255) / 256)), block(256);
        hipLaunchKernelGGL((alpha_map_kernel<D>), grid, block, 0, s, a, pad, m);
        return hipGetLastError();
    }
    const int rpw = rays_per_wave(a.n_rays);
    dim3 grid((unsigned)((a.n_rays + rpw - 1) / rpw)), block(64);
    if (n1 * n1 * n1 * D::C * (int64_t)sizeof(float) <= (int64_t)UINT32_MAX)
        hipLaunchKernelGGL((alpha_tile_kernel<D, true>), grid, block, 0, s, a, pad, rpw, pk, m);
    else
        hipLaunchKernelGGL((alpha_tile_kernel<D, false>), grid, block, 0, s, a, pad, rpw, pk, m);
    return hipGetLastError();
}

hipError_t launch_alpha_map(int variant, const RenderArgs& a, int pad, float* map,
                            const float* pack_pre, hipStream_t s) {
    if (a.n_rays == 0) return hipSuccess;
    switch (variant) {
        case 0: return launch_alpha_map_d<DecB32>(a, pad, map, pack_pre, s);
        case 1: return launch_alpha_map_d<DecB64>(a, pad, map, pack_pre, s);
        case 2: return launch_alpha_map_d<DecB64C64>(a, pad, map, pack_pre, s);
        case 3: return launch_alpha_map_d<DecR32>(a, pad, map, pack_pre, s);
        case 4: return launch_alpha_map_d<DecR64>(a, pad, map, pack_pre, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace diver
