// Fused render kernel for decoders whose density comes from the integrated feature alone
// (canonical form with no pre-layers): traversal and density run per lane, colour decoding runs
// in dense 64-row matrix-core batches.
//
// One wavefront owns a packet of 64 rays (one per lane).  It alternates two phases:
//   A. every live lane advances the exact DDA one voxel per iteration (lanes never wait on each
//      other's hits); at an occupied voxel it integrates only what density needs (channel 0 for
//      softplus(f[0]) decoders), composites transmittance (compensated log-sum, early stop at
//      T < t_stop) and pushes the row (entry, exit, weight = T*alpha, owner lane) into the
//      wave's LDS queue.  The phase ends when >= 64 rows are queued or every lane is done.
//   B. the wave takes up to 64 queued rows, one per lane: full FP32 segment integral of the
//      row's voxel, then the colour decoder on the matrix cores, one 32-row column tile at a
//      time (or FP32 per lane), and an LDS atomic add of weight * colour into the owner ray's
//      accumulator.  The per-ray view term W_v PE(d) + b_c is computed once per ray in FP32 and
//      kept in LDS.
// Empty-space skipping: the workgroup keeps the 8^3-block occupancy bitmap in LDS; a voxel in an
// empty block never reads the fine mask.  Skipping changes no arithmetic of the walk, so the
// intervals are bit-identical to the drop-in traversal kernels.
#include "common.h"
#include "decoder.h"
#include "kernels.h"
#include "mfma.h"
#include "raystate.h"

namespace diver {

constexpr int kQueue = 96;       // rows per wave queue (phase A stops at >= 32, adds <= 64)
constexpr int kFill = 32;        // phase A hands over to phase B at this many queued rows
constexpr int kRowFloats = 8;    // p_in (3), p_out (3), weight, owner (as int bits)
constexpr int kMaxLdsBlockWords = 8192;

__host__ __device__ constexpr int wave_lds_floats(int H) { return kQueue * kRowFloats + 64 * 3 + 64 * H; }

template <class D, int PREC>
__device__ __forceinline__ void shade_tile(const RenderArgs& a, const float* queue,
                                           const float* hvs, float* acc, int t, int nbatch,
                                           const u32x4_t (&b)[D::Z / 16], int lane) {
    constexpr int H = D::H, Z = D::Z;
    constexpr int KC0 = Z / 16, MT = H / 32, KCH = H / 16;
    const u32x4_t* pk = reinterpret_cast<const u32x4_t*>(a.pack);
    const int c = lane & 31;
    const int half = lane & 32;
    const int row = 32 * t + c;
    const bool live = row < nbatch;
    const int owner = live ? __builtin_bit_cast(int, queue[row * kRowFloats + 7]) : 0;
    // head: acc = W_c z + (W_v PE(d) + b_c) of the row's ray
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
#pragma unroll
        for (int kc = 0; kc < KC0; ++kc) y[mt] = mfma32<PREC>(pk[(mt * KC0 + kc) * 64 + lane], b[kc], y[mt]);
    }
    int pofs = MT * KC0;
    static_assert(D::NPOST <= 1, "one post layer at most");
    if constexpr (D::NPOST == 1) {
        u32x4_t e[KCH];
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) d_to_b_relu<PREC>(y[mt], e[2 * mt], e[2 * mt + 1]);
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
            f32x16_t s;
#pragma unroll
            for (int i = 0; i < 16; ++i)
                s[i] = a.W[D::post_off(0) + H * H + 32 * mt + 8 * (i >> 2) + (half >> 3) + (i & 3)];
#pragma unroll
            for (int kc = 0; kc < KCH; ++kc) s = mfma32<PREC>(pk[(pofs + mt * KCH + kc) * 64 + lane], e[kc], s);
            y[mt] = s;
        }
        pofs += MT * KCH;
    }
    u32x4_t e[KCH];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) d_to_b_relu<PREC>(y[mt], e[2 * mt], e[2 * mt + 1]);
    f32x16_t o = {};
#pragma unroll
    for (int kc = 0; kc < KCH; ++kc) o = mfma32<PREC>(pk[(pofs + kc) * 64 + lane], e[kc], o);
    // output rows 0..2 sit in registers 0..2 of lanes 0..31, column = row of the tile
    if (lane < 32 && live) {
        const float* bo = a.W + D::out_off + 3 * H;
        const float wgt = queue[row * kRowFloats + 6];
        atomicAdd(&acc[owner * 3 + 0], wgt * sigmoidf(o[0] + bo[0]));
        atomicAdd(&acc[owner * 3 + 1], wgt * sigmoidf(o[1] + bo[1]));
        atomicAdd(&acc[owner * 3 + 2], wgt * sigmoidf(o[2] + bo[2]));
    }
}

template <class D, int PREC, bool RT, bool DCH0>
__global__ void __launch_bounds__(256, 3) render_queue_kernel(RenderArgs a) {
    static_assert(D::NPRE == 0, "queue kernel needs density from the integrated feature");
    constexpr int C = D::C, H = D::H, Z = D::Z;
    extern __shared__ uint32_t smem[];
    const int n = a.n;
    const int nb = n >> 3;
    const int nwords = (nb * nb * nb + 31) >> 5;
    const bool lds_blocks = nwords <= kMaxLdsBlockWords;
    uint32_t* sblocks = smem;
    float* wave_mem = reinterpret_cast<float*>(smem + (lds_blocks ? ((nwords + 3) & ~3) : 0));
    for (int i = threadIdx.x; lds_blocks && i < nwords; i += blockDim.x) sblocks[i] = a.blocks[i];
    __syncthreads();
    const uint32_t* blocks = lds_blocks ? sblocks : a.blocks;

    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    float* queue = wave_mem + wid * wave_lds_floats(H);
    float* acc = queue + kQueue * kRowFloats;
    float* hvs = acc + 64 * 3;  // [64][H] per-ray view term
    acc[lane * 3 + 0] = 0.0f;
    acc[lane * 3 + 1] = 0.0f;
    acc[lane * 3 + 2] = 0.0f;

    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t r = wave * 64 + lane;
    bool valid;
    int64_t out_idx = r;
    float ox, oy, oz, dx, dy, dz;
    if (a.cam.width > 0) {
        int x, y;
        valid = (r < a.n_rays) && camera_pixel(a.cam, r, x, y);
        if (!valid) { x = 0; y = 0; }
        out_idx = (int64_t)y * a.cam.width + x;
        camera_ray(a.cam, x, y, ox, oy, oz, dx, dy, dz);
    } else {
        valid = r < a.n_rays;
        int64_t rr = valid ? r : 0;
        ox = a.o[3 * rr]; oy = a.o[3 * rr + 1]; oz = a.o[3 * rr + 2];
        dx = a.d[3 * rr]; dy = a.d[3 * rr + 1]; dz = a.d[3 * rr + 2];
    }
    {
        float hv[H];
        head_view<D>(a.W, dx, dy, dz, hv);
#pragma unroll
        for (int j = 0; j < H; j += 4)
            *reinterpret_cast<float4*>(hvs + lane * H + j) = make_float4(hv[j], hv[j + 1], hv[j + 2], hv[j + 3]);
    }
    RayWalk ray;
    if constexpr (RT)
        ray.init_rt(ox, oy, oz, dx, dy, dz, a.xyzmin, a.voxel_size, n, a.octrees, valid, blocks);
    else
        ray.init(ox, oy, oz, dx, dy, dz, a.xyzmin, a.voxel_size, (float)n, valid);
    OccCache occ;

    Kahan S;
    float T = 1.0f;
    int used = 0;
    int qn = 0;  // queued rows (wave-uniform)
    const float nf = (float)n;

    while (true) {
        // ---------------- phase A: walk, density, transmittance, enqueue ------------------------
        while (qn < kFill) {
            if (!wave_any(ray.alive)) break;
            bool push = false;
            float px = 0.f, py = 0.f, pz = 0.f, qx = 0.f, qy = 0.f, qz = 0.f, wgt = 0.f;
            if (ray.alive) {
                if (ray.steps >= ray.max_steps) {
                    ray.alive = false;
                } else {
                    ++ray.steps;
                    float nx, ny, nz;
                    ray.step(ray.lx, ray.ly, ray.lz, nx, ny, nz);
                    bool out = RT ? outside_open(nx, ny, nz, nf) : outside(nx, ny, nz, nf);
                    if (out) {
                        ray.alive = false;
                    } else {
                        int vx, vy, vz;
                        segment_voxel1(ray.lx, ray.ly, ray.lz, nx, ny, nz, vx, vy, vz);
                        px = ray.lx; py = ray.ly; pz = ray.lz;
                        ray.lx = nx; ray.ly = ny; ray.lz = nz;
                        if (occ.lookup(a.mask, blocks, a.bricks, n, vx, vy, vz)) {
                            qx = nx; qy = ny; qz = nz;
                            float w[8];
                            int64_t off[8];
                            segment_corners<C>(px, py, pz, qx, qy, qz, n, w, off);
                            float s;
                            if constexpr (DCH0) {
                                s = 0.0f;
#pragma unroll
                                for (int j = 0; j < 8; ++j) s += w[j] * a.voxels[off[j]];
                            } else {
                                float f[C];
                                gather_integral<C>(a.voxels, w, off, f);
                                s = a.W[D::bs_off];
#pragma unroll
                                for (int i = 0; i < C; ++i) s += a.W[D::ws_off + i] * f[i];
                            }
                            float sigma = density_act<D>(s);
                            float alpha = -expm1f(-sigma);
                            used += 1;
                            if (alpha >= a.alpha_skip) {
                                wgt = T * alpha;
                                push = true;
                                S.add(sigma);
                                T = __expf(-S.s);
                                if (T < a.t_stop) ray.alive = false;
                            }
                        }
                    }
                }
            }
            uint64_t pm = __builtin_amdgcn_ballot_w64(push);
            if (push) {
                int slot = qn + __builtin_popcountll(pm & ((1ull << lane) - 1ull));
                float* rw = queue + slot * kRowFloats;
                *reinterpret_cast<float4*>(rw) = make_float4(px, py, pz, qx);
                *reinterpret_cast<float4*>(rw + 4) =
                    make_float4(qy, qz, wgt, __builtin_bit_cast(float, lane));
            }
            qn += __builtin_popcountll(pm);
        }
        if (qn == 0) break;  // nothing queued and every lane done

        // ---------------- phase B: shade up to 64 queued rows -----------------------------------
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        const int nbatch = qn < 64 ? qn : 64;
        const bool mine = lane < nbatch;
        float z[Z];
        if (mine) {
            float4 r0 = *reinterpret_cast<const float4*>(queue + lane * kRowFloats);
            float4 r1 = *reinterpret_cast<const float4*>(queue + lane * kRowFloats + 4);
            float w[8];
            int64_t off[8];
            segment_corners<C>(r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, n, w, off);
            gather_integral<C>(a.voxels, w, off, z);
        } else {
#pragma unroll
            for (int i = 0; i < Z; ++i) z[i] = 0.0f;
        }
        if constexpr (PREC == kPrecFP32) {
            if (mine) {
                const int owner = __builtin_bit_cast(int, queue[lane * kRowFloats + 7]);
                float hvr[H];
#pragma unroll
                for (int j = 0; j < H; ++j) hvr[j] = hvs[owner * H + j];
                float rgb[3];
                colour<D>(a.W, z, hvr, rgb);
                const float wgt = queue[lane * kRowFloats + 6];
                atomicAdd(&acc[owner * 3 + 0], wgt * rgb[0]);
                atomicAdd(&acc[owner * 3 + 1], wgt * rgb[1]);
                atomicAdd(&acc[owner * 3 + 2], wgt * rgb[2]);
            }
        } else {
            u32x4_t b0[Z / 16], b1[Z / 16];
            rows_to_b<PREC, Z>(z, b0, b1);
            shade_tile<D, PREC>(a, queue, hvs, acc, 0, nbatch, b0, lane);
            if (nbatch > 32) shade_tile<D, PREC>(a, queue, hvs, acc, 1, nbatch, b1, lane);
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
        qn = rest;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    if (valid) {
        float4 outv = make_float4(acc[lane * 3], acc[lane * 3 + 1], acc[lane * 3 + 2], T);
        reinterpret_cast<float4*>(a.rgba)[out_idx] = outv;
        if (a.count) a.count[out_idx] = used;
    }
}

static size_t render_queue_lds_bytes(int n, int H) {
    int nb = n >> 3;
    int64_t nwords = ((int64_t)nb * nb * nb + 31) >> 5;
    size_t b = nwords <= kMaxLdsBlockWords ? (size_t)((nwords + 3) & ~3) * 4 : 0;
    return b + 4 * (size_t)wave_lds_floats(H) * sizeof(float);
}

template <class D, int PREC, bool RT>
static hipError_t launch_q3(const RenderArgs& a, bool dch0, hipStream_t s) {
    int64_t waves = (a.n_rays + 63) / 64;
    dim3 grid((unsigned)((waves + 3) / 4)), block(256);
    size_t lds = render_queue_lds_bytes(a.n, D::H);
    if (dch0)
        hipLaunchKernelGGL((render_queue_kernel<D, PREC, RT, true>), grid, block, lds, s, a);
    else
        hipLaunchKernelGGL((render_queue_kernel<D, PREC, RT, false>), grid, block, lds, s, a);
    return hipGetLastError();
}

template <class D, bool RT>
static hipError_t launch_q2(const RenderArgs& a, int prec, bool dch0, hipStream_t s) {
    if (prec == kPrecBF16) return launch_q3<D, kPrecBF16, RT>(a, dch0, s);
    if (prec == kPrecFP16) return launch_q3<D, kPrecFP16, RT>(a, dch0, s);
    return launch_q3<D, kPrecFP32, RT>(a, dch0, s);
}

template <class D>
static hipError_t launch_q1(const RenderArgs& a, int prec, bool dch0, hipStream_t s) {
    return a.traversal == 1 ? launch_q2<D, true>(a, prec, dch0, s)
                            : launch_q2<D, false>(a, prec, dch0, s);
}

hipError_t launch_render_queue(int variant, const RenderArgs& a, int prec, bool dch0,
                               hipStream_t s) {
    if (a.n_rays == 0) return hipSuccess;
    switch (variant) {
        case 0: return launch_q1<DecB32>(a, prec, dch0, s);
        case 1: return launch_q1<DecB64>(a, prec, dch0, s);
        case 2: return launch_q1<DecB64C64>(a, prec, dch0, s);
        default: return hipErrorInvalidValue;
    }
}

// The fused kernels' walk on its own: every occupied interval of each ray, in masked_intersect's
// layout ([rays][max_n][6], count per ray).  Same step, entry and occupancy code as the queue
// kernel's phase A, so its output is what the fused renderer composites.
template <bool RT>
__global__ void __launch_bounds__(256) walk_kernel(RenderArgs a, int max_n, float* __restrict__ isect,
                                                   int32_t* __restrict__ num) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= a.n_rays) return;
    const int n = a.n;
    RayWalk ray;
    if constexpr (RT)
        ray.init_rt(a.o[3 * r], a.o[3 * r + 1], a.o[3 * r + 2], a.d[3 * r], a.d[3 * r + 1],
                    a.d[3 * r + 2], a.xyzmin, a.voxel_size, n, a.octrees, true, a.blocks);
    else
        ray.init(a.o[3 * r], a.o[3 * r + 1], a.o[3 * r + 2], a.d[3 * r], a.d[3 * r + 1],
                 a.d[3 * r + 2], a.xyzmin, a.voxel_size, (float)n, true);
    const float nf = (float)n;
    OccCache occ;
    int count = 0;
    while (ray.alive && ray.steps < ray.max_steps) {
        ++ray.steps;
        float nx, ny, nz;
        ray.step(ray.lx, ray.ly, ray.lz, nx, ny, nz);
        if (RT ? outside_open(nx, ny, nz, nf) : outside(nx, ny, nz, nf)) break;
        int vx, vy, vz;
        segment_voxel1(ray.lx, ray.ly, ray.lz, nx, ny, nz, vx, vy, vz);
        if (occ.lookup(a.mask, a.blocks, a.bricks, n, vx, vy, vz)) {
            if (count < max_n) {
                float* o = isect + (r * max_n + count) * 6;
                o[0] = ray.lx; o[1] = ray.ly; o[2] = ray.lz;
                o[3] = nx; o[4] = ny; o[5] = nz;
            }
            ++count;
        }
        ray.lx = nx;
        ray.ly = ny;
        ray.lz = nz;
    }
    num[r] = count;
    for (int k = count; k < max_n; ++k) {
        float* o = isect + (r * max_n + k) * 6;
        o[0] = 0.f; o[1] = 0.f; o[2] = 0.f; o[3] = 0.f; o[4] = 0.f; o[5] = 0.f;
    }
}

hipError_t launch_walk(const RenderArgs& a, int max_n, float* isect, int32_t* num,
                       hipStream_t s) {
    if (a.n_rays == 0) return hipSuccess;
    dim3 grid((unsigned)((a.n_rays + 255) / 256)), block(256);
    if (a.traversal == 1)
        hipLaunchKernelGGL(walk_kernel<true>, grid, block, 0, s, a, max_n, isect, num);
    else
        hipLaunchKernelGGL(walk_kernel<false>, grid, block, 0, s, a, max_n, isect, num);
    return hipGetLastError();
}

__global__ void block_bits_kernel(const uint8_t* __restrict__ mask, int n,
                                  uint32_t* __restrict__ bits) {
    int nb = n >> 3;
    int64_t bid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (bid >= (int64_t)nb * nb * nb) return;
    int bx = (int)(bid % nb), by = (int)((bid / nb) % nb), bz = (int)(bid / ((int64_t)nb * nb));
    bool any = false;
    for (int z = 0; z < 8 && !any; ++z)
        for (int y = 0; y < 8 && !any; ++y) {
            const uint8_t* row = mask + (((int64_t)(bz * 8 + z) * n + (by * 8 + y)) * n + bx * 8);
            for (int x = 0; x < 8; ++x) any |= row[x] != 0;
        }
    if (any) atomicOr(&bits[bid >> 5], 1u << (bid & 31));
}

// 64-bit occupancy word per 4^3 brick: bit (z&3)*16 + (y&3)*4 + (x&3).
__global__ void brick_bits_kernel(const uint8_t* __restrict__ mask, int n,
                                  unsigned long long* __restrict__ bricks) {
    int n4 = n >> 2;
    int64_t br = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (br >= (int64_t)n4 * n4 * n4) return;
    int bx = (int)(br % n4), by = (int)((br / n4) % n4), bz = (int)(br / ((int64_t)n4 * n4));
    unsigned long long w = 0;
    for (int z = 0; z < 4; ++z)
        for (int y = 0; y < 4; ++y) {
            const uint8_t* row = mask + (((int64_t)(bz * 4 + z) * n + (by * 4 + y)) * n + bx * 4);
            for (int x = 0; x < 4; ++x)
                if (row[x]) w |= 1ull << ((z << 4) | (y << 2) | x);
        }
    bricks[br] = w;
}

hipError_t launch_brick_bits(const uint8_t* mask, int n, unsigned long long* bricks,
                             hipStream_t s) {
    int n4 = n >> 2;
    int64_t nbr = (int64_t)n4 * n4 * n4;
    hipLaunchKernelGGL(brick_bits_kernel, dim3((unsigned)((nbr + 255) / 256)), dim3(256), 0, s,
                       mask, n, bricks);
    return hipGetLastError();
}

__global__ void zero_words_kernel(uint32_t* __restrict__ p, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

hipError_t launch_block_bits(const uint8_t* mask, int n, uint32_t* bits, hipStream_t s) {
    int nb = n >> 3;
    int64_t nblk = (int64_t)nb * nb * nb;
    int64_t nwords = (nblk + 31) >> 5;
    hipLaunchKernelGGL(zero_words_kernel, dim3((unsigned)((nwords + 255) / 256)), dim3(256), 0, s,
                       bits, nwords);
    hipLaunchKernelGGL(block_bits_kernel, dim3((unsigned)((nblk + 255) / 256)), dim3(256), 0, s,
                       mask, n, bits);
    return hipGetLastError();
}

}  // namespace diver
