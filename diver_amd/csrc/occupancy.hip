// Occupancy structures and the fused renderer's walk on its own.
//
//   - walk_kernel: every occupied interval of each ray with the fused renderer's step, entry and
//     occupancy code, in masked_intersect's output layout (the walk the renderer composites);
//   - block_bits_kernel: one bit per 8^3 block (any voxel occupied), kept in LDS by the renderer;
//   - brick_bits_kernel: the mask bit-packed per 4^3 brick (one 64-bit word per brick), read by
//     the renderer's per-lane brick cache.
//   - tail_bits_kernel: the block bits dilated by one block, from which each walk bounds where
//     it can end.
// None of them changes which occupied voxels are visited: the block and brick bits replace
// byte-mask reads, and the tail bits only end a walk where nothing occupied remains ahead.
#include "common.h"
#include "kernels.h"
#include "raystate.h"

namespace diver {

// The fused renderer's walk on its own: every occupied interval of each ray, in masked_intersect's
// layout ([rays][max_n][6], count per ray).  Same step, entry (box entry; real-time: the octree
// descent) and occupancy code as the tile renderer, so its output is what that kernel
// composites.
template <bool RT>
__global__ void __launch_bounds__(256) walk_kernel(RenderArgs a, int max_n, float* __restrict__ isect,
                                                   int32_t* __restrict__ num) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= a.n_rays) return;
    const int n = a.n;
    RayWalk ray;
    if constexpr (RT)
        ray.init_rt(a.o[3 * r], a.o[3 * r + 1], a.o[3 * r + 2], a.d[3 * r], a.d[3 * r + 1],
                    a.d[3 * r + 2], a.xyzmin, a.voxel_size, n, a.octrees, true);
    else
        ray.init(a.o[3 * r], a.o[3 * r + 1], a.o[3 * r + 2], a.d[3 * r], a.d[3 * r + 1],
                 a.d[3 * r + 2], a.xyzmin, a.voxel_size, (float)n, true);
    const float nf = (float)n;
    OccCache occ;
    int count = 0;
    const float tail = tail_limit(a.tail, n, ray);
    while (ray.alive && ray.steps < ray.max_steps) {
        ++ray.steps;
        float nx, ny, nz;
        ray.step(ray.lx, ray.ly, ray.lz, nx, ny, nz);
        if (RT ? outside_open(nx, ny, nz, nf) : outside(nx, ny, nz, nf)) break;
        if (past_tail(ray, nx, ny, nz, tail)) break;
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

// Tail bits: one thread per 32-bit word of blocks; a block's bit is set when any block of its
// 3x3x3 neighbourhood (clipped to the grid) is occupied.
__global__ void tail_bits_kernel(const uint32_t* __restrict__ blocks, int n, uint32_t* __restrict__ tail) {
    const int m = n / kTailCell, bpc = kTailCell / 8, nb = n >> 3;
    const int64_t ncell = (int64_t)m * m * m;
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= (ncell + 31) >> 5) return;
    uint32_t word = 0;
    for (int k = 0; k < 32; ++k) {
        const int64_t c = w * 32 + k;
        if (c >= ncell) break;
        const int cx = (int)(c % m), cy = (int)((c / m) % m), cz = (int)(c / ((int64_t)m * m));
        bool any = false;
        for (int bz = max(bpc * (cz - 1), 0); bz <= min(bpc * (cz + 2) - 1, nb - 1) && !any; ++bz)
            for (int by = max(bpc * (cy - 1), 0); by <= min(bpc * (cy + 2) - 1, nb - 1) && !any; ++by)
                for (int bx = max(bpc * (cx - 1), 0); bx <= min(bpc * (cx + 2) - 1, nb - 1); ++bx) {
                    const int64_t b = ((int64_t)bz * nb + by) * nb + bx;
                    any |= ((blocks[b >> 5] >> (b & 31)) & 1u) != 0;
                }
        if (any) word |= 1u << k;
    }
    tail[w] = word;
}

hipError_t launch_tail_bits(const uint32_t* blocks, int n, uint32_t* tail, hipStream_t s) {
    const int64_t m = n / kTailCell;
    const int64_t nwords = (m * m * m + 31) >> 5;
    hipLaunchKernelGGL(tail_bits_kernel, dim3((unsigned)((nwords + 255) / 256)), dim3(256), 0, s,
                       blocks, n, tail);
    return hipGetLastError();
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
