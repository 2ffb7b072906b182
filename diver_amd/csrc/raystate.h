// Per-ray traversal state shared by the fused forward and backward kernels, so the backward's
// recompute walks exactly the intervals the forward composited.
#pragma once
#include "common.h"
#include "kernels.h"

namespace diver {

struct RayWalk {
    float lx, ly, lz;  // current point (grid units)
    float dx, dy, dz;
    float rx, ry, rz;  // RN(1/d) per axis, for div_recip
    float sx, sy, sz;
    int steps;
    int max_steps;
    bool alive;
    bool fast;         // every |d| >= kMinFastDir: divisions by reciprocal + correction

    __device__ __forceinline__ void set_dir(float dx_, float dy_, float dz_) {
        dx = dx_;
        dy = dy_;
        dz = dz_;
        rx = 1.0f / dx;
        ry = 1.0f / dy;
        rz = 1.0f / dz;
        fast = fabsf(dx) >= kMinFastDir && fabsf(dy) >= kMinFastDir && fabsf(dz) >= kMinFastDir;
        sx = (float)sgnf(dx);
        sy = (float)sgnf(dy);
        sz = (float)sgnf(dz);
    }

    // The voxel of the segment from the current point to (nx, ny, nz), int(min(l + 1e-4,
    // n + 1e-4)) per axis (segment_voxel1): positions only move the way the direction points
    // along each axis (every step adds d * t with t > 0, or snaps to the next plane ahead), so the
    // smaller coordinate is the entry's where the direction points up and the exit's otherwise.
    __device__ __forceinline__ void voxel(float nx, float ny, float nz, int& vx, int& vy,
                                          int& vz) const {
#pragma clang fp contract(off)
        vx = (int)((dx > 0.0f ? lx : nx) + 1e-4f);
        vy = (int)((dy > 0.0f ? ly : ny) + 1e-4f);
        vz = (int)((dz > 0.0f ? lz : nz) + 1e-4f);
    }

    // one exact DDA step from (px,py,pz)
    __device__ __forceinline__ void step(float px, float py, float pz, float& qx, float& qy,
                                         float& qz) const {
        if (fast)
            dda_step_r(px, py, pz, dx, dy, dz, rx, ry, rz, sx, sy, sz, qx, qy, qz);
        else
            dda_step(px, py, pz, dx, dy, dz, sx, sy, sz, qx, qy, qz);
    }

    // World-space ray -> grid units, box entry as masked_intersect does it.
    __device__ __forceinline__ void init(float ox, float oy, float oz, float dx_, float dy_,
                                         float dz_, float xyzmin, float voxel_size, float n,
                                         bool valid) {
#pragma clang fp contract(off)
        set_dir(dx_, dy_, dz_);
        float gx = (ox - xyzmin) / voxel_size;
        float gy = (oy - xyzmin) / voxel_size;
        float gz = (oz - xyzmin) / voxel_size;
        lx = gx;
        ly = gy;
        lz = gz;
        bool inside = (gx >= 0) & (gy >= 0) & (gz >= 0) & (gx <= n) & (gy <= n) & (gz <= n);
        alive = valid;
        if (valid && !inside) {
            float tmin;
            alive = box_entry(gx, gy, gz, dx, dy, dz, n, lx, ly, lz, tmin);
        }
        steps = 0;
        max_steps = max_dda_steps(n);
    }

    // Real-time entry (aabb_intersect, ray_march.cu:50-148): box entry, then a DDA over the
    // occupancy pyramid (levels N/64, N/16, N/4, coarse first), descending x4 at each occupied
    // cell; the walk starts at the fine-level entry of the first occupied N/4 cell, or the ray is
    // dead.  Bit-identical to aabb_intersect's start point.
    __device__ __forceinline__ void init_rt(float ox, float oy, float oz, float dx_, float dy_,
                                            float dz_, float xyzmin, float voxel_size, int n,
                                            const uint8_t* __restrict__ octrees, bool valid) {
#pragma clang fp contract(off)
        const float nf = (float)n;
        set_dir(dx_, dy_, dz_);
        steps = 0;
        max_steps = max_dda_steps(nf);
        alive = false;
        if (!valid) return;
        float px = (ox - xyzmin) / voxel_size;
        float py = (oy - xyzmin) / voxel_size;
        float pz = (oz - xyzmin) / voxel_size;
        bool inside = (px >= 0) & (py >= 0) & (pz >= 0) & (px <= nf) & (py <= nf) & (pz <= nf);
        if (!inside) {
            float tmin;
            if (!box_entry(px, py, pz, dx, dy, dz, nf, px, py, pz, tmin)) return;
        }
        px /= 64.0f;
        py /= 64.0f;
        pz /= 64.0f;
        float qx = px, qy = py, qz = pz;
        int m = n / 64;
        // cell indices fit 32 bits: the levels hold at most (n/4)^3 + (n/16)^3 + (n/64)^3 cells,
        // below 2^31 for n <= 4096 (callers check n)
        int off = 0, total = 0;
        for (int k = n / 64; k < n; k *= 4) total += k * k * k;
        const int bound = max_dda_steps(nf) + 16;
        for (int it = 0; it < bound && m < n; ++it) {
            qx = px;
            qy = py;
            qz = pz;
            step(qx, qy, qz, px, py, pz);
            if (outside(px, py, pz, (float)m)) return;
            int vx, vy, vz;
            segment_voxel1(qx, qy, qz, px, py, pz, vx, vy, vz);
            // |vx|, |vy|, |vz| <= m + 1 <= 1025: the 24-bit multiplies are exact (full rate)
            const int m3 = __mul24(__mul24(m, m), m);
            const int ci = __mul24(__mul24(vz, m) + vy, m) + vx;
            const int gi = off + ci;
            const bool occ = (ci >= 0 && ci < m3 && gi < total) ? octrees[gi] != 0 : false;
            if (occ) {
                off += m3;
                qx *= 4.0f;
                qy *= 4.0f;
                qz *= 4.0f;
                px = qx;
                py = qy;
                pz = qz;
                m *= 4;
            }
        }
        if (m < n) return;
        lx = qx;
        ly = qy;
        lz = qz;
        alive = true;
    }

    // Advance to the next occupied voxel; on success (p, q) is its (entry, exit).  OPEN selects
    // the real-time exit test (leave at >= N, ray_march.cu:203-204) instead of > N.
    template <bool OPEN = false>
    __device__ __forceinline__ bool next(const uint8_t* __restrict__ mask, int n, float& px,
                                         float& py, float& pz, float& qx, float& qy, float& qz) {
        const float nf = (float)n;
        const int64_t ntot = (int64_t)n * n * n;
        while (alive && steps < max_steps) {
            ++steps;
            float nx, ny, nz;
            step(lx, ly, lz, nx, ny, nz);
            if (OPEN ? outside_open(nx, ny, nz, nf) : outside(nx, ny, nz, nf)) break;
            int vx, vy, vz;
            segment_voxel1(lx, ly, lz, nx, ny, nz, vx, vy, vz);
            int64_t ci = ((int64_t)vz * n + vy) * n + vx;
            bool occ = (ci >= 0 && ci < ntot) ? mask[ci] != 0 : false;
            px = lx; py = ly; pz = lz;
            lx = nx; ly = ny; lz = nz;
            if (occ) {
                qx = nx; qy = ny; qz = nz;
                return true;
            }
        }
        alive = false;
        return false;
    }
};

// Occupancy lookups of the fine walk through a per-lane cache of one 4^3 brick (64 bits of the
// bit-packed mask): the 8^3-block bitmap (LDS) answers empty blocks, and the brick word is read
// from memory only when the walk enters a new brick of an occupied block.  Gives exactly
// mask[ci] for every voxel inside the grid; indices outside it (which the reference wraps) read
// the byte mask as the reference does.
struct OccCache {
    int brick = -1;
    uint32_t lo = 0, hi = 0;
    bool block_empty = false;  // the cached brick's 8^3 block has no occupied voxel
    uint32_t shift = 2;        // the region a cheap step may stay in: the 8^3 block (3) or the brick (2)
    uint32_t key = 0;          // that region's coordinates, 10 bits each: (z << 20) | (y << 10) | x

    // occupancy of a voxel inside the grid, through the cached brick
    __device__ __forceinline__ bool lookup_in(const uint32_t* blocks,
                                              const unsigned long long* __restrict__ bricks, int n,
                                              uint32_t ux, uint32_t uy, uint32_t uz) {
        const uint32_t n4 = (uint32_t)n >> 2, n8 = (uint32_t)n >> 3;
        const int br = (int)(__umul24(__umul24(uz >> 2, n4) + (uy >> 2), n4) + (ux >> 2));
        if (br != brick) {
            brick = br;
            const uint32_t bid = __umul24(__umul24(uz >> 3, n8) + (uy >> 3), n8) + (ux >> 3);
            // both words in flight together; the brick word of an empty block is zero
            const uint32_t bw = blocks[bid >> 5];
            const unsigned long long w = bricks[(uint32_t)br];
            block_empty = !((bw >> (bid & 31u)) & 1u);
            lo = (uint32_t)w;
            hi = (uint32_t)(w >> 32);
            shift = block_empty ? 3u : 2u;
            key = ((uz >> shift) << 20) | ((uy >> shift) << 10) | (ux >> shift);
        }
        const uint32_t bit = ((uz & 3u) << 4) | ((uy & 3u) << 2) | (ux & 3u);
        const uint32_t word = bit < 32u ? lo : hi;
        return (word >> (bit & 31u)) & 1u;
    }

    __device__ __forceinline__ bool lookup(const uint8_t* __restrict__ mask,
                                           const uint32_t* blocks,
                                           const unsigned long long* __restrict__ bricks, int n,
                                           int vx, int vy, int vz) {
        const bool inr = (unsigned)vx < (unsigned)n && (unsigned)vy < (unsigned)n && (unsigned)vz < (unsigned)n;
        // a voxel index outside the grid (the reference reads the byte mask at its flattened,
        // wrapped index) is rare: the whole wave takes the general branch then
        if (__builtin_expect(__builtin_amdgcn_ballot_w64(!inr) == 0, 1))
            return lookup_in(blocks, bricks, n, (uint32_t)vx, (uint32_t)vy, (uint32_t)vz);
        if (inr) return lookup_in(blocks, bricks, n, (uint32_t)vx, (uint32_t)vy, (uint32_t)vz);
        const int64_t ntot = (int64_t)n * n * n;
        int64_t ci = ((int64_t)vz * n + vy) * n + vx;
        return (ci >= 0 && ci < ntot) ? mask[ci] != 0 : false;
    }
};

// RayWalk::next with the occupancy read through an OccCache (8^3-block bits and 4^3-brick words
// instead of one mask byte per voxel): the same steps, the same intervals.
template <bool OPEN = false>
__device__ __forceinline__ bool next_cached(RayWalk& ray, OccCache& occ,
                                            const uint8_t* __restrict__ mask,
                                            const uint32_t* __restrict__ blocks,
                                            const unsigned long long* __restrict__ bricks, int n,
                                            float& px, float& py, float& pz, float& qx, float& qy,
                                            float& qz) {
    const float nf = (float)n;
    while (ray.alive && ray.steps < ray.max_steps) {
        ++ray.steps;
        float nx, ny, nz;
        ray.step(ray.lx, ray.ly, ray.lz, nx, ny, nz);
        if (OPEN ? outside_open(nx, ny, nz, nf) : outside(nx, ny, nz, nf)) break;
        int vx, vy, vz;
        segment_voxel1(ray.lx, ray.ly, ray.lz, nx, ny, nz, vx, vy, vz);
        const bool occd = occ.lookup(mask, blocks, bricks, n, vx, vy, vz);
        px = ray.lx; py = ray.ly; pz = ray.lz;
        ray.lx = nx; ray.ly = ny; ray.lz = nz;
        if (occd) {
            qx = nx; qy = ny; qz = nz;
            return true;
        }
    }
    ray.alive = false;
    return false;
}

// Corner weights of the segment mean of the trilinear field (integrator.py:92-111 arranged as
// mlp_eval.cu:83-101) and the 8 corner offsets into a [(n+1)^3][C] grid.  I is the offset type:
// uint32_t when the caller has checked (n+1)^3 * C < 2^32 (cheaper index arithmetic), else
// int64_t.
// The same with the cell (cx, cy, cz) = (min(int(p + 1e-4), int(q + 1e-4)) per axis) given: the
// walk's segment voxel, int(min(p + 1e-4, q + 1e-4)), is that cell (int() is monotonic).
template <int C, typename I = int64_t>
__device__ __forceinline__ void segment_corners_cell(int cx, int cy, int cz, float px, float py,
                                                     float pz, float qx, float qy, float qz, int n,
                                                     float (&w)[8], I (&off)[8]);

template <int C, typename I = int64_t>
__device__ __forceinline__ void segment_corners(float px, float py, float pz, float qx, float qy,
                                                float qz, int n, float (&w)[8], I (&off)[8]) {
    int cx, cy, cz;
    {
#pragma clang fp contract(off)
        int ax = (int)(px + 1e-4f), bx = (int)(qx + 1e-4f);
        int ay = (int)(py + 1e-4f), by = (int)(qy + 1e-4f);
        int az = (int)(pz + 1e-4f), bz = (int)(qz + 1e-4f);
        cx = ax < bx ? ax : bx;
        cy = ay < by ? ay : by;
        cz = az < bz ? az : bz;
    }
    segment_corners_cell<C, I>(cx, cy, cz, px, py, pz, qx, qy, qz, n, w, off);
}

template <int C, typename I>
__device__ __forceinline__ void segment_corners_cell(int cx, int cy, int cz, float px, float py,
                                                     float pz, float qx, float qy, float qz, int n,
                                                     float (&w)[8], I (&off)[8]) {
    float x0 = px - cx, x1 = qx - cx, y0 = py - cy, y1 = qy - cy, z0 = pz - cz, z1 = qz - cz;
    float z01 = z0 + z1, y01 = y0 + y1, x01 = x0 + x1;
    w[7] = (2.0f * x0 * y0 * z0 + 2.0f * x1 * y1 * z1 + x01 * y01 * z01) * (1.0f / 12.0f);
    float w2 = (y01 * z01 + y0 * z0 + y1 * z1) * (1.0f / 6.0f);
    w[5] = (x01 * z01 + x0 * z0 + x1 * z1) * (1.0f / 6.0f) - w[7];
    float w4 = (x01 * y01 + x0 * y0 + x1 * y1) * (1.0f / 6.0f);
    x01 *= 0.5f;
    y01 *= 0.5f;
    z01 *= 0.5f;
    w[4] = z01 - w[5] - w2;
    w[0] = 1.0f - w[4] - y01 - x01 + w4;
    w[6] = w2 - w[7];
    w[2] = y01 - w[6] - w4;
    w[1] = x01 - w[5] - w4;
    w[3] = w4 - w[7];
    const int cx1 = min(cx + 1, n), cy1 = min(cy + 1, n), cz1 = min(cz + 1, n);
    if constexpr (sizeof(I) == 4) {
        // Corner index of (cx, cy, cz) with full-rate 24-bit multiplies ((n + 1)^3 C < 2^32 puts
        // every operand below 2^24), then the +x/+y/+z neighbours by adding their strides (0 where
        // the index is clamped at n).
        const uint32_t np1 = (uint32_t)(n + 1);
        const uint32_t base = __umul24(__umul24((uint32_t)cz, np1) + (uint32_t)cy, np1) + (uint32_t)cx;
        const uint32_t dx = (uint32_t)(cx1 - cx);
        const uint32_t dy = cy1 > cy ? np1 : 0u;
        const uint32_t dz = cz1 > cz ? np1 * np1 : 0u;
#pragma unroll
        for (int j = 0; j < 8; ++j)
            off[j] = (I)((base + ((j & 1) ? dx : 0u) + ((j & 2) ? dy : 0u) + ((j & 4) ? dz : 0u)) * (uint32_t)C);
    } else {
        const I sz = (I)(n + 1) * (I)(n + 1) * (I)C, sy = (I)(n + 1) * (I)C;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            int gx = (j & 1) ? cx1 : cx;
            int gy = ((j >> 1) & 1) ? cy1 : cy;
            int gz = (j >> 2) ? cz1 : cz;
            off[j] = (I)gz * sz + (I)gy * sy + (I)gx * (I)C;
        }
    }
}

// f[i] = sum over corners j = 0..7 (in that order) of w[j] * voxels[off[j] + i].  The channels go
// in chunks of 4: the 8 loads of a chunk (8 corners x 16 B) are issued before the first is used,
// and the chunks are kept apart so the loads in flight fit the register budget without spills.
template <int C, typename I>
__device__ __forceinline__ void gather_integral(const float* __restrict__ voxels, const float (&w)[8],
                                                const I (&off)[8], float (&f)[C]) {
    static_assert(C % 4 == 0, "channel count must be a multiple of 4");
#pragma unroll
    for (int c = 0; c < C / 4; ++c) {
        float4 q[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) q[j] = reinterpret_cast<const float4*>(voxels + off[j])[c];
        float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            s0 += w[j] * q[j].x;
            s1 += w[j] * q[j].y;
            s2 += w[j] * q[j].z;
            s3 += w[j] * q[j].w;
        }
        f[4 * c + 0] = s0;
        f[4 * c + 1] = s1;
        f[4 * c + 2] = s2;
        f[4 * c + 3] = s3;
        __builtin_amdgcn_sched_barrier(0);
    }
}

// Pinhole camera ray of pixel (x, y) (FP32): pixel centres at +0.5, camera space
// ((x + 0.5 - cx) / fx, -(y + 0.5 - cy) / fy, -1), rotated by c2w and normalised, from the camera
// centre.  With fx = fy = focal and (cx, cy) = (W/2, H/2) this is utils/dataset/ray_utils.py:5-56
// (the Blender camera); otherwise the Tanks and Temples / BlendedMVS camera of
// utils/dataset/tanks.py:120-169 (whose poses and origins are converted on the host).
__device__ __forceinline__ void camera_ray(const CameraDesc& cam, int x, int y, float& ox, float& oy,
                                           float& oz, float& dx, float& dy, float& dz) {
#pragma clang fp contract(off)
    float cx = ((float)x + 0.5f - cam.cx) / cam.fx;
    float cy = -(((float)y + 0.5f - cam.cy) / cam.fy);
    float cz = -1.0f;
    const float* m = cam.c2w;  // row-major 3x4
    float wx = m[0] * cx + m[1] * cy + m[2] * cz;
    float wy = m[4] * cx + m[5] * cy + m[6] * cz;
    float wz = m[8] * cx + m[9] * cy + m[10] * cz;
    float nrm = sqrtf(wx * wx + wy * wy + wz * wz);
    dx = wx / nrm;
    dy = wy / nrm;
    dz = wz / nrm;
    ox = m[3];
    oy = m[7];
    oz = m[11];
}

// Ray r of a camera render in tile order: 8x8 pixel tiles, left to right within a row of tiles,
// rows of tiles from the image centre outwards (centre, centre + 1, centre - 1, centre + 2, ...).
// Workgroups start in id order, so the tiles whose rays usually cross the most of the scene start
// first; the order changes no pixel's value.
__device__ __forceinline__ bool camera_pixel(const CameraDesc& cam, int64_t r, int& x, int& y) {
    const int tiles_x = (cam.width + 7) >> 3;
    const int tiles_y = (cam.height + 7) >> 3;
    const int64_t tile = r >> 6;
    const int in = (int)(r & 63);
    const int k = (int)(tile / tiles_x);  // k-th row of tiles in centre-out order
    const int c = (tiles_y - 1) >> 1, below = c, above = tiles_y - 1 - c;
    const int m = below < above ? below : above;
    int ty;
    if (k == 0)
        ty = c;
    else if (k <= 2 * m)
        ty = (k & 1) ? c + ((k + 1) >> 1) : c - (k >> 1);
    else
        ty = above > below ? c + m + (k - 2 * m) : c - m - (k - 2 * m);
    x = (int)(tile % tiles_x) * 8 + (in & 7);
    y = ty * 8 + (in >> 3);
    return k < tiles_y && x < cam.width && y < cam.height;
}

__device__ __forceinline__ bool wave_any(bool p) { return __builtin_amdgcn_ballot_w64(p) != 0; }

// Address of feature element `elem` of the grid.  B32: the grid's bytes fit 32 bits, so the
// offset is formed in 32 bits and the load takes the grid pointer as its scalar base.
template <bool B32>
__device__ __forceinline__ const float* feature_at(const float* __restrict__ voxels, uint32_t elem) {
    if constexpr (B32)
        return reinterpret_cast<const float*>(reinterpret_cast<const char*>(voxels) + elem * 4u);
    else
        return voxels + elem;
}

// FP32 feature integral of gather_integral with feature_at addressing.
template <int C, bool B32>
__device__ __forceinline__ void gather_integral_at(const float* __restrict__ voxels, const float (&w)[8],
                                                   const uint32_t (&off)[8], float (&f)[C]) {
    static_assert(C % 4 == 0, "channel count must be a multiple of 4");
#pragma unroll
    for (int c = 0; c < C / 4; ++c) {
        float4 q[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) q[j] = reinterpret_cast<const float4*>(feature_at<B32>(voxels, off[j]))[c];
        float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            s0 += w[j] * q[j].x;
            s1 += w[j] * q[j].y;
            s2 += w[j] * q[j].z;
            s3 += w[j] * q[j].w;
        }
        f[4 * c + 0] = s0;
        f[4 * c + 1] = s1;
        f[4 * c + 2] = s2;
        f[4 * c + 3] = s3;
        __builtin_amdgcn_sched_barrier(0);
    }
}

// ---- where a walk can end ----------------------------------------------------------------------
// The tail bits are the 8^3-block occupancy dilated by one block: a block's bit is set when any
// block of its 3x3x3 neighbourhood holds an occupied voxel.  A walk's positions stay within a
// small fraction of a voxel of its ray's line (each step rounds once; one coordinate is snapped to
// a plane), and the voxel it reads is within 1e-4 of a position, so a voxel the walk visits lies
// within a few voxels of the line.  Once the line has left the last set block it meets, every
// such voxel lies in an unset block's neighbourhood and is empty: nothing after that point is
// composited, and the walk may end there with every output (intervals, colour, transmittance,
// counts) unchanged.
//
// tail_limit runs a DDA over the (n/8)^3 blocks backwards, from the line's exit from the grid
// towards the walk's start point, up to the first set block (the last one the walk meets), and
// returns a bound on p.d past which the walk can end: kTailMargin voxels beyond that block's exit,
// or -inf when no set block lies between the start and the exit (the walk can end at its first
// step).  Without tail bits (NULL), for a dead ray, or should rounding carry the DDA out of the
// grid before it reaches the start block, it returns +inf.
constexpr int kTailCell = 8;         // voxels per block edge (the blocks of the block bits)
constexpr float kTailMargin = 4.0f;  // voxels past the last set block's exit

__device__ __forceinline__ float tail_limit(const uint32_t* __restrict__ tail, int n,
                                            const RayWalk& ray) {
    const float inf = __builtin_huge_valf();
    if (!tail || !ray.alive) return inf;
    constexpr float cs = (float)kTailCell, ics = 1.0f / kTailCell;
    const int m = n / kTailCell;
    const float nf = (float)n;
    const float px = ray.lx, py = ray.ly, pz = ray.lz;
    const float dx = ray.dx, dy = ray.dy, dz = ray.dz;
    // the line's exit from the grid box, p0 + t_exit d
    const float ax = dx > 0.0f ? (nf - px) / dx : dx < 0.0f ? -px / dx : inf;
    const float ay = dy > 0.0f ? (nf - py) / dy : dy < 0.0f ? -py / dy : inf;
    const float az = dz > 0.0f ? (nf - pz) / dz : dz < 0.0f ? -pz / dz : inf;
    const float t_exit = fmaxf(fminf(ax, fminf(ay, az)), 0.0f);
    const float qx = fmaf(dx, t_exit, px), qy = fmaf(dy, t_exit, py), qz = fmaf(dz, t_exit, pz);
    // walk the blocks backwards from the exit (direction -d); s is the backward parameter
    int cx = min(max((int)floorf(qx * ics), 0), m - 1);
    int cy = min(max((int)floorf(qy * ics), 0), m - 1);
    int cz = min(max((int)floorf(qz * ics), 0), m - 1);
    float sx = dx < 0.0f ? ((cx + 1) * cs - qx) / -dx : dx > 0.0f ? (cx * cs - qx) / -dx : inf;
    float sy = dy < 0.0f ? ((cy + 1) * cs - qy) / -dy : dy > 0.0f ? (cy * cs - qy) / -dy : inf;
    float sz = dz < 0.0f ? ((cz + 1) * cs - qz) / -dz : dz > 0.0f ? (cz * cs - qz) / -dz : inf;
    const float ex = dx != 0.0f ? cs / fabsf(dx) : inf;
    const float ey = dy != 0.0f ? cs / fabsf(dy) : inf;
    const float ez = dz != 0.0f ? cs / fabsf(dz) : inf;
    const int stx = dx < 0.0f ? 1 : -1, sty = dy < 0.0f ? 1 : -1, stz = dz < 0.0f ? 1 : -1;
    float s_in = 0.0f;
    for (int it = 0; it < 3 * m + 3; ++it) {
        const int c = __mul24(__mul24(cz, m) + cy, m) + cx;
        if ((tail[c >> 5] >> (c & 31)) & 1u) {
            // the last set block the line meets: its forward exit is at t_exit - s_in
            const float dd = dx * dx + dy * dy + dz * dz;
            return (px * dx + py * dy + pz * dz) + (t_exit - s_in) * dd + kTailMargin * sqrtf(dd);
        }
        const float s_out = fminf(sx, fminf(sy, sz));
        if (s_out >= t_exit) return -inf;  // this block holds the start point: nothing set ahead
        if (s_out == sx) {
            cx += stx;
            sx += ex;
        } else if (s_out == sy) {
            cy += sty;
            sy += ey;
        } else {
            cz += stz;
            sz += ez;
        }
        s_in = s_out;
        if ((unsigned)cx >= (unsigned)m || (unsigned)cy >= (unsigned)m || (unsigned)cz >= (unsigned)m)
            break;
    }
    return inf;  // left the grid before the start block (rounding): no bound
}

// The walk has passed its tail limit at position (nx, ny, nz).
__device__ __forceinline__ bool past_tail(const RayWalk& ray, float nx, float ny, float nz,
                                          float limit) {
    return fmaf(nz, ray.dz, fmaf(ny, ray.dy, nx * ray.dx)) > limit;
}

constexpr int kSkip = 2;           // cheap steps per walk iteration over voxels known to be empty (0: off)

// Cheap steps over voxels already known to be empty.  Before each full walk iteration a lane
// takes exact DDA steps for as long as the new voxel lies inside the grid, in its cached 4^3
// brick (or in the cached brick's 8^3 block when that block is empty) and has its occupancy bit
// clear: the full iteration would only have stepped over such a voxel.  The region test compares
// the voxel's region coordinates, packed 10 bits an axis, with the cached region's (n <= 4096).
// A step that fails the test is not taken here; the full iteration takes it again (same operands,
// same result).  Up to kSkip steps per call (more lets lanes that cannot skip wait longer); lanes
// stop one by one.
template <bool RT, bool AF>
__device__ __forceinline__ void skip_empty(RayWalk& ray, const OccCache& occ, int n) {
    const float nf = (float)n;
    bool sk = ray.alive && occ.brick >= 0;
    const uint32_t sh = occ.shift;
#pragma unroll 1
    for (int k = 0; k < kSkip; ++k) {
        // a wave whose every ray takes the reciprocal step needs no step bound: each step moves
        // one axis to its next integer plane ahead, so a walk leaves the grid after <= 3(N+1)
        if constexpr (!AF) sk = sk && ray.steps < ray.max_steps;
        if (!wave_any(sk)) break;
        float nx, ny, nz;
        dda_step_r(ray.lx, ray.ly, ray.lz, ray.dx, ray.dy, ray.dz, ray.rx, ray.ry, ray.rz, ray.sx,
                   ray.sy, ray.sz, nx, ny, nz);
        if (!AF && !ray.fast)
            dda_step(ray.lx, ray.ly, ray.lz, ray.dx, ray.dy, ray.dz, ray.sx, ray.sy, ray.sz, nx, ny,
                     nz);
        int vx, vy, vz;
        ray.voxel(nx, ny, nz, vx, vy, vz);
        const uint32_t ux = (uint32_t)vx, uy = (uint32_t)vy, uz = (uint32_t)vz;
        // (ux | uy | uz) < n implies each is below n (and is exact when n is a power of two)
        const bool inr = (ux | uy | uz) < (uint32_t)n;
        const uint32_t key = ((uz >> sh) << 20) | ((uy >> sh) << 10) | (ux >> sh);
        const uint32_t bit = ((uz & 3u) << 4) | ((uy & 3u) << 2) | (ux & 3u);
        // an empty block's brick word is zero, so the bit test covers both regions
        const bool set = ((bit < 32u ? occ.lo : occ.hi) >> (bit & 31u)) & 1u;
        sk = sk && inr && key == occ.key && !set &&
             !(RT ? outside_open(nx, ny, nz, nf) : outside(nx, ny, nz, nf));
        ray.lx = sk ? nx : ray.lx;
        ray.ly = sk ? ny : ray.ly;
        ray.lz = sk ? nz : ray.lz;
        if constexpr (!AF) ray.steps += sk ? 1 : 0;
    }
}

struct Kahan {
    float s = 0.0f, c = 0.0f;
    __device__ __forceinline__ void add(float x) {
#pragma clang fp contract(off)
        float y = x - c;
        float t = s + y;
        c = (t - s) - y;
        s = t;
    }
};


}  // namespace diver
