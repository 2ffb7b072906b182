// Shared device helpers for the DIVeR render path on gfx950 (MI355X).
//
// The voxel DDA below is the reference's traversal recurrence
// (utils/ray_voxel_intersection/src/ray_voxel_intersect.cu:234-288): every kernel that emits
// voxel intervals uses it, so hit lists are bit-identical across the drop-in ops and the fused
// renderer.  Contraction is pinned: the reference's nvcc build fuses `o + d*tmin` and `p + d*t`
// into FMAs (--fmad=true default), so those are explicit fmaf() calls here and every other
// operation is compiled with contraction off.  Divisions are IEEE round-to-nearest.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace diver {

constexpr int kMaxHit = 8;  // intervals per real-time march round (ray_march.cu:9)

__device__ __forceinline__ int sgnf(float x) {
    int t = x < 0 ? -1 : 0;
    return x > 0 ? 1 : t;
}

__device__ __forceinline__ float clampf(float f, float a, float b) { return fmaxf(a, fminf(f, b)); }

// Bound on DDA iterations: a ray advances one axis to a new integer plane per step, so a valid
// ray takes at most 3N+3 steps inside [0,N]^3.  The bound only matters for zero/NaN directions,
// on which the reference never terminates.
__host__ __device__ __forceinline__ int max_dda_steps(float n) {
    float m = n > 0 ? n : 0;
    if (!(m < 1e7f)) m = 1e7f;
    return 3 * (int)ceilf(m) + 8;
}

// Slab entry into [0,n]^3 (grid units).  Returns true on a hit; (px,py,pz) is the clamped entry.
__device__ __forceinline__ bool box_entry(float ox, float oy, float oz, float dx, float dy, float dz,
                                          float n, float& px, float& py, float& pz, float& tentry) {
#pragma clang fp contract(off)
    float t0 = (-ox) / dx;
    float t1 = (n - ox) / dx;
    float tmin = fminf(t0, t1);
    float tmax = fmaxf(t0, t1);
    t0 = (-oy) / dy;
    t1 = (n - oy) / dy;
    tmin = fmaxf(tmin, fminf(t0, t1));
    tmax = fminf(tmax, fmaxf(t0, t1));
    t0 = (-oz) / dz;
    t1 = (n - oz) / dz;
    tmin = fmaxf(tmin, fminf(t0, t1));
    tmax = fminf(tmax, fmaxf(t0, t1));
    px = clampf(__builtin_fmaf(dx, tmin, ox), 0.0f, n);
    py = clampf(__builtin_fmaf(dy, tmin, oy), 0.0f, n);
    pz = clampf(__builtin_fmaf(dz, tmin, oz), 0.0f, n);
    tentry = tmin;
    return !(tmin > tmax);
}

// One DDA step: next voxel-boundary crossing from (px,py,pz), snapped onto the crossed plane.
// The reference's floor(p * s + 1) * s (s = -1, 0 or 1) is floor(fma(p, s, 1)) * s here: p * s is
// exact, so the fused form rounds once, as the reference's add does.
__device__ __forceinline__ float dda_step(float px, float py, float pz, float dx, float dy, float dz,
                                          float sx, float sy, float sz, float& qx, float& qy,
                                          float& qz) {
#pragma clang fp contract(off)
    float bx = floorf(__builtin_fmaf(px, sx, 1.0f)) * sx;
    float by = floorf(__builtin_fmaf(py, sy, 1.0f)) * sy;
    float bz = floorf(__builtin_fmaf(pz, sz, 1.0f)) * sz;
    float tx = (bx - px) / dx;
    float ty = (by - py) / dy;
    float tz = (bz - pz) / dz;
    float tn = fminf(tx, fminf(ty, tz));
    float nx = __builtin_fmaf(dx, tn, px);
    float ny = __builtin_fmaf(dy, tn, py);
    float nz = __builtin_fmaf(dz, tn, pz);
    if (tn == tx) {
        nx = bx;
    } else if (tn == ty) {
        ny = by;
    } else {
        nz = bz;
    }
    qx = nx;
    qy = ny;
    qz = nz;
    return tn;
}

// x / d, correctly rounded, from y = RN(1/d): q = RN(x*y), r = x - q*d (exact by FMA),
// q' = RN(q + r*y) (Markstein's correction).  Valid for normal d with |1/d| finite and moderate
// |x|; callers use it only for direction components with |d| >= kMinFastDir.
constexpr float kMinFastDir = 1e-30f;
__device__ __forceinline__ float div_recip(float x, float d, float y) {
#pragma clang fp contract(off)
    float q = x * y;
    float r = __builtin_fmaf(-q, d, x);
    return __builtin_fmaf(r, y, q);
}

// dda_step with the three divisions done by div_recip (bit-identical to dda_step).
__device__ __forceinline__ float dda_step_r(float px, float py, float pz, float dx, float dy,
                                            float dz, float rx, float ry, float rz, float sx,
                                            float sy, float sz, float& qx, float& qy, float& qz) {
#pragma clang fp contract(off)
    float bx = floorf(__builtin_fmaf(px, sx, 1.0f)) * sx;
    float by = floorf(__builtin_fmaf(py, sy, 1.0f)) * sy;
    float bz = floorf(__builtin_fmaf(pz, sz, 1.0f)) * sz;
    float tx = div_recip(bx - px, dx, rx);
    float ty = div_recip(by - py, dy, ry);
    float tz = div_recip(bz - pz, dz, rz);
    float tn = fminf(tx, fminf(ty, tz));
    float nx = __builtin_fmaf(dx, tn, px);
    float ny = __builtin_fmaf(dy, tn, py);
    float nz = __builtin_fmaf(dz, tn, pz);
    if (tn == tx) {
        nx = bx;
    } else if (tn == ty) {
        ny = by;
    } else {
        nz = bz;
    }
    qx = nx;
    qy = ny;
    qz = nz;
    return tn;
}

// Voxel of the segment [l, p]: trunc(min(l + 1e-4, p + 1e-4) * scale) per axis.
__device__ __forceinline__ void segment_voxel(float lx, float ly, float lz, float px, float py,
                                              float pz, float scale, int& vx, int& vy, int& vz) {
#pragma clang fp contract(off)
    vx = (int)(fminf(lx + 1e-4f, px + 1e-4f) * scale);
    vy = (int)(fminf(ly + 1e-4f, py + 1e-4f) * scale);
    vz = (int)(fminf(lz + 1e-4f, pz + 1e-4f) * scale);
}

__device__ __forceinline__ void segment_voxel1(float lx, float ly, float lz, float px, float py,
                                               float pz, int& vx, int& vy, int& vz) {
#pragma clang fp contract(off)
    vx = (int)fminf(lx + 1e-4f, px + 1e-4f);
    vy = (int)fminf(ly + 1e-4f, py + 1e-4f);
    vz = (int)fminf(lz + 1e-4f, pz + 1e-4f);
}

__device__ __forceinline__ bool outside(float px, float py, float pz, float n) {
    return (px < 0) | (py < 0) | (pz < 0) | (px > n) | (py > n) | (pz > n);
}

__device__ __forceinline__ bool outside_open(float px, float py, float pz, float n) {
    return (px < 0) | (py < 0) | (pz < 0) | (px >= n) | (py >= n) | (pz >= n);
}

}  // namespace diver
