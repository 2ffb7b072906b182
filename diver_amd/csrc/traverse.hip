// Drop-in traversal kernels (one ray per lane):
//   ray_voxel_intersect  <- utils/ray_voxel_intersection/src/ray_voxel_intersect.cu:29-151
//   masked_intersect     <- utils/ray_voxel_intersection/src/ray_voxel_intersect.cu:192-319
//   aabb_intersect       <- utils/ray_march/src/ray_march.cu:24-149
//   ray_march            <- utils/ray_march/src/ray_march.cu:151-230
// Unlike the reference these kernels (a) write every output element, padding included, so the
// caller's buffers need no pre-fill, (b) never write past max_n entries, (c) bound every DDA loop,
// and (d) launch on the caller's stream without a device-wide synchronisation.
#include "common.h"
#include "kernels.h"

namespace diver {

__global__ void __launch_bounds__(256) ray_voxel_intersect_kernel(
    int64_t batch, int max_n, float xyzmin, float voxel_num, float voxel_size,
    const float* __restrict__ o, const float* __restrict__ v, float* __restrict__ isect,
    int32_t* __restrict__ num, float* __restrict__ tns) {
#pragma clang fp contract(off)
    int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= batch) return;
    const int max_steps = max_dda_steps(voxel_num);
    float* ib = isect ? isect + (int64_t)3 * max_n * b : nullptr;
    float* tb = tns ? tns + (int64_t)max_n * b : nullptr;
    float ox = (o[3 * b + 0] - xyzmin) / voxel_size;
    float oy = (o[3 * b + 1] - xyzmin) / voxel_size;
    float oz = (o[3 * b + 2] - xyzmin) / voxel_size;
    float dx = v[3 * b + 0], dy = v[3 * b + 1], dz = v[3 * b + 2];
    float px = ox, py = oy, pz = oz;
    bool inside = (ox >= 0) & (oy >= 0) & (oz >= 0) & (ox <= voxel_num) & (oy <= voxel_num) &
                  (oz <= voxel_num);
    int count = 0;
    float t_now = 0.0f;
    bool hit = true;
    if (!inside) {
        float tmin;
        hit = box_entry(ox, oy, oz, dx, dy, dz, voxel_num, px, py, pz, tmin);
        t_now = tmin;
    }
    if (ib && max_n > 0) {
        ib[0] = px;
        ib[1] = py;
        ib[2] = pz;
    }
    if (hit) {
        if (tb && max_n > 0) tb[0] = t_now * voxel_size;
        count = 1;
        float sx = (float)sgnf(dx), sy = (float)sgnf(dy), sz = (float)sgnf(dz);
        for (int it = 0; it < max_steps; ++it) {
            float qx, qy, qz;
            float tn = dda_step(px, py, pz, dx, dy, dz, sx, sy, sz, qx, qy, qz);
            t_now += tn;
            px = qx;
            py = qy;
            pz = qz;
            if (outside(px, py, pz, voxel_num)) break;
            if (count < max_n) {
                if (ib) {
                    ib[3 * count + 0] = px;
                    ib[3 * count + 1] = py;
                    ib[3 * count + 2] = pz;
                }
                if (tb) tb[count] = t_now * voxel_size;
            }
            count += 1;
        }
    }
    num[b] = count;
    // padding, as the reference glue initialises it (zeros / -1)
    int first_pad = hit ? (count < max_n ? count : max_n) : 0;
    if (ib) {
        for (int k = (hit ? first_pad : 1); k < max_n; ++k) {
            ib[3 * k + 0] = 0.0f;
            ib[3 * k + 1] = 0.0f;
            ib[3 * k + 2] = 0.0f;
        }
    }
    if (tb) {
        for (int k = first_pad; k < max_n; ++k) tb[k] = -1.0f;
    }
}

__global__ void __launch_bounds__(256) masked_intersect_kernel(
    int64_t batch, int max_n, float xyzmin, float voxel_num, float voxel_size, float mask_scale,
    const float* __restrict__ o, const float* __restrict__ v, const uint8_t* __restrict__ mask,
    float* __restrict__ isect, int32_t* __restrict__ num, float* __restrict__ tns) {
#pragma clang fp contract(off)
    int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= batch) return;
    const int max_steps = max_dda_steps(voxel_num);
    const int mask_size = (int)(voxel_num * mask_scale);
    const int64_t mask_total = (int64_t)mask_size * mask_size * mask_size;
    float* ib = isect ? isect + (int64_t)6 * max_n * b : nullptr;
    float* tb = tns ? tns + (int64_t)max_n * b : nullptr;
    float ox = (o[3 * b + 0] - xyzmin) / voxel_size;
    float oy = (o[3 * b + 1] - xyzmin) / voxel_size;
    float oz = (o[3 * b + 2] - xyzmin) / voxel_size;
    float dx = v[3 * b + 0], dy = v[3 * b + 1], dz = v[3 * b + 2];
    float lx = ox, ly = oy, lz = oz;
    bool inside = (ox >= 0) & (oy >= 0) & (oz >= 0) & (ox <= voxel_num) & (oy <= voxel_num) &
                  (oz <= voxel_num);
    int count = 0;
    float t_now = 0.0f;
    bool hit = true;
    if (!inside) {
        float tmin;
        hit = box_entry(ox, oy, oz, dx, dy, dz, voxel_num, lx, ly, lz, tmin);
        t_now = tmin;
    }
    if (hit) {
        float sx = (float)sgnf(dx), sy = (float)sgnf(dy), sz = (float)sgnf(dz);
        for (int it = 0; it < max_steps; ++it) {
            float px, py, pz;
            float tn = dda_step(lx, ly, lz, dx, dy, dz, sx, sy, sz, px, py, pz);
            t_now += tn;
            if (outside(px, py, pz, voxel_num)) break;
            int vx, vy, vz;
            segment_voxel(lx, ly, lz, px, py, pz, mask_scale, vx, vy, vz);
            int64_t ci = (int64_t)vz * mask_size * mask_size + (int64_t)vy * mask_size + vx;
            bool occ = (ci >= 0 && ci < mask_total) ? mask[ci] != 0 : false;
            if (occ) {
                if (count < max_n) {
                    if (ib) {
                        float* r = ib + 6 * count;
                        r[0] = lx;
                        r[1] = ly;
                        r[2] = lz;
                        r[3] = px;
                        r[4] = py;
                        r[5] = pz;
                    }
                    if (tb) tb[count] = t_now * voxel_size;
                }
                count += 1;
            }
            lx = px;
            ly = py;
            lz = pz;
        }
    }
    num[b] = count;
    int first_pad = count < max_n ? count : max_n;
    if (ib) {
        for (int k = first_pad; k < max_n; ++k) {
            float* r = ib + 6 * k;
            r[0] = 0.0f; r[1] = 0.0f; r[2] = 0.0f; r[3] = 0.0f; r[4] = 0.0f; r[5] = 0.0f;
        }
    }
    if (tb) {
        for (int k = first_pad; k < max_n; ++k) tb[k] = -1.0f;
    }
}

__global__ void __launch_bounds__(256) aabb_intersect_kernel(
    int64_t batch, float* __restrict__ coords, const float* __restrict__ v,
    const float* __restrict__ center, uint8_t* __restrict__ finish, float* __restrict__ rgba,
    const uint8_t* __restrict__ octrees, int voxel_num) {
#pragma clang fp contract(off)
    int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= batch) return;
    const float n = (float)voxel_num;
    float* o = coords + batch * 6 * (kMaxHit - 1) + b * 6 + 3;
    rgba[4 * b + 0] = 0.0f;
    rgba[4 * b + 1] = 0.0f;
    rgba[4 * b + 2] = 0.0f;
    rgba[4 * b + 3] = 1.0f;
    finish[b] = 0;
    float px = center[0], py = center[1], pz = center[2];
    float dx = v[3 * b + 0], dy = v[3 * b + 1], dz = v[3 * b + 2];
    bool inside = (px >= 0) & (py >= 0) & (pz >= 0) & (px <= n) & (py <= n) & (pz <= n);
    if (!inside) {
        float tmin;
        if (!box_entry(px, py, pz, dx, dy, dz, n, px, py, pz, tmin)) {
            o[0] = -1.0f;
            finish[b] = 1;
            return;
        }
    }
    px /= 64.0f;
    py /= 64.0f;
    pz /= 64.0f;
    float lx = px, ly = py, lz = pz;
    float sx = (float)sgnf(dx), sy = (float)sgnf(dy), sz = (float)sgnf(dz);
    int mask_size = voxel_num / 64;
    int64_t oct_total = 0;
    for (int m = voxel_num / 64; m < voxel_num; m *= 4) oct_total += (int64_t)m * m * m;
    int64_t off = 0;
    const int max_steps = max_dda_steps(n) + 16;
    for (int it = 0; it < max_steps && mask_size < voxel_num; ++it) {
        lx = px;
        ly = py;
        lz = pz;
        dda_step(lx, ly, lz, dx, dy, dz, sx, sy, sz, px, py, pz);
        if (outside(px, py, pz, (float)mask_size)) {
            o[0] = -1.0f;
            finish[b] = 1;
            return;
        }
        int vx, vy, vz;
        segment_voxel1(lx, ly, lz, px, py, pz, vx, vy, vz);
        int64_t m3 = (int64_t)mask_size * mask_size * mask_size;
        int64_t ci = (int64_t)vz * mask_size * mask_size + (int64_t)vy * mask_size + vx;
        int64_t gi = off + ci;
        bool occ = (ci >= 0 && ci < m3 && gi < oct_total) ? octrees[gi] != 0 : false;
        if (occ) {
            off += m3;
            lx *= 4.0f;
            ly *= 4.0f;
            lz *= 4.0f;
            px = lx;
            py = ly;
            pz = lz;
            mask_size *= 4;
        }
    }
    if (mask_size < voxel_num) {
        o[0] = -1.0f;
        finish[b] = 1;
        return;
    }
    o[0] = lx;
    o[1] = ly;
    o[2] = lz;
}

// One real-time march round.  `stride` reproduces the reference's grid-stride ray assignment
// (256 blocks x 1024 threads, ray_march.cu:161,228): a thread that meets an unfinished ray whose
// resume point is negative leaves its loop (ray_march.cu:171-173), so the later rays of that
// thread are not marched in this round.  Here ray b checks its predecessors b - k*stride; the
// predicate reads only each predecessor's finish flag and resume x, which a non-breaking
// predecessor can only overwrite with another non-negative value, so the check is race-free.
__global__ void __launch_bounds__(256) ray_march_kernel(
    int64_t batch, float* __restrict__ coords, const float* __restrict__ v,
    const uint8_t* __restrict__ mask, const uint8_t* __restrict__ finish, int voxel_num,
    int64_t stride) {
#pragma clang fp contract(off)
    int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= batch) return;
    if (finish[b]) return;
    const int64_t last_base = (int64_t)(kMaxHit - 1) * 6 * batch + 3;
    for (int64_t p = b - stride; p >= 0; p -= stride) {
        if (!finish[p] && coords[last_base + p * 6] < 0.0f) return;
    }
    const float n = (float)voxel_num;
    const int64_t total = (int64_t)voxel_num * voxel_num * voxel_num;
    float lx = coords[last_base + b * 6], ly = coords[last_base + b * 6 + 1],
          lz = coords[last_base + b * 6 + 2];
    coords[b * 6] = -1.0f;
    if (lx < 0.0f) return;
    float dx = v[3 * b + 0], dy = v[3 * b + 1], dz = v[3 * b + 2];
    float sx = (float)sgnf(dx), sy = (float)sgnf(dy), sz = (float)sgnf(dz);
    int hit_num = 0;
    const int max_steps = max_dda_steps(n);
    for (int it = 0; it < max_steps && hit_num < kMaxHit; ++it) {
        float px, py, pz;
        dda_step(lx, ly, lz, dx, dy, dz, sx, sy, sz, px, py, pz);
        if (outside_open(px, py, pz, n)) {
            coords[b * 6 + (int64_t)hit_num * 6 * batch] = -1.0f;
            break;
        }
        int vx, vy, vz;
        segment_voxel1(lx, ly, lz, px, py, pz, vx, vy, vz);
        int64_t ci = (int64_t)vz * voxel_num * voxel_num + (int64_t)vy * voxel_num + vx;
        bool occ = (ci >= 0 && ci < total) ? mask[ci] != 0 : false;
        if (occ) {
            float* r = coords + b * 6 + (int64_t)hit_num * 6 * batch;
            r[0] = lx;
            r[1] = ly;
            r[2] = lz;
            r[3] = px;
            r[4] = py;
            r[5] = pz;
            hit_num += 1;
        }
        lx = px;
        ly = py;
        lz = pz;
    }
}

// ------------------------------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------------------------------

static inline unsigned grid_for(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }

hipError_t launch_ray_voxel_intersect(int64_t batch, int max_n, float xyzmin, float voxel_num,
                                      float voxel_size, const float* o, const float* v,
                                      float* isect, int32_t* num, float* tns, hipStream_t s) {
    if (batch == 0) return hipSuccess;
    hipLaunchKernelGGL(ray_voxel_intersect_kernel, dim3(grid_for(batch, 256)), dim3(256), 0, s,
                       batch, max_n, xyzmin, voxel_num, voxel_size, o, v, isect, num, tns);
    return hipGetLastError();
}

hipError_t launch_masked_intersect(int64_t batch, int max_n, float xyzmin, float voxel_num,
                                   float voxel_size, float mask_scale, const float* o,
                                   const float* v, const uint8_t* mask, float* isect, int32_t* num,
                                   float* tns, hipStream_t s) {
    if (batch == 0) return hipSuccess;
    hipLaunchKernelGGL(masked_intersect_kernel, dim3(grid_for(batch, 256)), dim3(256), 0, s, batch,
                       max_n, xyzmin, voxel_num, voxel_size, mask_scale, o, v, mask, isect, num,
                       tns);
    return hipGetLastError();
}

hipError_t launch_aabb_intersect(int64_t batch, float* coords, const float* v, const float* center,
                                 uint8_t* finish, float* rgba, const uint8_t* octrees,
                                 int voxel_num, hipStream_t s) {
    if (batch == 0) return hipSuccess;
    hipLaunchKernelGGL(aabb_intersect_kernel, dim3(grid_for(batch, 256)), dim3(256), 0, s, batch,
                       coords, v, center, finish, rgba, octrees, voxel_num);
    return hipGetLastError();
}

hipError_t launch_ray_march(int64_t batch, float* coords, const float* v, const uint8_t* mask,
                            const uint8_t* finish, int voxel_num, int64_t stride, hipStream_t s) {
    if (batch == 0) return hipSuccess;
    hipLaunchKernelGGL(ray_march_kernel, dim3(grid_for(batch, 256)), dim3(256), 0, s, batch, coords,
                       v, mask, finish, voxel_num, stride);
    return hipGetLastError();
}

}  // namespace diver
