// Recompute pass of the render adjoint.
//
// The forward fused kernel keeps no per-interval samples: only (rgb, T) and the number of
// intervals composited per ray.  The backward re-walks each ray (same walk, same FP32 integral,
// same compensated transmittance, so the same values) and writes one row per composited
// interval at the ray's row offset:
//   [p_in (3), p_out (3), s = w_s.F + b_s, T_before, alpha, F (C)]
// With the block/brick bits the walk writes each row's segment only; a second kernel gathers every
// row's F (and s, sigma, alpha) in parallel, and a third runs each ray's rows in order for
// T_before, so no feature load sits on the walk's serial chain (the walk's length is the forward's count, not
// a transmittance test).  The decoder and compositing adjoints then run over the rows
// (adjoint_mlp.hip), and the grid gradient is scattered by the segment-integral adjoint.
#include "common.h"
#include "decoder.h"
#include "kernels.h"
#include "raystate.h"

namespace diver {

// The walk of one lane's ray, writing the segment [p_in, p_out] of each composited interval (the
// first `todo` occupied voxels).  AF: every live ray of the wave takes the reciprocal-division
// step (no step bound needed, as in the fused renderer's walk).
template <class D, bool RT, bool AF>
__device__ __forceinline__ void rows_walk(RayWalk& ray, const RenderArgs& a, int todo, float* out) {
    constexpr int RW = 9 + D::C;
    const int n = a.n;
    const float nf = (float)n;
    OccCache occ;
    int k = 0;
    if (todo <= 0) ray.alive = false;
    while (wave_any(ray.alive)) {
        if constexpr (kSkip > 0) skip_empty<RT, AF>(ray, occ, n);
        if (!ray.alive) continue;
        if (!AF && ray.steps >= ray.max_steps) {
            ray.alive = false;
            continue;
        }
        if constexpr (!AF) ++ray.steps;
        float nx, ny, nz;
        if constexpr (AF)
            dda_step_r(ray.lx, ray.ly, ray.lz, ray.dx, ray.dy, ray.dz, ray.rx, ray.ry, ray.rz, ray.sx,
                       ray.sy, ray.sz, nx, ny, nz);
        else
            ray.step(ray.lx, ray.ly, ray.lz, nx, ny, nz);
        if (RT ? outside_open(nx, ny, nz, nf) : outside(nx, ny, nz, nf)) {
            ray.alive = false;
            continue;
        }
        int vx, vy, vz;
        ray.voxel(nx, ny, nz, vx, vy, vz);
        const float px = ray.lx, py = ray.ly, pz = ray.lz;
        ray.lx = nx;
        ray.ly = ny;
        ray.lz = nz;
        if (!occ.lookup(a.mask, a.blocks, a.bricks, n, vx, vy, vz)) continue;
        float* o = out + (int64_t)k * RW;
        o[0] = px; o[1] = py; o[2] = pz;
        o[3] = nx; o[4] = ny; o[5] = nz;
        if (++k == todo) ray.alive = false;
    }
}

// One wave per workgroup; lanes [0, rpw) take consecutive rays (rays_per_wave: small batches
// spread over more waves).  The walk reads occupancy through the block/brick bits (the forward's
// lookup); grids without them take rows_kernel_mask.
template <class D, bool RT>
__global__ void __launch_bounds__(64) rows_kernel(RenderArgs a, const int32_t* __restrict__ count,
                                                  const int64_t* __restrict__ row_off,
                                                  float* __restrict__ rows, int rpw) {
    static_assert(D::NPRE == 0, "rows are written for decoders without pre-layers");
    constexpr int C = D::C;
    constexpr int RW = 9 + C;
    const int64_t r = (int64_t)blockIdx.x * rpw + threadIdx.x;
    const bool valid = (int)threadIdx.x < rpw && r < a.n_rays;
    const int64_t rr = valid ? r : 0;
    const int n = a.n;
    RayWalk ray;
    if constexpr (RT)
        ray.init_rt(a.o[3 * rr], a.o[3 * rr + 1], a.o[3 * rr + 2], a.d[3 * rr], a.d[3 * rr + 1],
                    a.d[3 * rr + 2], a.xyzmin, a.voxel_size, n, a.octrees, valid);
    else
        ray.init(a.o[3 * rr], a.o[3 * rr + 1], a.o[3 * rr + 2], a.d[3 * rr], a.d[3 * rr + 1],
                 a.d[3 * rr + 2], a.xyzmin, a.voxel_size, (float)n, valid);
    const int todo = valid ? count[rr] : 0;
    float* out = rows + row_off[rr] * RW;
    if (!wave_any(ray.alive && !ray.fast))
        rows_walk<D, RT, true>(ray, a, todo, out);
    else
        rows_walk<D, RT, false>(ray, a, todo, out);
}

// The same over the byte mask (grids without block/brick bits).
template <class D, bool RT>
__global__ void __launch_bounds__(64) rows_kernel_mask(RenderArgs a, const int32_t* __restrict__ count,
                                                       const int64_t* __restrict__ row_off,
                                                       float* __restrict__ rows, int rpw) {
    static_assert(D::NPRE == 0, "rows are written for decoders without pre-layers");
    constexpr int C = D::C;
    constexpr int RW = 9 + C;
    const int64_t r = (int64_t)blockIdx.x * rpw + threadIdx.x;
    if ((int)threadIdx.x >= rpw || r >= a.n_rays) return;
    const int n = a.n;
    RayWalk ray;
    if constexpr (RT)
        ray.init_rt(a.o[3 * r], a.o[3 * r + 1], a.o[3 * r + 2], a.d[3 * r], a.d[3 * r + 1],
                    a.d[3 * r + 2], a.xyzmin, a.voxel_size, n, a.octrees, true);
    else
        ray.init(a.o[3 * r], a.o[3 * r + 1], a.o[3 * r + 2], a.d[3 * r], a.d[3 * r + 1],
                 a.d[3 * r + 2], a.xyzmin, a.voxel_size, (float)n, true);
    const int todo = count[r];
    float* out = rows + row_off[r] * RW;
    Kahan S;
    float T = 1.0f;
    for (int k = 0; k < todo; ++k) {
        float px, py, pz, qx, qy, qz;
        if (!ray.next<RT>(a.mask, n, px, py, pz, qx, qy, qz)) break;
        float w[8];
        int64_t off[8];
        segment_corners<C>(px, py, pz, qx, qy, qz, n, w, off);
        float f[C];
        gather_integral<C>(a.voxels, w, off, f);
        float s = a.W[D::bs_off];
#pragma unroll
        for (int i = 0; i < C; ++i) s += a.W[D::ws_off + i] * f[i];
        float sigma, alpha;
        density_alpha<D>(s, sigma, alpha);
        float* o = out + (int64_t)k * RW;
        o[0] = px; o[1] = py; o[2] = pz;
        o[3] = qx; o[4] = qy; o[5] = qz;
        o[6] = s;
        o[7] = T;
        o[8] = alpha;
#pragma unroll
        for (int i = 0; i < C; ++i) o[9 + i] = f[i];
        S.add(sigma);
        T = __expf(-S.s);
    }
}

// F of each row from its segment (gather_integral's corner weights, order and roundings): 8 lanes
// per row, 4 channels per lane per pass, so a row's 8 corner loads of one channel quad are in
// flight together and neighbouring lanes read neighbouring 16-byte pieces of a corner row.  For
// DCH0 the lane holding channel 0 also writes the row's s = F[0] (the fused forward's channel-0
// sum: the same corners, order and roundings), sigma (in the T_before slot, until
// rows_transmittance_kernel replaces it) and alpha.  The row count K = row_off[R - 1] +
// count[R - 1] is read on the device; the grid strides over it.
template <class D, bool DCH0>
__global__ void __launch_bounds__(256) rows_features_kernel(int64_t R, const int64_t* __restrict__ row_off,
                                                            const int32_t* __restrict__ count, int n,
                                                            const float* __restrict__ voxels,
                                                            float* __restrict__ rows) {
    constexpr int C = D::C;
    constexpr int RW = 9 + C;
    const int64_t K = row_off[R - 1] + count[R - 1];
    const int sub = (int)(threadIdx.x & 7);
    for (int64_t row = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 3; row < K;
         row += ((int64_t)gridDim.x * blockDim.x) >> 3) {
        float* o = rows + row * RW;
        float w[8];
        int64_t off[8];
        segment_corners<C>(o[0], o[1], o[2], o[3], o[4], o[5], n, w, off);
#pragma unroll
        for (int c4 = sub; c4 < C / 4; c4 += 8) {
            float4 q[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) q[j] = reinterpret_cast<const float4*>(voxels + off[j])[c4];
            float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                s0 += w[j] * q[j].x;
                s1 += w[j] * q[j].y;
                s2 += w[j] * q[j].z;
                s3 += w[j] * q[j].w;
            }
            o[9 + 4 * c4 + 0] = s0;
            o[9 + 4 * c4 + 1] = s1;
            o[9 + 4 * c4 + 2] = s2;
            o[9 + 4 * c4 + 3] = s3;
            if (DCH0 && c4 == 0) {
                float sigma, alpha;
                density_alpha<D>(s0, sigma, alpha);
                o[6] = s0;
                o[7] = sigma;
                o[8] = alpha;
            }
        }
    }
}

// T_before of each row, one lane per ray over its rows in walk order with the forward's
// compensated sum of sigma, so every value equals what the forward composited.  DCH0: s, sigma
// and alpha are in the rows already (sigma in the T_before slot), read 8 rows ahead of the sum;
// otherwise s = w_s.F + b_s, sigma and alpha are worked out here.
template <class D, bool DCH0>
__global__ void __launch_bounds__(256) rows_transmittance_kernel(int64_t R,
                                                                 const int64_t* __restrict__ row_off,
                                                                 const int32_t* __restrict__ count,
                                                                 const float* __restrict__ W,
                                                                 float* __restrict__ rows) {
    constexpr int C = D::C;
    constexpr int RW = 9 + C;
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R) return;
    float* o = rows + row_off[r] * RW;
    const int m = count[r];
    Kahan S;
    float T = 1.0f;
    if constexpr (DCH0) {
        for (int k0 = 0; k0 < m; k0 += 8) {
            float sg[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) sg[j] = k0 + j < m ? o[(int64_t)(k0 + j) * RW + 7] : 0.0f;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                if (k0 + j < m) {
                    o[(int64_t)(k0 + j) * RW + 7] = T;
                    S.add(sg[j]);
                    T = __expf(-S.s);
                }
            }
        }
    } else {
        for (int k = 0; k < m; ++k, o += RW) {
            float sv = W[D::bs_off];
#pragma unroll
            for (int i = 0; i < C; ++i) sv += W[D::ws_off + i] * o[9 + i];
            float sigma, alpha;
            density_alpha<D>(sv, sigma, alpha);
            o[6] = sv;
            o[7] = T;
            o[8] = alpha;
            S.add(sigma);
            T = __expf(-S.s);
        }
    }
}

// Per-ray suffix sums of the compositing adjoint: for row k of ray r (rows [off_r, off_r + cnt_r)),
// suffix[k] = sum over the ray's rows j > k of w_j c_j, plus bg_r.  One lane per ray runs backwards
// over its rows with FP64 sums (the rows of a ray are a few to a few hundred).
__global__ void __launch_bounds__(256) ray_suffix_kernel(int64_t R, const int64_t* __restrict__ off,
                                                         const int32_t* __restrict__ cnt,
                                                         const float* __restrict__ w,
                                                         const float* __restrict__ c,
                                                         const float* __restrict__ bg,
                                                         float* __restrict__ out) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R) return;
    const int64_t o = off[r];
    const int n = cnt[r];
    const double b = bg ? (double)bg[r] : 0.0;
    double s0 = b, s1 = b, s2 = b;
#pragma unroll 4
    for (int k = n - 1; k >= 0; --k) {
        const int64_t i = o + k;
        out[3 * i + 0] = (float)s0;
        out[3 * i + 1] = (float)s1;
        out[3 * i + 2] = (float)s2;
        const double wk = (double)w[i];
        s0 += wk * (double)c[3 * i + 0];
        s1 += wk * (double)c[3 * i + 1];
        s2 += wk * (double)c[3 * i + 2];
    }
}

hipError_t launch_ray_suffix(int64_t R, const int64_t* off, const int32_t* cnt, const float* w,
                             const float* c, const float* bg, float* out, hipStream_t s) {
    if (R <= 0) return hipSuccess;
    hipLaunchKernelGGL(ray_suffix_kernel, dim3((unsigned)((R + 255) / 256)), dim3(256), 0, s, R, off,
                       cnt, w, c, bg, out);
    return hipGetLastError();
}

template <class D>
static hipError_t launch_rows_fill(const RenderArgs& a, bool dch0, const int32_t* count,
                                   const int64_t* off, float* rows, hipStream_t s);

template <class D>
static hipError_t launch_rows_d(const RenderArgs& a, bool dch0, const int32_t* count,
                                const int64_t* off, float* rows, hipStream_t s) {
    const int rpw = rays_per_wave(a.n_rays);
    dim3 grid((unsigned)((a.n_rays + rpw - 1) / rpw)), block(64);
    const bool rt = a.traversal == 1;
    if (!(a.blocks && a.bricks)) {
        if (rt)
            hipLaunchKernelGGL((rows_kernel_mask<D, true>), grid, block, 0, s, a, count, off, rows, rpw);
        else
            hipLaunchKernelGGL((rows_kernel_mask<D, false>), grid, block, 0, s, a, count, off, rows, rpw);
        return hipGetLastError();
    }
    if (rt)
        hipLaunchKernelGGL((rows_kernel<D, true>), grid, block, 0, s, a, count, off, rows, rpw);
    else
        hipLaunchKernelGGL((rows_kernel<D, false>), grid, block, 0, s, a, count, off, rows, rpw);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_rows_fill<D>(a, dch0, count, off, rows, s);
}

// Entry and exit of each row copied from the forward's segments [R][cap][6]: 8 lanes per ray.
template <class D>
__global__ void __launch_bounds__(256) rows_segments_kernel(int64_t R, const int32_t* __restrict__ count,
                                                            const int64_t* __restrict__ row_off,
                                                            const float* __restrict__ seg, int cap,
                                                            float* __restrict__ rows,
                                                            int64_t* __restrict__ ray) {
    constexpr int RW = 9 + D::C;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t r = t >> 3;
    if (r >= R) return;
    const int m = count[r] < cap ? count[r] : cap;
    const float2* in = reinterpret_cast<const float2*>(seg + r * (int64_t)cap * 6);
    float* out = rows + row_off[r] * RW;
    for (int k = (int)(t & 7); k < m; k += 8) {
        const float2 a0 = in[3 * k], a1 = in[3 * k + 1], a2 = in[3 * k + 2];
        float* o = out + (int64_t)k * RW;
        o[0] = a0.x;
        o[1] = a0.y;
        o[2] = a1.x;
        o[3] = a1.y;
        o[4] = a2.x;
        o[5] = a2.y;
        if (ray) ray[row_off[r] + k] = r;
    }
}

template <class D>
static hipError_t launch_rows_segments_d(const RenderArgs& a, bool dch0, const int32_t* count,
                                         const int64_t* off, const float* seg, int cap, float* rows,
                                         int64_t* ray, hipStream_t s) {
    hipLaunchKernelGGL((rows_segments_kernel<D>), dim3((unsigned)((a.n_rays * 8 + 255) / 256)), dim3(256),
                       0, s, a.n_rays, count, off, seg, cap, rows, ray);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_rows_fill<D>(a, dch0, count, off, rows, s);
}

hipError_t launch_rows_from_segments(int variant, const RenderArgs& a, bool dch0, const int32_t* count,
                                     const int64_t* off, const float* segments, int seg_cap,
                                     float* rows, int64_t* ray, hipStream_t s) {
    if (a.n_rays == 0) return hipSuccess;
    switch (variant) {
        case 0: return launch_rows_segments_d<DecB32>(a, dch0, count, off, segments, seg_cap, rows, ray, s);
        case 1: return launch_rows_segments_d<DecB64>(a, dch0, count, off, segments, seg_cap, rows, ray, s);
        case 2: return launch_rows_segments_d<DecB64C64>(a, dch0, count, off, segments, seg_cap, rows, ray, s);
        default: return hipErrorInvalidValue;
    }
}

// The rows' features, s, sigma, alpha and transmittance once their entry and exit are written.
template <class D>
static hipError_t launch_rows_fill(const RenderArgs& a, bool dch0, const int32_t* count,
                                   const int64_t* off, float* rows, hipStream_t s) {
    hipError_t e;
    // 2048 x 256 lanes: 65,536 rows per pass over the grid, enough to fill the GPU
    const dim3 rgrid((unsigned)((a.n_rays + 255) / 256));
    if (dch0) {
        hipLaunchKernelGGL((rows_features_kernel<D, true>), dim3(2048), dim3(256), 0, s, a.n_rays, off,
                           count, a.n, a.voxels, rows);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        hipLaunchKernelGGL((rows_transmittance_kernel<D, true>), rgrid, dim3(256), 0, s, a.n_rays, off,
                           count, a.W, rows);
    } else {
        hipLaunchKernelGGL((rows_features_kernel<D, false>), dim3(2048), dim3(256), 0, s, a.n_rays, off,
                           count, a.n, a.voxels, rows);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        hipLaunchKernelGGL((rows_transmittance_kernel<D, false>), rgrid, dim3(256), 0, s, a.n_rays, off,
                           count, a.W, rows);
    }
    return hipGetLastError();
}

hipError_t launch_rows(int variant, const RenderArgs& a, bool dch0, const int32_t* count,
                       const int64_t* off, float* rows, hipStream_t s) {
    if (a.n_rays == 0) return hipSuccess;
    switch (variant) {
        case 0: return launch_rows_d<DecB32>(a, dch0, count, off, rows, s);
        case 1: return launch_rows_d<DecB64>(a, dch0, count, off, rows, s);
        case 2: return launch_rows_d<DecB64C64>(a, dch0, count, off, rows, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace diver
