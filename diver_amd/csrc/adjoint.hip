// Recompute pass of the render adjoint.
//
// The forward fused kernel keeps no per-interval samples: only (rgb, T) and the number of
// intervals composited per ray.  The backward re-walks each ray (same walk, same FP32 integral,
// same compensated transmittance, so the same values) and writes one row per composited
// interval at the ray's row offset:
//   [p_in (3), p_out (3), s = w_s.F + b_s, T_before, alpha, F (C)]
// For decoders whose density is softplus(F[0]) the walk evaluates channel 0 only, as the forward
// does, and a second kernel gathers every row's F in parallel.  The decoder and compositing
// adjoints then run over the rows (adjoint_mlp.hip), and the grid gradient is scattered by the
// segment-integral adjoint.
#include "common.h"
#include "decoder.h"
#include "kernels.h"
#include "raystate.h"

namespace diver {

// The walk of one lane's ray, writing the row header [p_in, p_out, s, T_before, alpha] of each
// composited interval (and F unless DCH0, whose density needs channel 0 only: the features are
// then gathered afterwards by rows_features_kernel, in parallel over all rows instead of on the
// walk's serial chain).  AF: every live ray of the wave takes the reciprocal-division step (no
// step bound needed, as in the fused renderer's walk).
template <class D, bool RT, bool DCH0, bool AF>
__device__ __forceinline__ void rows_walk(RayWalk& ray, const RenderArgs& a, int todo, float* out) {
    constexpr int C = D::C;
    constexpr int RW = 9 + C;
    const int n = a.n;
    const float nf = (float)n;
    OccCache occ;
    Kahan S;
    float T = 1.0f;
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
        float w[8];
        int64_t off[8];
        segment_corners<C>(px, py, pz, nx, ny, nz, n, w, off);
        float* o = out + (int64_t)k * RW;
        float s;
        if constexpr (DCH0) {
            // the fused forward's channel-0 integral: the same corners, order and roundings
            s = 0.0f;
#pragma unroll
            for (int j = 0; j < 8; ++j) s += w[j] * a.voxels[off[j]];
        } else {
            float f[C];
            gather_integral<C>(a.voxels, w, off, f);
            s = a.W[D::bs_off];
#pragma unroll
            for (int i = 0; i < C; ++i) s += a.W[D::ws_off + i] * f[i];
#pragma unroll
            for (int i = 0; i < C; ++i) o[9 + i] = f[i];
        }
        float sigma, alpha;
        density_alpha<D>(s, sigma, alpha);
        o[0] = px; o[1] = py; o[2] = pz;
        o[3] = nx; o[4] = ny; o[5] = nz;
        o[6] = s;
        o[7] = T;
        o[8] = alpha;
        S.add(sigma);
        T = __expf(-S.s);
        if (++k == todo) ray.alive = false;
    }
}

// One wave per workgroup; lanes [0, rpw) take consecutive rays (rays_per_wave: small batches
// spread over more waves).  The walk reads occupancy through the block/brick bits (the forward's
// lookup); grids without them take rows_kernel_mask.
template <class D, bool RT, bool DCH0>
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
        rows_walk<D, RT, DCH0, true>(ray, a, todo, out);
    else
        rows_walk<D, RT, DCH0, false>(ray, a, todo, out);
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
// flight together and neighbouring lanes read neighbouring 16-byte pieces of a corner row.  The
// row count K = row_off[R - 1] + count[R - 1] is read on the device; the grid strides over it.
template <int C>
__global__ void __launch_bounds__(256) rows_features_kernel(int64_t R, const int64_t* __restrict__ row_off,
                                                            const int32_t* __restrict__ count, int n,
                                                            const float* __restrict__ voxels,
                                                            float* __restrict__ rows) {
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
    if (rt && dch0)
        hipLaunchKernelGGL((rows_kernel<D, true, true>), grid, block, 0, s, a, count, off, rows, rpw);
    else if (rt)
        hipLaunchKernelGGL((rows_kernel<D, true, false>), grid, block, 0, s, a, count, off, rows, rpw);
    else if (dch0)
        hipLaunchKernelGGL((rows_kernel<D, false, true>), grid, block, 0, s, a, count, off, rows, rpw);
    else
        hipLaunchKernelGGL((rows_kernel<D, false, false>), grid, block, 0, s, a, count, off, rows, rpw);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || !dch0) return e;
    // 2048 x 256 lanes: 65,536 rows per pass over the grid, enough to fill the GPU
    hipLaunchKernelGGL((rows_features_kernel<D::C>), dim3(2048), dim3(256), 0, s, a.n_rays, off, count,
                       a.n, a.voxels, rows);
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
