// Fused render kernel: traversal + per-voxel segment integral + decoder + front-to-back
// compositing in one launch, one wavefront per 64-ray packet, one ray per lane.
//
// Per round every live lane advances the reference's exact DDA recurrence to its next occupied
// voxel (so hit lists are bit-identical to masked_intersect), integrates that voxel's 8 corner
// features in FP32, decodes density and colour, and composites.  Transmittance is exp(-S) with S a
// compensated (Kahan) FP32 sum of densities (log(1 - alpha) = -sigma exactly), and a ray stops once
// T < t_stop.  The colour decoder runs either per lane in FP32 or on the matrix cores
// (v_mfma_f32_32x32x16 bf16/f16, FP32 accumulate) over the wave's 64 rows; the per-ray view term
// W_v PE(d) + b_c is hoisted out of the interval loop in FP32.
//
// Rays come either from explicit (o, d) arrays or from a pinhole camera; camera rays are
// generated per 8x8 pixel tile (utils/dataset/ray_utils.py:5-56 conventions) so a wavefront
// packet is spatially coherent.
#include "common.h"
#include "decoder.h"
#include "kernels.h"
#include "mfma.h"

namespace diver {

__device__ __forceinline__ bool wave_any(bool p) { return __builtin_amdgcn_ballot_w64(p) != 0; }

// Pinhole camera ray of pixel (x, y) (FP32; same arithmetic as launch_camera_rays' kernel).
__device__ __forceinline__ void camera_ray(const CameraDesc& cam, int x, int y, float& ox, float& oy,
                                           float& oz, float& dx, float& dy, float& dz) {
#pragma clang fp contract(off)
    float cx = ((float)x + 0.5f - 0.5f * (float)cam.width) / cam.focal;
    float cy = -(((float)y + 0.5f - 0.5f * (float)cam.height) / cam.focal);
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

// Ray r of a camera render in tile order: 8x8 tiles, row-major over the image.
__device__ __forceinline__ bool camera_pixel(const CameraDesc& cam, int64_t r, int& x, int& y) {
    int tiles_x = (cam.width + 7) >> 3;
    int64_t tile = r >> 6;
    int in = (int)(r & 63);
    x = (int)(tile % tiles_x) * 8 + (in & 7);
    y = (int)(tile / tiles_x) * 8 + (in >> 3);
    return x < cam.width && y < cam.height;
}

__global__ void camera_rays_kernel(CameraDesc cam, int64_t n, float* __restrict__ o,
                                   float* __restrict__ d) {
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    int x = (int)(p % cam.width), y = (int)(p / cam.width);
    float ox, oy, oz, dx, dy, dz;
    camera_ray(cam, x, y, ox, oy, oz, dx, dy, dz);
    o[3 * p] = ox; o[3 * p + 1] = oy; o[3 * p + 2] = oz;
    d[3 * p] = dx; d[3 * p + 1] = dy; d[3 * p + 2] = dz;
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

template <class D, int PREC>
__global__ void __launch_bounds__(256) render_kernel(RenderArgs a) {
    constexpr int C = D::C, H = D::H, Z = D::Z;
    const int lane = threadIdx.x & 63;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t r = wave * 64 + lane;

    // ---- ray setup ---------------------------------------------------------------------------
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
    const float n = (float)a.n;
    float lx, ly, lz;
    bool alive;
    {
#pragma clang fp contract(off)
        float gx = (ox - a.xyzmin) / a.voxel_size;
        float gy = (oy - a.xyzmin) / a.voxel_size;
        float gz = (oz - a.xyzmin) / a.voxel_size;
        lx = gx; ly = gy; lz = gz;
        bool inside = (gx >= 0) & (gy >= 0) & (gz >= 0) & (gx <= n) & (gy <= n) & (gz <= n);
        alive = valid;
        if (valid && !inside) {
            float tmin;
            alive = box_entry(gx, gy, gz, dx, dy, dz, n, lx, ly, lz, tmin);
        }
    }
    const float sx = (float)sgnf(dx), sy = (float)sgnf(dy), sz = (float)sgnf(dz);
    const int max_steps = max_dda_steps(n);
    int steps = 0;

    // per-ray view term of the colour head (FP32)
    float hv[H];
    head_view<D>(a.W, dx, dy, dz, hv);
    f32x16_t init0[H / 32], init1[H / 32];
    if constexpr (PREC != kPrecFP32) rows_to_d<H>(hv, init0, init1);

    // matrix-core weights (A operands) and post-layer biases stay in registers for the whole ray
    constexpr int KC0 = Z / 16, MT = H / 32, KCH = H / 16;
    constexpr int NPK = (PREC == kPrecFP32) ? 1 : (MT * KC0 + D::NPOST * MT * KCH + KCH);
    u32x4_t wa[NPK];
    f32x16_t pbias[MT];
    if constexpr (PREC != kPrecFP32) {
        const u32x4_t* pk = reinterpret_cast<const u32x4_t*>(a.pack);
#pragma unroll
        for (int i = 0; i < NPK; ++i) wa[i] = pk[i * 64 + lane];
        if constexpr (D::NPOST > 0) {
#pragma unroll
            for (int mt = 0; mt < MT; ++mt)
#pragma unroll
                for (int i = 0; i < 16; ++i)
                    pbias[mt][i] = a.W[D::post_off(0) + H * H + 32 * mt + 8 * (i >> 2) +
                                       4 * (lane >> 5) + (i & 3)];
        }
    }

    Kahan S;
    float T = 1.0f, cr = 0.0f, cg = 0.0f, cb = 0.0f;
    int used = 0;
    const int64_t sz_ = (int64_t)(a.n + 1) * (a.n + 1) * C, sy_ = (int64_t)(a.n + 1) * C;
    const int64_t ntot = (int64_t)a.n * a.n * a.n;

    while (true) {
        // ---- traverse to the next occupied voxel ---------------------------------------------
        bool has = false;
        float px = 0.f, py = 0.f, pz = 0.f, qx = 0.f, qy = 0.f, qz = 0.f;
        if (alive) {
            while (steps < max_steps) {
                ++steps;
                float nx, ny, nz;
                dda_step(lx, ly, lz, dx, dy, dz, sx, sy, sz, nx, ny, nz);
                if (outside(nx, ny, nz, n)) { steps = max_steps; break; }
                int vx, vy, vz;
                segment_voxel1(lx, ly, lz, nx, ny, nz, vx, vy, vz);
                int64_t ci = ((int64_t)vz * a.n + vy) * a.n + vx;
                bool occ = (ci >= 0 && ci < ntot) ? a.mask[ci] != 0 : false;
                if (occ) {
                    px = lx; py = ly; pz = lz;
                    qx = nx; qy = ny; qz = nz;
                    lx = nx; ly = ny; lz = nz;
                    has = true;
                    break;
                }
                lx = nx; ly = ny; lz = nz;
            }
            if (!has) alive = false;
        }
        if (!wave_any(has)) break;

        // ---- segment integral (FP32) ----------------------------------------------------------
        float f[C];
#pragma unroll
        for (int i = 0; i < C; ++i) f[i] = 0.0f;
        float sigma = 0.0f;
        float z[Z];
        if (has) {
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
            float x0 = px - cx, x1 = qx - cx, y0 = py - cy, y1 = qy - cy, z0 = pz - cz, z1 = qz - cz;
            float z01 = z0 + z1, y01 = y0 + y1, x01 = x0 + x1;
            float w[8];
            w[7] = (2.0f * x0 * y0 * z0 + 2.0f * x1 * y1 * z1 + x01 * y01 * z01) * (1.0f / 12.0f);
            float w2 = (y01 * z01 + y0 * z0 + y1 * z1) * (1.0f / 6.0f);
            w[5] = (x01 * z01 + x0 * z0 + x1 * z1) * (1.0f / 6.0f) - w[7];
            float w4 = (x01 * y01 + x0 * y0 + x1 * y1) * (1.0f / 6.0f);
            x01 *= 0.5f; y01 *= 0.5f; z01 *= 0.5f;
            w[4] = z01 - w[5] - w2;
            w[0] = 1.0f - w[4] - y01 - x01 + w4;
            w[6] = w2 - w[7];
            w[2] = y01 - w[6] - w4;
            w[1] = x01 - w[5] - w4;
            w[3] = w4 - w[7];
            const int cx1 = min(cx + 1, a.n), cy1 = min(cy + 1, a.n), cz1 = min(cz + 1, a.n);
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                int gx = (j & 1) ? cx1 : cx;
                int gy = ((j >> 1) & 1) ? cy1 : cy;
                int gz = (j >> 2) ? cz1 : cz;
                const float4* src =
                    reinterpret_cast<const float4*>(a.voxels + gz * sz_ + gy * sy_ + (int64_t)gx * C);
#pragma unroll
                for (int i = 0; i < C / 4; ++i) {
                    float4 q = src[i];
                    f[4 * i + 0] += w[j] * q.x;
                    f[4 * i + 1] += w[j] * q.y;
                    f[4 * i + 2] += w[j] * q.z;
                    f[4 * i + 3] += w[j] * q.w;
                }
            }
            sigma = density_act<D>(pre_and_sigma<D>(a.W, f, z));
        } else {
#pragma unroll
            for (int i = 0; i < Z; ++i) z[i] = 0.0f;
        }

        // ---- colour -----------------------------------------------------------------------------
        float rgb[3] = {0.f, 0.f, 0.f};
        if constexpr (PREC == kPrecFP32) {
            if (has) colour<D>(a.W, z, hv, rgb);
        } else {
            static_assert(D::NPRE == 0 || PREC == kPrecFP32, "matrix-core path needs NPRE == 0");
            static_assert(D::NPOST <= 1, "one post layer at most");
            u32x4_t b0[KC0], b1[KC0];
            rows_to_b<PREC, Z>(z, b0, b1);
            f32x16_t acc0[MT], acc1[MT];
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) {
                acc0[mt] = init0[mt];
                acc1[mt] = init1[mt];
#pragma unroll
                for (int kc = 0; kc < KC0; ++kc) {
                    acc0[mt] = mfma32<PREC>(wa[mt * KC0 + kc], b0[kc], acc0[mt]);
                    acc1[mt] = mfma32<PREC>(wa[mt * KC0 + kc], b1[kc], acc1[mt]);
                }
            }
            int pofs = MT * KC0;
            if constexpr (D::NPOST == 1) {
                u32x4_t c0[KCH], c1[KCH];
#pragma unroll
                for (int mt = 0; mt < MT; ++mt) {
                    d_to_b_relu<PREC>(acc0[mt], c0[2 * mt], c0[2 * mt + 1]);
                    d_to_b_relu<PREC>(acc1[mt], c1[2 * mt], c1[2 * mt + 1]);
                }
#pragma unroll
                for (int mt = 0; mt < MT; ++mt) {
                    f32x16_t s0 = pbias[mt], s1 = pbias[mt];
#pragma unroll
                    for (int kc = 0; kc < KCH; ++kc) {
                        s0 = mfma32<PREC>(wa[pofs + mt * KCH + kc], c0[kc], s0);
                        s1 = mfma32<PREC>(wa[pofs + mt * KCH + kc], c1[kc], s1);
                    }
                    acc0[mt] = s0;
                    acc1[mt] = s1;
                }
                pofs += MT * KCH;
            }
            // output layer: 3 rows padded to one 32-row tile
            u32x4_t c0[KCH], c1[KCH];
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) {
                d_to_b_relu<PREC>(acc0[mt], c0[2 * mt], c0[2 * mt + 1]);
                d_to_b_relu<PREC>(acc1[mt], c1[2 * mt], c1[2 * mt + 1]);
            }
            f32x16_t o0 = {}, o1 = {};
#pragma unroll
            for (int kc = 0; kc < KCH; ++kc) {
                o0 = mfma32<PREC>(wa[pofs + kc], c0[kc], o0);
                o1 = mfma32<PREC>(wa[pofs + kc], c1[kc], o1);
            }
            // rows 0..2 of the output tile sit in registers 0..2 of lanes 0..31; tile 1's must
            // move to lanes 32..63 (the owners of those rows)
            float t0 = o0[0], t1 = o0[1], t2 = o0[2];
            float u0 = o1[0], u1 = o1[1], u2 = o1[2];
            swap32f(t0, u0);
            swap32f(t1, u1);
            swap32f(t2, u2);
            // lanes 0..31: t = tile0 rows (own); lanes 32..63: t = tile 1 values from lanes 0..31
            const float* bo = a.W + D::out_off + 3 * H;
            rgb[0] = sigmoidf(t0 + bo[0]);
            rgb[1] = sigmoidf(t1 + bo[1]);
            rgb[2] = sigmoidf(t2 + bo[2]);
        }

        // ---- composite --------------------------------------------------------------------------
        if (has) {
            used += 1;
            float alpha = -expm1f(-sigma);
            float wgt = T * alpha;
            cr += wgt * rgb[0];
            cg += wgt * rgb[1];
            cb += wgt * rgb[2];
            S.add(sigma);
            T = __expf(-S.s);
            if (T < a.t_stop) alive = false;
        }
    }
    if (valid) {
        float4 outv = make_float4(cr, cg, cb, T);
        reinterpret_cast<float4*>(a.rgba)[out_idx] = outv;
        if (a.count) a.count[out_idx] = used;
    }
}

template <class D>
static hipError_t launch_render_t(const RenderArgs& a, int prec, hipStream_t s) {
    int64_t waves = (a.n_rays + 63) / 64;
    dim3 grid((unsigned)((waves + 3) / 4)), block(256);
    if (prec == kPrecFP32) {
        hipLaunchKernelGGL((render_kernel<D, kPrecFP32>), grid, block, 0, s, a);
    } else if constexpr (D::NPRE == 0) {
        if (prec == kPrecBF16)
            hipLaunchKernelGGL((render_kernel<D, kPrecBF16>), grid, block, 0, s, a);
        else
            hipLaunchKernelGGL((render_kernel<D, kPrecFP16>), grid, block, 0, s, a);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_render(int variant, const RenderArgs& a, int prec, hipStream_t s) {
    if (a.n_rays == 0) return hipSuccess;
    switch (variant) {
        case 0: return launch_render_t<DecB32>(a, prec, s);
        case 1: return launch_render_t<DecB64>(a, prec, s);
        case 2: return launch_render_t<DecB64C64>(a, prec, s);
        case 3: return launch_render_t<DecR32>(a, prec, s);
        case 4: return launch_render_t<DecR64>(a, prec, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_camera_rays(const CameraDesc& cam, float* o, float* d, hipStream_t s) {
    int64_t n = (int64_t)cam.width * cam.height;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(camera_rays_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, cam,
                       n, o, d);
    return hipGetLastError();
}

}  // namespace diver
