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
#include "raystate.h"

namespace diver {

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

template <class D, int PREC, bool RT>
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
    RayWalk ray;
    if constexpr (RT)
        ray.init_rt(ox, oy, oz, dx, dy, dz, a.xyzmin, a.voxel_size, a.n, a.octrees, valid);
    else
        ray.init(ox, oy, oz, dx, dy, dz, a.xyzmin, a.voxel_size, (float)a.n, valid);

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
    while (true) {
        // ---- traverse to the next occupied voxel ---------------------------------------------
        float px = 0.f, py = 0.f, pz = 0.f, qx = 0.f, qy = 0.f, qz = 0.f;
        bool has = ray.next<RT>(a.mask, a.n, px, py, pz, qx, qy, qz);
        if (!wave_any(has)) break;

        // ---- segment integral (FP32) ----------------------------------------------------------
        float f[C];
        float sigma = 0.0f, alpha = 0.0f;
        float z[Z];
        if (has) {
            float w[8];
            int64_t off[8];
            segment_corners<C>(px, py, pz, qx, qy, qz, a.n, w, off);
            gather_integral<C>(a.voxels, w, off, f);
            density_alpha<D>(pre_and_sigma<D>(a.W, f, z), sigma, alpha);
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
            if (alpha < a.alpha_skip) continue;
            float wgt = T * alpha;
            cr += wgt * rgb[0];
            cg += wgt * rgb[1];
            cb += wgt * rgb[2];
            S.add(sigma);
            T = __expf(-S.s);
            if (T < a.t_stop) ray.alive = false;
        }
    }
    if (valid) {
        float4 outv = make_float4(cr, cg, cb, T);
        reinterpret_cast<float4*>(a.rgba)[out_idx] = outv;
        if (a.count) a.count[out_idx] = used;
    }
}

template <class D, bool RT>
static hipError_t launch_render_t(const RenderArgs& a, int prec, hipStream_t s) {
    int64_t waves = (a.n_rays + 63) / 64;
    dim3 grid((unsigned)((waves + 3) / 4)), block(256);
    if (prec == kPrecFP32) {
        hipLaunchKernelGGL((render_kernel<D, kPrecFP32, RT>), grid, block, 0, s, a);
    } else if constexpr (D::NPRE == 0) {
        if (prec == kPrecBF16)
            hipLaunchKernelGGL((render_kernel<D, kPrecBF16, RT>), grid, block, 0, s, a);
        else
            hipLaunchKernelGGL((render_kernel<D, kPrecFP16, RT>), grid, block, 0, s, a);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <class D>
static hipError_t launch_render_d(const RenderArgs& a, int prec, hipStream_t s) {
    return a.traversal == 1 ? launch_render_t<D, true>(a, prec, s)
                            : launch_render_t<D, false>(a, prec, s);
}

hipError_t launch_render(int variant, const RenderArgs& a, int prec, hipStream_t s) {
    if (a.n_rays == 0) return hipSuccess;
    switch (variant) {
        case 0: return launch_render_d<DecB32>(a, prec, s);
        case 1: return launch_render_d<DecB64>(a, prec, s);
        case 2: return launch_render_d<DecB64C64>(a, prec, s);
        case 3: return launch_render_d<DecR32>(a, prec, s);
        case 4: return launch_render_d<DecR64>(a, prec, s);
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
