// Occupancy extraction: the alpha map of prune.py:83-143 in one launch.
//
// One lane per ray.  The lane walks the masked_intersect recurrence (offline walk, exit > N) to
// each occupied voxel, integrates the voxel's features over the segment in FP32, evaluates the
// decoder's density (pre-layers included, FP32) and forms the rendering weight of
// model/diver.py:194-202, w_i = alpha_i * T_i with T_i the transmittance in front of the interval
// (no early termination).  The weight goes to the voxel min(int(p_in + 1e-4), int(p_out + 1e-4))
// at flat index x + y N + z N^2 (prune.py:118-129) unless that voxel lies within `pad` of the
// border, and the map keeps the maximum (prune.py:132-133).  Weights are non-negative, so the
// maximum is an integer atomicMax on their bit patterns (sign bit cleared): exact and
// independent of the order in which lanes arrive, so repeated runs give the same bits.
#include "common.h"
#include "decoder.h"
#include "kernels.h"
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

template <class D>
static hipError_t launch_alpha_map_d(const RenderArgs& a, int pad, float* map, hipStream_t s) {
    dim3 grid((unsigned)((a.n_rays + 255) / 256)), block(256);
    hipLaunchKernelGGL((alpha_map_kernel<D>), grid, block, 0, s, a, pad,
                       reinterpret_cast<unsigned int*>(map));
    return hipGetLastError();
}

hipError_t launch_alpha_map(int variant, const RenderArgs& a, int pad, float* map, hipStream_t s) {
    if (a.n_rays == 0) return hipSuccess;
    switch (variant) {
        case 0: return launch_alpha_map_d<DecB32>(a, pad, map, s);
        case 1: return launch_alpha_map_d<DecB64>(a, pad, map, s);
        case 2: return launch_alpha_map_d<DecB64C64>(a, pad, map, s);
        case 3: return launch_alpha_map_d<DecR32>(a, pad, map, s);
        case 4: return launch_alpha_map_d<DecR64>(a, pad, map, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace diver
