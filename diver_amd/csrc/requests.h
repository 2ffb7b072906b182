// Launch-wide ray ids over a table of requests (camera views or explicit ray batches).
//
// Request i owns ids [ray_start_i, ray_start_i + n_rays_i).  A camera request enumerates its
// pixels in 8x8 tiles (64 ids per tile, ids past the image edge are padding); an explicit request
// maps id k to its k-th ray.
#pragma once
#include "kernels.h"
#include "raystate.h"

namespace diver {

// The ray with id g (g < p.total): its request, output index and world-space origin/direction.
// Returns false for camera padding ids.
__device__ __forceinline__ bool ray_of(const BatchArgs& p, int64_t g, int& req, int64_t& out_idx,
                                       float& ox, float& oy, float& oz, float& dx, float& dy,
                                       float& dz) {
    int ri = 0;
    for (int i = 1; i < p.n_req; ++i)
        if (g >= p.req[i].ray_start) ri = i;
    const ReqDesc& rq = p.req[ri];
    const int64_t local = g - rq.ray_start;
    req = ri;
    if (rq.cam.width > 0) {
        int x, y;
        if (!camera_pixel(rq.cam, local, x, y)) return false;
        out_idx = (int64_t)y * rq.cam.width + x;
        camera_ray(rq.cam, x, y, ox, oy, oz, dx, dy, dz);
        return true;
    }
    out_idx = local;
    ox = rq.o[3 * local];
    oy = rq.o[3 * local + 1];
    oz = rq.o[3 * local + 2];
    dx = rq.d[3 * local];
    dy = rq.d[3 * local + 1];
    dz = rq.d[3 * local + 2];
    return true;
}

}  // namespace diver
