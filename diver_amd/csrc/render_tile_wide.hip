// The fused render kernel (render_tile.h) for grids whose features exceed 32-bit byte offsets
// (such as 256^3 x 64 channels, 4.3 GB): the same kernels with 64-bit feature addressing, in a
// translation unit of their own so both sets compile in parallel.
#include "render_tile.h"

namespace diver {

hipError_t launch_tiles_wide(int variant, const BatchArgs& p, int prec, bool dch0, hipStream_t s) {
    return launch_tiles<false>(variant, p, prec, dch0, s);
}

}  // namespace diver
