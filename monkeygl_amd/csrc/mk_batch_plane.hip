// mk_batch_plane.hip -- the MPR/CPR variant of the batch kernels, compiled
// with -ffp-contract=off like mk_plane.hip so that a batched plane or CPR
// image is bit-identical to its one-request twin.
#include <hip/hip_runtime.h>

#include "mk_batch_dev.h"

namespace {

using namespace mkb;

__global__ void __launch_bounds__(256) k_batch_plane_cpr(const uint8_t* __restrict__ block, uint32_t tile_base_off,
                                                         uint32_t desc_off) {
    const BatchDesc* d;
    int x, y;
    if (!batch_item<V_PLANE_CPR>(block, tile_base_off, desc_off, d, x, y)) return;
    if (x >= d->width || y >= d->height) return;
    const bool nn = d->flags & F_NEAREST;
    if (d->kind == MK_ITEM_PLANE) {
        if (d->sub == MK_MPR_AVERAGE) {
            if (nn) mkp::plane_pixel<MK_MPR_AVERAGE, true>(d->pa, x, y);
            else mkp::plane_pixel<MK_MPR_AVERAGE, false>(d->pa, x, y);
        } else if (d->sub == MK_MPR_MIP) {
            if (nn) mkp::plane_pixel<MK_MPR_MIP, true>(d->pa, x, y);
            else mkp::plane_pixel<MK_MPR_MIP, false>(d->pa, x, y);
        } else {
            if (nn) mkp::plane_pixel<MK_MPR_MINIP, true>(d->pa, x, y);
            else mkp::plane_pixel<MK_MPR_MINIP, false>(d->pa, x, y);
        }
    } else {
        if (nn) mkp::cpr_pixel<true>(d->ca, x, y);
        else mkp::cpr_pixel<false>(d->ca, x, y);
    }
}

}  // namespace

namespace mkb {
const void* plane_cpr_kernel() { return (const void*)k_batch_plane_cpr; }
}  // namespace mkb
