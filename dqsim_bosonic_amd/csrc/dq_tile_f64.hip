// dq_tile_f64.hip -- complex128, exact arithmetic, instantiations of the register-resident tile kernel.
// Compiled with -ffp-contract=off: every product and sum rounded separately, as the reference.
#include "dq_tile_launch.hpp"

using namespace dq;

namespace {

template <int B>
hipError_t launch(unsigned grid, hipStream_t stream, void *amps, const TilePass &pd, const TileOp *ops,
                  const void *mats, uint64_t ntiles, double *partials) {
    using T = double;
    hipLaunchKernelGGL((k_tile_reg<T, B, 0>), dim3(grid), dim3(1 << (B - 4)), 0, stream,
                       static_cast<cplx<T> *>(amps), pd, ops, static_cast<const cplx<T> *>(mats), ntiles, partials);
    return hipGetLastError();
}

}  // namespace

hipError_t dq_launch_tile_reg_f64(int b, unsigned grid, hipStream_t stream, void *amps, const TilePass &pd,
                                   const TileOp *ops, const void *mats, uint64_t ntiles, double *partials) {
    switch (b) {
    case 9: return launch<9>(grid, stream, amps, pd, ops, mats, ntiles, partials);
    case 12: return launch<12>(grid, stream, amps, pd, ops, mats, ntiles, partials);
    case 13: return launch<13>(grid, stream, amps, pd, ops, mats, ntiles, partials);
    default: return hipErrorInvalidValue;
    }
}
