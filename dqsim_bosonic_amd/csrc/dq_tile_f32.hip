// dq_tile_f32.hip -- complex64 instantiations of the register-resident tile kernel.
// Compiled with -ffp-contract=fast (the complex64 path makes no bit-exactness claim).
#include "dq_tile_launch.hpp"

using namespace dq;

namespace {

template <int B>
hipError_t launch(unsigned grid, hipStream_t stream, void *amps, const TilePass &pd, const TileOp *ops,
                  const void *mats, uint64_t ntiles, double *partials) {
    using T = float;
    hipLaunchKernelGGL((k_tile_reg_c64<B, 1>), dim3(grid), dim3(1 << (B - 4)), 0, stream,
                       static_cast<cplx<T> *>(amps), pd, ops, static_cast<const cplx<T> *>(mats), ntiles, partials);
    return hipGetLastError();
}

}  // namespace

hipError_t dq_launch_tile_reg_f32(int b, unsigned grid, hipStream_t stream, void *amps, const TilePass &pd,
                                   const TileOp *ops, const void *mats, uint64_t ntiles, double *partials) {
    switch (b) {
    case 9: return launch<9>(grid, stream, amps, pd, ops, mats, ntiles, partials);
    case 12: return launch<12>(grid, stream, amps, pd, ops, mats, ntiles, partials);
    case 13: return launch<13>(grid, stream, amps, pd, ops, mats, ntiles, partials);
    default: return hipErrorInvalidValue;
    }
}
