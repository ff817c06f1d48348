// dq_tile_f64_fma.hip -- complex128, contracted arithmetic, instantiations of the register-resident tile kernel.
// Compiled with -ffp-contract=fast: products fused into adds (FMA); values agree with the exact kernel to rounding.
#include "dq_tile_launch.hpp"

using namespace dq;

namespace {

template <int B>
hipError_t launch(unsigned grid, hipStream_t stream, void *amps, const TilePass &pd, const TileOp *ops,
                  const void *mats, uint64_t ntiles, double *partials) {
    using T = double;
    hipLaunchKernelGGL((k_tile_reg<T, B, 1>), dim3(grid), dim3(1 << (B - 4)), 0, stream,
                       static_cast<cplx<T> *>(amps), pd, ops, static_cast<const cplx<T> *>(mats), ntiles, partials);
    return hipGetLastError();
}

template <int B>
hipError_t launch8(unsigned grid, hipStream_t stream, void *amps, const TilePass &pd, const TileOp *ops,
                   const void *mats, uint64_t ntiles, double *partials) {
    using T = double;
    hipLaunchKernelGGL((k_tile_reg8<T, B, 1>), dim3(grid), dim3(1 << (B - 3)), 0, stream,
                       static_cast<cplx<T> *>(amps), pd, ops, static_cast<const cplx<T> *>(mats), ntiles, partials);
    return hipGetLastError();
}

}  // namespace

hipError_t dq_launch_tile_reg8_f64_fma(int b, unsigned grid, hipStream_t stream, void *amps, const TilePass &pd,
                                       const TileOp *ops, const void *mats, uint64_t ntiles, double *partials) {
    switch (b) {
    case 9: return launch8<9>(grid, stream, amps, pd, ops, mats, ntiles, partials);
    case 12: return launch8<12>(grid, stream, amps, pd, ops, mats, ntiles, partials);
    case 13: return launch8<13>(grid, stream, amps, pd, ops, mats, ntiles, partials);
    default: return hipErrorInvalidValue;
    }
}

hipError_t dq_launch_tile_reg_f64_fma(int b, unsigned grid, hipStream_t stream, void *amps, const TilePass &pd,
                                   const TileOp *ops, const void *mats, uint64_t ntiles, double *partials) {
    switch (b) {
    case 9: return launch<9>(grid, stream, amps, pd, ops, mats, ntiles, partials);
    case 12: return launch<12>(grid, stream, amps, pd, ops, mats, ntiles, partials);
    case 13: return launch<13>(grid, stream, amps, pd, ops, mats, ntiles, partials);
    default: return hipErrorInvalidValue;
    }
}
