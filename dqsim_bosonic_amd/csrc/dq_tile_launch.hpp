// dq_tile_launch.hpp -- host launchers of the register-resident tile kernel.
// Each precision is compiled in its own translation unit (dq_tile_f64.hip:
// exact; dq_tile_f64_fma.hip: contracted; dq_tile_f32.hip); the engine calls
// these host functions.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dq_tile.hpp"

hipError_t dq_launch_tile_reg_f64(int b, unsigned grid, hipStream_t stream, void *amps, const dq::TilePass &pd,
                                  const dq::TileOp *ops, const void *mats, uint64_t ntiles,
                                  double *partials);
hipError_t dq_launch_tile_reg_f64_fma(int b, unsigned grid, hipStream_t stream, void *amps, const dq::TilePass &pd,
                                      const dq::TileOp *ops, const void *mats, uint64_t ntiles, double *partials);
// Eight amplitudes per thread (k_tile_reg8): layouts of 3 register bits.
hipError_t dq_launch_tile_reg8_f64_fma(int b, unsigned grid, hipStream_t stream, void *amps, const dq::TilePass &pd,
                                       const dq::TileOp *ops, const void *mats, uint64_t ntiles, double *partials);
hipError_t dq_launch_tile_reg_f32(int b, unsigned grid, hipStream_t stream, void *amps, const dq::TilePass &pd,
                                  const dq::TileOp *ops, const void *mats, uint64_t ntiles,
                                  double *partials);
hipError_t dq_launch_tile_reg_f32_mfma(int b, unsigned grid, hipStream_t stream, void *amps, const dq::TilePass &pd,
                                  const dq::TileOp *ops, const void *mats, uint64_t ntiles,
                                  double *partials);
