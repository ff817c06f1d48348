// dq_state.hpp -- the statevector handle behind dq_state_* (include/dqsim_gpu.h),
// shared by the translation units that launch work on a state.
#pragma once
#include <hip/hip_runtime.h>

#include <utility>
#include <vector>

#include "../../include/dqsim_gpu.h"
#include "dq_plan.hpp"

struct dq_state {
    int n = 0;
    int dtype = DQ_C128;
    int device = 0;
    size_t elem = 16;
    void *amps = nullptr;
    hipStream_t stream = nullptr;
    int fuse = 2;
    int tile_bits = 12;
    int reg_tiles = 1;  // register-resident tile kernel for tiles of >= 9 qubits
    int last_grid = 0;  // grid of the last register-tile launch (partials it wrote)
    int low_bits = 4;
    // qubits that are 0 in every amplitude of the state in memory (the queue
    // not yet applied): all of them after a reset, none after an upload;
    // register sweeps visit only the tiles where these qubits are 0
    uint64_t zero_q = 0;
    int zero_tiles = 1;  // use zero_q to skip tiles (dq_set_zero_tiles)
    std::vector<dq::HostOp> queue;
    // upload buffers
    void *d_upload = nullptr;
    size_t d_upload_cap = 0;
    unsigned char *h_upload = nullptr;
    size_t h_upload_cap = 0;
    hipEvent_t upload_done = nullptr;
    // adjoint sweep buffers (grow-only, reused across calls): device tables and
    // partials, and a pinned staging copy of the tables for one upload
    void *d_adj = nullptr;
    size_t d_adj_cap = 0;
    unsigned char *h_adj = nullptr;
    size_t h_adj_cap = 0;
    // reduction buffers
    double *d_partials = nullptr;
    double *h_result = nullptr;  // pinned, 64 doubles
    double *d_result = nullptr;
    // stats
    uint64_t sweeps = 0, launches = 0, gates = 0;
    uint64_t fused_blocks = 0, fused_gates = 0;  // fusion level 3 statistics
    double fuse_min_cost = 128.0;                // fusion level 3 cost threshold
    int pair_fusion = 1;                         // fusion levels 2-3: merge gates on <= 2 qubits
    int reg_bits = 4;                            // complex128 register-tile layouts at fusion levels 2-3
    // optional per-launch timing of the sweep kernels (HIP events on `stream`)
    int timing = 0;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_pending;
    std::vector<hipEvent_t> ev_pool;
    double timed_ms = 0.0;
    uint64_t timed_launches = 0;
    // amplitudes the sweep kernels have read and written (each once) so far,
    // and the count when timing was last switched
    uint64_t swept_amps = 0, swept_amps0 = 0;
};
