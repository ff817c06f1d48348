// dq_batch.hip -- many independent circuits in one call (batching form).
//
// No reference function: the reference runs circuits one at a time (its
// closest loop is the per-shot loop of src/monolithic/statevector/
// simulator.rs:250-262).  A batch holds instances (n qubits, a gate list and
// a diagonal observable sum_t c_t Z(zmask_t)); dq_batch_run returns <H> for
// each.
//
// Two device paths, chosen per instance by size:
//   * LDS-resident: when the whole state fits one workgroup's LDS (2^n
//     amplitudes <= 128 KiB: n <= 13 complex128, n <= 14 complex64) the
//     circuit runs start to finish inside one workgroup: |0..0> is written
//     to LDS, every gate is an LDS sweep, the expectation is reduced in the
//     same workgroup and one double goes back to HBM.  One launch per qubit
//     count covers all such instances (one workgroup each).  The host
//     compiles each gate list first: single-qubit gates are multiplied per
//     qubit and applied up to three at a time in registers (one LDS sweep
//     per three qubits); runs of cx / swap / x (maps of basis states that are
//     affine over GF(2)) compose into one permutation sweep; other gates use
//     the engine's structure-aware op kinds (diagonal, monomial, controlled,
//     dense).
//   * Engine: larger instances run on the statevector engine (LDS-tiled HBM
//     sweeps), several at once from worker threads, each worker with its own
//     states and stream, so that small-to-medium instances overlap on the
//     GPU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <atomic>
#include <cstring>
#include <exception>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/dqsim_gpu.h"
#include "dq_device.hpp"
#include "dq_plan.hpp"
#include "dq_state.hpp"
#include "dq_tile.hpp"

using namespace dq;

namespace {

constexpr int K_B1Q = 16;   // up to 3 single-qubit 2x2 gates on distinct qubits, applied in registers
constexpr int K_BLIN = 17;  // basis permutation new[y] = old[src(y)], src affine over GF(2)
constexpr int kMaxLdsBytes = 128 * 1024;

struct BatchInst {
    int op_begin, op_count;
    int term_begin, term_count;
};

int lds_max_qubits(int dtype) { return dtype == DQ_C128 ? 13 : 14; }
constexpr int kPerThread = 32;  // amplitudes per thread in the gather sweeps (N / threads <= 32)
unsigned batch_threads(int n) { return std::min(512u, std::max(256u, (1u << n) / kPerThread)); }

// Up to three 2x2 gates on distinct qubits bit[0..k-1]; matrix j at mat + 4j.
template <typename T>
__device__ __forceinline__ void b1q(cplx<T> *__restrict__ sh, int n, const DevOp &op,
                                    const cplx<T> *__restrict__ mats, uint32_t tid, uint32_t nthr) {
    const int k = op.k;
    const uint32_t ngroups = 1u << (n - k);
    cplx<T> u[3][4];
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int e = 0; e < 4; ++e) u[j][e] = j < k ? mats[op.mat + 4 * j + e] : cplx<T>{T(0), T(0)};
    uint32_t off[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        off[r] = 0;
#pragma unroll
        for (int j = 0; j < 3; ++j)
            if (j < k && ((r >> j) & 1)) off[r] |= 1u << op.bit[j];
    }
    for (uint32_t g = tid; g < ngroups; g += nthr) {
        const uint32_t base = insert_zeros<uint32_t>(g, op.sbit, k);
        cplx<T> v[8];
#pragma unroll
        for (int r = 0; r < 8; ++r)
            if (r < (1 << k)) v[r] = sh[base | off[r]];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            if (j >= k) break;
#pragma unroll
            for (int r = 0; r < 8; ++r) {
                if (r >= (1 << k) || ((r >> j) & 1)) continue;
                const cplx<T> x = v[r], y = v[r | (1 << j)];
                v[r] = cadd(cmul(u[j][0], x), cmul(u[j][1], y));
                v[r | (1 << j)] = cadd(cmul(u[j][2], x), cmul(u[j][3], y));
            }
        }
#pragma unroll
        for (int r = 0; r < 8; ++r)
            if (r < (1 << k)) sh[base | off[r]] = v[r];
    }
}

// new[y] = old[src(y)], bit b of src(y) = parity(y & rows[b]) ^ bit b of c.
// Each thread reads all of its sources before the barrier, then writes.
template <typename T>
__device__ __forceinline__ void blin(cplx<T> *__restrict__ sh, int n, const DevOp &op,
                                     const int *__restrict__ aux, uint32_t tid, uint32_t nthr) {
    const uint32_t N = 1u << n;
    const int *rows = aux + op.aux;
    const uint32_t c = uint32_t(rows[n]);
    cplx<T> v[kPerThread];  // N / nthr <= kPerThread (batch_threads)
#pragma unroll
    for (int j = 0; j < kPerThread; ++j) {
        const uint32_t y = tid + uint32_t(j) * nthr;
        if (y < N) {
            uint32_t src = c;
            for (int b = 0; b < n; ++b) src ^= uint32_t(__popc(y & uint32_t(rows[b])) & 1) << b;
            v[j] = sh[src];
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kPerThread; ++j) {
        const uint32_t y = tid + uint32_t(j) * nthr;
        if (y < N) sh[y] = v[j];
    }
}

// Monomial or dense op on 4-5 qubits: each thread computes its (up to 16)
// output amplitudes from LDS into registers, then writes them after a
// barrier.  Same sums as the dense row loop (acc = 0; acc += U[r][c] v[c]).
template <typename T>
__device__ __forceinline__ void bbig(cplx<T> *__restrict__ sh, int n, const DevOp &op,
                                     const cplx<T> *__restrict__ mats, const int *__restrict__ aux,
                                     uint32_t tid, uint32_t nthr) {
    const uint32_t N = 1u << n;
    const int k = op.k, dim = 1 << k;
    uint32_t gm = 0;
    for (int j = 0; j < k; ++j) gm |= 1u << op.bit[j];
    cplx<T> v[kPerThread];
#pragma unroll
    for (int j = 0; j < kPerThread; ++j) {
        const uint32_t y = tid + uint32_t(j) * nthr;
        if (y >= N) continue;
        const int r = gather_bits<uint32_t>(y, op.bit, k);
        const uint32_t base = y & ~gm;
        if (op.kind == K_MONO) {
            const int c = aux[op.aux + r];
            v[j] = cmul(mats[op.mat + r], sh[base | group_offset<uint32_t>(c, op.bit, k)]);
        } else {
            cplx<T> acc{T(0), T(0)};
            for (int c = 0; c < dim; ++c)
                acc = cadd(acc, cmul(mats[op.mat + r * dim + c], sh[base | group_offset<uint32_t>(c, op.bit, k)]));
            v[j] = acc;
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kPerThread; ++j) {
        const uint32_t y = tid + uint32_t(j) * nthr;
        if (y < N) sh[y] = v[j];
    }
}

__device__ __forceinline__ double batch_block_sum(double v, double *scratch) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) scratch[w] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0)
        for (int j = 0; j < int(blockDim.x >> 6); ++j) r += scratch[j];
    return r;
}

// One workgroup per instance: the whole circuit in LDS, then <H>.
template <typename T>
__global__ __launch_bounds__(512) void k_batch_lds(int n, const int *__restrict__ ids,
                                                    const BatchInst *__restrict__ insts,
                                                    const DevOp *__restrict__ ops,
                                                    const cplx<T> *__restrict__ mats,
                                                    const int *__restrict__ aux,
                                                    const uint64_t *__restrict__ zmask,
                                                    const double *__restrict__ zcoef,
                                                    double *__restrict__ out) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ double scratch[8];
    cplx<T> *sh = reinterpret_cast<cplx<T> *>(smem_raw);
    const int id = ids[blockIdx.x];
    const BatchInst bi = insts[id];
    const uint32_t N = 1u << n, tid = threadIdx.x, nthr = blockDim.x;
    for (uint32_t i = tid; i < N; i += nthr) sh[i] = cplx<T>{T(i == 0 ? 1 : 0), T(0)};
    __syncthreads();
    for (int o = 0; o < bi.op_count; ++o) {
        const DevOp op = ops[bi.op_begin + o];
        if (op.kind == K_B1Q)
            b1q<T>(sh, n, op, mats, tid, nthr);
        else if (op.kind == K_BLIN)
            blin<T>(sh, n, op, aux, tid, nthr);
        else if ((op.kind == K_DENSEK || op.kind == K_MONO) && op.k >= 4)
            bbig<T>(sh, n, op, mats, aux, tid, nthr);
        else
            apply_op<T, uint32_t, 3>(sh, n, op, mats, aux, tid, nthr);
        __syncthreads();
    }
    double acc = 0.0;
    for (uint32_t i = tid; i < N; i += nthr) {
        const double re = double(sh[i].re), im = double(sh[i].im);
        double e = 0.0;
        for (int t = 0; t < bi.term_count; ++t) {
            const double c = zcoef[bi.term_begin + t];
            e += (__popcll(uint64_t(i) & zmask[bi.term_begin + t]) & 1) ? -c : c;
        }
        acc += (re * re + im * im) * e;
    }
    const double r = batch_block_sum(acc, scratch);
    if (tid == 0) out[id] = r;
}

// ---- persistent path: one workgroup per instance, instances from a queue ----
//
// Instances too large for LDS but small enough that one of them cannot fill
// the GPU (14..kPersistMaxQubits qubits) run here: a single launch of as many
// workgroups as the GPU holds at once; each workgroup takes the next
// instance from a device counter, keeps its state in its own slot of a
// global arena and applies the instance's planned tile passes (the engine's
// register tile sweep, walking all tiles of the state itself), then reduces
// <H> and writes one double.  Instances are queued largest first.  Every
// workgroup leaves the loop when the counter passes the last instance.
constexpr int kPersistTileBits = 12;
constexpr int kPersistMaxQubits = 20;
constexpr int kPersistThreads = 1 << (kPersistTileBits - 4);

struct PersistItem {
    int pass_begin, pass_count;
    int term_begin, term_count;
    int n;
    int id;  // output index
};

__global__ void k_batch_reset_counter(unsigned *counter) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *counter = 0u;
}

template <typename T>
__global__ __launch_bounds__(kPersistThreads, 2) void k_batch_persist(
    const PersistItem *__restrict__ items, int n_items, unsigned *__restrict__ counter, cplx<T> *__restrict__ arena,
    uint64_t slot_amps, const TilePass *__restrict__ passes, const TileOp *__restrict__ ops,
    const cplx<T> *__restrict__ mats, const uint64_t *__restrict__ zmask, const double *__restrict__ zcoef,
    double *__restrict__ out) {
    __shared__ int s_item;
    __shared__ double s_red[kPersistThreads / 64];
    extern __shared__ __align__(16) unsigned char tile_lds[];
    cplx<T> *amps = arena + uint64_t(blockIdx.x) * slot_amps;
    const unsigned t = threadIdx.x;
    for (;;) {
        __syncthreads();  // every thread has read s_item of the previous instance
        if (t == 0) s_item = int(atomicAdd(counter, 1u));
        __syncthreads();
        const int it = s_item;
        if (it >= n_items) break;
        const PersistItem item = items[it];
        const uint64_t N = uint64_t(1) << item.n;
        for (uint64_t i = t; i < N; i += kPersistThreads) amps[i] = cplx<T>{T(i == 0 ? 1 : 0), T(0)};
        // the stores of this workgroup are visible to its loads of the next pass
        __threadfence();
        __syncthreads();
        for (int p = 0; p < item.pass_count; ++p) {
            const TilePass &pd = passes[item.pass_begin + p];
            tile_sweep<T, kPersistTileBits, 1, 4>(amps, pd, ops, mats, uint64_t(1) << pd.n_outer, nullptr, 0, 1,
                                                  tile_lds);
            __threadfence();
            __syncthreads();
        }
        double acc = 0.0;
        for (uint64_t i = t; i < N; i += kPersistThreads) {
            const double re = double(amps[i].re), im = double(amps[i].im);
            double e = 0.0;
            for (int q = 0; q < item.term_count; ++q) {
                const double c = zcoef[item.term_begin + q];
                e += (__popcll(i & zmask[item.term_begin + q]) & 1) ? -c : c;
            }
            acc += (re * re + im * im) * e;
        }
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
        if ((t & 63) == 0) s_red[t >> 6] = acc;
        __syncthreads();
        if (t == 0) {
            double r = 0.0;
            for (int w = 0; w < kPersistThreads / 64; ++w) r += s_red[w];
            out[item.id] = r;
        }
    }
}

}  // namespace

struct BGate {
    int k = 1;
    int q[5] = {0, 0, 0, 0, 0};
    std::vector<double> u;  // 2 * 4^k doubles, row-major, q[0] = MSB
};

struct BInstance {
    int n = 0;
    std::vector<BGate> gates;
    std::vector<uint64_t> zmask;
    std::vector<double> zcoef;
    bool has_obs = false;
};

// Choice of the batch autotuner for one (width, dtype): the path and, on the
// engine path, the sweep configuration.
struct BatchTune {
    int path = 2;  // 1: persistent workgroup-per-instance launch, 2: engine
    int fusion = 2, tile_bits = 12, reg_bits = 4;
    double persist_ms = 0.0, engine_ms = 0.0;  // measured per instance of the proxy (persist: per workgroup slot)
};

struct dq_batch {
    int dtype = DQ_C128;
    int workers = 8;
    int autotune = 1;
    std::vector<BInstance> inst;
    uint64_t lds_instances = 0, engine_instances = 0, persist_instances = 0;
    double lds_ms = 0.0;    // GPU time of the captured graph (LDS and persistent launches)
    double engine_s = 0.0;  // wall time of the engine path
    double bytes[3] = {0, 0, 0};  // algorithmic HBM bytes of the last run per path (LDS: 0)
    // persistent-path arena (grow-only)
    void *arena = nullptr;
    size_t arena_bytes = 0;
    ~dq_batch() {
        if (arena) (void)hipFree(arena);
    }
};

namespace {

int bfail(int code, const std::string &msg) { return set_error(code, msg); }

// ---- host compile of one instance for the LDS path ----

struct LdsProgram {
    std::vector<DevOp> ops;
    std::vector<double> mats;
    std::vector<int> aux;
};

void mul2(const double *a, const double *b, double *c) {  // c = a * b (2x2 complex)
    for (int r = 0; r < 2; ++r)
        for (int col = 0; col < 2; ++col) {
            double re = 0, im = 0;
            for (int k = 0; k < 2; ++k) {
                const double ar = a[2 * (2 * r + k)], ai = a[2 * (2 * r + k) + 1];
                const double br = b[2 * (2 * k + col)], bi = b[2 * (2 * k + col) + 1];
                re += ar * br - ai * bi;
                im += ar * bi + ai * br;
            }
            c[2 * (2 * r + col)] = re;
            c[2 * (2 * r + col) + 1] = im;
        }
}

// If the gate maps basis states to basis states (entries exactly 0 or 1) by
// an affine map over GF(2) of its qubits, returns the inverse map on the full
// n-bit index as rows/const (bit b of inverse(y) = parity(y & rows[b]) ^ c_b).
bool affine_inverse(const BGate &g, int n, std::vector<uint32_t> &rows, uint32_t &c) {
    const int k = g.k, dim = 1 << k;
    std::vector<int> f(dim, -1);  // forward map on local indices
    for (int col = 0; col < dim; ++col) {
        for (int r = 0; r < dim; ++r) {
            const double re = g.u[2 * (size_t(r) * dim + col)], im = g.u[2 * (size_t(r) * dim + col) + 1];
            if (im != 0.0) return false;
            if (re == 1.0) {
                if (f[col] >= 0) return false;
                f[col] = r;
            } else if (re != 0.0) {
                return false;
            }
        }
        if (f[col] < 0) return false;
    }
    std::vector<int> inv(dim, -1);
    for (int col = 0; col < dim; ++col) {
        if (inv[f[col]] >= 0) return false;
        inv[f[col]] = col;
    }
    // local bit j (MSB-first index position k-1-j) <-> qubit q[j]
    const int c0 = inv[0];
    std::vector<int> colmap(k);  // inv(e_i) ^ inv(0) for local unit vectors
    for (int i = 0; i < k; ++i) colmap[i] = inv[1 << i] ^ c0;
    for (int y = 0; y < dim; ++y) {
        int v = c0;
        for (int i = 0; i < k; ++i)
            if ((y >> i) & 1) v ^= colmap[i];
        if (v != inv[y]) return false;  // not affine
    }
    rows.assign(size_t(n), 0);
    c = 0;
    for (int b = 0; b < n; ++b) rows[size_t(b)] = 1u << b;  // identity outside the gate
    for (int j = 0; j < k; ++j) {
        const int lbit = k - 1 - j;  // local bit of qubit q[j]
        uint32_t row = 0;
        for (int i = 0; i < k; ++i)
            if ((colmap[i] >> lbit) & 1) row |= 1u << g.q[k - 1 - i];
        rows[size_t(g.q[j])] = row;
        if ((c0 >> lbit) & 1) c |= 1u << g.q[j];
    }
    return true;
}

void compile_lds(const BInstance &in, LdsProgram &pg) {
    const int n = in.n;
    std::vector<std::vector<double>> pend(static_cast<size_t>(n));  // pending 2x2 per qubit (empty = none)
    std::vector<uint32_t> lin_rows;                    // open permutation (src form)
    uint32_t lin_c = 0;
    bool lin_open = false;
    auto emit_lin = [&]() {
        if (!lin_open) return;
        DevOp d;
        std::memset(&d, 0, sizeof d);
        d.kind = K_BLIN;
        d.aux = int(pg.aux.size());
        for (int b = 0; b < n; ++b) pg.aux.push_back(int(lin_rows[size_t(b)]));
        pg.aux.push_back(int(lin_c));
        pg.ops.push_back(d);
        lin_open = false;
    };
    auto flush_1q = [&]() {
        std::vector<int> qs;
        for (int q = 0; q < n; ++q)
            if (!pend[size_t(q)].empty()) qs.push_back(q);
        if (!qs.empty()) emit_lin();
        for (size_t i = 0; i < qs.size(); i += 3) {
            DevOp d;
            std::memset(&d, 0, sizeof d);
            d.kind = K_B1Q;
            d.k = int(std::min<size_t>(3, qs.size() - i));
            d.mat = int(pg.mats.size() / 2);
            for (int j = 0; j < d.k; ++j) {
                const int q = qs[i + size_t(j)];
                d.bit[j] = q;
                d.sbit[j] = q;
                pg.mats.insert(pg.mats.end(), pend[size_t(q)].begin(), pend[size_t(q)].end());
                pend[size_t(q)].clear();
            }
            std::sort(d.sbit, d.sbit + d.k);
            pg.ops.push_back(d);
        }
    };
    for (const BGate &g : in.gates) {
        if (g.k == 1) {
            auto &p = pend[size_t(g.q[0])];
            if (p.empty()) {
                p = g.u;
            } else {
                std::vector<double> m(8);
                mul2(g.u.data(), p.data(), m.data());  // later gate on the left
                p = m;
            }
            continue;
        }
        std::vector<uint32_t> grows;
        uint32_t gc = 0;
        if (affine_inverse(g, n, grows, gc)) {
            // pending single-qubit gates on this gate's qubits must come first
            bool touches = false;
            for (int j = 0; j < g.k; ++j) touches |= !pend[size_t(g.q[j])].empty();
            if (touches) flush_1q();
            if (!lin_open) {
                lin_rows.assign(size_t(n), 0);
                for (int b = 0; b < n; ++b) lin_rows[size_t(b)] = 1u << b;
                lin_c = 0;
                lin_open = true;
            }
            // src'(y) = src(ginv(y)): rows'[b] = xor of grows[j] over j in rows[b]
            std::vector<uint32_t> nr(size_t(n), 0);
            uint32_t nc = lin_c;
            for (int b = 0; b < n; ++b) {
                for (int j = 0; j < n; ++j)
                    if ((lin_rows[size_t(b)] >> j) & 1) nr[size_t(b)] ^= grows[size_t(j)];
                nc ^= uint32_t(__builtin_popcount(gc & lin_rows[size_t(b)]) & 1) << b;
            }
            lin_rows = nr;
            lin_c = nc;
            continue;
        }
        flush_1q();
        emit_lin();
        HostOp h;
        h.k = g.k;
        for (int j = 0; j < g.k; ++j) h.q[j] = g.q[j];
        classify(h, g.u.data());
        DevOp d;
        std::memset(&d, 0, sizeof d);
        d.kind = h.kind;
        d.k = h.k;
        d.mat = int(pg.mats.size() / 2);
        pg.mats.insert(pg.mats.end(), h.mat.begin(), h.mat.end());
        d.aux = int(pg.aux.size());
        pg.aux.insert(pg.aux.end(), h.perm.begin(), h.perm.end());
        for (int j = 0; j < h.k; ++j) d.bit[j] = d.sbit[j] = h.q[j];
        std::sort(d.sbit, d.sbit + h.k);
        if (h.kind == K_CTRL1) {
            d.cmask = h.cmask;
            d.cval = h.cval;
        }
        pg.ops.push_back(d);
    }
    flush_1q();
    emit_lin();
}

void default_observable(BInstance &in) {  // sum_{i=0}^{n-2} Z_i Z_{i+1}
    in.zmask.clear();
    in.zcoef.clear();
    for (int i = 0; i + 1 < in.n; ++i) {
        in.zmask.push_back((uint64_t(3)) << i);
        in.zcoef.push_back(1.0);
    }
}

#define B_TRY(expr)                                                                 \
    do {                                                                            \
        hipError_t e_ = (expr);                                                     \
        if (e_ != hipSuccess) return bfail(DQ_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// ---- the captured part of a run: LDS-resident and persistent instances ----

// Device tables of the LDS path, laid out in one buffer.
template <typename T>
struct LdsTables {
    std::vector<int> order;                  // instance ids grouped by width
    std::vector<std::pair<int, size_t>> by_n;  // (width, count) in order
    std::vector<unsigned char> host;
    size_t o_ids = 0, o_inst = 0, o_ops = 0, o_mats = 0, o_aux = 0, o_zm = 0, o_zc = 0, total = 0;
};

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

template <typename T>
void build_lds(const dq_batch *b, const std::vector<int> &ids, LdsTables<T> &tb) {
    std::vector<BatchInst> insts(b->inst.size());
    std::vector<DevOp> ops;
    std::vector<T> mats;
    std::vector<int> aux;
    std::vector<uint64_t> zm;
    std::vector<double> zc;
    std::map<int, std::vector<int>> by_n;
    for (int id : ids) {
        const BInstance &in = b->inst[size_t(id)];
        LdsProgram pg;
        compile_lds(in, pg);
        BatchInst bi;
        bi.op_begin = int(ops.size());
        bi.op_count = int(pg.ops.size());
        const int mbase = int(mats.size() / 2), abase = int(aux.size());
        for (DevOp d : pg.ops) {
            d.mat += mbase;
            d.aux += abase;
            ops.push_back(d);
        }
        for (double v : pg.mats) mats.push_back(T(v));
        aux.insert(aux.end(), pg.aux.begin(), pg.aux.end());
        bi.term_begin = int(zm.size());
        bi.term_count = int(in.zmask.size());
        zm.insert(zm.end(), in.zmask.begin(), in.zmask.end());
        zc.insert(zc.end(), in.zcoef.begin(), in.zcoef.end());
        insts[size_t(id)] = bi;
        by_n[in.n].push_back(id);
    }
    for (auto &kv : by_n) {
        tb.order.insert(tb.order.end(), kv.second.begin(), kv.second.end());
        tb.by_n.emplace_back(kv.first, kv.second.size());
    }
    tb.o_ids = 0;
    tb.o_inst = align256(tb.o_ids + std::max<size_t>(tb.order.size(), 1) * sizeof(int));
    tb.o_ops = align256(tb.o_inst + std::max<size_t>(insts.size(), 1) * sizeof(BatchInst));
    tb.o_mats = align256(tb.o_ops + std::max<size_t>(ops.size(), 1) * sizeof(DevOp));
    tb.o_aux = align256(tb.o_mats + std::max<size_t>(mats.size(), 2) * sizeof(T));
    tb.o_zm = align256(tb.o_aux + std::max<size_t>(aux.size(), 1) * sizeof(int));
    tb.o_zc = align256(tb.o_zm + std::max<size_t>(zm.size(), 1) * sizeof(uint64_t));
    tb.total = align256(tb.o_zc + std::max<size_t>(zc.size(), 1) * sizeof(double));
    tb.host.assign(tb.total, 0);
    unsigned char *h = tb.host.data();
    if (!tb.order.empty()) std::memcpy(h + tb.o_ids, tb.order.data(), tb.order.size() * sizeof(int));
    if (!insts.empty()) std::memcpy(h + tb.o_inst, insts.data(), insts.size() * sizeof(BatchInst));
    if (!ops.empty()) std::memcpy(h + tb.o_ops, ops.data(), ops.size() * sizeof(DevOp));
    if (!mats.empty()) std::memcpy(h + tb.o_mats, mats.data(), mats.size() * sizeof(T));
    if (!aux.empty()) std::memcpy(h + tb.o_aux, aux.data(), aux.size() * sizeof(int));
    if (!zm.empty()) std::memcpy(h + tb.o_zm, zm.data(), zm.size() * sizeof(uint64_t));
    if (!zc.empty()) std::memcpy(h + tb.o_zc, zc.data(), zc.size() * sizeof(double));
}

// Tile passes of one instance for the persistent path (the engine's planner
// at fusion level 2 on 12-qubit register tiles, the queue cut every 4096
// gates as the engine flushes it).  Offsets are local to the instance.
struct PersistPlan {
    std::vector<TilePass> passes;
    std::vector<int> pass_lay;  // first layout-table entry of each pass (in lays)
    std::vector<TileOp> ops;
    std::vector<double> mats;
    std::vector<uint32_t> lays;
    double swept_amps = 0.0;  // amplitudes the passes read and write (tiles visited x tile size)
    bool ok = true;
};

PersistPlan plan_persist(const BInstance &in, unsigned elem) {
    PersistPlan pp;
    PlanConfig cfg;
    cfg.n = in.n;
    cfg.fuse = 2;
    cfg.tile_bits = kPersistTileBits;
    cfg.low_bits = 4;
    cfg.reg_tiles = 1;
    cfg.reg_bits = 4;
    if (effective_tile_bits(&cfg) != kPersistTileBits) {
        pp.ok = false;
        return pp;
    }
    std::vector<HostOp> all;
    all.reserve(in.gates.size());
    for (const BGate &g : in.gates) {
        HostOp op;
        op.k = g.k;
        for (int j = 0; j < g.k; ++j) op.q[j] = g.q[j];
        classify(op, g.u.data());
        all.push_back(std::move(op));
    }
    // the slot starts as |0...0>: qubits no pass has touched yet are 0
    // everywhere, and each pass visits only the tiles where they are 0
    KnownBits kb;
    kb.zero = ~uint64_t(0) >> (64 - in.n);
    for (size_t c0 = 0; c0 < all.size(); c0 += 4096) {
        std::vector<HostOp> queue(all.begin() + long(c0), all.begin() + long(std::min(all.size(), c0 + 4096)));
        for (const Pass &p : schedule(&cfg, queue, true)) {
            if (!p.S) {
                pp.ok = false;
                return pp;
            }
            RegPlan rp;
            plan_pass(&cfg, queue, p, pp.mats, rp, kb.zero);
            advance_known(queue, p, kb);
            pp.swept_amps += std::ldexp(1.0, rp.b + rp.pd.n_outer);
            if (rp.b != kPersistTileBits) {
                pp.ok = false;
                return pp;
            }
            rp.pd.op_begin = int(pp.ops.size());
            rp.pd.op_count = int(rp.ops.size());
            rp.pd.red_bit = -1;
            pp.pass_lay.push_back(int(pp.lays.size()));
            for (int rb : rp.layouts) append_layout_table(pp.lays, rb, rp.b, elem, 4);
            pp.ops.insert(pp.ops.end(), rp.ops.begin(), rp.ops.end());
            pp.passes.push_back(rp.pd);
        }
    }
    return pp;
}

// Plans every instance of `ids` on host threads.
std::vector<PersistPlan> plan_all_persist(const dq_batch *b, const std::vector<int> &ids, unsigned elem) {
    std::vector<PersistPlan> plans(ids.size());
    std::atomic<size_t> next{0};
    auto work = [&]() {
        for (size_t i = next.fetch_add(1); i < ids.size(); i = next.fetch_add(1))
            plans[i] = plan_persist(b->inst[size_t(ids[i])], elem);
    };
    const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (unsigned w = 0; w + 1 < nt; ++w) th.emplace_back(work);
    work();
    for (auto &t : th) t.join();
    return plans;
}

template <typename T>
struct PersistTables {
    int n_items = 0;
    int max_n = 0;
    std::vector<unsigned char> host;
    size_t o_items = 0, o_passes = 0, o_ops = 0, o_mats = 0, o_zm = 0, o_zc = 0, total = 0;
    double bytes = 0.0;  // algorithmic: one read and one write of the state per pass
};

// Concatenates the plans, largest instances (passes x 2^n) first.
template <typename T>
void build_persist(const dq_batch *b, const std::vector<int> &ids, std::vector<PersistPlan> &plans,
                   PersistTables<T> &tb) {
    std::vector<size_t> order(ids.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = i;
    auto work = [&](size_t i) { return double(plans[i].passes.size()) * std::ldexp(1.0, b->inst[size_t(ids[i])].n); };
    std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return work(x) > work(y); });
    std::vector<PersistItem> items;
    std::vector<TilePass> passes;
    std::vector<TileOp> ops;
    std::vector<T> mats;
    std::vector<uint32_t> lays;
    std::vector<int> pass_lay;  // global lays entry of every pass
    std::vector<uint64_t> zm;
    std::vector<double> zc;
    for (size_t oi : order) {
        const PersistPlan &pp = plans[oi];
        const BInstance &in = b->inst[size_t(ids[oi])];
        PersistItem it;
        it.pass_begin = int(passes.size());
        it.pass_count = int(pp.passes.size());
        it.term_begin = int(zm.size());
        it.term_count = int(in.zmask.size());
        it.n = in.n;
        it.id = ids[oi];
        items.push_back(it);
        const int ops_base = int(ops.size()), mats_base = int(mats.size() / 2), lays_base = int(lays.size());
        for (size_t p = 0; p < pp.passes.size(); ++p) {
            TilePass pd = pp.passes[p];
            pd.op_begin += ops_base;
            passes.push_back(pd);
            pass_lay.push_back(lays_base + pp.pass_lay[p]);
        }
        for (TileOp op : pp.ops) {
            op.mat += mats_base;
            ops.push_back(op);
        }
        for (double v : pp.mats) mats.push_back(T(v));
        lays.insert(lays.end(), pp.lays.begin(), pp.lays.end());
        zm.insert(zm.end(), in.zmask.begin(), in.zmask.end());
        zc.insert(zc.end(), in.zcoef.begin(), in.zcoef.end());
        tb.max_n = std::max(tb.max_n, in.n);
        tb.bytes += (2.0 * pp.swept_amps + 2.0 * std::ldexp(1.0, in.n)) * double(sizeof(cplx<T>));
    }
    tb.n_items = int(items.size());
    // the layout tables follow the matrices; a pass finds its own at lay_off bytes from the matrices
    const size_t mats_b = align256(std::max<size_t>(mats.size(), 2) * sizeof(T));
    for (size_t p = 0; p < passes.size(); ++p)
        passes[p].lay_off = int(mats_b + size_t(pass_lay[p]) * sizeof(uint32_t));
    tb.o_items = 0;
    tb.o_passes = align256(std::max<size_t>(items.size(), 1) * sizeof(PersistItem));
    tb.o_ops = align256(tb.o_passes + std::max<size_t>(passes.size(), 1) * sizeof(TilePass));
    tb.o_mats = align256(tb.o_ops + std::max<size_t>(ops.size(), 1) * sizeof(TileOp));
    tb.o_zm = align256(tb.o_mats + mats_b + std::max<size_t>(lays.size(), 1) * sizeof(uint32_t));
    tb.o_zc = align256(tb.o_zm + std::max<size_t>(zm.size(), 1) * sizeof(uint64_t));
    tb.total = align256(tb.o_zc + std::max<size_t>(zc.size(), 1) * sizeof(double));
    tb.host.assign(tb.total, 0);
    unsigned char *h = tb.host.data();
    if (!items.empty()) std::memcpy(h + tb.o_items, items.data(), items.size() * sizeof(PersistItem));
    if (!passes.empty()) std::memcpy(h + tb.o_passes, passes.data(), passes.size() * sizeof(TilePass));
    if (!ops.empty()) std::memcpy(h + tb.o_ops, ops.data(), ops.size() * sizeof(TileOp));
    if (!mats.empty()) std::memcpy(h + tb.o_mats, mats.data(), mats.size() * sizeof(T));
    if (!lays.empty()) std::memcpy(h + tb.o_mats + mats_b, lays.data(), lays.size() * sizeof(uint32_t));
    if (!zm.empty()) std::memcpy(h + tb.o_zm, zm.data(), zm.size() * sizeof(uint64_t));
    if (!zc.empty()) std::memcpy(h + tb.o_zc, zc.data(), zc.size() * sizeof(double));
}

// Workgroups of the persistent launch: as many as are resident at once.
template <typename T>
int persist_grid(int n_items) {
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    const size_t shmem = sizeof(cplx<T>) << kPersistTileBits;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_batch_persist<T>, kPersistThreads, shmem) !=
        hipSuccess)
        return 0;
    return std::max(1, std::min(n_items, std::max(1, per_cu) * cus));
}

// Runs the LDS-resident and persistent instances as one captured graph:
// upload of both table sets, counter reset (a kernel, not a memset), one
// LDS launch per width, the persistent launch, download of the results.
template <typename T>
int run_graph(dq_batch *b, const std::vector<int> &lds_ids, const std::vector<int> &persist_ids, double *out,
              int grid_override = 0) {
    if (lds_ids.empty() && persist_ids.empty()) return DQ_OK;
    LdsTables<T> lt;
    build_lds<T>(b, lds_ids, lt);
    PersistTables<T> pt;
    if (!persist_ids.empty()) {
        std::vector<PersistPlan> plans = plan_all_persist(b, persist_ids, unsigned(sizeof(cplx<T>)));
        for (const PersistPlan &pp : plans)
            if (!pp.ok) return bfail(DQ_ERR_ARG, "an instance does not fit the persistent batch path");
        build_persist<T>(b, persist_ids, plans, pt);
    }
    const int grid = pt.n_items ? (grid_override > 0 ? std::min(grid_override, pt.n_items) : persist_grid<T>(pt.n_items))
                                : 0;
    if (pt.n_items && grid <= 0) return bfail(DQ_ERR_DEVICE, "persistent batch: occupancy query failed");
    const size_t slot_amps = pt.n_items ? size_t(1) << pt.max_n : 0;
    const size_t arena_need = size_t(grid) * slot_amps * sizeof(cplx<T>);
    if (arena_need > b->arena_bytes) {
        if (b->arena) (void)hipFree(b->arena);
        b->arena = nullptr;
        b->arena_bytes = 0;
        B_TRY(hipMalloc(&b->arena, arena_need));
        b->arena_bytes = arena_need;
    }
    // one device buffer: LDS tables | persistent tables | results | counter
    const size_t o_lds = 0, o_per = align256(lt.total), o_out = align256(o_per + pt.total);
    const size_t o_cnt = align256(o_out + b->inst.size() * sizeof(double)), total = o_cnt + 256;
    unsigned char *d = nullptr, *h = nullptr;
    B_TRY(hipMalloc(&d, total));
    struct Free {
        unsigned char *d, *h;
        ~Free() {
            if (d) (void)hipFree(d);
            if (h) (void)hipHostFree(h);
        }
    } guard{d, nullptr};
    B_TRY(hipHostMalloc(reinterpret_cast<void **>(&h), total, hipHostMallocDefault));
    guard.h = h;
    std::memcpy(h + o_lds, lt.host.data(), lt.total);
    if (pt.total) std::memcpy(h + o_per, pt.host.data(), pt.total);
    if (!lds_ids.empty())
        B_TRY(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_batch_lds<T>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, kMaxLdsBytes));
    hipStream_t stream;
    B_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } sguard{stream};
    hipGraph_t graph = nullptr;
    B_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    hipError_t cap = hipMemcpyAsync(d, h, o_out, hipMemcpyHostToDevice, stream);
    if (cap == hipSuccess) {
        hipLaunchKernelGGL(k_batch_reset_counter, dim3(1), dim3(64), 0, stream, reinterpret_cast<unsigned *>(d + o_cnt));
        size_t first = 0;
        for (const auto &wn : lt.by_n) {
            const int n = wn.first;
            const size_t shmem = (size_t(1) << n) * sizeof(cplx<T>);
            hipLaunchKernelGGL(k_batch_lds<T>, dim3(unsigned(wn.second)), dim3(batch_threads(n)), shmem, stream, n,
                               reinterpret_cast<const int *>(d + o_lds + lt.o_ids) + first,
                               reinterpret_cast<const BatchInst *>(d + o_lds + lt.o_inst),
                               reinterpret_cast<const DevOp *>(d + o_lds + lt.o_ops),
                               reinterpret_cast<const cplx<T> *>(d + o_lds + lt.o_mats),
                               reinterpret_cast<const int *>(d + o_lds + lt.o_aux),
                               reinterpret_cast<const uint64_t *>(d + o_lds + lt.o_zm),
                               reinterpret_cast<const double *>(d + o_lds + lt.o_zc),
                               reinterpret_cast<double *>(d + o_out));
            first += wn.second;
        }
        if (pt.n_items) {
            unsigned char *dp = d + o_per;
            hipLaunchKernelGGL(k_batch_persist<T>, dim3(unsigned(grid)), dim3(kPersistThreads),
                               sizeof(cplx<T>) << kPersistTileBits, stream,
                               reinterpret_cast<const PersistItem *>(dp + pt.o_items), pt.n_items,
                               reinterpret_cast<unsigned *>(d + o_cnt), static_cast<cplx<T> *>(b->arena),
                               uint64_t(slot_amps), reinterpret_cast<const TilePass *>(dp + pt.o_passes),
                               reinterpret_cast<const TileOp *>(dp + pt.o_ops),
                               reinterpret_cast<const cplx<T> *>(dp + pt.o_mats),
                               reinterpret_cast<const uint64_t *>(dp + pt.o_zm),
                               reinterpret_cast<const double *>(dp + pt.o_zc), reinterpret_cast<double *>(d + o_out));
        }
        cap = hipGetLastError();
        if (cap == hipSuccess)
            cap = hipMemcpyAsync(h + o_out, d + o_out, b->inst.size() * sizeof(double), hipMemcpyDeviceToHost, stream);
    }
    const hipError_t end = hipStreamEndCapture(stream, &graph);
    struct GraphGuard {
        hipGraph_t g = nullptr;
        hipGraphExec_t e = nullptr;
        ~GraphGuard() {
            if (e) (void)hipGraphExecDestroy(e);
            if (g) (void)hipGraphDestroy(g);
        }
    } gguard{graph, nullptr};
    if (cap != hipSuccess) return bfail(DQ_ERR_DEVICE, std::string("batch graph capture: ") + hipGetErrorString(cap));
    B_TRY(end);
    B_TRY(hipGraphInstantiate(&gguard.e, graph, nullptr, nullptr, 0));
    hipEvent_t e0, e1;
    B_TRY(hipEventCreate(&e0));
    B_TRY(hipEventCreate(&e1));
    B_TRY(hipEventRecord(e0, stream));
    B_TRY(hipGraphLaunch(gguard.e, stream));
    B_TRY(hipEventRecord(e1, stream));
    B_TRY(hipStreamSynchronize(stream));
    float ms = 0;
    B_TRY(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b->lds_ms = ms;
    b->bytes[1] = pt.bytes;
    const double *res = reinterpret_cast<const double *>(h + o_out);
    for (int id : lds_ids) out[id] = res[size_t(id)];
    for (int id : persist_ids) out[id] = res[size_t(id)];
    return DQ_OK;
}

// ---- the autotuner ----
//
// Per (width, dtype), chosen once per process and cached: the path of the
// instances of that width (persistent launch or engine) and, on the engine
// path, the sweep configuration (tile qubits 12 or 13; complex128: 3 or 4
// register bits; complex64: fusion level 2 or 3).  Each candidate runs a
// proxy -- the first instance of that width cut to its first 12 n gates
// (four layers of the batch ansatz) -- and is timed: the engine candidates
// by wall clock from reset to synchronisation (planning included), the
// persistent path by the GPU time of one workgroup on the proxy, divided by
// the number of workgroups that run at once.  The smallest time per
// instance wins.
std::mutex tune_mu;
std::map<std::pair<int, int>, BatchTune> tune_cache;

int engine_config(dq_state *s, int dtype, const BatchTune &t) {
    if (int rc = dq_set_options(s, t.fusion, t.tile_bits, 4)) return rc;
    if (dtype == DQ_C128)
        if (int rc = dq_set_reg_bits(s, t.reg_bits)) return rc;
    return DQ_OK;
}

int run_on_state(dq_state *s, const BInstance &in) {
    if (int rc = dq_state_reset(s)) return rc;
    for (const BGate &g : in.gates) {
        const int rc = g.k == 1 ? dq_apply_1q(s, g.u.data(), g.q[0], nullptr, nullptr, 0)
                                : dq_apply_nq(s, g.u.data(), g.q, g.k);
        if (rc) return rc;
    }
    return dq_synchronize(s);
}

template <typename T>
int tune_width(dq_batch *b, int n, const BInstance &first, BatchTune &out) {
    {
        std::lock_guard<std::mutex> lk(tune_mu);
        auto it = tune_cache.find({n, b->dtype});
        if (it != tune_cache.end()) {
            out = it->second;
            return DQ_OK;
        }
    }
    BInstance proxy = first;
    if (proxy.gates.size() > size_t(12 * n)) proxy.gates.resize(size_t(12 * n));
    BatchTune best;
    best.path = 2;
    double best_ms = 1e300;
    {
        dq_state *s = nullptr;
        if (int rc = dq_state_create(n, b->dtype, &s)) return rc;
        std::vector<BatchTune> cands;
        for (int tb : {12, 13}) {
            if (b->dtype == DQ_C128) {
                for (int rbits : {4, 3}) {
                    BatchTune c;
                    c.tile_bits = tb;
                    c.reg_bits = rbits;
                    cands.push_back(c);
                }
            } else {
                for (int fu : {2, 3}) {
                    BatchTune c;
                    c.tile_bits = tb;
                    c.fusion = fu;
                    cands.push_back(c);
                }
            }
        }
        int rc = DQ_OK;
        for (BatchTune &c : cands) {
            if ((rc = engine_config(s, b->dtype, c))) break;
            if ((rc = run_on_state(s, proxy))) break;  // warm-up
            double t = 1e300;
            for (int rep = 0; rep < 2 && rc == DQ_OK; ++rep) {
                const auto t0 = std::chrono::steady_clock::now();
                rc = run_on_state(s, proxy);
                t = std::min(t, std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
            }
            if (rc) break;
            c.engine_ms = t;
            if (t < best_ms) {
                best_ms = t;
                best = c;
            }
        }
        dq_state_destroy(s);
        if (rc) return rc;
    }
    best.engine_ms = best_ms;
    if (n > lds_max_qubits(b->dtype) && n <= kPersistMaxQubits && n >= kPersistTileBits) {
        // one workgroup on the proxy; per-instance time = its time / the workgroups resident at once
        dq_batch tmp;
        tmp.dtype = b->dtype;
        tmp.inst.push_back(proxy);
        std::vector<double> res(1);
        if (int rc = run_graph<T>(&tmp, {}, {0}, res.data(), 1)) return rc;
        int dev = 0, cus = 0, per_cu = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_batch_persist<T>, kPersistThreads,
                                                           sizeof(cplx<T>) << kPersistTileBits);
        const double per_inst = tmp.lds_ms / double(std::max(1, per_cu * cus));
        best.persist_ms = tmp.lds_ms;
        if (per_inst < best_ms) best.path = 1;
    }
    std::lock_guard<std::mutex> lk(tune_mu);
    tune_cache[{n, b->dtype}] = best;
    out = best;
    return DQ_OK;
}

// Engine path: worker threads, each with its own states (one per qubit count).
int run_engine(dq_batch *b, const std::vector<int> &ids, double *out,
               const std::map<int, BatchTune> &tuned = std::map<int, BatchTune>()) {
    if (ids.empty()) return DQ_OK;
    std::atomic<size_t> next{0};
    std::mutex mu;
    int first_rc = DQ_OK;
    std::string first_msg;
    auto worker = [&]() {
        std::map<int, dq_state *> states;
        auto record = [&](int rc, const std::string &msg) {
            std::lock_guard<std::mutex> lk(mu);
            if (first_rc == DQ_OK) {
                first_rc = rc;
                first_msg = msg;
            }
        };
        for (;;) {
            {
                std::lock_guard<std::mutex> lk(mu);
                if (first_rc != DQ_OK) break;
            }
            const size_t i = next.fetch_add(1);
            if (i >= ids.size()) break;
            const BInstance &in = b->inst[size_t(ids[i])];
            dq_state *&s = states[in.n];
            int rc = DQ_OK;
            if (!s) {
                rc = dq_state_create(in.n, b->dtype, &s);
                auto tn = tuned.find(in.n);
                if (rc == DQ_OK && tn != tuned.end()) rc = engine_config(s, b->dtype, tn->second);
            }
            const uint64_t sw0 = s->swept_amps;
            if (rc == DQ_OK) rc = dq_state_reset(s);
            for (size_t gi = 0; rc == DQ_OK && gi < in.gates.size(); ++gi) {
                const BGate &g = in.gates[gi];
                rc = g.k == 1 ? dq_apply_1q(s, g.u.data(), g.q[0], nullptr, nullptr, 0)
                              : dq_apply_nq(s, g.u.data(), g.q, g.k);
            }
            double re = 0, im = 0;
            if (rc == DQ_OK) {
                const int nt = int(in.zmask.size());
                std::vector<uint64_t> xm(size_t(std::max(nt, 1)), 0);
                std::vector<double> c(2 * size_t(std::max(nt, 1)), 0.0);
                for (int t = 0; t < nt; ++t) c[2 * size_t(t)] = in.zcoef[size_t(t)];
                rc = nt ? dq_pauli_inner(s, s, nt, xm.data(), in.zmask.data(), c.data(), &re, &im) : DQ_OK;
            }
            if (rc != DQ_OK) {
                record(rc, dq_last_error());
                break;
            }
            out[ids[i]] = re;
            {
                // one read and one write of each amplitude a sweep visits, the
                // reset's write and the expectation's read
                const double elem = b->dtype == DQ_C128 ? 16.0 : 8.0;
                const double swept = double(s->swept_amps - sw0);
                std::lock_guard<std::mutex> lk(mu);
                b->bytes[2] += (2.0 * swept + 2.0 * std::ldexp(1.0, in.n)) * elem;
            }
        }
        for (auto &kv : states)
            if (kv.second) dq_state_destroy(kv.second);
    };
    const int nw = std::max(1, std::min<int>(b->workers, int(ids.size())));
    std::vector<std::thread> th;
    for (int w = 0; w < nw; ++w) th.emplace_back(worker);
    for (auto &t : th) t.join();
    if (first_rc != DQ_OK) return bfail(first_rc, first_msg);
    return DQ_OK;
}

}  // namespace

extern "C" {

int dq_batch_create(int dtype, dq_batch **out) {
    if (!out) return bfail(DQ_ERR_ARG, "null output pointer");
    *out = nullptr;
    if (dtype != DQ_C128 && dtype != DQ_C64) return bfail(DQ_ERR_ARG, "unknown dtype");
    dq_batch *b = new dq_batch();
    b->dtype = dtype;
    *out = b;
    return DQ_OK;
}

void dq_batch_destroy(dq_batch *b) { delete b; }

int dq_batch_add_instance(dq_batch *b, int n_qubits, int *index) {
    if (!b) return bfail(DQ_ERR_ARG, "null batch");
    if (n_qubits < 1 || n_qubits > 34) return bfail(DQ_ERR_ARG, "n_qubits must be in [1, 34]");
    BInstance in;
    in.n = n_qubits;
    b->inst.push_back(std::move(in));
    if (index) *index = int(b->inst.size() - 1);
    return DQ_OK;
}

int dq_batch_apply(dq_batch *b, int index, const double *u, const int *qubits, int k) {
    if (!b || !u || !qubits) return bfail(DQ_ERR_ARG, "null argument");
    if (index < 0 || size_t(index) >= b->inst.size()) return bfail(DQ_ERR_ARG, "instance index out of range");
    BInstance &in = b->inst[size_t(index)];
    if (k < 1 || k > 5) return bfail(DQ_ERR_ARG, "gate arity must be 1..5");
    BGate g;
    g.k = k;
    uint64_t seen = 0;
    for (int j = 0; j < k; ++j) {
        if (qubits[j] < 0 || qubits[j] >= in.n)
            return bfail(DQ_ERR_ARG, "qubit index " + std::to_string(qubits[j]) + " out of range for " +
                                         std::to_string(in.n) + "-qubit instance");
        if ((seen >> qubits[j]) & 1) return bfail(DQ_ERR_ARG, "gate qubits must be distinct");
        seen |= uint64_t(1) << qubits[j];
        g.q[j] = qubits[j];
    }
    g.u.assign(u, u + 2 * (size_t(1) << (2 * k)));
    in.gates.push_back(std::move(g));
    return DQ_OK;
}

int dq_batch_set_observable(dq_batch *b, int index, int n_terms, const uint64_t *zmask, const double *coeff) {
    if (!b) return bfail(DQ_ERR_ARG, "null batch");
    if (index < 0 || size_t(index) >= b->inst.size()) return bfail(DQ_ERR_ARG, "instance index out of range");
    if (n_terms < 0 || (n_terms > 0 && (!zmask || !coeff))) return bfail(DQ_ERR_ARG, "bad observable terms");
    BInstance &in = b->inst[size_t(index)];
    for (int t = 0; t < n_terms; ++t)
        if (in.n < 64 && (zmask[t] >> in.n)) return bfail(DQ_ERR_ARG, "observable term acts outside the instance");
    in.zmask.assign(zmask, zmask + n_terms);
    in.zcoef.assign(coeff, coeff + n_terms);
    in.has_obs = true;
    return DQ_OK;
}

int dq_batch_set_workers(dq_batch *b, int workers) {
    if (!b) return bfail(DQ_ERR_ARG, "null batch");
    if (workers < 1 || workers > 64) return bfail(DQ_ERR_ARG, "workers must be in [1, 64]");
    b->workers = workers;
    return DQ_OK;
}

int dq_batch_run(dq_batch *b, double *expect) {
    if (!b || !expect) return bfail(DQ_ERR_ARG, "null argument");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return bfail(DQ_ERR_DEVICE, "no HIP device available");
    std::vector<int> small, medium, large;
    std::map<int, int> first_of;  // width -> first instance
    for (size_t i = 0; i < b->inst.size(); ++i) {
        BInstance &in = b->inst[i];
        if (!in.has_obs) default_observable(in);
        if (!first_of.count(in.n)) first_of[in.n] = int(i);
    }
    // the autotuner's choice per width above the LDS limit
    std::map<int, BatchTune> tuned;
    for (const auto &kv : first_of) {
        const int n = kv.first;
        if (n <= lds_max_qubits(b->dtype)) continue;
        BatchTune t;
        if (b->autotune) {
            const int rc = b->dtype == DQ_C128 ? tune_width<double>(b, n, b->inst[size_t(kv.second)], t)
                                               : tune_width<float>(b, n, b->inst[size_t(kv.second)], t);
            if (rc) return rc;
        } else {
            t.path = (n <= kPersistMaxQubits) ? 1 : 2;
        }
        tuned[n] = t;
    }
    b->bytes[0] = b->bytes[1] = b->bytes[2] = 0.0;
    for (size_t i = 0; i < b->inst.size(); ++i) {
        const int n = b->inst[i].n;
        if (n <= lds_max_qubits(b->dtype)) small.push_back(int(i));
        else if (tuned[n].path == 1) medium.push_back(int(i));
        else large.push_back(int(i));
    }
    b->lds_instances = small.size();
    b->persist_instances = medium.size();
    b->engine_instances = large.size();
    // the engine workers take the large instances while the captured graph
    // runs the LDS-resident and persistent ones
    int rc = DQ_OK;
    std::string msg;
    std::thread graph_thread([&]() {
        rc = b->dtype == DQ_C128 ? run_graph<double>(b, small, medium, expect)
                                 : run_graph<float>(b, small, medium, expect);
        if (rc != DQ_OK) msg = dq_last_error();
    });
    const auto t0 = std::chrono::steady_clock::now();
    const int rc2 = run_engine(b, large, expect, tuned);
    b->engine_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    const std::string msg2 = rc2 != DQ_OK ? std::string(dq_last_error()) : std::string();
    graph_thread.join();
    if (rc2 != DQ_OK) return bfail(rc2, msg2);
    if (rc != DQ_OK) return bfail(rc, msg);
    return DQ_OK;
}

int dq_batch_set_autotune(dq_batch *b, int on) {
    if (!b) return bfail(DQ_ERR_ARG, "null batch");
    if (on != 0 && on != 1) return bfail(DQ_ERR_ARG, "autotune flag must be 0 or 1");
    b->autotune = on;
    return DQ_OK;
}

int dq_batch_path_stats(const dq_batch *b, uint64_t *counts, double *graph_ms, double *engine_s, double *bytes) {
    if (!b) return bfail(DQ_ERR_ARG, "null batch");
    if (counts) {
        counts[0] = b->lds_instances;
        counts[1] = b->persist_instances;
        counts[2] = b->engine_instances;
    }
    if (graph_ms) *graph_ms = b->lds_ms;
    if (engine_s) *engine_s = b->engine_s;
    if (bytes)
        for (int k = 0; k < 3; ++k) bytes[k] = b->bytes[k];
    return DQ_OK;
}

int dq_batch_tuned(int n_qubits, int dtype, int *path, int *fusion, int *tile_bits, int *reg_bits) {
    std::lock_guard<std::mutex> lk(tune_mu);
    auto it = tune_cache.find({n_qubits, dtype});
    if (it == tune_cache.end()) return bfail(DQ_ERR_ARG, "no tuning for this width and dtype");
    if (path) *path = it->second.path;
    if (fusion) *fusion = it->second.fusion;
    if (tile_bits) *tile_bits = it->second.tile_bits;
    if (reg_bits) *reg_bits = it->second.reg_bits;
    return DQ_OK;
}

int dq_batch_stats(const dq_batch *b, uint64_t *lds_instances, uint64_t *engine_instances, double *lds_ms) {
    if (!b) return bfail(DQ_ERR_ARG, "null batch");
    if (lds_instances) *lds_instances = b->lds_instances;
    if (engine_instances) *engine_instances = b->engine_instances;
    if (lds_ms) *lds_ms = b->lds_ms;
    return DQ_OK;
}

}  // extern "C"
