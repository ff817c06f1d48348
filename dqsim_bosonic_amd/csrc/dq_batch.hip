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

struct dq_batch {
    int dtype = DQ_C128;
    int workers = 8;
    std::vector<BInstance> inst;
    uint64_t lds_instances = 0, engine_instances = 0;
    double lds_ms = 0.0;
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

template <typename T>
int run_lds(dq_batch *b, const std::vector<int> &ids, double *out, hipStream_t stream) {
    if (ids.empty()) return DQ_OK;
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
    std::vector<int> order;
    for (auto &kv : by_n) order.insert(order.end(), kv.second.begin(), kv.second.end());
    auto align = [](size_t x) { return (x + 255) & ~size_t(255); };
    const size_t o_ids = 0, o_inst = align(o_ids + order.size() * sizeof(int));
    const size_t o_ops = align(o_inst + insts.size() * sizeof(BatchInst));
    const size_t o_mats = align(o_ops + std::max<size_t>(ops.size(), 1) * sizeof(DevOp));
    const size_t o_aux = align(o_mats + std::max<size_t>(mats.size(), 2) * sizeof(T));
    const size_t o_zm = align(o_aux + std::max<size_t>(aux.size(), 1) * sizeof(int));
    const size_t o_zc = align(o_zm + std::max<size_t>(zm.size(), 1) * sizeof(uint64_t));
    const size_t o_out = align(o_zc + std::max<size_t>(zc.size(), 1) * sizeof(double));
    const size_t total = align(o_out + b->inst.size() * sizeof(double));
    std::vector<unsigned char> h(total, 0);
    std::memcpy(h.data() + o_ids, order.data(), order.size() * sizeof(int));
    std::memcpy(h.data() + o_inst, insts.data(), insts.size() * sizeof(BatchInst));
    if (!ops.empty()) std::memcpy(h.data() + o_ops, ops.data(), ops.size() * sizeof(DevOp));
    if (!mats.empty()) std::memcpy(h.data() + o_mats, mats.data(), mats.size() * sizeof(T));
    if (!aux.empty()) std::memcpy(h.data() + o_aux, aux.data(), aux.size() * sizeof(int));
    if (!zm.empty()) std::memcpy(h.data() + o_zm, zm.data(), zm.size() * sizeof(uint64_t));
    if (!zc.empty()) std::memcpy(h.data() + o_zc, zc.data(), zc.size() * sizeof(double));
    unsigned char *d = nullptr;
    B_TRY(hipMalloc(&d, total));
    struct Free {
        unsigned char *p;
        ~Free() { (void)hipFree(p); }
    } guard{d};
    B_TRY(hipMemcpyAsync(d, h.data(), total, hipMemcpyHostToDevice, stream));
    B_TRY(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_batch_lds<T>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, kMaxLdsBytes));
    hipEvent_t e0, e1;
    B_TRY(hipEventCreate(&e0));
    B_TRY(hipEventCreate(&e1));
    B_TRY(hipEventRecord(e0, stream));
    size_t first = 0;
    for (auto &kv : by_n) {
        const int n = kv.first;
        const size_t shmem = (size_t(1) << n) * sizeof(cplx<T>);
        hipLaunchKernelGGL(k_batch_lds<T>, dim3(unsigned(kv.second.size())), dim3(batch_threads(n)), shmem, stream,
                           n, reinterpret_cast<const int *>(d + o_ids) + first,
                           reinterpret_cast<const BatchInst *>(d + o_inst), reinterpret_cast<const DevOp *>(d + o_ops),
                           reinterpret_cast<const cplx<T> *>(d + o_mats), reinterpret_cast<const int *>(d + o_aux),
                           reinterpret_cast<const uint64_t *>(d + o_zm), reinterpret_cast<const double *>(d + o_zc),
                           reinterpret_cast<double *>(d + o_out));
        B_TRY(hipGetLastError());
        first += kv.second.size();
    }
    B_TRY(hipEventRecord(e1, stream));
    std::vector<double> res(b->inst.size());
    B_TRY(hipMemcpyAsync(res.data(), d + o_out, res.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
    B_TRY(hipStreamSynchronize(stream));
    float ms = 0;
    B_TRY(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b->lds_ms = ms;
    for (int id : ids) out[id] = res[size_t(id)];
    return DQ_OK;
}

// Engine path: worker threads, each with its own states (one per qubit count).
int run_engine(dq_batch *b, const std::vector<int> &ids, double *out) {
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
            if (!s) rc = dq_state_create(in.n, b->dtype, &s);
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
    std::vector<int> small, large;
    for (size_t i = 0; i < b->inst.size(); ++i) {
        BInstance &in = b->inst[i];
        if (!in.has_obs) default_observable(in);
        (in.n <= lds_max_qubits(b->dtype) ? small : large).push_back(int(i));
    }
    b->lds_instances = small.size();
    b->engine_instances = large.size();
    hipStream_t stream;
    B_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    // the LDS launches go first (asynchronous), then the engine workers
    int rc = DQ_OK;
    std::string msg;
    std::thread lds_thread([&]() {
        rc = b->dtype == DQ_C128 ? run_lds<double>(b, small, expect, stream) : run_lds<float>(b, small, expect, stream);
        if (rc != DQ_OK) msg = dq_last_error();
    });
    const int rc2 = run_engine(b, large, expect);
    const std::string msg2 = rc2 != DQ_OK ? std::string(dq_last_error()) : std::string();
    lds_thread.join();
    (void)hipStreamDestroy(stream);
    if (rc2 != DQ_OK) return bfail(rc2, msg2);
    if (rc != DQ_OK) return bfail(rc, msg);
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
