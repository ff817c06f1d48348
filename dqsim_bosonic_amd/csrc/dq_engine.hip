// dq_engine.hip -- gfx950 statevector engine behind the C ABI of
// include/dqsim_gpu.h.  Replaces the Rayon loops of the reference's
// src/engine.rs (apply_one_qubit :24-101, apply_n_qubit :109-181,
// measure_qubit :193-257, marginal_probs :368-390) and the fidelity fold of
// src/monolithic/statevector/simulator.rs:184-191.
//
// Design:
//  * Gate calls are classified on the host (diagonal / monomial /
//    controlled-2x2 / dense) and queued on the handle.
//  * At a flush the queue is cut into passes.  A pass picks a set S of
//    `tile_bits` qubits (the low `low_bits` qubits always included, so each
//    tile is made of contiguous 2^low_bits-amplitude runs and loads coalesce)
//    and takes, in program order, every queued op whose qubits fit in S and do
//    not touch a qubit held back by an earlier op left for a later pass
//    (fusion level 2, the default).  Ops on disjoint qubits commute, so the
//    result equals program order in exact arithmetic; the rounding of the
//    reordered products may differ in the last bit.  Fusion level 1 keeps
//    strict program order (each pass is a prefix of the queue) and so keeps
//    the reference's rounding bit for bit; level 0 issues one sweep per op.
//  * One pass = one kernel = one HBM sweep: each workgroup loads a 2^tile_bits
//    amplitude tile into LDS, applies all of the pass's ops in order with the
//    reference's arithmetic, and writes the tile back.
//  * Reductions (p1, marginals, inner products) are two-level: per-block tree
//    sums (wavefront shuffles + LDS) into a fixed number of partials, then one
//    block sums the partials in a fixed order; results are deterministic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <mutex>
#include <vector>

#include "../../include/dqsim_gpu.h"
#include "dq_device.hpp"
#include "dq_tile.hpp"
#include "dq_tile_launch.hpp"
#include "dq_plan.hpp"
#include "dq_state.hpp"
#include <type_traits>

using namespace dq;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

}  // namespace

// The one error channel of the library (dq_last_error), also used by the
// batch entry points in dq_batch.hip.
int dq::set_error(int code, const std::string &msg) { return fail(code, msg); }

namespace {

#define HIP_TRY(expr)                                                                  \
    do {                                                                               \
        hipError_t _e = (expr);                                                        \
        if (_e != hipSuccess)                                                          \
            return fail(DQ_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

constexpr int kThreads = 256;       // workgroup size of every kernel
constexpr int kReduceBlocks = 1024;  // partial sums per reduction (fixed => deterministic)

// ------------------------------------------------------------------------
// kernels
// ------------------------------------------------------------------------

template <typename T>
__global__ __launch_bounds__(kThreads) void k_fill_basis(cplx<T> *amps, uint64_t count) {
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += stride)
        amps[i] = cplx<T>{T(i == 0 ? 1 : 0), T(0)};
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_from_f64(cplx<T> *dst, const cplx<double> *src,
                                                       uint64_t count) {
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += stride)
        dst[i] = cplx<T>{T(src[i].re), T(src[i].im)};
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_to_f64(cplx<double> *dst, const cplx<T> *src,
                                                     uint64_t count) {
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += stride)
        dst[i] = cplx<double>{double(src[i].re), double(src[i].im)};
}

// One op over the whole state in global memory (unfused mode).
template <typename T>
__global__ __launch_bounds__(kThreads) void k_direct_op(cplx<T> *amps, int n, const DevOp *ops,
                                                        int op_index, const cplx<T> *mats,
                                                        const int *perms) {
    const DevOp op = ops[op_index];
    const uint64_t tid = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const uint64_t nthr = uint64_t(gridDim.x) * blockDim.x;
    apply_op<T, uint64_t>(amps, n, op, mats, perms, tid, nthr);
}

// Tile pass descriptor, passed by value (kernarg).
struct PassDesc {
    int b;         // tile bits
    int n_outer;   // n - b
    int op_begin;  // first op of the pass
    int op_count;
    int tq[16];    // tile bit -> qubit (tq[j] = j for j < low_bits)
    int oq[48];    // outer bit -> qubit
};

// One HBM sweep: load tile -> apply ops in LDS -> store tile.
// Tile-local index i = s * 256 + tid; low 8 bits from the thread, the rest
// from the slot s, so each thread's global offsets split into a per-thread
// part and a per-slot part computed once per launch.
template <typename T, int NSLOT>
__global__ __launch_bounds__(kThreads) void k_tile_pass(cplx<T> *__restrict__ amps, PassDesc pd,
                                                        const DevOp *__restrict__ ops,
                                                        const cplx<T> *__restrict__ mats,
                                                        const int *__restrict__ perms,
                                                        uint64_t ntiles) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cplx<T> *sh = reinterpret_cast<cplx<T> *>(smem_raw);
    const int tid = threadIdx.x;
    const int b = pd.b;
    const int lowb = b < 8 ? b : 8;
    const bool active = tid < (1 << lowb);
    uint64_t off_lo = 0;
    for (int j = 0; j < lowb; ++j)
        if ((tid >> j) & 1) off_lo |= uint64_t(1) << pd.tq[j];
    uint64_t off_hi[NSLOT];
#pragma unroll
    for (int s = 0; s < NSLOT; ++s) {
        uint64_t o = 0;
        for (int j = 8; j < b; ++j)
            if ((s >> (j - 8)) & 1) o |= uint64_t(1) << pd.tq[j];
        off_hi[s] = o;
    }
    for (uint64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        uint64_t base = 0;
        for (int j = 0; j < pd.n_outer; ++j)
            if ((t >> j) & 1) base |= uint64_t(1) << pd.oq[j];
        if (active) {
#pragma unroll
            for (int s = 0; s < NSLOT; ++s) sh[s * kThreads + tid] = amps[base + off_lo + off_hi[s]];
        }
        __syncthreads();
        for (int o = 0; o < pd.op_count; ++o) {
            const DevOp op = ops[pd.op_begin + o];
            apply_op<T, uint32_t>(sh, b, op, mats, perms, uint32_t(tid), uint32_t(kThreads));
            __syncthreads();
        }
        if (active) {
#pragma unroll
            for (int s = 0; s < NSLOT; ++s) amps[base + off_lo + off_hi[s]] = sh[s * kThreads + tid];
        }
        // The next tile's loads write only this thread's own slots, which it
        // has just read, so no barrier is needed here.
    }
}

// ---- reductions ----

__device__ __forceinline__ double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Fixed-shape block reduction: wave butterflies, then wave 0 sums the four
// wave totals in order.  Result valid in thread 0.
__device__ __forceinline__ double block_sum(double v, double *scratch) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) scratch[w] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0)
        for (int j = 0; j < kThreads / 64; ++j) r += scratch[j];
    __syncthreads();
    return r;
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_prob_one(const cplx<T> *__restrict__ amps, int n, int q,
                                                       double *__restrict__ partials) {
    __shared__ double scratch[kThreads / 64];
    const uint64_t half = uint64_t(1) << q;
    const uint64_t count = uint64_t(1) << (n - 1);
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    double s = 0.0;
    uint64_t g = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    // four loads in flight per thread, summed in the plain loop's order
    for (; g + 3 * stride < count; g += 4 * stride) {
        cplx<T> a[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const uint64_t gu = g + uint64_t(u) * stride;
            a[u] = amps[((gu >> q) << (q + 1)) | half | (gu & (half - 1))];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const double re = double(a[u].re), im = double(a[u].im);
            s += re * re + im * im;
        }
    }
    for (; g < count; g += stride) {
        const uint64_t i = ((g >> q) << (q + 1)) | half | (g & (half - 1));
        const cplx<T> a = amps[i];
        const double re = double(a.re), im = double(a.im);
        s += re * re + im * im;
    }
    const double r = block_sum(s, scratch);
    if (threadIdx.x == 0) partials[blockIdx.x] = r;
}

// p1 of the state that the queued elementwise ops (diagonals, collapses)
// would produce, without writing it: the same grid-stride order and block
// sums as k_prob_one, each element computed with the ops' own arithmetic, so
// the result equals k_prob_one on the updated state.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_prob_one_virtual(const cplx<T> *__restrict__ amps, int n, int q,
                                                               const DevOp *__restrict__ ops, int nops,
                                                               const cplx<T> *__restrict__ mats,
                                                               double *__restrict__ partials) {
    __shared__ double scratch[kThreads / 64];
    const uint64_t half = uint64_t(1) << q;
    const uint64_t count = uint64_t(1) << (n - 1);
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    auto apply = [&](uint64_t i, cplx<T> a) {
        for (int o = 0; o < nops; ++o) {
            const DevOp &op = ops[o];
            if (op.kind == K_DIAG) {
                a = cmul(mats[op.mat + gather_bits<uint64_t>(i, op.bit, op.k)], a);
            } else {  // K_COLLAPSE
                if (int((i >> op.bit[0]) & 1) == op.outcome) {
                    a.re = T(double(a.re) / op.norm);
                    a.im = T(double(a.im) / op.norm);
                } else {
                    a = cplx<T>{T(0), T(0)};
                }
            }
        }
        return a;
    };
    double s = 0.0;
    uint64_t g = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    // as in k_prob_one: four loads in flight, summed in the plain loop's order
    for (; g + 3 * stride < count; g += 4 * stride) {
        uint64_t iu[4];
        cplx<T> a[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const uint64_t gu = g + uint64_t(u) * stride;
            iu[u] = ((gu >> q) << (q + 1)) | half | (gu & (half - 1));
            a[u] = amps[iu[u]];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const cplx<T> b = apply(iu[u], a[u]);
            const double re = double(b.re), im = double(b.im);
            s += re * re + im * im;
        }
    }
    for (; g < count; g += stride) {
        const uint64_t i = ((g >> q) << (q + 1)) | half | (g & (half - 1));
        const cplx<T> a = apply(i, amps[i]);
        const double re = double(a.re), im = double(a.im);
        s += re * re + im * im;
    }
    const double r = block_sum(s, scratch);
    if (threadIdx.x == 0) partials[blockIdx.x] = r;
}

// Sums `count` partials (one block, fixed order) -> out[slot].
__global__ __launch_bounds__(kThreads) void k_sum_partials(const double *__restrict__ partials, int count,
                                                           int stride_elems, double *__restrict__ out) {
    __shared__ double scratch[kThreads / 64];
    for (int c = 0; c < stride_elems; ++c) {
        double s = 0.0;
        for (int i = threadIdx.x; i < count; i += blockDim.x) s += partials[size_t(i) * stride_elems + c];
        const double r = block_sum(s, scratch);
        if (threadIdx.x == 0) out[c] = r;
    }
}

// out[bin] = sum over chunks of partials[chunk * bins + bin], ascending chunks.
__global__ __launch_bounds__(kThreads) void k_sum_chunks(const double *__restrict__ partials, int chunks,
                                                         uint64_t bins, double *__restrict__ out) {
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t j = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; j < bins; j += stride) {
        double s = 0.0;
        for (int c = 0; c < chunks; ++c) s += partials[uint64_t(c) * bins + j];
        out[j] = s;
    }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_inner(const cplx<T> *__restrict__ amps,
                                                    const cplx<double> *__restrict__ other, uint64_t count,
                                                    double *__restrict__ partials) {
    __shared__ double scratch[kThreads / 64];
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    double sr = 0.0, si = 0.0;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += stride) {
        const double ar = double(amps[i].re), ai = -double(amps[i].im);
        const cplx<double> b = other[i];
        sr += ar * b.re - ai * b.im;
        si += ar * b.im + ai * b.re;
    }
    const double r0 = block_sum(sr, scratch);
    const double r1 = block_sum(si, scratch);
    if (threadIdx.x == 0) {
        partials[2 * size_t(blockIdx.x)] = r0;
        partials[2 * size_t(blockIdx.x) + 1] = r1;
    }
}

// ---- Pauli sums (observables and adjoint gradients) ----
//
// A Pauli string on n qubits is (x, z): X on the bits of x, Z on the bits of
// z, Y where both are set, with Y = i X Z.  P|y> = i^{#Y} (-1)^{popcount(z&y)}
// |y ^ x>, so (P psi)[r] = i^{#Y} (-1)^{popcount(z & (r ^ x))} psi[r ^ x].

struct PauliTerm {
    uint64_t x, z;
    double cre, cim;  // coefficient times i^{#Y}
};

// dst[r] = sum_k c_k (P_k src)[r]; terms in the order given.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_apply_pauli_sum(cplx<T> *__restrict__ dst, const cplx<T> *__restrict__ src,
                                                              uint64_t count, const PauliTerm *__restrict__ terms,
                                                              int nterms) {
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t r = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; r < count; r += stride) {
        double ar = 0.0, ai = 0.0;
        uint64_t y = r;
        cplx<T> v{};
        for (int k = 0; k < nterms; ++k) {
            const PauliTerm t = terms[k];
            // consecutive terms with the same X part share the partner
            // amplitude (a Z-string run: one load for the whole run)
            if (k == 0 || t.x != terms[k - 1].x) {
                y = r ^ t.x;
                v = src[y];
            }
            const double sgn = (__popcll(t.z & y) & 1) ? -1.0 : 1.0;
            const double vr = sgn * double(v.re), vi = sgn * double(v.im);
            ar += t.cre * vr - t.cim * vi;
            ai += t.cre * vi + t.cim * vr;
        }
        dst[r] = cplx<T>{T(ar), T(ai)};
    }
}

// partials[block] = per-block sum over r of conj(bra[r]) * sum_k c_k (P_k ket)[r].
template <typename T>
__global__ __launch_bounds__(kThreads) void k_pauli_inner(const cplx<T> *__restrict__ bra, const cplx<T> *__restrict__ ket,
                                                          uint64_t count, const PauliTerm *__restrict__ terms, int nterms,
                                                          double *__restrict__ partials) {
    __shared__ double scratch[kThreads / 64];
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    double sr = 0.0, si = 0.0;
    for (uint64_t r = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; r < count; r += stride) {
        const double br = double(bra[r].re), bi = -double(bra[r].im);
        double ar = 0.0, ai = 0.0;
        uint64_t y = r;
        cplx<T> v{};
        for (int k = 0; k < nterms; ++k) {
            const PauliTerm t = terms[k];
            if (k == 0 || t.x != terms[k - 1].x) {  // as in k_apply_pauli_sum
                y = r ^ t.x;
                v = ket[y];
            }
            const double sgn = (__popcll(t.z & y) & 1) ? -1.0 : 1.0;
            const double vr = sgn * double(v.re), vi = sgn * double(v.im);
            ar += t.cre * vr - t.cim * vi;
            ai += t.cre * vi + t.cim * vr;
        }
        sr += br * ar - bi * ai;
        si += br * ai + bi * ar;
    }
    const double r0 = block_sum(sr, scratch);
    const double r1 = block_sum(si, scratch);
    if (threadIdx.x == 0) {
        partials[2 * size_t(blockIdx.x)] = r0;
        partials[2 * size_t(blockIdx.x) + 1] = r1;
    }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_copy_state(cplx<T> *__restrict__ dst, const cplx<T> *__restrict__ src,
                                                         uint64_t count) {
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += stride) dst[i] = src[i];
}

// Tensor product with interleaved qubits (merge_blocks of the reference's
// pblock model, src/distributed/pblock/model.rs:30-58):
// dst[i] = a[bits of i at apos] * b[bits of i at bpos], one complex product
// in num-complex order per amplitude.
struct TensorPos {
    int na, nb;
    int apos[40], bpos[40];
};

template <typename T>
__global__ __launch_bounds__(kThreads) void k_tensor(cplx<T> *__restrict__ dst, const cplx<T> *__restrict__ a,
                                                     const cplx<T> *__restrict__ b, TensorPos tp, uint64_t count) {
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += stride) {
        uint64_t ia = 0, ib = 0;
        for (int k = 0; k < tp.na; ++k) ia |= ((i >> tp.apos[k]) & 1) << k;
        for (int k = 0; k < tp.nb; ++k) ib |= ((i >> tp.bpos[k]) & 1) << k;
        dst[i] = cmul(a[ia], b[ib]);
    }
}

// Differences between two states of the same size (any dtypes): partials of
// max |a_i - b_i|, Re/Im of sum conj(a_i) b_i and sum | |a_i|^2 - |b_i|^2 |.
template <typename TA, typename TB>
__global__ __launch_bounds__(kThreads) void k_compare(const cplx<TA> *__restrict__ a, const cplx<TB> *__restrict__ b,
                                                      uint64_t count, double *__restrict__ partials) {
    __shared__ double scratch[kThreads / 64];
    __shared__ double mx[kThreads];
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    double m = 0.0, sr = 0.0, si = 0.0, tv = 0.0;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += stride) {
        const double ar = double(a[i].re), ai = double(a[i].im);
        const double br = double(b[i].re), bi = double(b[i].im);
        const double dr = ar - br, di = ai - bi;
        m = fmax(m, sqrt(dr * dr + di * di));
        sr += ar * br + ai * bi;
        si += ar * bi - ai * br;
        tv += fabs((ar * ar + ai * ai) - (br * br + bi * bi));
    }
    mx[threadIdx.x] = m;
    const double r0 = block_sum(sr, scratch);
    const double r1 = block_sum(si, scratch);
    const double r2 = block_sum(tv, scratch);
    __syncthreads();
    if (threadIdx.x == 0) {
        double mm = 0.0;
        for (int t = 0; t < kThreads; ++t) mm = fmax(mm, mx[t]);
        partials[4 * size_t(blockIdx.x)] = mm;
        partials[4 * size_t(blockIdx.x) + 1] = r0;
        partials[4 * size_t(blockIdx.x) + 2] = r1;
        partials[4 * size_t(blockIdx.x) + 3] = r2;
    }
}

// Full distribution over all n qubits in a given order: out[j(i)] = |psi_i|^2.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_probs_perm(const cplx<T> *__restrict__ amps, int n,
                                                         const int *__restrict__ qubits,
                                                         double *__restrict__ out, int identity) {
    const uint64_t count = uint64_t(1) << n;
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += stride) {
        const double re = double(amps[i].re), im = double(amps[i].im);
        uint64_t j = i;
        if (!identity) {
            j = 0;
            for (int b = 0; b < n; ++b) j = (j << 1) | ((i >> qubits[b]) & 1);
        }
        out[j] = re * re + im * im;
    }
}

// Marginal with few terms per bin: one thread per bin sums its terms in
// ascending index order (the reference's order, src/engine.rs:384-387).
template <typename T>
__global__ __launch_bounds__(kThreads) void k_marginal_seq(const cplx<T> *__restrict__ amps, int n,
                                                           const int *__restrict__ qubits, int k,
                                                           const int *__restrict__ rest, double *__restrict__ out) {
    const uint64_t bins = uint64_t(1) << k;
    const uint64_t terms = uint64_t(1) << (n - k);
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t j = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; j < bins; j += stride) {
        uint64_t base = 0;
        for (int b = 0; b < k; ++b)
            if ((j >> (k - 1 - b)) & 1) base |= uint64_t(1) << qubits[b];
        double s = 0.0;
        for (uint64_t r = 0; r < terms; ++r) {
            uint64_t i = base;
            for (int b = 0; b < n - k; ++b)
                if ((r >> b) & 1) i |= uint64_t(1) << rest[b];
            const double re = double(amps[i].re), im = double(amps[i].im);
            s += re * re + im * im;
        }
        out[j] = s;
    }
}

// Marginal with many terms per bin: block (bin, chunk) tree-sums a chunk of
// the bin's terms into partials[chunk * bins + bin].
template <typename T>
__global__ __launch_bounds__(kThreads) void k_marginal_tree(const cplx<T> *__restrict__ amps, int n,
                                                            const int *__restrict__ qubits, int k,
                                                            const int *__restrict__ rest, int chunks,
                                                            double *__restrict__ partials) {
    __shared__ double scratch[kThreads / 64];
    const uint64_t bins = uint64_t(1) << k;
    const uint64_t terms = uint64_t(1) << (n - k);
    const uint64_t bin = blockIdx.x % bins;
    const uint64_t chunk = blockIdx.x / bins;
    uint64_t base = 0;
    for (int b = 0; b < k; ++b)
        if ((bin >> (k - 1 - b)) & 1) base |= uint64_t(1) << qubits[b];
    const uint64_t per = terms / chunks;
    double s = 0.0;
    for (uint64_t r = chunk * per + threadIdx.x; r < (chunk + 1) * per; r += blockDim.x) {
        uint64_t i = base;
        for (int b = 0; b < n - k; ++b)
            if ((r >> b) & 1) i |= uint64_t(1) << rest[b];
        const double re = double(amps[i].re), im = double(amps[i].im);
        s += re * re + im * im;
    }
    const double v = block_sum(s, scratch);
    if (threadIdx.x == 0) partials[chunk * bins + bin] = v;
}

// ------------------------------------------------------------------------
// host side
// ------------------------------------------------------------------------

}  // namespace


namespace {

uint64_t amp_count(const dq_state *s) { return uint64_t(1) << s->n; }

PlanConfig plan_config(const dq_state *s) {
    PlanConfig c;
    c.n = s->n;
    c.fuse = s->fuse;
    c.tile_bits = s->tile_bits;
    c.low_bits = s->low_bits;
    c.reg_tiles = s->reg_tiles;
    c.fuse_min_cost = s->fuse_min_cost;
    c.pair_fusion = s->pair_fusion;
    // 8 amplitudes per thread for the contracted complex128 sweep; the exact
    // (fusion level 1) and complex64 sweeps keep 16
    c.reg_bits = (s->dtype == DQ_C128 && s->fuse >= 2) ? s->reg_bits : 4;
    return c;
}

int grid_for(uint64_t work) {
    uint64_t g = (work + kThreads - 1) / kThreads;
    if (g > 65536) g = 65536;
    if (g < 1) g = 1;
    return int(g);
}

int check_qubit(const dq_state *s, int q) {
    if (q < 0 || q >= s->n) return fail(DQ_ERR_ARG, "qubit index " + std::to_string(q) +
                                                        " out of range for " + std::to_string(s->n) +
                                                        "-qubit state");
    return DQ_OK;
}

// ---- upload + launch ----

struct Packed {
    std::vector<DevOp> ops;
    std::vector<double> mats;  // complex entries as (re, im)
    std::vector<int> perms;
};

// Converts a host op to a device op with positions through `pos` (qubit ->
// position in the index space of the launch).
DevOp to_dev(const HostOp &h, const int *pos, Packed &pk) {
    DevOp d;
    std::memset(&d, 0, sizeof d);
    d.kind = h.kind;
    d.k = h.k;
    d.mat = int(pk.mats.size() / 2);
    pk.mats.insert(pk.mats.end(), h.mat.begin(), h.mat.end());
    d.aux = int(pk.perms.size());
    pk.perms.insert(pk.perms.end(), h.perm.begin(), h.perm.end());
    for (int j = 0; j < h.k; ++j) d.bit[j] = pos[h.q[j]];
    for (int j = 0; j < h.k; ++j) d.sbit[j] = d.bit[j];
    std::sort(d.sbit, d.sbit + h.k);
    d.outcome = h.outcome;
    d.norm = h.norm;
    d.cmask = d.cval = 0;
    for (int j = 0; j < h.k - 1 && h.kind == K_CTRL1; ++j) {
        const uint64_t gbit = uint64_t(1) << h.q[j];
        d.cmask |= uint64_t(1) << pos[h.q[j]];
        if (h.cval & gbit) d.cval |= uint64_t(1) << pos[h.q[j]];
    }
    return d;
}

int ensure_device_buffer(dq_state *s, size_t bytes) {
    if (bytes <= s->d_upload_cap) return DQ_OK;
    HIP_TRY(hipStreamSynchronize(s->stream));
    if (s->d_upload) HIP_TRY(hipFree(s->d_upload));
    s->d_upload = nullptr;
    size_t cap = std::max<size_t>(bytes, 1 << 20);
    if (hipMalloc(&s->d_upload, cap) != hipSuccess) {
        s->d_upload_cap = 0;
        return fail(DQ_ERR_OOM, "hipMalloc of op buffer failed");
    }
    s->d_upload_cap = cap;
    return DQ_OK;
}

int ensure_host_buffer(dq_state *s, size_t bytes) {
    if (bytes <= s->h_upload_cap) return DQ_OK;
    HIP_TRY(hipStreamSynchronize(s->stream));
    if (s->h_upload) HIP_TRY(hipHostFree(s->h_upload));
    s->h_upload = nullptr;
    size_t cap = std::max<size_t>(bytes, 1 << 20);
    if (hipHostMalloc(reinterpret_cast<void **>(&s->h_upload), cap, hipHostMallocDefault) != hipSuccess) {
        s->h_upload_cap = 0;
        return fail(DQ_ERR_OOM, "hipHostMalloc of staging buffer failed");
    }
    s->h_upload_cap = cap;
    return DQ_OK;
}

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Copies ops/mats/perms to the device; returns device pointers.
int upload(dq_state *s, const Packed &pk, DevOp **d_ops, void **d_mats, int **d_perms) {
    const size_t ops_b = align_up(pk.ops.size() * sizeof(DevOp), 256);
    const size_t real_b = (s->dtype == DQ_C128) ? 8 : 4;
    const size_t mats_b = align_up(pk.mats.size() * real_b, 256);
    const size_t perms_b = align_up(pk.perms.size() * sizeof(int), 256);
    const size_t total = ops_b + mats_b + perms_b + 256;
    // the previous upload must have been consumed before the staging area is reused
    HIP_TRY(hipEventSynchronize(s->upload_done));
    int rc = ensure_host_buffer(s, total);
    if (rc) return rc;
    rc = ensure_device_buffer(s, total);
    if (rc) return rc;
    unsigned char *h = s->h_upload;
    std::memcpy(h, pk.ops.data(), pk.ops.size() * sizeof(DevOp));
    if (s->dtype == DQ_C128) {
        std::memcpy(h + ops_b, pk.mats.data(), pk.mats.size() * 8);
    } else {
        float *f = reinterpret_cast<float *>(h + ops_b);
        for (size_t i = 0; i < pk.mats.size(); ++i) f[i] = float(pk.mats[i]);
    }
    std::memcpy(h + ops_b + mats_b, pk.perms.data(), pk.perms.size() * sizeof(int));
    HIP_TRY(hipMemcpyAsync(s->d_upload, h, total, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipEventRecord(s->upload_done, s->stream));
    unsigned char *d = static_cast<unsigned char *>(s->d_upload);
    *d_ops = reinterpret_cast<DevOp *>(d);
    *d_mats = d + ops_b;
    *d_perms = reinterpret_cast<int *>(d + ops_b + mats_b);
    return DQ_OK;
}

hipEvent_t take_event(dq_state *s) {
    if (!s->ev_pool.empty()) {
        hipEvent_t e = s->ev_pool.back();
        s->ev_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Brackets one sweep launch with events when timing is on.
struct SweepTimer {
    dq_state *s;
    hipEvent_t a = nullptr, b = nullptr;
    explicit SweepTimer(dq_state *st) : s(st) {
        if (!s->timing) return;
        a = take_event(s);
        b = take_event(s);
        if (a) (void)hipEventRecord(a, s->stream);
    }
    ~SweepTimer() {
        if (!a || !b) return;
        (void)hipEventRecord(b, s->stream);
        s->ev_pending.emplace_back(a, b);
    }
};

int collect_timing(dq_state *s) {
    if (s->ev_pending.empty()) return DQ_OK;
    HIP_TRY(hipStreamSynchronize(s->stream));
    for (auto &pr : s->ev_pending) {
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, pr.first, pr.second));
        s->timed_ms += ms;
        s->timed_launches++;
        s->ev_pool.push_back(pr.first);
        s->ev_pool.push_back(pr.second);
    }
    s->ev_pending.clear();
    return DQ_OK;
}

template <typename T>
int launch_tile(dq_state *s, const PassDesc &pd, const DevOp *ops, const void *mats, const int *perms) {
    SweepTimer timer(s);
    const uint64_t ntiles = uint64_t(1) << pd.n_outer;
    const int nslot = pd.b > 8 ? 1 << (pd.b - 8) : 1;
    const size_t shmem = sizeof(cplx<T>) * size_t(nslot) * kThreads;
    // persistent-style grid: enough workgroups for every CU several times over
    const uint64_t grid = std::min<uint64_t>(ntiles, 8192);
    const cplx<T> *m = static_cast<const cplx<T> *>(mats);
    cplx<T> *a = static_cast<cplx<T> *>(s->amps);
#define DQ_TILE_CASE(NS)                                                                       \
    case NS:                                                                                   \
        if (shmem > 65536)                                                                     \
            HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_tile_pass<T, NS>),   \
                                        hipFuncAttributeMaxDynamicSharedMemorySize, int(shmem))); \
        hipLaunchKernelGGL((k_tile_pass<T, NS>), dim3(grid), dim3(kThreads), shmem, s->stream, a, pd, \
                           ops, m, perms, ntiles);                                             \
        break;
    switch (nslot) {
        DQ_TILE_CASE(1)
        DQ_TILE_CASE(2)
        DQ_TILE_CASE(4)
        DQ_TILE_CASE(8)
        DQ_TILE_CASE(16)
        DQ_TILE_CASE(32)
    default:
        return fail(DQ_ERR_ARG, "unsupported tile size");
    }
#undef DQ_TILE_CASE
    HIP_TRY(hipGetLastError());
    s->swept_amps += amp_count(s);
    s->launches++;
    s->sweeps++;
    return DQ_OK;
}

template <typename T>
int launch_direct(dq_state *s, const DevOp *ops, int index, const void *mats, const int *perms) {
    SweepTimer timer(s);
    const uint64_t work = amp_count(s) / 2;
    hipLaunchKernelGGL(k_direct_op<T>, dim3(grid_for(work)), dim3(kThreads), 0, s->stream,
                       static_cast<cplx<T> *>(s->amps), s->n, ops, index,
                       static_cast<const cplx<T> *>(mats), perms);
    HIP_TRY(hipGetLastError());
    s->swept_amps += amp_count(s);
    s->launches++;
    s->sweeps++;
    return DQ_OK;
}

// Greedy pass builder (see the file header).
template <typename T>
int launch_reg(dq_state *s, const RegPlan &rp, const TileOp *ops, const void *mats) {
    SweepTimer timer(s);
    const uint64_t ntiles = uint64_t(1) << rp.pd.n_outer;
    // a power of two: the kernel derives each tile's offset from blockIdx.x
    // and the tile's step count.  At least 8 tiles per workgroup where that
    // still leaves 512 workgroups (two per CU), so that each workgroup's
    // next-tile prefetch overlaps its work also on small states (measured on
    // 21-24 qubit states: 6% more throughput than one tile per workgroup);
    // at most 16384 workgroups.
    const uint64_t grid = std::min<uint64_t>(std::min<uint64_t>(ntiles, 16384), std::max<uint64_t>(512, ntiles / 8));
    if (grid & (grid - 1)) return fail(DQ_ERR_ARG, "tile grid is not a power of two");
    // fusion level 1 keeps the reference's rounding: the exact kernel;
    // level 2 already reorders commuting ops and uses the contracted kernel.
    auto *fn = std::is_same<T, double>::value
                   ? (s->fuse >= 2 ? (rp.reg_bits == 3 ? dq_launch_tile_reg8_f64_fma : dq_launch_tile_reg_f64_fma)
                                   : dq_launch_tile_reg_f64)
                   : (rp.fused_blocks > 0 ? dq_launch_tile_reg_f32_mfma : dq_launch_tile_reg_f32);
    hipError_t e = fn(rp.b, unsigned(grid), s->stream, s->amps, rp.pd, ops, mats, ntiles, s->d_partials);
    s->last_grid = int(grid);
    if (e == hipErrorInvalidValue) return fail(DQ_ERR_ARG, "unsupported tile size " + std::to_string(rp.b));
    HIP_TRY(e);
    s->swept_amps += ntiles << rp.b;
    s->launches++;
    s->sweeps++;
    return DQ_OK;
}

// red_qubit >= 0: the last pass also sums |amp|^2 over that qubit being 1
// into per-block partials (s->last_grid of them).
template <typename T>
int flush_reg(dq_state *s, const std::vector<Pass> &passes, int red_qubit = -1) {
    const PlanConfig cfg = plan_config(s);
    std::vector<RegPlan> plans(passes.size());
    std::vector<double> mats;
    size_t total_ops = 0;
    // qubits still 0 everywhere before each pass; a pass's sweep skips the
    // tiles where any outer one of them is 1 (they hold zeros and stay zero)
    KnownBits kb;
    kb.zero = s->zero_tiles ? s->zero_q : 0;
    for (size_t i = 0; i < passes.size(); ++i) {
        plan_pass(&cfg, s->queue, passes[i], mats, plans[i], kb.zero);
        if (s->zero_tiles) advance_known(s->queue, passes[i], kb);
        s->fused_blocks += uint64_t(plans[i].fused_blocks);
        s->fused_gates += uint64_t(plans[i].fused_gates);
        plans[i].pd.op_begin = int(total_ops);
        plans[i].pd.op_count = int(plans[i].ops.size());
        plans[i].pd.red_bit = -1;
        if (red_qubit >= 0 && i + 1 == passes.size()) {
            // an outer bit past n_outer reads 0 in every visited tile: p1 = +0.0
            for (int ob = 0; ob < s->n - plans[i].b; ++ob)
                if (plans[i].pd.oq[ob] == red_qubit) plans[i].pd.red_bit = kOuterBit + ob;
            for (int tb = 0; tb < plans[i].b; ++tb)
                if (plans[i].pd.tq[tb] == red_qubit) plans[i].pd.red_bit = tb;
        }
        total_ops += plans[i].ops.size();
    }
    const size_t real_b = sizeof(T);
    const size_t ops_b = align_up(std::max<size_t>(total_ops, 1) * sizeof(TileOp), 256);
    const size_t mats_b = align_up(std::max<size_t>(mats.size(), 2) * real_b, 256);
    // layout tables follow the matrix pool; each pass finds its own at pd.lay_off
    std::vector<uint32_t> lays;
    for (RegPlan &rp : plans) {
        rp.pd.lay_off = int(mats_b + lays.size() * sizeof(uint32_t));
        for (int rb : rp.layouts) append_layout_table(lays, rb, rp.b, unsigned(2 * sizeof(T)), rp.reg_bits);
    }
    const size_t lays_b = align_up(std::max<size_t>(lays.size(), 1) * sizeof(uint32_t), 256);
    const size_t total = ops_b + mats_b + lays_b;
    HIP_TRY(hipEventSynchronize(s->upload_done));
    if (int rc = ensure_host_buffer(s, total)) return rc;
    if (int rc = ensure_device_buffer(s, total)) return rc;
    unsigned char *h = s->h_upload;
    size_t off = 0;
    for (const RegPlan &rp : plans) {
        std::memcpy(h + off, rp.ops.data(), rp.ops.size() * sizeof(TileOp));
        off += rp.ops.size() * sizeof(TileOp);
    }
    T *hm = reinterpret_cast<T *>(h + ops_b);
    for (size_t i = 0; i < mats.size(); ++i) hm[i] = T(mats[i]);
    std::memcpy(h + ops_b + mats_b, lays.data(), lays.size() * sizeof(uint32_t));
    HIP_TRY(hipMemcpyAsync(s->d_upload, h, total, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipEventRecord(s->upload_done, s->stream));
    const TileOp *d_ops = reinterpret_cast<const TileOp *>(s->d_upload);
    const void *d_mats = static_cast<unsigned char *>(s->d_upload) + ops_b;
    s->zero_q = 0;  // unknown should a launch fail part way
    for (const RegPlan &rp : plans)
        if (int rc = launch_reg<T>(s, rp, d_ops, d_mats)) return rc;
    s->zero_q = kb.zero;
    s->queue.clear();
    return DQ_OK;
}

template <typename T>
int flush_t(dq_state *s, int red_qubit = -1) {
    if (s->queue.empty()) return DQ_OK;
    const int n = s->n;
    Packed pk;
    std::vector<PassDesc> descs;
    int rc = DQ_OK;
    const PlanConfig cfg = plan_config(s);
    if (s->fuse && s->reg_tiles && effective_tile_bits(&cfg) >= 9) {
        return flush_reg<T>(s, schedule(&cfg, s->queue, true), red_qubit);
    }
    for (const HostOp &h : s->queue) s->zero_q &= ~op_support(h);
    if (s->fuse && n >= 1) {
        std::vector<Pass> passes = schedule(&cfg, s->queue);
        for (const Pass &p : passes) {
            PassDesc pd;
            std::memset(&pd, 0, sizeof pd);
            int pos[64];
            int tb = 0, ob = 0;
            for (int q = 0; q < n; ++q) {
                if ((p.S >> q) & 1) {
                    pd.tq[tb] = q;
                    pos[q] = tb++;
                } else {
                    pd.oq[ob++] = q;
                    pos[q] = -1;
                }
            }
            pd.b = tb;
            pd.n_outer = ob;
            pd.op_begin = int(pk.ops.size());
            pd.op_count = int(p.ops.size());
            for (int idx : p.ops) pk.ops.push_back(to_dev(s->queue[idx], pos, pk));
            descs.push_back(pd);
        }
    } else {
        int pos[64];
        for (int q = 0; q < 64; ++q) pos[q] = q;
        for (const HostOp &h : s->queue) pk.ops.push_back(to_dev(h, pos, pk));
    }
    DevOp *d_ops;
    void *d_mats;
    int *d_perms;
    rc = upload(s, pk, &d_ops, &d_mats, &d_perms);
    if (rc) return rc;
    if (s->fuse) {
        for (const PassDesc &pd : descs) {
            rc = launch_tile<T>(s, pd, d_ops, d_mats, d_perms);
            if (rc) return rc;
        }
    } else {
        for (size_t i = 0; i < pk.ops.size(); ++i) {
            rc = launch_direct<T>(s, d_ops, int(i), d_mats, d_perms);
            if (rc) return rc;
        }
    }
    s->queue.clear();
    return DQ_OK;
}

int flush(dq_state *s) {
    return s->dtype == DQ_C128 ? flush_t<double>(s) : flush_t<float>(s);
}

// True when the register-tile path handles the flush (it can fuse a reduction).
bool reg_path(const dq_state *s) {
    const PlanConfig cfg = plan_config(s);
    return s->fuse && s->reg_tiles && effective_tile_bits(&cfg) >= 9;
}

bool all_elementwise(const dq_state *s, size_t count) {
    for (size_t i = 0; i < count; ++i)
        if (s->queue[i].kind != K_DIAG && s->queue[i].kind != K_COLLAPSE) return false;
    return true;
}
bool all_elementwise(const dq_state *s) { return all_elementwise(s, s->queue.size()); }

// Whether a queued op is unitary: every entry of U^H U - I within 1e-13 (the
// C ABI takes any matrix, so this is checked, not assumed; a NaN or infinite
// entry fails).  Selector-fused ops and collapses are not.
bool queued_unitary(const HostOp &h) {
    if (h.kind == K_COLLAPSE || h.nsel > 0 || h.k < 1 || h.k > 5) return false;
    constexpr double tol = 1e-13;
    const int dim = 1 << h.k;
    if (h.kind == K_DIAG || h.kind == K_MONO) {
        if (h.mat.size() != size_t(2 * dim)) return false;
        for (int r = 0; r < dim; ++r)
            if (!(std::fabs(h.mat[2 * r] * h.mat[2 * r] + h.mat[2 * r + 1] * h.mat[2 * r + 1] - 1.0) <= tol)) return false;
        if (h.kind == K_DIAG) return true;
        if (h.perm.size() != size_t(dim)) return false;
        std::vector<char> seen(dim, 0);
        for (int c : h.perm) {
            if (c < 0 || c >= dim || seen[c]) return false;
            seen[c] = 1;
        }
        return true;
    }
    const std::vector<double> *u = nullptr;
    int d = dim;
    if (h.dense.size() == size_t(2) * dim * dim) {
        u = &h.dense;
    } else if ((h.kind == K_DENSE1 || h.kind == K_CTRL1) && h.mat.size() == 8) {
        u = &h.mat;  // the 2x2 (a controlled 2x2 is unitary when its block is)
        d = 2;
    } else if (h.mat.size() == size_t(2) * dim * dim) {
        u = &h.mat;
    }
    if (!u) return false;
    const double *m = u->data();
    for (int a = 0; a < d; ++a)
        for (int b = a; b < d; ++b) {
            double re = 0.0, im = 0.0;  // column a^H . column b
            for (int r = 0; r < d; ++r) {
                const double *x = m + 2 * (size_t(r) * d + a), *y = m + 2 * (size_t(r) * d + b);
                re += x[0] * y[0] + x[1] * y[1];
                im += x[0] * y[1] - x[1] * y[0];
            }
            if (!(std::fabs(re - (a == b ? 1.0 : 0.0)) <= tol && std::fabs(im) <= tol)) return false;  // NaN fails
        }
    return true;
}

// Queue length whose ops decide p1 of `qubit`: unitaries at the end of the
// queue that do not act on it leave its marginal unchanged (U acts as I on
// the qubit, so <P1> is the same before and after), so they are left out.
size_t marginal_prefix(const dq_state *s, int qubit) {
    size_t live = s->queue.size();
    while (live > 0) {
        const HostOp &h = s->queue[live - 1];
        if ((((h.qmask() | h.cmask) >> qubit) & 1) || !queued_unitary(h)) break;
        --live;
    }
    return live;
}

int enqueue(dq_state *s, HostOp &&op) {
    if (s->ev_pending.size() > 4096)
        if (int rc = collect_timing(s)) return rc;
    s->queue.push_back(std::move(op));
    s->gates++;
    if (s->queue.size() >= 4096) return flush(s);
    return DQ_OK;
}

int fetch_result(dq_state *s, int count) {
    HIP_TRY(hipMemcpyAsync(s->h_result, s->d_result, sizeof(double) * count, hipMemcpyDeviceToHost,
                           s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return DQ_OK;
}

}  // namespace

// ------------------------------------------------------------------------
// C ABI
// ------------------------------------------------------------------------

extern "C" {

const char *dq_last_error(void) { return g_err.c_str(); }

int dq_state_create(int n_qubits, int dtype, dq_state **out) {
    if (!out) return fail(DQ_ERR_ARG, "null output pointer");
    *out = nullptr;
    if (n_qubits < 0 || n_qubits > 40) return fail(DQ_ERR_ARG, "n_qubits must be in [0, 40]");
    if (dtype != DQ_C128 && dtype != DQ_C64) return fail(DQ_ERR_ARG, "unknown dtype");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(DQ_ERR_DEVICE, "no HIP device available");
    dq_state *s = new dq_state();
    s->n = n_qubits;
    s->dtype = dtype;
    s->elem = dtype == DQ_C128 ? 16 : 8;
    s->tile_bits = 12;
    HIP_TRY(hipGetDevice(&s->device));
    auto cleanup = [&](int code, const std::string &msg) {
        dq_state_destroy(s);
        return fail(code, msg);
    };
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess)
        return cleanup(DQ_ERR_DEVICE, "hipStreamCreate failed");
    if (hipEventCreateWithFlags(&s->upload_done, hipEventDisableTiming) != hipSuccess)
        return cleanup(DQ_ERR_DEVICE, "hipEventCreate failed");
    if (hipMalloc(&s->amps, s->elem * amp_count(s)) != hipSuccess) {
        s->amps = nullptr;
        return cleanup(DQ_ERR_OOM, "hipMalloc of " + std::to_string(s->elem * amp_count(s)) +
                                       " bytes for the statevector failed");
    }
    if (hipMalloc(reinterpret_cast<void **>(&s->d_partials), sizeof(double) * 2 * kReduceBlocks * 64) !=
            hipSuccess ||
        hipMalloc(reinterpret_cast<void **>(&s->d_result), sizeof(double) * 64) != hipSuccess)
        return cleanup(DQ_ERR_OOM, "hipMalloc of reduction buffers failed");
    if (hipHostMalloc(reinterpret_cast<void **>(&s->h_result), sizeof(double) * 64, hipHostMallocDefault) !=
        hipSuccess)
        return cleanup(DQ_ERR_OOM, "hipHostMalloc failed");
    // KFD allocations are write-combined staging; make sure the event exists
    // in a signalled state before the first upload waits on it.
    if (hipEventRecord(s->upload_done, s->stream) != hipSuccess)
        return cleanup(DQ_ERR_DEVICE, "hipEventRecord failed");
    int rc = dq_state_reset(s);
    if (rc) {
        std::string m = g_err;
        return cleanup(rc, m);
    }
    *out = s;
    return DQ_OK;
}

void dq_state_destroy(dq_state *s) {
    if (!s) return;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    for (auto &pr : s->ev_pending) {
        (void)hipEventDestroy(pr.first);
        (void)hipEventDestroy(pr.second);
    }
    for (hipEvent_t e : s->ev_pool) (void)hipEventDestroy(e);
    if (s->amps) (void)hipFree(s->amps);
    if (s->d_upload) (void)hipFree(s->d_upload);
    if (s->h_upload) (void)hipHostFree(s->h_upload);
    if (s->d_partials) (void)hipFree(s->d_partials);
    if (s->d_result) (void)hipFree(s->d_result);
    if (s->h_result) (void)hipHostFree(s->h_result);
    if (s->d_adj) (void)hipFree(s->d_adj);
    if (s->h_adj) (void)hipHostFree(s->h_adj);
    if (s->upload_done) (void)hipEventDestroy(s->upload_done);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

int dq_state_num_qubits(const dq_state *s) { return s ? s->n : -1; }

int dq_state_reset(dq_state *s) {
    if (!s) return fail(DQ_ERR_ARG, "null state");
    s->queue.clear();
    s->zero_q = 0;  // known again once the fill has completed
    const uint64_t count = amp_count(s);
    if (s->dtype == DQ_C128)
        hipLaunchKernelGGL(k_fill_basis<double>, dim3(grid_for(count)), dim3(kThreads), 0, s->stream,
                           static_cast<cplx<double> *>(s->amps), count);
    else
        hipLaunchKernelGGL(k_fill_basis<float>, dim3(grid_for(count)), dim3(kThreads), 0, s->stream,
                           static_cast<cplx<float> *>(s->amps), count);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s->stream));
    s->zero_q = s->n && s->zero_tiles ? ~uint64_t(0) >> (64 - s->n) : 0;
    return DQ_OK;
}

int dq_state_set(dq_state *s, const double *src) {
    if (!s || !src) return fail(DQ_ERR_ARG, "null argument");
    s->queue.clear();
    s->zero_q = 0;
    const uint64_t count = amp_count(s);
    HIP_TRY(hipStreamSynchronize(s->stream));
    if (s->dtype == DQ_C128) {
        HIP_TRY(hipMemcpy(s->amps, src, 16 * count, hipMemcpyHostToDevice));
        return DQ_OK;
    }
    void *tmp = nullptr;
    if (hipMalloc(&tmp, 16 * count) != hipSuccess) return fail(DQ_ERR_OOM, "hipMalloc failed");
    hipError_t e = hipMemcpy(tmp, src, 16 * count, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_from_f64<float>, dim3(grid_for(count)), dim3(kThreads), 0, s->stream,
                           static_cast<cplx<float> *>(s->amps), static_cast<const cplx<double> *>(tmp), count);
        e = hipStreamSynchronize(s->stream);
    }
    (void)hipFree(tmp);
    if (e != hipSuccess) return fail(DQ_ERR_DEVICE, hipGetErrorString(e));
    return DQ_OK;
}

int dq_apply_1q(dq_state *s, const double u[8], int target, const int *ctrl_wires,
                const unsigned char *ctrl_vals, int n_ctrl) {
    if (!s || !u) return fail(DQ_ERR_ARG, "null argument");
    if (int rc = check_qubit(s, target)) return rc;
    if (n_ctrl < 0 || n_ctrl > 4) return fail(DQ_ERR_ARG, "at most 4 controls are supported");
    HostOp op;
    if (n_ctrl == 0) {
        op.k = 1;
        op.q[0] = target;
        classify(op, u);
        return enqueue(s, std::move(op));
    }
    op.kind = K_CTRL1;
    op.k = n_ctrl + 1;
    uint64_t seen = uint64_t(1) << target;
    for (int c = 0; c < n_ctrl; ++c) {
        if (int rc = check_qubit(s, ctrl_wires[c])) return rc;
        const uint64_t m = uint64_t(1) << ctrl_wires[c];
        if (seen & m) return fail(DQ_ERR_ARG, "control and target qubits must be distinct");
        seen |= m;
        op.q[c] = ctrl_wires[c];
        op.cmask |= m;
        if (ctrl_vals && ctrl_vals[c]) op.cval |= m;
    }
    op.q[n_ctrl] = target;
    op.mat.assign(u, u + 8);
    op.dense.assign(u, u + 8);
    return enqueue(s, std::move(op));
}

int dq_apply_nq(dq_state *s, const double *u, const int *qubits, int k) {
    if (!s || !u || !qubits) return fail(DQ_ERR_ARG, "null argument");
    if (k < 1 || k > 5) return fail(DQ_ERR_ARG, "gate arity must be 1..5 (got " + std::to_string(k) + ")");
    if (k > s->n) return fail(DQ_ERR_ARG, "gate arity exceeds the number of qubits");
    HostOp op;
    op.k = k;
    uint64_t seen = 0;
    for (int j = 0; j < k; ++j) {
        if (int rc = check_qubit(s, qubits[j])) return rc;
        const uint64_t m = uint64_t(1) << qubits[j];
        if (seen & m) return fail(DQ_ERR_ARG, "gate qubits must be distinct");
        seen |= m;
        op.q[j] = qubits[j];
    }
    classify(op, u);
    return enqueue(s, std::move(op));
}

int dq_collapse(dq_state *s, int qubit, int outcome, double norm) {
    if (!s) return fail(DQ_ERR_ARG, "null state");
    if (int rc = check_qubit(s, qubit)) return rc;
    if (outcome != 0 && outcome != 1) return fail(DQ_ERR_ARG, "outcome must be 0 or 1");
    HostOp op;
    op.kind = K_COLLAPSE;
    op.k = 1;
    op.q[0] = qubit;
    op.outcome = outcome;
    op.norm = norm;
    s->queue.push_back(std::move(op));
    if (s->queue.size() >= 4096) return flush(s);
    return DQ_OK;
}

int dq_prob_one(dq_state *s, int qubit, double *p1) {
    if (!s || !p1) return fail(DQ_ERR_ARG, "null argument");
    if (int rc = check_qubit(s, qubit)) return rc;
    // a qubit 0 in every amplitude in memory, on which no queued op acts
    // other than diagonally, is still 0 everywhere: the sum over its 1-half
    // is a sum of exact zeros, +0.0, with nothing flushed or read
    if (s->zero_tiles && ((s->zero_q >> qubit) & 1)) {
        bool touched = false;
        for (const HostOp &h : s->queue) touched = touched || ((op_support(h) >> qubit) & 1);
        if (!touched) {
            *p1 = 0.0;
            return DQ_OK;
        }
    }
    // ops that decide p1 (the queue up to trailing unitaries on other qubits,
    // such as the x of a reset of another qubit); the rest stay queued
    const size_t live = marginal_prefix(s, qubit);
    const bool elementwise = live > 0 && all_elementwise(s, live);
    if (elementwise) {
        // a pending collapse of this qubit to 0, with only diagonal ops and
        // collapses after it, leaves every amplitude with the bit set at an
        // exact zero (unitaries on other qubits only combine those zeros):
        // the sum is +0.0 without reading the state (a reset right after a
        // measurement that gave 0)
        for (size_t i = live; i-- > 0;) {
            const HostOp &h = s->queue[i];
            if (h.kind != K_COLLAPSE || h.q[0] != qubit) continue;
            if (h.outcome == 0) {
                *p1 = 0.0;
                return DQ_OK;
            }
            break;
        }
    }
    const uint64_t count = amp_count(s) / 2;
    const int blocks = std::min<int>(kReduceBlocks, grid_for(count));
    int nparts = blocks;
    if (elementwise && live <= 64) {
        // only diagonals/collapses decide p1: reduce over the state they
        // would produce and keep the whole queue for the next sweep
        Packed pk;
        int pos[64];
        for (int q = 0; q < 64; ++q) pos[q] = q;
        for (size_t i = 0; i < live; ++i) pk.ops.push_back(to_dev(s->queue[i], pos, pk));
        DevOp *d_ops;
        void *d_mats;
        int *d_perms;
        if (int rc = upload(s, pk, &d_ops, &d_mats, &d_perms)) return rc;
        if (s->dtype == DQ_C128)
            hipLaunchKernelGGL(k_prob_one_virtual<double>, dim3(blocks), dim3(kThreads), 0, s->stream,
                               static_cast<const cplx<double> *>(s->amps), s->n, qubit, d_ops, int(pk.ops.size()),
                               static_cast<const cplx<double> *>(d_mats), s->d_partials);
        else
            hipLaunchKernelGGL(k_prob_one_virtual<float>, dim3(blocks), dim3(kThreads), 0, s->stream,
                               static_cast<const cplx<float> *>(s->amps), s->n, qubit, d_ops, int(pk.ops.size()),
                               static_cast<const cplx<float> *>(d_mats), s->d_partials);
        s->launches++;
    } else if (live > 0 && reg_path(s)) {
        // the flush's last sweep sums p1 as it writes the state
        const int rc = s->dtype == DQ_C128 ? flush_t<double>(s, qubit) : flush_t<float>(s, qubit);
        if (rc) return rc;
        nparts = s->last_grid;
    } else {
        // nothing queued decides p1 (live == 0): the state in memory has the
        // same marginal, and the queue stays for the next sweep
        if (live > 0)
            if (int rc = flush(s)) return rc;
        if (s->dtype == DQ_C128)
            hipLaunchKernelGGL(k_prob_one<double>, dim3(blocks), dim3(kThreads), 0, s->stream,
                               static_cast<const cplx<double> *>(s->amps), s->n, qubit, s->d_partials);
        else
            hipLaunchKernelGGL(k_prob_one<float>, dim3(blocks), dim3(kThreads), 0, s->stream,
                               static_cast<const cplx<float> *>(s->amps), s->n, qubit, s->d_partials);
        s->launches++;
    }
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kThreads), 0, s->stream, s->d_partials, nparts, 1,
                       s->d_result);
    HIP_TRY(hipGetLastError());
    s->launches++;
    if (int rc = fetch_result(s, 1)) return rc;
    *p1 = s->h_result[0];
    return DQ_OK;
}

}  // extern "C"

// Work space of dq_marginal_probs_noalloc (see below); 4 MiB + 4 KiB holds
// the marginal over up to 18 qubits (19 when each bin sums <= 64 amplitudes).
constexpr size_t kNoallocScratchBytes = (size_t(4) << 20) + 4096;
static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

namespace {

// Argument checks shared by the marginal entry points.
int check_marginal_args(dq_state *s, const int *qubits, int k, const double *out) {
    if (!s || !out || (k > 0 && !qubits)) return fail(DQ_ERR_ARG, "null argument");
    if (k < 0 || k > s->n) return fail(DQ_ERR_ARG, "number of qubits out of range");
    if (k > 30) return fail(DQ_ERR_ARG, "marginal over more than 30 qubits is not supported");
    uint64_t seen = 0;
    for (int j = 0; j < k; ++j) {
        if (int rc = check_qubit(s, qubits[j])) return rc;
        const uint64_t m = uint64_t(1) << qubits[j];
        if (seen & m) return fail(DQ_ERR_ARG, "qubits must be distinct");
        seen |= m;
    }
    return DQ_OK;
}

// Device work space of one marginal: the qubit table (128 ints), the output
// bins and, when a bin has more than 64 terms, per-(bin, chunk) partials.
struct MarginalShape {
    uint64_t bins, terms;
    int chunks;
    size_t meta_b, out_b, part_b;
    size_t bytes() const { return meta_b + out_b + part_b; }
};

MarginalShape marginal_shape(int n, int k) {
    MarginalShape m;
    m.bins = uint64_t(1) << k;
    m.terms = uint64_t(1) << (n - k);
    m.chunks = 1;
    if (m.terms > 64) {
        // enough blocks to fill the device, at least 256 terms per block
        uint64_t want = std::max<uint64_t>(1, 4096 / m.bins);
        while (uint64_t(m.chunks) * 2 <= want && m.terms / (uint64_t(m.chunks) * 2) >= 256) m.chunks *= 2;
    }
    m.meta_b = 128 * sizeof(int);
    m.out_b = m.bins * sizeof(double);
    m.part_b = (m.terms > 64) ? m.bins * uint64_t(m.chunks) * sizeof(double) : 0;
    return m;
}

// The marginal of a flushed state into `out` (host), with `dbuf` (device,
// at least shape.bytes(), 256-byte aligned) as the work space.  Every bin
// and partial it reads is written first in the same call, so dbuf needs no
// clearing.  Synchronises the state's stream before returning.
int marginal_into(dq_state *s, const int *qubits, int k, double *out, void *dbuf, const MarginalShape &sh) {
    const int n = s->n;
    int meta[128] = {0};
    uint64_t seen = 0;
    for (int j = 0; j < k; ++j) {
        meta[j] = qubits[j];
        seen |= uint64_t(1) << qubits[j];
    }
    int nrest = 0;
    for (int q = 0; q < n; ++q)
        if (!((seen >> q) & 1)) meta[64 + nrest++] = q;
    int *d_meta = static_cast<int *>(dbuf);
    double *d_out = reinterpret_cast<double *>(static_cast<unsigned char *>(dbuf) + sh.meta_b);
    double *d_part = reinterpret_cast<double *>(static_cast<unsigned char *>(dbuf) + sh.meta_b + sh.out_b);
    const uint64_t bins = sh.bins;
    hipError_t e = hipMemcpyAsync(d_meta, meta, sh.meta_b, hipMemcpyHostToDevice, s->stream);
    bool identity = (k == n);
    for (int j = 0; j < k && identity; ++j) identity = (qubits[j] == n - 1 - j);
    if (e == hipSuccess) {
        if (k == n) {
            if (s->dtype == DQ_C128)
                hipLaunchKernelGGL(k_probs_perm<double>, dim3(grid_for(bins)), dim3(kThreads), 0, s->stream,
                                   static_cast<const cplx<double> *>(s->amps), n, d_meta, d_out, int(identity));
            else
                hipLaunchKernelGGL(k_probs_perm<float>, dim3(grid_for(bins)), dim3(kThreads), 0, s->stream,
                                   static_cast<const cplx<float> *>(s->amps), n, d_meta, d_out, int(identity));
        } else if (sh.terms <= 64) {
            if (s->dtype == DQ_C128)
                hipLaunchKernelGGL(k_marginal_seq<double>, dim3(grid_for(bins)), dim3(kThreads), 0, s->stream,
                                   static_cast<const cplx<double> *>(s->amps), n, d_meta, k, d_meta + 64, d_out);
            else
                hipLaunchKernelGGL(k_marginal_seq<float>, dim3(grid_for(bins)), dim3(kThreads), 0, s->stream,
                                   static_cast<const cplx<float> *>(s->amps), n, d_meta, k, d_meta + 64, d_out);
        } else {
            const unsigned nblk = unsigned(bins * uint64_t(sh.chunks));
            if (s->dtype == DQ_C128)
                hipLaunchKernelGGL(k_marginal_tree<double>, dim3(nblk), dim3(kThreads), 0, s->stream,
                                   static_cast<const cplx<double> *>(s->amps), n, d_meta, k, d_meta + 64, sh.chunks,
                                   d_part);
            else
                hipLaunchKernelGGL(k_marginal_tree<float>, dim3(nblk), dim3(kThreads), 0, s->stream,
                                   static_cast<const cplx<float> *>(s->amps), n, d_meta, k, d_meta + 64, sh.chunks,
                                   d_part);
            hipLaunchKernelGGL(k_sum_chunks, dim3(grid_for(bins)), dim3(kThreads), 0, s->stream, d_part, sh.chunks,
                               bins, d_out);
        }
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(out, d_out, sh.out_b, hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) {
        (void)hipStreamSynchronize(s->stream);
        return fail(DQ_ERR_DEVICE, std::string("marginal_probs: ") + hipGetErrorString(e));
    }
    s->launches += 2;
    return DQ_OK;
}

// ---- the allocation-free marginal (dq_marginal_probs_noalloc) ----
//
// Its work space is this file-scope device array, one copy per GPU for the
// life of the process.  Host code never passes the array's name to a kernel
// or a copy (on the host that name is only a stand-in): the array's device
// address is looked up once per device, with that device current, and kept
// in noalloc_ptr.  Calls on one device are serialised by noalloc_mu, since
// they share the array.
constexpr int kMaxDevices = 64;
std::mutex noalloc_mu[kMaxDevices];
void *noalloc_ptr[kMaxDevices] = {nullptr};

// Device address of `scratch` on the state's device (looked up on the first
// call there; the caller holds noalloc_mu[device]).
int noalloc_scratch(dq_state *s, void **ptr) {
    const int dev = s->device;
    if (!noalloc_ptr[dev]) {
        int cur = 0;
        HIP_TRY(hipGetDevice(&cur));
        if (cur != dev) HIP_TRY(hipSetDevice(dev));
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != dev) HIP_TRY(hipSetDevice(cur));
        if (e != hipSuccess || !p)
            return fail(DQ_ERR_DEVICE, std::string("hipGetSymbolAddress of the no-allocation work space: ") +
                                           hipGetErrorString(e));
        noalloc_ptr[dev] = p;
    }
    *ptr = noalloc_ptr[dev];
    return DQ_OK;
}

}  // namespace

extern "C" {

int dq_marginal_probs(dq_state *s, const int *qubits, int k, double *out) {
    if (int rc = check_marginal_args(s, qubits, k, out)) return rc;
    if (int rc = flush(s)) return rc;
    const MarginalShape sh = marginal_shape(s->n, k);
    void *dbuf = nullptr;
    if (hipMalloc(&dbuf, sh.bytes()) != hipSuccess) return fail(DQ_ERR_OOM, "hipMalloc for marginal probabilities failed");
    const int rc = marginal_into(s, qubits, k, out, dbuf, sh);
    (void)hipFree(dbuf);
    return rc;
}

int dq_marginal_probs_noalloc(dq_state *s, const int *qubits, int k, double *out) {
    if (int rc = check_marginal_args(s, qubits, k, out)) return rc;
    if (!s->queue.empty())
        return fail(DQ_ERR_ARG, "marginal_probs_noalloc: gates are still queued; call dq_synchronize first");
    const MarginalShape sh = marginal_shape(s->n, k);
    if (sh.bytes() > kNoallocScratchBytes)
        return fail(DQ_ERR_ARG, "marginal_probs_noalloc: " + std::to_string(k) + " qubits need " +
                                    std::to_string(sh.bytes()) + " bytes of work space, more than its maximum of " +
                                    std::to_string(kNoallocScratchBytes));
    if (s->device < 0 || s->device >= kMaxDevices) return fail(DQ_ERR_ARG, "device index out of range");
    std::lock_guard<std::mutex> lock(noalloc_mu[s->device]);
    void *dbuf = nullptr;
    if (int rc = noalloc_scratch(s, &dbuf)) return rc;
    return marginal_into(s, qubits, k, out, dbuf, sh);
}

int dq_inner(dq_state *s, const double *other, double *re, double *im) {
    if (!s || !other || !re || !im) return fail(DQ_ERR_ARG, "null argument");
    if (int rc = flush(s)) return rc;
    const uint64_t count = amp_count(s);
    void *tmp = nullptr;
    if (hipMalloc(&tmp, 16 * count) != hipSuccess) return fail(DQ_ERR_OOM, "hipMalloc for inner product failed");
    hipError_t e = hipMemcpyAsync(tmp, other, 16 * count, hipMemcpyHostToDevice, s->stream);
    const int blocks = std::min<int>(kReduceBlocks, grid_for(count));
    if (e == hipSuccess) {
        if (s->dtype == DQ_C128)
            hipLaunchKernelGGL(k_inner<double>, dim3(blocks), dim3(kThreads), 0, s->stream,
                               static_cast<const cplx<double> *>(s->amps), static_cast<const cplx<double> *>(tmp),
                               count, s->d_partials);
        else
            hipLaunchKernelGGL(k_inner<float>, dim3(blocks), dim3(kThreads), 0, s->stream,
                               static_cast<const cplx<float> *>(s->amps), static_cast<const cplx<double> *>(tmp),
                               count, s->d_partials);
        hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kThreads), 0, s->stream, s->d_partials, blocks, 2,
                           s->d_result);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    (void)hipFree(tmp);
    if (e != hipSuccess) return fail(DQ_ERR_DEVICE, std::string("inner: ") + hipGetErrorString(e));
    if (int rc = fetch_result(s, 2)) return rc;
    *re = s->h_result[0];
    *im = s->h_result[1];
    return DQ_OK;
}

int dq_copy_to_host(dq_state *s, double *dst) {
    if (!s || !dst) return fail(DQ_ERR_ARG, "null argument");
    if (int rc = flush(s)) return rc;
    const uint64_t count = amp_count(s);
    HIP_TRY(hipStreamSynchronize(s->stream));
    if (s->dtype == DQ_C128) {
        HIP_TRY(hipMemcpy(dst, s->amps, 16 * count, hipMemcpyDeviceToHost));
        return DQ_OK;
    }
    void *tmp = nullptr;
    if (hipMalloc(&tmp, 16 * count) != hipSuccess) return fail(DQ_ERR_OOM, "hipMalloc failed");
    hipLaunchKernelGGL(k_to_f64<float>, dim3(grid_for(count)), dim3(kThreads), 0, s->stream,
                       static_cast<cplx<double> *>(tmp), static_cast<const cplx<float> *>(s->amps), count);
    hipError_t e = hipStreamSynchronize(s->stream);
    if (e == hipSuccess) e = hipMemcpy(dst, tmp, 16 * count, hipMemcpyDeviceToHost);
    (void)hipFree(tmp);
    if (e != hipSuccess) return fail(DQ_ERR_DEVICE, hipGetErrorString(e));
    return DQ_OK;
}

int dq_synchronize(dq_state *s) {
    if (!s) return fail(DQ_ERR_ARG, "null state");
    if (int rc = flush(s)) return rc;
    HIP_TRY(hipStreamSynchronize(s->stream));
    return DQ_OK;
}

namespace {

// Builds the device term list; Y phases folded into the coefficients.
int pauli_terms(const dq_state *s, int n_terms, const uint64_t *xmask, const uint64_t *zmask, const double *coeff,
                std::vector<PauliTerm> &out) {
    if (n_terms < 0 || n_terms > 4096) return fail(DQ_ERR_ARG, "number of Pauli terms must be in [0, 4096]");
    if (n_terms > 0 && (!xmask || !zmask || !coeff)) return fail(DQ_ERR_ARG, "null Pauli term arrays");
    const uint64_t full = s->n >= 64 ? ~uint64_t(0) : ((uint64_t(1) << s->n) - 1);
    out.resize(size_t(n_terms));
    for (int k = 0; k < n_terms; ++k) {
        if ((xmask[k] | zmask[k]) & ~full) return fail(DQ_ERR_ARG, "Pauli term acts on a qubit outside the state");
        const int ny = __builtin_popcountll(xmask[k] & zmask[k]) & 3;
        double re = coeff[2 * k], im = coeff[2 * k + 1];
        for (int j = 0; j < ny; ++j) {  // multiply by i
            const double t = re;
            re = -im;
            im = t;
        }
        out[k] = PauliTerm{xmask[k], zmask[k], re, im};
    }
    return DQ_OK;
}

int upload_terms(dq_state *s, const std::vector<PauliTerm> &terms, PauliTerm **d_terms) {
    const size_t bytes = std::max<size_t>(terms.size(), 1) * sizeof(PauliTerm);
    HIP_TRY(hipEventSynchronize(s->upload_done));
    if (int rc = ensure_host_buffer(s, bytes)) return rc;
    if (int rc = ensure_device_buffer(s, bytes)) return rc;
    std::memcpy(s->h_upload, terms.data(), terms.size() * sizeof(PauliTerm));
    HIP_TRY(hipMemcpyAsync(s->d_upload, s->h_upload, bytes, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipEventRecord(s->upload_done, s->stream));
    *d_terms = static_cast<PauliTerm *>(s->d_upload);
    return DQ_OK;
}

// Makes `b`'s pending work visible on `a`'s stream (both flushed first).
int join(dq_state *a, dq_state *b) {
    if (int rc = flush(a)) return rc;
    if (a == b) return DQ_OK;
    if (int rc = flush(b)) return rc;
    HIP_TRY(hipStreamSynchronize(b->stream));
    return DQ_OK;
}

}  // namespace

int dq_state_copy(dq_state *dst, dq_state *src) {
    if (!dst || !src) return fail(DQ_ERR_ARG, "null state");
    if (dst->n != src->n || dst->dtype != src->dtype) return fail(DQ_ERR_ARG, "states differ in size or dtype");
    if (dst == src) return flush(dst);
    if (int rc = join(dst, src)) return rc;
    dst->queue.clear();
    dst->zero_q = src->zero_q;
    const uint64_t count = amp_count(dst);
    if (dst->dtype == DQ_C128)
        hipLaunchKernelGGL(k_copy_state<double>, dim3(grid_for(count)), dim3(kThreads), 0, dst->stream,
                           static_cast<cplx<double> *>(dst->amps), static_cast<const cplx<double> *>(src->amps), count);
    else
        hipLaunchKernelGGL(k_copy_state<float>, dim3(grid_for(count)), dim3(kThreads), 0, dst->stream,
                           static_cast<cplx<float> *>(dst->amps), static_cast<const cplx<float> *>(src->amps), count);
    HIP_TRY(hipGetLastError());
    dst->launches++;
    // src may be modified on its own stream right after this call returns
    HIP_TRY(hipStreamSynchronize(dst->stream));
    return DQ_OK;
}

int dq_state_tensor(dq_state *dst, dq_state *a, const int *a_pos, dq_state *b, const int *b_pos) {
    if (!dst || !a || !b || (a->n && !a_pos) || (b->n && !b_pos)) return fail(DQ_ERR_ARG, "null argument");
    if (dst == a || dst == b) return fail(DQ_ERR_ARG, "dq_state_tensor needs a destination distinct from its factors");
    if (dst->n != a->n + b->n) return fail(DQ_ERR_ARG, "destination must have the factors' qubit count combined");
    if (dst->dtype != a->dtype || dst->dtype != b->dtype) return fail(DQ_ERR_ARG, "states differ in dtype");
    TensorPos tp;
    tp.na = a->n;
    tp.nb = b->n;
    uint64_t seen = 0;
    for (int k = 0; k < a->n + b->n; ++k) {
        const int p = k < a->n ? a_pos[k] : b_pos[k - a->n];
        if (p < 0 || p >= dst->n || ((seen >> p) & 1))
            return fail(DQ_ERR_ARG, "a_pos and b_pos must name every destination qubit once");
        seen |= uint64_t(1) << p;
        (k < a->n ? tp.apos[k] : tp.bpos[k - a->n]) = p;
    }
    if (int rc = join(dst, a)) return rc;
    if (int rc = join(dst, b)) return rc;
    dst->queue.clear();
    // a product amplitude is zero wherever either factor's is
    dst->zero_q = 0;
    for (int k = 0; k < a->n; ++k)
        if ((a->zero_q >> k) & 1) dst->zero_q |= uint64_t(1) << tp.apos[k];
    for (int k = 0; k < b->n; ++k)
        if ((b->zero_q >> k) & 1) dst->zero_q |= uint64_t(1) << tp.bpos[k];
    const uint64_t count = amp_count(dst);
    if (dst->dtype == DQ_C128)
        hipLaunchKernelGGL(k_tensor<double>, dim3(grid_for(count)), dim3(kThreads), 0, dst->stream,
                           static_cast<cplx<double> *>(dst->amps), static_cast<const cplx<double> *>(a->amps),
                           static_cast<const cplx<double> *>(b->amps), tp, count);
    else
        hipLaunchKernelGGL(k_tensor<float>, dim3(grid_for(count)), dim3(kThreads), 0, dst->stream,
                           static_cast<cplx<float> *>(dst->amps), static_cast<const cplx<float> *>(a->amps),
                           static_cast<const cplx<float> *>(b->amps), tp, count);
    HIP_TRY(hipGetLastError());
    dst->launches++;
    // a and b may be modified or freed right after this call returns
    HIP_TRY(hipStreamSynchronize(dst->stream));
    return DQ_OK;
}

int dq_apply_pauli_sum(dq_state *dst, dq_state *src, int n_terms, const uint64_t *xmask, const uint64_t *zmask,
                       const double *coeff) {
    if (!dst || !src) return fail(DQ_ERR_ARG, "null state");
    if (dst == src) return fail(DQ_ERR_ARG, "dq_apply_pauli_sum needs distinct source and destination states");
    if (dst->n != src->n || dst->dtype != src->dtype) return fail(DQ_ERR_ARG, "states differ in size or dtype");
    std::vector<PauliTerm> terms;
    if (int rc = pauli_terms(src, n_terms, xmask, zmask, coeff, terms)) return rc;
    if (int rc = join(dst, src)) return rc;
    dst->queue.clear();
    dst->zero_q = 0;
    PauliTerm *d_terms;
    if (int rc = upload_terms(dst, terms, &d_terms)) return rc;
    const uint64_t count = amp_count(dst);
    if (dst->dtype == DQ_C128)
        hipLaunchKernelGGL(k_apply_pauli_sum<double>, dim3(grid_for(count)), dim3(kThreads), 0, dst->stream,
                           static_cast<cplx<double> *>(dst->amps), static_cast<const cplx<double> *>(src->amps), count,
                           d_terms, n_terms);
    else
        hipLaunchKernelGGL(k_apply_pauli_sum<float>, dim3(grid_for(count)), dim3(kThreads), 0, dst->stream,
                           static_cast<cplx<float> *>(dst->amps), static_cast<const cplx<float> *>(src->amps), count,
                           d_terms, n_terms);
    HIP_TRY(hipGetLastError());
    dst->launches++;
    // src may be modified on its own stream right after this call returns
    HIP_TRY(hipStreamSynchronize(dst->stream));
    return DQ_OK;
}

int dq_pauli_inner(dq_state *bra, dq_state *ket, int n_terms, const uint64_t *xmask, const uint64_t *zmask,
                   const double *coeff, double *re, double *im) {
    if (!bra || !ket || !re || !im) return fail(DQ_ERR_ARG, "null argument");
    if (bra->n != ket->n || bra->dtype != ket->dtype) return fail(DQ_ERR_ARG, "states differ in size or dtype");
    std::vector<PauliTerm> terms;
    if (int rc = pauli_terms(ket, n_terms, xmask, zmask, coeff, terms)) return rc;
    if (int rc = join(ket, bra)) return rc;
    PauliTerm *d_terms;
    if (int rc = upload_terms(ket, terms, &d_terms)) return rc;
    const uint64_t count = amp_count(ket);
    const int blocks = std::min<int>(kReduceBlocks, grid_for(count));
    if (ket->dtype == DQ_C128)
        hipLaunchKernelGGL(k_pauli_inner<double>, dim3(blocks), dim3(kThreads), 0, ket->stream,
                           static_cast<const cplx<double> *>(bra->amps), static_cast<const cplx<double> *>(ket->amps),
                           count, d_terms, n_terms, ket->d_partials);
    else
        hipLaunchKernelGGL(k_pauli_inner<float>, dim3(blocks), dim3(kThreads), 0, ket->stream,
                           static_cast<const cplx<float> *>(bra->amps), static_cast<const cplx<float> *>(ket->amps),
                           count, d_terms, n_terms, ket->d_partials);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kThreads), 0, ket->stream, ket->d_partials, blocks, 2,
                       ket->d_result);
    HIP_TRY(hipGetLastError());
    ket->launches += 2;
    if (int rc = fetch_result(ket, 2)) return rc;
    *re = ket->h_result[0];
    *im = ket->h_result[1];
    return DQ_OK;
}

// ---- top-k of the distribution (lazy results; no reference counterpart) ----
//
// The k most probable basis states without copying the distribution to the
// host: a radix select over the bit patterns of p = re^2 + im^2 (as for
// probabilities(); non-negative doubles order like their bit patterns).
// Each level histograms 11 bits of the keys that share the prefix chosen so
// far (a read of the state), until the keys at or above the threshold are
// few enough to collect; ties at the k-th value go to the smaller indices.

extern "C++" {
namespace {

constexpr int kTopBins = 2048;

__global__ void k_zero_bytes(unsigned char *__restrict__ p, size_t n) {
    for (size_t i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

template <typename T>
__device__ __forceinline__ uint64_t prob_key(const cplx<T> &a, double &p) {
    const double re = double(a.re), im = double(a.im);
    p = re * re + im * im;
    return uint64_t(__double_as_longlong(p));
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_topk_hist(const cplx<T> *__restrict__ amps, uint64_t count, int shift,
                                                        int pshift, uint64_t prefix, unsigned mask,
                                                        unsigned *__restrict__ hist) {
    __shared__ unsigned h[kTopBins];
    for (int i = threadIdx.x; i < kTopBins; i += kThreads) h[i] = 0;
    __syncthreads();
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += stride) {
        double p;
        const uint64_t key = prob_key(amps[i], p);
        if (pshift >= 64 || (key >> pshift) == prefix) atomicAdd(&h[unsigned(key >> shift) & mask], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kTopBins; i += kThreads)
        if (h[i]) atomicAdd(&hist[i], h[i]);
}

// Collects (index, p) of the keys >= lo (and == lo only below index `below`
// when tie_only is set).
template <typename T>
__global__ __launch_bounds__(kThreads) void k_topk_collect(const cplx<T> *__restrict__ amps, uint64_t count,
                                                           uint64_t lo, int tie_only, uint64_t below,
                                                           uint64_t *__restrict__ idx, double *__restrict__ val,
                                                           unsigned long long *__restrict__ counter, uint64_t cap) {
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += stride) {
        double p;
        const uint64_t key = prob_key(amps[i], p);
        const bool take = tie_only ? (key == lo && i < below) : (key >= lo);
        if (take) {
            const unsigned long long pos = atomicAdd(counter, 1ull);
            if (pos < cap) {
                idx[pos] = i;
                val[pos] = p;
            }
        }
    }
}

// Counts the keys equal to `key_eq` at indices below `below`.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_topk_count_ties(const cplx<T> *__restrict__ amps, uint64_t count,
                                                              uint64_t key_eq, uint64_t below,
                                                              unsigned long long *__restrict__ counter) {
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    unsigned long long c = 0;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count && i < below; i += stride) {
        double p;
        c += prob_key(amps[i], p) == key_eq;
    }
    if (c) atomicAdd(counter, c);
}

template <typename T>
int top_k_impl(dq_state *s, uint64_t k, uint64_t *idx_out, double *p_out) {
    const uint64_t count = amp_count(s);
    const cplx<T> *amps = static_cast<const cplx<T> *>(s->amps);
    const uint64_t cap = std::max<uint64_t>(uint64_t(1) << 22, 2 * k);
    const size_t hist_b = kTopBins * sizeof(unsigned), cnt_b = 256;
    const size_t idx_b = cap * sizeof(uint64_t), val_b = cap * sizeof(double);
    unsigned char *d = nullptr;
    if (hipMalloc(&d, hist_b + cnt_b + idx_b + val_b) != hipSuccess) return fail(DQ_ERR_OOM, "hipMalloc for top-k failed");
    struct Free {
        unsigned char *p;
        ~Free() { (void)hipFree(p); }
    } guard{d};
    unsigned *d_hist = reinterpret_cast<unsigned *>(d);
    unsigned long long *d_cnt = reinterpret_cast<unsigned long long *>(d + hist_b);
    uint64_t *d_idx = reinterpret_cast<uint64_t *>(d + hist_b + cnt_b);
    double *d_val = reinterpret_cast<double *>(d + hist_b + cnt_b + idx_b);
    const int blocks = std::min<int>(4096, grid_for(count));
    std::vector<unsigned> hist(kTopBins);
    auto zero = [&](void *p, size_t b) -> int {
        hipLaunchKernelGGL(k_zero_bytes, dim3(1), dim3(kThreads), 0, s->stream, static_cast<unsigned char *>(p), b);
        HIP_TRY(hipGetLastError());
        return DQ_OK;
    };
    auto collect = [&](uint64_t lo, int tie_only, uint64_t below, std::vector<std::pair<uint64_t, double>> &out) -> int {
        if (int rc = zero(d_cnt, sizeof(unsigned long long))) return rc;
        hipLaunchKernelGGL(k_topk_collect<T>, dim3(blocks), dim3(kThreads), 0, s->stream, amps, count, lo, tie_only,
                           below, d_idx, d_val, d_cnt, cap);
        HIP_TRY(hipGetLastError());
        unsigned long long got = 0;
        HIP_TRY(hipMemcpyAsync(&got, d_cnt, sizeof got, hipMemcpyDeviceToHost, s->stream));
        HIP_TRY(hipStreamSynchronize(s->stream));
        const uint64_t m = std::min<uint64_t>(got, cap);
        std::vector<uint64_t> ii(m);
        std::vector<double> vv(m);
        HIP_TRY(hipMemcpyAsync(ii.data(), d_idx, m * sizeof(uint64_t), hipMemcpyDeviceToHost, s->stream));
        HIP_TRY(hipMemcpyAsync(vv.data(), d_val, m * sizeof(double), hipMemcpyDeviceToHost, s->stream));
        HIP_TRY(hipStreamSynchronize(s->stream));
        for (uint64_t j = 0; j < m; ++j) out.emplace_back(ii[j], vv[j]);
        return DQ_OK;
    };
    uint64_t above = 0, prefix = 0;  // keys known above the bracket; key bits fixed so far
    int pshift = 64;
    std::vector<std::pair<uint64_t, double>> cand;
    static const int shifts[6] = {53, 42, 31, 20, 9, 0};
    for (int level = 0; level < 6; ++level) {
        const int shift = shifts[level], width = pshift - shift;
        if (int rc = zero(d_hist, hist_b)) return rc;
        hipLaunchKernelGGL(k_topk_hist<T>, dim3(blocks), dim3(kThreads), 0, s->stream, amps, count, shift, pshift,
                           prefix, (1u << width) - 1u, d_hist);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(hist.data(), d_hist, hist_b, hipMemcpyDeviceToHost, s->stream));
        HIP_TRY(hipStreamSynchronize(s->stream));
        int b = (1 << width) - 1;
        uint64_t acc = above;
        for (; b > 0; --b) {
            if (acc + hist[size_t(b)] >= k) break;
            acc += hist[size_t(b)];
        }
        above = acc;
        prefix = (prefix << width) | uint64_t(b);
        pshift = shift;
        const uint64_t lo = prefix << shift;
        if (above + hist[size_t(b)] <= cap) {  // few enough: collect every key >= lo
            if (int rc = collect(lo, 0, 0, cand)) return rc;
            break;
        }
        if (shift == 0) {  // key fully resolved: all keys above it, then the smallest-index ties
            if (int rc = collect(lo + 1, 0, 0, cand)) return rc;
            const uint64_t need = k - cand.size();
            uint64_t lo_i = 0, hi_i = count;  // smallest bound with `need` ties below it
            while (lo_i < hi_i) {
                const uint64_t mid = lo_i + (hi_i - lo_i) / 2;
                if (int rc = zero(d_cnt, sizeof(unsigned long long))) return rc;
                hipLaunchKernelGGL(k_topk_count_ties<T>, dim3(blocks), dim3(kThreads), 0, s->stream, amps, count, lo,
                                   mid, d_cnt);
                HIP_TRY(hipGetLastError());
                unsigned long long c = 0;
                HIP_TRY(hipMemcpyAsync(&c, d_cnt, sizeof c, hipMemcpyDeviceToHost, s->stream));
                HIP_TRY(hipStreamSynchronize(s->stream));
                if (c >= need) hi_i = mid;
                else lo_i = mid + 1;
            }
            if (int rc = collect(lo, 1, lo_i, cand)) return rc;
            break;
        }
    }
    std::sort(cand.begin(), cand.end(), [](const std::pair<uint64_t, double> &x, const std::pair<uint64_t, double> &y) {
        return x.second != y.second ? x.second > y.second : x.first < y.first;
    });
    if (cand.size() < k) return fail(DQ_ERR_DEVICE, "top-k selection collected too few candidates");
    for (uint64_t j = 0; j < k; ++j) {
        idx_out[j] = cand[j].first;
        p_out[j] = cand[j].second;
    }
    return DQ_OK;
}

}  // namespace
}  // extern "C++"

int dq_top_k(dq_state *s, uint64_t k, uint64_t *idx_out, double *p_out) {
    if (!s || !idx_out || !p_out) return fail(DQ_ERR_ARG, "null argument");
    if (k < 1 || k > (uint64_t(1) << 20) || k > amp_count(s))
        return fail(DQ_ERR_ARG, "k must be in [1, min(2^20, 2^n)]");
    if (int rc = flush(s)) return rc;
    return s->dtype == DQ_C128 ? top_k_impl<double>(s, k, idx_out, p_out) : top_k_impl<float>(s, k, idx_out, p_out);
}

int dq_compare_states(dq_state *a, dq_state *b, double *max_abs, double *inner_re, double *inner_im,
                      double *tvd) {
    if (!a || !b || !max_abs || !inner_re || !inner_im || !tvd) return fail(DQ_ERR_ARG, "null argument");
    if (a->n != b->n) return fail(DQ_ERR_ARG, "states differ in size");
    if (int rc = join(a, b)) return rc;
    const uint64_t count = amp_count(a);
    const int blocks = std::min<int>(kReduceBlocks, grid_for(count));
#define DQ_CMP(TA, TB)                                                                                   \
    hipLaunchKernelGGL((k_compare<TA, TB>), dim3(blocks), dim3(kThreads), 0, a->stream,                  \
                       static_cast<const cplx<TA> *>(a->amps), static_cast<const cplx<TB> *>(b->amps), count, \
                       a->d_partials)
    if (a->dtype == DQ_C128 && b->dtype == DQ_C128) DQ_CMP(double, double);
    else if (a->dtype == DQ_C128) DQ_CMP(double, float);
    else if (b->dtype == DQ_C128) DQ_CMP(float, double);
    else DQ_CMP(float, float);
#undef DQ_CMP
    HIP_TRY(hipGetLastError());
    std::vector<double> part(4 * size_t(blocks));
    HIP_TRY(hipMemcpyAsync(part.data(), a->d_partials, part.size() * sizeof(double), hipMemcpyDeviceToHost, a->stream));
    HIP_TRY(hipStreamSynchronize(a->stream));
    double m = 0, r = 0, i = 0, t = 0;
    for (int k = 0; k < blocks; ++k) {
        m = std::max(m, part[4 * size_t(k)]);
        r += part[4 * size_t(k) + 1];
        i += part[4 * size_t(k) + 2];
        t += part[4 * size_t(k) + 3];
    }
    *max_abs = m;
    *inner_re = r;
    *inner_im = i;
    *tvd = 0.5 * t;
    return DQ_OK;
}

int dq_set_options(dq_state *s, int fuse, int tile_bits, int low_bits) {
    if (!s) return fail(DQ_ERR_ARG, "null state");
    if (int rc = flush(s)) return rc;
    const int max_bits = 13;
    if (tile_bits != 0 && (tile_bits < 6 || tile_bits > max_bits))
        return fail(DQ_ERR_ARG, "tile_bits out of range");
    if (fuse < 0 || fuse > 3) return fail(DQ_ERR_ARG, "fusion level must be 0, 1, 2 or 3");
    if (fuse == 3 && s->dtype != DQ_C64) return fail(DQ_ERR_ARG, "fusion level 3 (MFMA blocks) needs a complex64 state");
    if (low_bits != 0 && (low_bits < 1 || low_bits > 8)) return fail(DQ_ERR_ARG, "low_bits out of range");
    s->fuse = fuse;
    if (tile_bits) s->tile_bits = tile_bits;
    if (low_bits) s->low_bits = low_bits;
    return DQ_OK;
}

int dq_set_timing(dq_state *s, int on) {
    if (!s) return fail(DQ_ERR_ARG, "null state");
    if (int rc = flush(s)) return rc;
    if (int rc = collect_timing(s)) return rc;
    s->timing = on ? 1 : 0;
    s->timed_ms = 0.0;
    s->timed_launches = 0;
    s->swept_amps0 = s->swept_amps;
    return DQ_OK;
}

int dq_get_timing(dq_state *s, double *sweep_ms, uint64_t *sweep_launches) {
    if (!s) return fail(DQ_ERR_ARG, "null state");
    if (int rc = flush(s)) return rc;
    if (int rc = collect_timing(s)) return rc;
    if (sweep_ms) *sweep_ms = s->timed_ms;
    if (sweep_launches) *sweep_launches = s->timed_launches;
    return DQ_OK;
}

int dq_set_zero_tiles(dq_state *s, int on) {
    if (!s) return fail(DQ_ERR_ARG, "null state");
    if (int rc = flush(s)) return rc;
    s->zero_tiles = on ? 1 : 0;
    if (!on) s->zero_q = 0;  // known again from the next reset
    return DQ_OK;
}

int dq_get_timed_amplitudes(dq_state *s, uint64_t *amps) {
    if (!s || !amps) return fail(DQ_ERR_ARG, "null argument");
    if (int rc = flush(s)) return rc;
    *amps = s->swept_amps - s->swept_amps0;
    return DQ_OK;
}

int dq_get_stats(const dq_state *s, uint64_t *sweeps, uint64_t *launches, uint64_t *gates) {
    if (!s) return fail(DQ_ERR_ARG, "null state");
    if (sweeps) *sweeps = s->sweeps;
    if (launches) *launches = s->launches;
    if (gates) *gates = s->gates;
    return DQ_OK;
}

int dq_set_pair_fusion(dq_state *s, int on) {
    if (!s) return fail(DQ_ERR_ARG, "null state");
    if (on != 0 && on != 1) return fail(DQ_ERR_ARG, "pair fusion flag must be 0 or 1");
    if (int rc = flush(s)) return rc;
    s->pair_fusion = on;
    return DQ_OK;
}

int dq_set_reg_bits(dq_state *s, int reg_bits) {
    if (!s) return fail(DQ_ERR_ARG, "null state");
    if (reg_bits != 3 && reg_bits != 4) return fail(DQ_ERR_ARG, "register bits must be 3 or 4");
    if (int rc = flush(s)) return rc;
    s->reg_bits = reg_bits;
    return DQ_OK;
}

int dq_set_fuse_min_cost(dq_state *s, double min_cost) {
    if (!s) return fail(DQ_ERR_ARG, "null state");
    if (!(min_cost >= 0.0)) return fail(DQ_ERR_ARG, "fusion cost threshold must be >= 0");
    if (int rc = flush(s)) return rc;
    s->fuse_min_cost = min_cost;
    return DQ_OK;
}

int dq_get_fused_stats(const dq_state *s, uint64_t *blocks, uint64_t *gates) {
    if (!s) return fail(DQ_ERR_ARG, "null state");
    if (blocks) *blocks = s->fused_blocks;
    if (gates) *gates = s->fused_gates;
    return DQ_OK;
}

}  // extern "C"
