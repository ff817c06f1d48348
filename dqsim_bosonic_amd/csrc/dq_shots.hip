// dq_shots.hip -- batched shots: StatevectorSimulator.simulate_shots on the GPU.
//
// Reference: src/monolithic/statevector/simulator.rs:215-322 runs every shot
// from |0...0> with its own ChaCha8Rng::seed_from_u64(base_seed + i) over the
// fused instruction list (src/types.rs:248-293), applying gates with the
// sequential kernels of src/engine.rs:265-360 (apply_one_qubit_seq,
// apply_n_qubit_seq, measure_qubit_seq) and returns the counts of the
// formatted classical registers.
//
// Here a program (gates, measure, reset, conditional) is built once on the
// host and every shot runs it:
//   * LDS path (2^n amplitudes fit one workgroup's LDS, n <= 13): one
//     workgroup per shot; the state lives in LDS; gates use the engine's exact
//     op classes (values equal to the reference's up to the sign of a zero);
//     a measurement sums p1 in ascending index order on one thread (the
//     reference's sequential sum, so p1 is bit-identical), draws from the
//     shot's ChaCha8 stream generated on the device, collapses in parallel.
//     Per-shot classical bits go back packed in 64-bit words.
//   * Engine path (larger n): the shots run one after another on one
//     statevector with the engine's sweeps; p1 then comes from the engine's
//     tree reduction (it can differ from the sequential sum in the last bits,
//     which changes an outcome only when a draw falls within that distance of
//     p1).
// Compiled without contraction (exact products and sums).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/dqsim_gpu.h"
#include "dq_device.hpp"
#include "dq_plan.hpp"

using namespace dq;

namespace {

enum ShotKind : int { SO_GATE = 0, SO_MEASURE = 1, SO_RESET = 2, SO_COND = 3 };

struct ShotOp {
    int kind;
    int a, b;   // GATE: device op index; MEASURE: qubit, cbit; RESET: qubit; COND: creg_base, creg_size
    int skip;   // COND: ops that form the conditioned instruction (skipped when false)
    uint64_t value;  // COND: creg_value
};

constexpr int kShotLdsQubits = 13;   // 128 KiB of complex128
constexpr int kMaxCbitMem = 1024;    // classical bits addressable on the LDS path

// ---- device ChaCha8 (rand_chacha 0.3.1 as called by the reference) ----
// Same algorithm as the host RNG of core.cpp: seed_from_u64 (PCG32 XSH-RR
// per 4-byte chunk), 8-round ChaCha, 64-bit block counter in words 12-13,
// stream 0, four blocks buffered, next_u64 = w[i] | w[i+1] << 32.
struct DevRng {
    uint32_t key[8];
    uint64_t counter;
    uint32_t buf[64];
    int idx;
};

__device__ __forceinline__ uint32_t rotl32(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }

__device__ void rng_seed(DevRng &r, uint64_t state) {
    for (int c = 0; c < 8; ++c) {
        state = state * 6364136223846793005ull + 11634580027462260723ull;
        const uint32_t xs = uint32_t(((state >> 18) ^ state) >> 27);
        const uint32_t rot = uint32_t(state >> 59);
        r.key[c] = (xs >> rot) | (xs << ((32 - rot) & 31));
    }
    r.counter = 0;
    r.idx = 64;
}

__device__ void rng_block(const DevRng &r, uint64_t ctr, uint32_t *out) {
    uint32_t in[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u};
    for (int i = 0; i < 8; ++i) in[4 + i] = r.key[i];
    in[12] = uint32_t(ctr);
    in[13] = uint32_t(ctr >> 32);
    in[14] = 0;
    in[15] = 0;
    uint32_t x[16];
    for (int i = 0; i < 16; ++i) x[i] = in[i];
#define DQ_QR(a, b, c, d)                       \
    x[a] += x[b]; x[d] = rotl32(x[d] ^ x[a], 16); \
    x[c] += x[d]; x[b] = rotl32(x[b] ^ x[c], 12); \
    x[a] += x[b]; x[d] = rotl32(x[d] ^ x[a], 8);  \
    x[c] += x[d]; x[b] = rotl32(x[b] ^ x[c], 7);
    for (int rd = 0; rd < 8; rd += 2) {
        DQ_QR(0, 4, 8, 12) DQ_QR(1, 5, 9, 13) DQ_QR(2, 6, 10, 14) DQ_QR(3, 7, 11, 15)
        DQ_QR(0, 5, 10, 15) DQ_QR(1, 6, 11, 12) DQ_QR(2, 7, 8, 13) DQ_QR(3, 4, 9, 14)
    }
#undef DQ_QR
    for (int i = 0; i < 16; ++i) out[i] = x[i] + in[i];
}

__device__ uint64_t rng_next_u64(DevRng &r) {
    if (r.idx >= 64) {
        for (int b = 0; b < 4; ++b) rng_block(r, r.counter + uint64_t(b), r.buf + 16 * b);
        r.counter += 4;
        r.idx = 0;
    }
    // the buffer holds 64 words and every draw takes two: a draw never straddles a refill
    const uint64_t v = uint64_t(r.buf[r.idx]) | (uint64_t(r.buf[r.idx + 1]) << 32);
    r.idx += 2;
    return v;
}

__device__ __forceinline__ double rng_f64(DevRng &r) { return double(rng_next_u64(r) >> 11) * (1.0 / 9007199254740992.0); }

// measure_qubit_seq (src/engine.rs:335-360): p1 summed in ascending index
// order by thread 0, then the draw, then the collapse by all threads.
__device__ int shot_measure(cplx<double> *sh, int n, int q, DevRng *rng, double *bcast, uint32_t tid, uint32_t nthr) {
    __syncthreads();
    const uint32_t N = 1u << n, bit = 1u << q;
    if (tid == 0) {
        double p1 = 0.0;
        for (uint32_t i = 0; i < N; ++i)
            if (i & bit) {
                const cplx<double> v = sh[i];
                p1 = p1 + (v.re * v.re + v.im * v.im);
            }
        const int outcome = rng_f64(*rng) < p1 ? 1 : 0;
        const double norm = outcome ? fmax(sqrt(p1), 1e-15) : fmax(sqrt(fmax(1.0 - p1, 0.0)), 1e-15);
        bcast[0] = double(outcome);
        bcast[1] = norm;
    }
    __syncthreads();
    const int outcome = int(bcast[0]);
    const double norm = bcast[1];
    for (uint32_t i = tid; i < N; i += nthr) {
        const int v = (i & bit) ? 1 : 0;
        if (v == outcome) {
            sh[i].re = sh[i].re / norm;
            sh[i].im = sh[i].im / norm;
        } else {
            sh[i] = cplx<double>{0.0, 0.0};
        }
    }
    __syncthreads();
    return outcome;
}

__global__ __launch_bounds__(512) void k_shots_lds(int n, uint64_t shots, uint64_t base_seed,
                                                   const ShotOp *__restrict__ prog, int nprog,
                                                   const DevOp *__restrict__ ops,
                                                   const cplx<double> *__restrict__ mats,
                                                   const int *__restrict__ aux, int cbit_mem, int words,
                                                   uint64_t *__restrict__ bits_out,
                                                   uint64_t *__restrict__ ticks_out) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ DevRng rng;
    __shared__ double bcast[2];
    __shared__ unsigned char cb[kMaxCbitMem];
    cplx<double> *sh = reinterpret_cast<cplx<double> *>(smem_raw);
    const uint32_t N = 1u << n, tid = threadIdx.x, nthr = blockDim.x;
    for (uint64_t shot = blockIdx.x; shot < shots; shot += gridDim.x) {
        const uint64_t t0 = wall_clock64();
        for (uint32_t i = tid; i < N; i += nthr) sh[i] = cplx<double>{i == 0 ? 1.0 : 0.0, 0.0};
        for (int i = int(tid); i < cbit_mem; i += int(nthr)) cb[i] = 0;
        if (tid == 0) rng_seed(rng, base_seed + shot);  // wrapping add, as base_seed.wrapping_add(i)
        __syncthreads();
        for (int pc = 0; pc < nprog; ++pc) {
            const ShotOp so = prog[pc];
            if (so.kind == SO_GATE) {
                apply_op<double, uint32_t, 3>(sh, n, ops[so.a], mats, aux, tid, nthr);
                __syncthreads();
            } else if (so.kind == SO_MEASURE) {
                const int o = shot_measure(sh, n, so.a, &rng, bcast, tid, nthr);
                if (tid == 0) cb[so.b] = (unsigned char)o;
                __syncthreads();
            } else if (so.kind == SO_RESET) {
                const int o = shot_measure(sh, n, so.a, &rng, bcast, tid, nthr);
                if (o == 1) {  // x on the qubit: 0 * a + 1 * b == b up to the sign of a zero
                    const uint32_t bit = 1u << so.a;
                    for (uint32_t g = tid; g < N / 2; g += nthr) {
                        const uint32_t i0 = ((g >> so.a) << (so.a + 1)) | (g & (bit - 1)), i1 = i0 | bit;
                        const cplx<double> x = sh[i0], y = sh[i1];
                        sh[i0] = cadd(cmul(cplx<double>{0.0, 0.0}, x), cmul(cplx<double>{1.0, 0.0}, y));
                        sh[i1] = cadd(cmul(cplx<double>{1.0, 0.0}, x), cmul(cplx<double>{0.0, 0.0}, y));
                    }
                    __syncthreads();
                }
            } else {  // SO_COND: every thread reads the same classical bits
                uint64_t actual = 0;
                for (int bb = 0; bb < so.b; ++bb) {
                    const int idx = so.a + bb;
                    const uint64_t v = idx < cbit_mem ? uint64_t(cb[idx]) : 0;
                    actual |= v << (bb & 63);
                }
                if (actual != so.value) pc += so.skip;
            }
        }
        if (tid < uint32_t(words)) {
            uint64_t w = 0;
            for (int bb = 0; bb < 64; ++bb) {
                const int idx = int(tid) * 64 + bb;
                if (idx < cbit_mem && cb[idx]) w |= uint64_t(1) << bb;
            }
            bits_out[shot * uint64_t(words) + tid] = w;
        }
        if (tid == 0 && ticks_out) ticks_out[shot] = wall_clock64() - t0;
        __syncthreads();
    }
}

}  // namespace

struct dq_shots {
    int n = 0;
    int num_cbits = 0;
    int cbit_mem = 0;  // classical bits the program touches (max index + 1, >= num_cbits)
    std::vector<ShotOp> prog;
    std::vector<HostOp> gates;  // SO_GATE payloads (op index into this vector)
    std::vector<std::vector<double>> raw;  // the gates' matrices as given (engine path)
    std::vector<std::vector<int>> raw_q;
    int open_conds = 0;
};

namespace {

int sfail(int code, const std::string &msg) { return set_error(code, msg); }

#define S_TRY(expr)                                                                                 \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess) return sfail(DQ_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

unsigned shot_threads(int n) { return std::min(512u, std::max(256u, (1u << n) / 16u)); }

int run_lds(dq_shots *S, uint64_t base_seed, uint64_t shots, uint64_t *bits, double *secs) {
    std::vector<DevOp> ops;
    std::vector<double> mats;
    std::vector<int> aux;
    for (const HostOp &h : S->gates) {
        DevOp d;
        std::memset(&d, 0, sizeof d);
        d.kind = h.kind;
        d.k = h.k;
        d.mat = int(mats.size() / 2);
        mats.insert(mats.end(), h.mat.begin(), h.mat.end());
        d.aux = int(aux.size());
        aux.insert(aux.end(), h.perm.begin(), h.perm.end());
        for (int j = 0; j < h.k; ++j) d.bit[j] = d.sbit[j] = h.q[j];
        std::sort(d.sbit, d.sbit + h.k);
        d.cmask = h.cmask;
        d.cval = h.cval;
        ops.push_back(d);
    }
    // 4-5 qubit dense/monomial gates would need the large register arrays
    // that this kernel leaves out: such programs take the engine path
    for (const DevOp &d : ops)
        if ((d.kind == K_DENSEK || d.kind == K_MONO) && d.k > 3) return 1;
    const int words = std::max(1, (S->cbit_mem + 63) / 64);
    auto align = [](size_t x) { return (x + 255) & ~size_t(255); };
    const size_t o_prog = 0, o_ops = align(std::max<size_t>(S->prog.size(), 1) * sizeof(ShotOp));
    const size_t o_mats = align(o_ops + std::max<size_t>(ops.size(), 1) * sizeof(DevOp));
    const size_t o_aux = align(o_mats + std::max<size_t>(mats.size(), 2) * sizeof(double));
    const size_t in_total = align(o_aux + std::max<size_t>(aux.size(), 1) * sizeof(int));
    std::vector<unsigned char> h(in_total, 0);
    if (!S->prog.empty()) std::memcpy(h.data() + o_prog, S->prog.data(), S->prog.size() * sizeof(ShotOp));
    if (!ops.empty()) std::memcpy(h.data() + o_ops, ops.data(), ops.size() * sizeof(DevOp));
    if (!mats.empty()) std::memcpy(h.data() + o_mats, mats.data(), mats.size() * sizeof(double));
    if (!aux.empty()) std::memcpy(h.data() + o_aux, aux.data(), aux.size() * sizeof(int));
    const size_t out_bytes = size_t(shots) * size_t(words) * sizeof(uint64_t);
    const size_t tick_bytes = secs ? size_t(shots) * sizeof(uint64_t) : 0;
    unsigned char *d = nullptr;
    S_TRY(hipMalloc(&d, in_total + out_bytes + tick_bytes + 256));
    struct Free {
        unsigned char *p;
        ~Free() { (void)hipFree(p); }
    } guard{d};
    hipStream_t stream;
    S_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    struct SFree {
        hipStream_t s;
        ~SFree() { (void)hipStreamDestroy(s); }
    } sguard{stream};
    S_TRY(hipMemcpyAsync(d, h.data(), in_total, hipMemcpyHostToDevice, stream));
    const size_t shmem = (size_t(1) << S->n) * sizeof(cplx<double>);
    S_TRY(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_shots_lds), hipFuncAttributeMaxDynamicSharedMemorySize,
                              int(std::max<size_t>(shmem, 1024))));
    const uint64_t grid = std::min<uint64_t>(shots, 65535);
    uint64_t *d_bits = reinterpret_cast<uint64_t *>(d + in_total);
    uint64_t *d_ticks = secs ? reinterpret_cast<uint64_t *>(d + in_total + out_bytes) : nullptr;
    hipLaunchKernelGGL(k_shots_lds, dim3(unsigned(grid)), dim3(shot_threads(S->n)), shmem, stream, S->n, shots,
                       base_seed, reinterpret_cast<const ShotOp *>(d + o_prog), int(S->prog.size()),
                       reinterpret_cast<const DevOp *>(d + o_ops), reinterpret_cast<const cplx<double> *>(d + o_mats),
                       reinterpret_cast<const int *>(d + o_aux), S->cbit_mem, words, d_bits, d_ticks);
    S_TRY(hipGetLastError());
    S_TRY(hipMemcpyAsync(bits, d_bits, out_bytes, hipMemcpyDeviceToHost, stream));
    std::vector<uint64_t> ticks(secs ? shots : 0);
    if (secs) S_TRY(hipMemcpyAsync(ticks.data(), d_ticks, tick_bytes, hipMemcpyDeviceToHost, stream));
    S_TRY(hipStreamSynchronize(stream));
    if (secs) {
        int dev = 0, khz = 0;
        S_TRY(hipGetDevice(&dev));
        S_TRY(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
        const double hz = khz > 0 ? double(khz) * 1e3 : 1e8;
        for (uint64_t i = 0; i < shots; ++i) secs[i] = double(ticks[i]) / hz;
    }
    return DQ_OK;
}

// Host ChaCha8 for the engine path (the same stream as DevRng).
struct HostRng {
    uint32_t key[8];
    uint64_t counter = 0;
    uint32_t buf[64];
    int idx = 64;
    explicit HostRng(uint64_t state) {
        for (int c = 0; c < 8; ++c) {
            state = state * 6364136223846793005ull + 11634580027462260723ull;
            const uint32_t xs = uint32_t(((state >> 18) ^ state) >> 27);
            const uint32_t rot = uint32_t(state >> 59);
            key[c] = (xs >> rot) | (xs << ((32 - rot) & 31));
        }
    }
    static uint32_t rl(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }
    void block(uint64_t ctr, uint32_t *out) const {
        uint32_t in[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u};
        for (int i = 0; i < 8; ++i) in[4 + i] = key[i];
        in[12] = uint32_t(ctr);
        in[13] = uint32_t(ctr >> 32);
        in[14] = in[15] = 0;
        uint32_t x[16];
        std::memcpy(x, in, sizeof x);
        auto qr = [&](int a, int b, int c, int d) {
            x[a] += x[b]; x[d] = rl(x[d] ^ x[a], 16);
            x[c] += x[d]; x[b] = rl(x[b] ^ x[c], 12);
            x[a] += x[b]; x[d] = rl(x[d] ^ x[a], 8);
            x[c] += x[d]; x[b] = rl(x[b] ^ x[c], 7);
        };
        for (int rd = 0; rd < 8; rd += 2) {
            qr(0, 4, 8, 12); qr(1, 5, 9, 13); qr(2, 6, 10, 14); qr(3, 7, 11, 15);
            qr(0, 5, 10, 15); qr(1, 6, 11, 12); qr(2, 7, 8, 13); qr(3, 4, 9, 14);
        }
        for (int i = 0; i < 16; ++i) out[i] = x[i] + in[i];
    }
    double f64() {
        if (idx >= 64) {
            for (int b = 0; b < 4; ++b) block(counter + uint64_t(b), buf + 16 * b);
            counter += 4;
            idx = 0;
        }
        const uint64_t v = uint64_t(buf[idx]) | (uint64_t(buf[idx + 1]) << 32);
        idx += 2;
        return double(v >> 11) * (1.0 / 9007199254740992.0);
    }
};

int run_engine(dq_shots *S, uint64_t base_seed, uint64_t shots, uint64_t *bits, double *secs) {
    dq_state *s = nullptr;
    if (int rc = dq_state_create(S->n, DQ_C128, &s)) return rc;
    struct Destroy {
        dq_state *s;
        ~Destroy() { dq_state_destroy(s); }
    } guard{s};
    if (int rc = dq_set_options(s, 1, 0, 0)) return rc;  // program order, exact arithmetic
    const int words = std::max(1, (S->cbit_mem + 63) / 64);
    std::vector<unsigned char> cb(size_t(std::max(S->cbit_mem, 1)));
    static const double X[8] = {0, 0, 1, 0, 1, 0, 0, 0};
    for (uint64_t shot = 0; shot < shots; ++shot) {
        const auto t0 = std::chrono::steady_clock::now();
        if (int rc = dq_state_reset(s)) return rc;
        std::fill(cb.begin(), cb.end(), 0);
        HostRng rng(base_seed + shot);
        auto measure = [&](int q, int &out) {
            double p1 = 0;
            if (int rc = dq_prob_one(s, q, &p1)) return rc;
            out = rng.f64() < p1 ? 1 : 0;
            const double norm = out ? std::max(std::sqrt(p1), 1e-15) : std::max(std::sqrt(std::max(1.0 - p1, 0.0)), 1e-15);
            return dq_collapse(s, q, out, norm);
        };
        for (size_t pc = 0; pc < S->prog.size(); ++pc) {
            const ShotOp &so = S->prog[pc];
            int rc = DQ_OK;
            if (so.kind == SO_GATE) {
                const auto &u = S->raw[size_t(so.a)];
                const auto &q = S->raw_q[size_t(so.a)];
                rc = q.size() == 1 ? dq_apply_1q(s, u.data(), q[0], nullptr, nullptr, 0)
                                   : dq_apply_nq(s, u.data(), q.data(), int(q.size()));
            } else if (so.kind == SO_MEASURE) {
                int o = 0;
                rc = measure(so.a, o);
                cb[size_t(so.b)] = (unsigned char)o;
            } else if (so.kind == SO_RESET) {
                int o = 0;
                rc = measure(so.a, o);
                if (rc == DQ_OK && o == 1) rc = dq_apply_1q(s, X, so.a, nullptr, nullptr, 0);
            } else {
                uint64_t actual = 0;
                for (int bb = 0; bb < so.b; ++bb) {
                    const int idx = so.a + bb;
                    actual |= uint64_t(idx < S->cbit_mem ? cb[size_t(idx)] : 0) << (bb & 63);
                }
                if (actual != so.value) pc += size_t(so.skip);
            }
            if (rc != DQ_OK) return rc;
        }
        if (int rc = dq_synchronize(s)) return rc;
        for (int w = 0; w < words; ++w) {
            uint64_t v = 0;
            for (int bb = 0; bb < 64; ++bb) {
                const int idx = w * 64 + bb;
                if (idx < S->cbit_mem && cb[size_t(idx)]) v |= uint64_t(1) << bb;
            }
            bits[shot * uint64_t(words) + uint64_t(w)] = v;
        }
        if (secs) secs[shot] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    return DQ_OK;
}

}  // namespace

extern "C" {

int dq_shots_create(int n_qubits, int num_cbits, dq_shots **out) {
    if (!out) return sfail(DQ_ERR_ARG, "null output pointer");
    *out = nullptr;
    if (n_qubits < 1 || n_qubits > 34) return sfail(DQ_ERR_ARG, "n_qubits must be in [1, 34]");
    if (num_cbits < 0) return sfail(DQ_ERR_ARG, "num_cbits must be >= 0");
    dq_shots *S = new dq_shots();
    S->n = n_qubits;
    S->num_cbits = num_cbits;
    S->cbit_mem = num_cbits;
    *out = S;
    return DQ_OK;
}

void dq_shots_destroy(dq_shots *S) { delete S; }

int dq_shots_add_gate(dq_shots *S, const double *u, const int *qubits, int k) {
    if (!S || !u || !qubits) return sfail(DQ_ERR_ARG, "null argument");
    if (k < 1 || k > 5) return sfail(DQ_ERR_ARG, "gate arity must be 1..5");
    uint64_t seen = 0;
    HostOp h;
    h.k = k;
    for (int j = 0; j < k; ++j) {
        if (qubits[j] < 0 || qubits[j] >= S->n) return sfail(DQ_ERR_ARG, "qubit index out of range");
        if ((seen >> qubits[j]) & 1) return sfail(DQ_ERR_ARG, "gate qubits must be distinct");
        seen |= uint64_t(1) << qubits[j];
        h.q[j] = qubits[j];
    }
    classify(h, u);
    ShotOp so{SO_GATE, int(S->gates.size()), 0, 0, 0};
    S->gates.push_back(std::move(h));
    S->raw.emplace_back(u, u + 2 * (size_t(1) << (2 * k)));
    S->raw_q.emplace_back(qubits, qubits + k);
    S->prog.push_back(so);
    return DQ_OK;
}

int dq_shots_add_measure(dq_shots *S, int qubit, int cbit) {
    if (!S) return sfail(DQ_ERR_ARG, "null argument");
    if (qubit < 0 || qubit >= S->n) return sfail(DQ_ERR_ARG, "qubit index out of range");
    if (cbit < 0) return sfail(DQ_ERR_ARG, "cbit must be >= 0");
    S->cbit_mem = std::max(S->cbit_mem, cbit + 1);
    S->prog.push_back(ShotOp{SO_MEASURE, qubit, cbit, 0, 0});
    return DQ_OK;
}

int dq_shots_add_reset(dq_shots *S, int qubit) {
    if (!S) return sfail(DQ_ERR_ARG, "null argument");
    if (qubit < 0 || qubit >= S->n) return sfail(DQ_ERR_ARG, "qubit index out of range");
    S->prog.push_back(ShotOp{SO_RESET, qubit, 0, 0, 0});
    return DQ_OK;
}

int dq_shots_add_conditional(dq_shots *S, int creg_base, int creg_size, uint64_t creg_value, int n_ops) {
    if (!S) return sfail(DQ_ERR_ARG, "null argument");
    if (creg_base < 0 || creg_size < 0 || n_ops < 0) return sfail(DQ_ERR_ARG, "bad conditional");
    S->prog.push_back(ShotOp{SO_COND, creg_base, creg_size, n_ops, creg_value});
    return DQ_OK;
}

int dq_shots_run(dq_shots *S, uint64_t base_seed, uint64_t shots, uint64_t *bits_out, double *shot_seconds) {
    if (!S || (!bits_out && shots)) return sfail(DQ_ERR_ARG, "null argument");
    for (size_t pc = 0; pc < S->prog.size(); ++pc)
        if (S->prog[pc].kind == SO_COND && pc + size_t(S->prog[pc].skip) >= S->prog.size())
            return sfail(DQ_ERR_ARG, "conditional extends past the end of the program");
    if (shots == 0) return DQ_OK;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return sfail(DQ_ERR_DEVICE, "no HIP device available");
    if (S->n <= kShotLdsQubits && S->cbit_mem <= kMaxCbitMem) {
        const int rc = run_lds(S, base_seed, shots, bits_out, shot_seconds);
        if (rc <= 0) return rc;  // 1: program needs the engine path
    }
    return run_engine(S, base_seed, shots, bits_out, shot_seconds);
}

int dq_shots_words(const dq_shots *S) { return S ? std::max(1, (S->cbit_mem + 63) / 64) : 0; }

}  // extern "C"
