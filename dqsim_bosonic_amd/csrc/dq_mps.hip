// dq_mps.hip -- matrix-product-state contraction on the GPU: the dense
// statevector of Mps::to_statevector and the p1 of Mps::measure
// (reference: src/monolithic/mps/simulator.rs:246-279).
//
// The reference contracts, for every basis index x separately, the row
// vector work = [1]; work = work * A_q[bit q of x] over q = 0..n-1, which
// costs O(2^n n chi^2).  Here the chain is cut once, at bond m:
//   L[c][x_lo] = (A_0[b_0] ... A_{m-1}[b_{m-1}])[0][c]      x_lo = bits 0..m-1
//   R[x_hi][c] = (A_m[b_m] ... A_{n-1}[b_{n-1}])[c][0]      x_hi = bits m..n-1
//   psi[x_hi 2^m + x_lo] = sum_c L[c][x_lo] R[x_hi][c]
// L grows one site at a time from the left (2^k chi_k entries after k sites),
// R from the right, and the join is a complex GEMM [2^(n-m) x chi] x
// [chi x 2^m] written straight into the statevector: 16 bytes per amplitude
// (HBM-bound for small chi) and 4 chi FP64 FMAs per amplitude.
// The probability that a qubit reads 1 runs the same join over the half of
// the index space where that bit is set and sums |psi|^2 without writing it.
// Compiled with contraction (the join's FMAs); the sum over c runs in a
// different association than the reference's left-to-right chain, so values
// agree to rounding, not bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/dqsim_gpu.h"
#include "dq_device.hpp"
#include "dq_state.hpp"

using namespace dq;

namespace {

using C = cplx<double>;

constexpr int kThreads = 256;
constexpr int kColReps = 4;                  // join micro-tile: columns (x_lo) per thread
constexpr int kRowReps = 4;                  // join micro-tile: rows (x_hi) per thread
constexpr int kTileCols = 16 * kColReps;     // join tile, x_lo
constexpr int kTileRows = 16 * kRowReps;     // join tile, x_hi
constexpr int kChunk = 8;                    // bond entries staged in LDS per step of the join
constexpr uint64_t kMaxBond = uint64_t(1) << 16;
constexpr int kMaxSites = 34;
constexpr uint64_t kMaxWorkBytes = uint64_t(16) << 30;  // L and R growth buffers

__device__ __forceinline__ C cmac(C acc, C a, C b) {
    acc.re = fma(a.re, b.re, acc.re);
    acc.re = fma(-a.im, b.im, acc.re);
    acc.im = fma(a.re, b.im, acc.im);
    acc.im = fma(a.im, b.re, acc.im);
    return acc;
}

// x with a 1 inserted at position `bit` (bit < 0: x unchanged).
__device__ __forceinline__ uint64_t with_one(uint64_t x, int bit) {
    if (bit < 0) return x;
    const uint64_t lo = x & ((uint64_t(1) << bit) - 1);
    return ((x >> bit) << (bit + 1)) | (uint64_t(1) << bit) | lo;
}

__device__ double block_sum(double v, double *scratch) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) scratch[wave] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < int(blockDim.x >> 6); ++w) s += scratch[w];
    return s;
}

__global__ void k_set_one(C *p) {
    if (threadIdx.x == 0 && blockIdx.x == 0) p[0] = C{1.0, 0.0};
}

// One site from the left: L'[j][x + b 2^k] = sum_c L[c][x] A[c][b][j], x < 2^k.
// Site tensor layout (left, state, right), row-major, as Tensor::idx of the
// reference (simulator.rs:49-52).
__global__ void k_grow_left(const C *__restrict__ in, C *__restrict__ out, const C *__restrict__ A, int k,
                            uint32_t ci, uint32_t co) {
    const uint64_t w = uint64_t(1) << k;
    const uint64_t total = 2 * w * co;
    for (uint64_t t = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; t < total;
         t += uint64_t(gridDim.x) * blockDim.x) {
        const uint64_t xp = t & (2 * w - 1);
        const uint64_t j = t >> (k + 1);
        const uint64_t x = xp & (w - 1), b = xp >> k;
        C acc{0.0, 0.0};
        for (uint32_t c = 0; c < ci; ++c) acc = cmac(acc, in[uint64_t(c) * w + x], A[(uint64_t(c) * 2 + b) * co + j]);
        out[t] = acc;  // = out[j * 2w + xp]
    }
}

// One site from the right: R'[2x + b][i] = sum_j A[i][b][j] R[x][j], x < w.
__global__ void k_grow_right(const C *__restrict__ in, C *__restrict__ out, const C *__restrict__ A, uint64_t w,
                             uint32_t ci, uint32_t co) {
    const uint64_t total = 2 * w * ci;
    for (uint64_t t = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; t < total;
         t += uint64_t(gridDim.x) * blockDim.x) {
        const uint64_t i = t % ci, xp = t / ci;
        const uint64_t x = xp >> 1, b = xp & 1;
        const C *__restrict__ a = A + (i * 2 + b) * co;
        const C *__restrict__ r = in + x * co;
        C acc{0.0, 0.0};
        for (uint32_t j = 0; j < co; ++j) acc = cmac(acc, a[j], r[j]);
        out[t] = acc;  // = out[xp * ci + i]
    }
}

// Statevector stores of the join are non-temporal: the 2^n amplitudes are
// written once and not read back by this kernel, and keeping them out of the
// caches keeps L and R there.
template <typename TO>
__device__ __forceinline__ void store_amp(cplx<TO> *psi, uint64_t i, C v) {
    typedef TO vec2 __attribute__((ext_vector_type(2)));
    vec2 x;
    x.x = TO(v.re);
    x.y = TO(v.im);
    __builtin_nontemporal_store(x, reinterpret_cast<vec2 *>(psi + i));
}

// Tiled join.  Workgroup = one tile of kTileRows rows (x_hi) by kTileCols
// columns (x_lo); thread (tx, ty) owns rows ty + 16a and columns tx + 16b.
// Rows / columns are effective indices: with fix_lo (fix_hi) >= 0 the x_lo
// (x_hi) of column (row) e is e with a 1 inserted at that bit.
// PROB = false: writes psi[(x_hi << m) | x_lo].  PROB = true: writes the
// tile's sum of |psi|^2 to partials[blockIdx.x].
template <typename TO, bool PROB>
__global__ __launch_bounds__(kThreads) void k_join(const C *__restrict__ L, const C *__restrict__ R,
                                                   cplx<TO> *__restrict__ psi, uint32_t chi, int m,
                                                   uint64_t tiles_lo, int fix_lo, int fix_hi,
                                                   double *__restrict__ partials) {
    __shared__ C Ls[kChunk][kTileCols];
    __shared__ C Rs[kTileRows][kChunk];
    __shared__ double scratch[kThreads / 64];
    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
    const uint64_t tl = blockIdx.x % tiles_lo, th = blockIdx.x / tiles_lo;
    const uint64_t col0 = tl * kTileCols, row0 = th * kTileRows;
    const uint64_t wl = uint64_t(1) << m;
    C acc[kRowReps][kColReps];
#pragma unroll
    for (int a = 0; a < kRowReps; ++a)
#pragma unroll
        for (int b = 0; b < kColReps; ++b) acc[a][b] = C{0.0, 0.0};
    for (uint32_t c0 = 0; c0 < chi; c0 += kChunk) {
        for (int e = tid; e < kChunk * kTileCols; e += kThreads) {
            const int cc = e / kTileCols, col = e % kTileCols;
            const uint32_t c = c0 + cc;
            Ls[cc][col] = c < chi ? L[uint64_t(c) * wl + with_one(col0 + col, fix_lo)] : C{0.0, 0.0};
        }
        for (int e = tid; e < kChunk * kTileRows; e += kThreads) {
            const int row = e / kChunk, cc = e % kChunk;
            const uint32_t c = c0 + cc;
            Rs[row][cc] = c < chi ? R[with_one(row0 + row, fix_hi) * chi + c] : C{0.0, 0.0};
        }
        __syncthreads();
        const int kc = int(min(uint32_t(kChunk), chi - c0));
        for (int cc = 0; cc < kc; ++cc) {
            C l[kColReps], r[kRowReps];
#pragma unroll
            for (int b = 0; b < kColReps; ++b) l[b] = Ls[cc][tx + 16 * b];
#pragma unroll
            for (int a = 0; a < kRowReps; ++a) r[a] = Rs[ty + 16 * a][cc];
#pragma unroll
            for (int a = 0; a < kRowReps; ++a)
#pragma unroll
                for (int b = 0; b < kColReps; ++b) acc[a][b] = cmac(acc[a][b], l[b], r[a]);
        }
        __syncthreads();
    }
    if constexpr (!PROB) {
#pragma unroll
        for (int a = 0; a < kRowReps; ++a) {
            const uint64_t xh = row0 + ty + 16 * a;
#pragma unroll
            for (int b = 0; b < kColReps; ++b) store_amp<TO>(psi, (xh << m) | (col0 + tx + 16 * b), acc[a][b]);
        }
    } else {
        double s = 0.0;
#pragma unroll
        for (int a = 0; a < kRowReps; ++a)
#pragma unroll
            for (int b = 0; b < kColReps; ++b) s += acc[a][b].re * acc[a][b].re + acc[a][b].im * acc[a][b].im;
        const double v = block_sum(s, scratch);
        if (tid == 0) partials[blockIdx.x] = v;
    }
}

// Join for index spaces too small to tile (fewer rows or columns than a tile):
// one thread per amplitude, grid-stride.
template <typename TO, bool PROB>
__global__ __launch_bounds__(kThreads) void k_join_small(const C *__restrict__ L, const C *__restrict__ R,
                                                         cplx<TO> *__restrict__ psi, uint32_t chi, int m,
                                                         uint64_t ncols, uint64_t nrows, int fix_lo, int fix_hi,
                                                         double *__restrict__ partials) {
    __shared__ double scratch[kThreads / 64];
    const uint64_t wl = uint64_t(1) << m;
    double s = 0.0;
    for (uint64_t t = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; t < ncols * nrows;
         t += uint64_t(gridDim.x) * blockDim.x) {
        const uint64_t xl = with_one(t % ncols, fix_lo), xh = with_one(t / ncols, fix_hi);
        C acc{0.0, 0.0};
        for (uint32_t c = 0; c < chi; ++c) acc = cmac(acc, L[uint64_t(c) * wl + xl], R[xh * chi + c]);
        if constexpr (!PROB)
            store_amp<TO>(psi, (xh << m) | xl, acc);
        else
            s += acc.re * acc.re + acc.im * acc.im;
    }
    if constexpr (PROB) {
        const double v = block_sum(s, scratch);
        if (threadIdx.x == 0) partials[blockIdx.x] = v;
    }
}

// Sum of `count` partials in a fixed order (deterministic).
__global__ __launch_bounds__(kThreads) void k_sum_partials(const double *__restrict__ partials, uint64_t count,
                                                           double *__restrict__ out) {
    __shared__ double scratch[kThreads / 64];
    double s = 0.0;
    for (uint64_t i = threadIdx.x; i < count; i += kThreads) s += partials[i];
    const double v = block_sum(s, scratch);
    if (threadIdx.x == 0) out[0] = v;
}

int mfail(int code, const std::string &msg) { return dq::set_error(code, "mps: " + msg); }

#define MPS_TRY(expr)                                                                      \
    do {                                                                                   \
        hipError_t _e = (expr);                                                            \
        if (_e != hipSuccess) return mfail(DQ_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

unsigned grid_of(uint64_t work) {
    uint64_t g = (work + kThreads - 1) / kThreads;
    return unsigned(std::max<uint64_t>(1, std::min<uint64_t>(g, 65536)));
}

}  // namespace

struct dq_mps {
    int n = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<uint64_t> bond;  // n + 1 bond dimensions, bond[0] = bond[n] = 1
    std::vector<uint64_t> off;   // complex offset of each site tensor in d_sites
    C *d_sites = nullptr;
    uint64_t sites_cap = 0;      // complex entries
    double *h_stage = nullptr;   // pinned upload staging
    uint64_t stage_cap = 0;      // bytes
    C *d_work = nullptr;
    uint64_t work_cap = 0;       // complex entries
    double *d_partials = nullptr;
    uint64_t part_cap = 0;       // doubles
    double *d_out = nullptr;
    double *h_out = nullptr;     // pinned
    int loaded = 0;
    int cut_req = -1;            // requested cut bond (-1: chosen from the bond dimensions)
    int cut = -1;                // cut of the grown halves; -1 when L and R are stale
    C *L = nullptr, *R = nullptr;
};

namespace {

// Cut bond: minimises the join (2^n chi_m) plus the growth of both halves,
// with both halves within the workspace limit and, when n >= 16, at least
// two tiles of rows and of columns (so that the half-space join of p1 still
// tiles).
int choose_cut(const dq_mps *M) {
    const int n = M->n;
    if (M->cut_req >= 0) return std::min(M->cut_req, n);
    double best = 0.0;
    int bm = -1;
    for (int m = 0; m <= n; ++m) {
        if (n >= 16 && ((uint64_t(1) << m) < 2 * uint64_t(kTileCols) ||
                        (uint64_t(1) << (n - m)) < 2 * uint64_t(kTileRows)))
            continue;
        double cost = std::ldexp(double(M->bond[m]), n), memL = 0.0, memR = 0.0;
        for (int k = 0; k < m; ++k) {
            cost += std::ldexp(double(M->bond[k]) * double(M->bond[k + 1]), k + 1);
            memL = std::max(memL, std::ldexp(double(M->bond[k + 1]), k + 1));
        }
        for (int k = n - 1; k >= m; --k) {
            cost += std::ldexp(double(M->bond[k]) * double(M->bond[k + 1]), n - k);
            memR = std::max(memR, std::ldexp(double(M->bond[k]), n - k));
        }
        if (32.0 * (memL + memR + 2.0) > double(kMaxWorkBytes)) continue;
        if (bm < 0 || cost < best) {
            best = cost;
            bm = m;
        }
    }
    return bm < 0 ? n / 2 : bm;
}

int ensure_work(dq_mps *M, uint64_t entries) {
    if (entries <= M->work_cap) return DQ_OK;
    if (M->d_work) (void)hipFree(M->d_work);
    M->d_work = nullptr;
    M->work_cap = 0;
    if (hipMalloc(reinterpret_cast<void **>(&M->d_work), entries * sizeof(C)) != hipSuccess) {
        M->d_work = nullptr;
        return mfail(DQ_ERR_OOM, "hipMalloc of " + std::to_string(entries * sizeof(C)) + " bytes of workspace failed");
    }
    M->work_cap = entries;
    return DQ_OK;
}

int ensure_partials(dq_mps *M, uint64_t count) {
    if (count <= M->part_cap) return DQ_OK;
    if (M->d_partials) (void)hipFree(M->d_partials);
    M->d_partials = nullptr;
    M->part_cap = 0;
    if (hipMalloc(reinterpret_cast<void **>(&M->d_partials), count * sizeof(double)) != hipSuccess) {
        M->d_partials = nullptr;
        return mfail(DQ_ERR_OOM, "hipMalloc of partial sums failed");
    }
    M->part_cap = count;
    return DQ_OK;
}

// Grows L (sites 0..m-1) and R (sites m..n-1) for the current tensors.
int grow(dq_mps *M) {
    const int n = M->n, m = choose_cut(M);
    if (M->cut == m) return DQ_OK;
    uint64_t maxL = 1, maxR = 1;
    for (int k = 0; k < m; ++k) maxL = std::max(maxL, M->bond[k + 1] << (k + 1));
    for (int k = n - 1; k >= m; --k) maxR = std::max(maxR, M->bond[k] << (n - k));
    if ((2 * (maxL + maxR)) * sizeof(C) > kMaxWorkBytes)
        return mfail(DQ_ERR_ARG, "bond dimensions too large for the contraction workspace");
    if (int rc = ensure_work(M, 2 * (maxL + maxR))) return rc;
    C *la = M->d_work, *lb = la + maxL, *ra = lb + maxL, *rb = ra + maxR;
    hipLaunchKernelGGL(k_set_one, dim3(1), dim3(64), 0, M->stream, la);
    hipLaunchKernelGGL(k_set_one, dim3(1), dim3(64), 0, M->stream, ra);
    for (int k = 0; k < m; ++k) {
        const uint64_t total = (M->bond[k + 1] << (k + 1));
        hipLaunchKernelGGL(k_grow_left, dim3(grid_of(total)), dim3(kThreads), 0, M->stream, la, lb,
                           M->d_sites + M->off[k], k, uint32_t(M->bond[k]), uint32_t(M->bond[k + 1]));
        std::swap(la, lb);
    }
    for (int k = n - 1; k >= m; --k) {
        const uint64_t w = uint64_t(1) << (n - 1 - k);
        hipLaunchKernelGGL(k_grow_right, dim3(grid_of(2 * w * M->bond[k])), dim3(kThreads), 0, M->stream, ra, rb,
                           M->d_sites + M->off[k], w, uint32_t(M->bond[k]), uint32_t(M->bond[k + 1]));
        std::swap(ra, rb);
    }
    MPS_TRY(hipGetLastError());
    M->L = la;
    M->R = ra;
    M->cut = m;
    return DQ_OK;
}

template <typename TO, bool PROB>
int join(dq_mps *M, cplx<TO> *psi, int fix_lo, int fix_hi, uint64_t *nblocks) {
    const int n = M->n, m = M->cut;
    const uint32_t chi = uint32_t(M->bond[m]);
    const uint64_t ncols = (uint64_t(1) << m) >> (fix_lo >= 0 ? 1 : 0);
    const uint64_t nrows = (uint64_t(1) << (n - m)) >> (fix_hi >= 0 ? 1 : 0);
    if (ncols >= uint64_t(kTileCols) && nrows >= uint64_t(kTileRows)) {
        const uint64_t tiles_lo = ncols / kTileCols, tiles = tiles_lo * (nrows / kTileRows);
        if (PROB)
            if (int rc = ensure_partials(M, tiles)) return rc;
        hipLaunchKernelGGL((k_join<TO, PROB>), dim3(unsigned(tiles)), dim3(kThreads), 0, M->stream, M->L, M->R,
                           psi, chi, m, tiles_lo, fix_lo, fix_hi, M->d_partials);
        *nblocks = tiles;
    } else {
        const unsigned g = std::min(grid_of(ncols * nrows), 1024u);
        if (PROB)
            if (int rc = ensure_partials(M, g)) return rc;
        hipLaunchKernelGGL((k_join_small<TO, PROB>), dim3(g), dim3(kThreads), 0, M->stream, M->L, M->R, psi, chi,
                           m, ncols, nrows, fix_lo, fix_hi, M->d_partials);
        *nblocks = g;
    }
    MPS_TRY(hipGetLastError());
    return DQ_OK;
}

}  // namespace

extern "C" {

int dq_mps_create(int n_sites, dq_mps **out) {
    if (!out) return mfail(DQ_ERR_ARG, "null output pointer");
    *out = nullptr;
    if (n_sites < 1 || n_sites > kMaxSites)
        return mfail(DQ_ERR_ARG, "n_sites must be in [1, " + std::to_string(kMaxSites) + "]");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return mfail(DQ_ERR_DEVICE, "no HIP device available");
    dq_mps *M = new dq_mps();
    M->n = n_sites;
    auto cleanup = [&](int code, const std::string &msg) {
        dq_mps_destroy(M);
        return mfail(code, msg);
    };
    if (hipGetDevice(&M->device) != hipSuccess) return cleanup(DQ_ERR_DEVICE, "hipGetDevice failed");
    if (hipStreamCreateWithFlags(&M->stream, hipStreamNonBlocking) != hipSuccess)
        return cleanup(DQ_ERR_DEVICE, "hipStreamCreate failed");
    if (hipMalloc(reinterpret_cast<void **>(&M->d_out), sizeof(double)) != hipSuccess)
        return cleanup(DQ_ERR_OOM, "hipMalloc failed");
    if (hipHostMalloc(reinterpret_cast<void **>(&M->h_out), sizeof(double), hipHostMallocDefault) != hipSuccess)
        return cleanup(DQ_ERR_OOM, "hipHostMalloc failed");
    *out = M;
    return DQ_OK;
}

void dq_mps_destroy(dq_mps *M) {
    if (!M) return;
    if (M->stream) (void)hipStreamSynchronize(M->stream);
    if (M->d_sites) (void)hipFree(M->d_sites);
    if (M->h_stage) (void)hipHostFree(M->h_stage);
    if (M->d_work) (void)hipFree(M->d_work);
    if (M->d_partials) (void)hipFree(M->d_partials);
    if (M->d_out) (void)hipFree(M->d_out);
    if (M->h_out) (void)hipHostFree(M->h_out);
    if (M->stream) (void)hipStreamDestroy(M->stream);
    delete M;
}

int dq_mps_load(dq_mps *M, const uint64_t *bond, const double *tensors) {
    if (!M || !bond || !tensors) return mfail(DQ_ERR_ARG, "null argument");
    const int n = M->n;
    if (bond[0] != 1 || bond[n] != 1) return mfail(DQ_ERR_ARG, "the outer bond dimensions must be 1");
    std::vector<uint64_t> off(n + 1, 0);
    for (int q = 0; q < n; ++q) {
        if (bond[q + 1] < 1 || bond[q + 1] > kMaxBond)
            return mfail(DQ_ERR_ARG, "bond dimension out of range [1, 65536] at bond " + std::to_string(q + 1));
        off[q + 1] = off[q] + 2 * bond[q] * bond[q + 1];
    }
    const uint64_t entries = off[n], bytes = entries * sizeof(C);
    MPS_TRY(hipStreamSynchronize(M->stream));
    if (entries > M->sites_cap) {
        if (M->d_sites) (void)hipFree(M->d_sites);
        M->d_sites = nullptr;
        M->sites_cap = 0;
        if (hipMalloc(reinterpret_cast<void **>(&M->d_sites), bytes) != hipSuccess) {
            M->d_sites = nullptr;
            return mfail(DQ_ERR_OOM, "hipMalloc of site tensors failed");
        }
        M->sites_cap = entries;
    }
    if (bytes > M->stage_cap) {
        if (M->h_stage) (void)hipHostFree(M->h_stage);
        M->h_stage = nullptr;
        M->stage_cap = 0;
        if (hipHostMalloc(reinterpret_cast<void **>(&M->h_stage), bytes, hipHostMallocDefault) != hipSuccess) {
            M->h_stage = nullptr;
            return mfail(DQ_ERR_OOM, "hipHostMalloc of the upload staging failed");
        }
        M->stage_cap = bytes;
    }
    std::copy(tensors, tensors + 2 * entries, M->h_stage);
    MPS_TRY(hipMemcpyAsync(M->d_sites, M->h_stage, bytes, hipMemcpyHostToDevice, M->stream));
    MPS_TRY(hipStreamSynchronize(M->stream));
    M->bond.assign(bond, bond + n + 1);
    M->off = std::move(off);
    M->loaded = 1;
    M->cut = -1;
    return DQ_OK;
}

int dq_mps_set_cut(dq_mps *M, int cut) {
    if (!M) return mfail(DQ_ERR_ARG, "null argument");
    if (cut < -1 || cut > M->n) return mfail(DQ_ERR_ARG, "cut must be -1 (automatic) or in [0, n_sites]");
    M->cut_req = cut;
    M->cut = -1;
    return DQ_OK;
}

int dq_mps_get_cut(dq_mps *M) { return M && M->loaded ? choose_cut(M) : -1; }

int dq_mps_to_state(dq_mps *M, dq_state *dst) {
    if (!M || !dst) return mfail(DQ_ERR_ARG, "null argument");
    if (!M->loaded) return mfail(DQ_ERR_ARG, "no tensors loaded");
    if (dst->n != M->n) return mfail(DQ_ERR_ARG, "state and MPS differ in qubit count");
    if (dst->device != M->device) return mfail(DQ_ERR_ARG, "state and MPS live on different devices");
    dst->queue.clear();
    dst->zero_q = 0;
    MPS_TRY(hipStreamSynchronize(dst->stream));
    if (int rc = grow(M)) return rc;
    uint64_t nb = 0;
    int rc = dst->dtype == DQ_C128
                 ? join<double, false>(M, static_cast<cplx<double> *>(dst->amps), -1, -1, &nb)
                 : join<float, false>(M, static_cast<cplx<float> *>(dst->amps), -1, -1, &nb);
    if (rc) return rc;
    MPS_TRY(hipStreamSynchronize(M->stream));
    return DQ_OK;
}

int dq_mps_prob_one(dq_mps *M, int qubit, double *p1) {
    if (!M || !p1) return mfail(DQ_ERR_ARG, "null argument");
    if (!M->loaded) return mfail(DQ_ERR_ARG, "no tensors loaded");
    if (qubit < 0 || qubit >= M->n) return mfail(DQ_ERR_ARG, "qubit index out of range");
    if (int rc = grow(M)) return rc;
    const int m = M->cut;
    uint64_t nb = 0;
    if (int rc = join<double, true>(M, nullptr, qubit < m ? qubit : -1, qubit < m ? -1 : qubit - m, &nb)) return rc;
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kThreads), 0, M->stream, M->d_partials, nb, M->d_out);
    MPS_TRY(hipGetLastError());
    MPS_TRY(hipMemcpyAsync(M->h_out, M->d_out, sizeof(double), hipMemcpyDeviceToHost, M->stream));
    MPS_TRY(hipStreamSynchronize(M->stream));
    *p1 = *M->h_out;
    return DQ_OK;
}

}  // extern "C"
