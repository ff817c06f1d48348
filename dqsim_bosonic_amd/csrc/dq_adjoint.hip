// dq_adjoint.hip -- the reverse sweep of the adjoint-gradient method as tile
// passes over both states (no reference counterpart: the adjoint form of
// the statevector path's gradient workload).
//
// For l = L-1 .. 0 the method needs g_l = Im <lam| G_l |psi> and then
// psi <- U_l psi, lam <- U_l lam (U_l the inverse of gate l).  Every factor
// acts on gate l's qubits only, so a run of steps whose qubits fit a set S of
// b qubits can be applied tile by tile: a workgroup loads the 2^b-amplitude
// tile of psi and of lam into LDS, runs the steps in order -- gate steps
// rewrite both tiles, gradient steps sum conj(lam) (G psi) over the tile into
// a per-tile partial -- and stores both tiles.  One pass reads and writes
// each state once, instead of one flush sweep per state plus one inner-product
// sweep per gradient.  Steps are taken out of order only across steps on
// disjoint qubits, which commute (and leave <lam|G|psi> unchanged, G acting
// on the step's qubits only).  Compiled with contraction (FMA): the gradients
// agree with the per-parameter inner products to rounding.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/dqsim_gpu.h"
#include "dq_device.hpp"
#include "dq_plan.hpp"
#include "dq_state.hpp"

using namespace dq;

namespace {

constexpr int kThreads = 256;       // reduction kernel
constexpr int kPassThreads = 256;   // pass kernel workgroup
constexpr int kMaxTileBits = 12;
constexpr int kLowBits = 5;  // qubits 0..4 in every tile: loads in 512-byte runs (complex128)
constexpr int kRegBits = 3;  // register-resident pass: amplitudes per thread and state = 2^kRegBits

enum AdjKind : int {
    A_GENERIC = 0,  // any gate (k <= 5): one output per amplitude through a barrier
    A_GRAD = 1,     // gradient partial of the tile
    A_DENSE1 = 2,   // 2x2 on bit[0], in place per pair
    A_DENSE2 = 3,   // 4x4 on bit[0] (MSB), bit[1], in place per quad
    A_DIAG = 4      // diagonal (k <= 2): elementwise phases
};

struct AdjStep {
    int kind;      // AdjKind
    int k;         // gate qubits
    int bit[5];    // tile bits, bit[0] = MSB of the gate index
    int mat;       // gate: offset (complex entries) into the matrix pool
    int term0;     // gradient: first term
    int nterms;    // gradient: term count
    int slot;      // gradient: partial-sum slot within the pass
};

struct AdjTerm {
    unsigned x, z;     // tile-local masks
    double cre, cim;   // coefficient times i^{#Y}
};

struct AdjPass {
    int b;           // tile bits
    int n_outer;     // state qubits outside the tile
    int tq[kMaxTileBits];  // tile bit -> qubit
    int oq[64];            // outer bit -> qubit
    int step0, nsteps;
};

template <typename T>
__device__ __forceinline__ cplx<T> cvt(cplx<double> u) {
    return cplx<T>{T(u.re), T(u.im)};
}

template <typename T>
__device__ __forceinline__ double block_sum2(double v, double *scratch) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) scratch[wave] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < kPassThreads / 64; ++w) s += scratch[w];
    return s;
}

// One pass over every tile.  partials[slot * ntiles + tile] receives the
// tile's Im sum of each gradient step.
template <typename T, int B>
__global__ __launch_bounds__(kPassThreads) void k_adjoint_pass(cplx<T> *__restrict__ psi, cplx<T> *__restrict__ lam,
                                                           AdjPass pd, const AdjStep *__restrict__ steps,
                                                           const cplx<double> *__restrict__ mats,
                                                           const AdjTerm *__restrict__ terms,
                                                           double *__restrict__ partials, uint64_t ntiles) {
    extern __shared__ unsigned char smem_raw[];
    constexpr int kAmps = 1 << B;
    constexpr int kPer = kAmps >= kPassThreads ? kAmps / kPassThreads : 1;  // amplitudes per thread (>= 1)
    cplx<T> *P = reinterpret_cast<cplx<T> *>(smem_raw);
    cplx<T> *Lm = P + kAmps;
    double *scratch = reinterpret_cast<double *>(Lm + kAmps);
    const int t = threadIdx.x;
    for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        uint64_t base = 0;
        for (int o = 0; o < pd.n_outer; ++o) base |= ((tile >> o) & 1) << pd.oq[o];
        auto gidx = [&](unsigned i) {
            uint64_t g = base;
#pragma unroll
            for (int j = 0; j < B; ++j) g |= uint64_t((i >> j) & 1) << pd.tq[j];
            return g;
        };
        __syncthreads();  // the previous tile's stores have read the LDS
#pragma unroll
        for (int r = 0; r < kPer; ++r) {
            const unsigned i = unsigned(t + r * kPassThreads);
            if constexpr (kAmps < kPassThreads) { if (i >= unsigned(kAmps)) break; }
            const uint64_t g = gidx(i);
            P[i] = psi[g];
            Lm[i] = lam[g];
        }
        __syncthreads();
        for (int s = 0; s < pd.nsteps; ++s) {
            const AdjStep st = steps[pd.step0 + s];
            if (st.kind == A_GRAD) {
                const int nt = st.nterms;
                double acc = 0.0;
                if (nt <= 4) {
                    // up to four terms held in registers (every generator here)
                    AdjTerm t0{0, 0, 0.0, 0.0}, t1 = t0, t2 = t0, t3 = t0;
                    if (nt > 0) t0 = terms[st.term0];
                    if (nt > 1) t1 = terms[st.term0 + 1];
                    if (nt > 2) t2 = terms[st.term0 + 2];
                    if (nt > 3) t3 = terms[st.term0 + 3];
#pragma unroll
                    for (int r = 0; r < kPer; ++r) {
                        const unsigned i = unsigned(t + r * kPassThreads);
                        if constexpr (kAmps < kPassThreads) { if (i >= unsigned(kAmps)) break; }
                        double gr = 0.0, gi = 0.0;
                        auto term = [&](const AdjTerm &tq) {
                            const unsigned y = i ^ tq.x;
                            const cplx<T> v = P[y];
                            const double sg = (__popc(tq.z & y) & 1) ? -1.0 : 1.0;
                            const double vr = sg * double(v.re), vi = sg * double(v.im);
                            gr += tq.cre * vr - tq.cim * vi;
                            gi += tq.cre * vi + tq.cim * vr;
                        };
                        if (nt > 0) term(t0);
                        if (nt > 1) term(t1);
                        if (nt > 2) term(t2);
                        if (nt > 3) term(t3);
                        const double lr = double(Lm[i].re), li = double(Lm[i].im);
                        acc += lr * gi - li * gr;  // Im(conj(lam) * g)
                    }
                } else {
                    for (int r = 0; r < kPer; ++r) {
                        const unsigned i = unsigned(t + r * kPassThreads);
                        if constexpr (kAmps < kPassThreads) { if (i >= unsigned(kAmps)) break; }
                        double gr = 0.0, gi = 0.0;
                        for (int q = 0; q < nt; ++q) {
                            const AdjTerm tq = terms[st.term0 + q];
                            const unsigned y = i ^ tq.x;
                            const cplx<T> v = P[y];
                            const double sg = (__popc(tq.z & y) & 1) ? -1.0 : 1.0;
                            const double vr = sg * double(v.re), vi = sg * double(v.im);
                            gr += tq.cre * vr - tq.cim * vi;
                            gi += tq.cre * vi + tq.cim * vr;
                        }
                        const double lr = double(Lm[i].re), li = double(Lm[i].im);
                        acc += lr * gi - li * gr;
                    }
                }
                const double v = block_sum2<T>(acc, scratch);
                if (t == 0) partials[uint64_t(st.slot) * ntiles + tile] = v;
                continue;
            }
            if (st.kind == A_DENSE1) {
                const int pb = st.bit[0];
                const cplx<T> u0 = cvt<T>(mats[st.mat]), u1 = cvt<T>(mats[st.mat + 1]);
                const cplx<T> u2 = cvt<T>(mats[st.mat + 2]), u3 = cvt<T>(mats[st.mat + 3]);
#pragma unroll
                for (unsigned jj = unsigned(t); jj < unsigned(kAmps / 2); jj += kPassThreads) {
                    const unsigned i0 = ((jj >> pb) << (pb + 1)) | (jj & ((1u << pb) - 1)), i1 = i0 | (1u << pb);
                    const cplx<T> a0 = P[i0], a1 = P[i1], b0 = Lm[i0], b1 = Lm[i1];
                    P[i0] = cadd(cmul(u0, a0), cmul(u1, a1));
                    P[i1] = cadd(cmul(u2, a0), cmul(u3, a1));
                    Lm[i0] = cadd(cmul(u0, b0), cmul(u1, b1));
                    Lm[i1] = cadd(cmul(u2, b0), cmul(u3, b1));
                }
                __syncthreads();
                continue;
            }
            if (st.kind == A_DENSE2) {
                const int ph = st.bit[0], pl = st.bit[1];  // matrix MSB, LSB
                const int lo = ph < pl ? ph : pl, hi = ph < pl ? pl : ph;
                cplx<T> u[16];
#pragma unroll
                for (int e = 0; e < 16; ++e) u[e] = cvt<T>(mats[st.mat + e]);
#pragma unroll
                for (unsigned jj = unsigned(t); jj < unsigned(kAmps / 4); jj += kPassThreads) {
                    unsigned b0 = ((jj >> lo) << (lo + 1)) | (jj & ((1u << lo) - 1));
                    b0 = ((b0 >> hi) << (hi + 1)) | (b0 & ((1u << hi) - 1));
                    const unsigned idx[4] = {b0, b0 | (1u << pl), b0 | (1u << ph), b0 | (1u << ph) | (1u << pl)};
                    cplx<T> a[4], c[4];
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        a[e] = P[idx[e]];
                        c[e] = Lm[idx[e]];
                    }
#pragma unroll
                    for (int row = 0; row < 4; ++row) {
                        cplx<T> x = cmul(u[4 * row], a[0]), y = cmul(u[4 * row], c[0]);
#pragma unroll
                        for (int col = 1; col < 4; ++col) {
                            x = cadd(x, cmul(u[4 * row + col], a[col]));
                            y = cadd(y, cmul(u[4 * row + col], c[col]));
                        }
                        P[idx[row]] = x;
                        Lm[idx[row]] = y;
                    }
                }
                __syncthreads();
                continue;
            }
            if (st.kind == A_DIAG) {
                const int k = st.k;
                const int dim = 1 << k;
                const cplx<T> d0 = cvt<T>(mats[st.mat]), d1 = cvt<T>(mats[st.mat + (dim + 1)]);
                const cplx<T> d2 = k == 2 ? cvt<T>(mats[st.mat + 2 * (dim + 1)]) : d0;
                const cplx<T> d3 = k == 2 ? cvt<T>(mats[st.mat + 3 * (dim + 1)]) : d0;
                const int b0 = st.bit[0], b1 = k == 2 ? st.bit[1] : 0;
#pragma unroll
                for (int r = 0; r < kPer; ++r) {
                    const unsigned i = unsigned(t + r * kPassThreads);
                    if constexpr (kAmps < kPassThreads) { if (i >= unsigned(kAmps)) break; }
                    const int row = k == 2 ? int(((i >> b0) & 1) << 1 | ((i >> b1) & 1)) : int((i >> b0) & 1);
                    const cplx<T> ph = row == 0 ? d0 : row == 1 ? d1 : row == 2 ? d2 : d3;
                    P[i] = cmul(ph, P[i]);
                    Lm[i] = cmul(ph, Lm[i]);
                }
                __syncthreads();
                continue;
            }
            // generic gate: each thread computes its kPer outputs of both
            // tiles, then (after a barrier) writes them back
            const int k = st.k, dim = 1 << k;
            unsigned gm = 0;
            for (int j2 = 0; j2 < k; ++j2) gm |= 1u << st.bit[j2];
            cplx<T> op_[kPer], ol[kPer];
#pragma unroll
            for (int r = 0; r < kPer; ++r) {
                const unsigned i = unsigned(t + r * kPassThreads);
                if constexpr (kAmps < kPassThreads) { if (i >= unsigned(kAmps)) break; }
                int row = 0;
                for (int j2 = 0; j2 < k; ++j2) row = (row << 1) | int((i >> st.bit[j2]) & 1);
                const unsigned b0 = i & ~gm;
                cplx<T> ap{T(0), T(0)}, al{T(0), T(0)};
                for (int c = 0; c < dim; ++c) {
                    unsigned off = 0;
                    for (int j2 = 0; j2 < k; ++j2)
                        if ((c >> (k - 1 - j2)) & 1) off |= 1u << st.bit[j2];
                    const cplx<T> uu = cvt<T>(mats[st.mat + row * dim + c]);
                    ap = cadd(ap, cmul(uu, P[b0 | off]));
                    al = cadd(al, cmul(uu, Lm[b0 | off]));
                }
                op_[r] = ap;
                ol[r] = al;
            }
            __syncthreads();
#pragma unroll
            for (int r = 0; r < kPer; ++r) {
                const unsigned i = unsigned(t + r * kPassThreads);
                if constexpr (kAmps < kPassThreads) { if (i >= unsigned(kAmps)) break; }
                P[i] = op_[r];
                Lm[i] = ol[r];
            }
            __syncthreads();
        }
#pragma unroll
        for (int r = 0; r < kPer; ++r) {
            const unsigned i = unsigned(t + r * kPassThreads);
            if constexpr (kAmps < kPassThreads) { if (i >= unsigned(kAmps)) break; }
            const uint64_t g = gidx(i);
            psi[g] = P[i];
            lam[g] = Lm[i];
        }
    }
}

// ---- register-resident pass ----
//
// Each thread holds 2^RB amplitudes of each tile (the register bits of the
// current layout, RB of the tile's B bits) and the 2^(B-RB) threads cover
// the other bits.  Dense gates act on register bits in registers; diagonal
// gates act on any bit (the phase follows from the amplitude's index);
// gradient steps need their X bits among the register bits (the partner
// amplitude is in the same thread) and reduce per wave with shuffles, so
// they need no barrier.  A relayout step moves the tile through LDS to a new
// set of register bits; it is the only step with barriers.

enum AdjRKind : int {
    R_DENSE1 = 0,
    R_DENSE2 = 1,
    R_DIAG = 2,
    R_GRAD = 3,
    R_RELAYOUT = 4,
    R_DTABLE = 5,
    R_DENSE1R = 6,  // 2x2 with real entries on register slot a
    R_GRAD1 = 7,    // one term, X (b = 0) or Y (b = 1) on register slot a alone
    R_ZGRADS = 8,   // nterms (<= 8) gradients of one Z-string term each, terms term0.., the slot of
                    // gradient g in its term's `slot`; b = 1: then the phase table (mask, mat) of R_DTABLE
    R_GRAD1R = 9,   // R_GRAD1 on slot a (b = 1 for Y, term term0, partial slot `slot`) followed by
                    // R_DENSE1R on the same slot (entries at mat): a rotation's gradient and its inverse
    R_ROT3 = 10,    // three R_GRAD1R, on register slots 0, 1, 2 in that order, entry term0 + m for slot
                    // m: 1 for Y in `z`, the real 2x2 at mats offset `xt`, partial slot `slot` (-1: no
                    // gradient; the identity and a zero coefficient stand in for an absent rotation)
    R_PERM2 = 11,   // a 4x4 of exact zeros and ones that only exchanges amplitudes: k = 0, cx (control
                    // slot a = the matrix MSB, target slot b); k = 1, exchange of slots a < b (swap)
    R_CTRL2 = 12    // a 4x4 that is the identity where the matrix MSB (slot a) is 0: the 2x2 block (at
                    // mats offset `mat`, 4 entries) on target slot b where it is 1 (crx, cry, cu3, ...)
};

struct AdjRStep {
    int kind;
    int a, b;          // DENSE1(R): register slot a; DENSE2: slots a (matrix MSB), b; GRAD: a = 1 if all terms
                       // are Z strings; GRAD1: slot a, b = 1 for Y
    int k;             // DIAG: 1 or 2
    int bits[2];       // DIAG: tile bits, MSB first
    int mat;           // matrix offset (complex entries)
    int term0, nterms, slot;  // GRAD
    unsigned mask;     // RELAYOUT: register tile bits of the new layout; DTABLE (and ZGRADS with b = 1):
                       // the table's tile bits
};

struct AdjRTerm {
    unsigned xr;       // X bits as register-slot mask
    unsigned xt, z;    // X bits and Z bits as tile-bit masks
    int slot;          // R_ZGRADS: the gradient's partial-sum slot
    double cre, cim;   // coefficient times i^{#Y}
};

// One DPP lane move of a double (two 32-bit moves); lanes that the row mask
// leaves out read 0.
template <int kCtrl, int kRowMask>
__device__ __forceinline__ double dpp_f64(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, int(b & 0xffffffff), kCtrl, kRowMask, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, int(b >> 32), kCtrl, kRowMask, 0xf, false);
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned>(lo));
}

// Sum of v over the 64 lanes of a wave, complete in lane 63, in a fixed
// order, with lane moves only (no LDS): pairs and quads by quad permutes,
// rows of 16 by rotations by 4 and 8, then row 0 into row 1 and row 2 into
// row 3 (broadcast of lane 15 of each row), then rows 0-1 into row 3
// (broadcast of lane 31).
__device__ __forceinline__ double wave_sum_lane63(double v) {
    v += dpp_f64<0xB1, 0xf>(v);   // quad_perm [1,0,3,2]
    v += dpp_f64<0x4E, 0xf>(v);   // quad_perm [2,3,0,1]
    v += dpp_f64<0x124, 0xf>(v);  // row_ror:4
    v += dpp_f64<0x128, 0xf>(v);  // row_ror:8
    v += dpp_f64<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
    v += dpp_f64<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
    return v;
}

// One 32-bit-word exchange step of a double pair across lane bit C (C = 5, 4:
// v_permlane32_swap / v_permlane16_swap; C = 3: DPP row shifts by 8 with bank
// masks): afterwards lanes whose bit C is 0 hold x of both lanes of the pair
// and lanes whose bit C is 1 hold y of both, so x + y sums x over the pair in
// the first and y in the second.
template <int C>
__device__ __forceinline__ double fold_pair(double x, double y) {
    const long long xb = __double_as_longlong(x), yb = __double_as_longlong(y);
    unsigned xl = unsigned(xb), xh = unsigned(xb >> 32), yl = unsigned(yb), yh = unsigned(yb >> 32);
    if constexpr (C == 5) {
        auto r = __builtin_amdgcn_permlane32_swap(xl, yl, false, false);
        xl = r[0];
        yl = r[1];
        r = __builtin_amdgcn_permlane32_swap(xh, yh, false, false);
        xh = r[0];
        yh = r[1];
    } else if constexpr (C == 4) {
        auto r = __builtin_amdgcn_permlane16_swap(xl, yl, false, false);
        xl = r[0];
        yl = r[1];
        r = __builtin_amdgcn_permlane16_swap(xh, yh, false, false);
        xh = r[0];
        yh = r[1];
    } else {
        const unsigned nxl = __builtin_amdgcn_update_dpp(xl, yl, 0x118, 0xF, 0xC, false);  // row_shr:8, banks 2-3
        const unsigned nyl = __builtin_amdgcn_update_dpp(yl, xl, 0x108, 0xF, 0x3, false);  // row_shl:8, banks 0-1
        const unsigned nxh = __builtin_amdgcn_update_dpp(xh, yh, 0x118, 0xF, 0xC, false);
        const unsigned nyh = __builtin_amdgcn_update_dpp(yh, xh, 0x108, 0xF, 0x3, false);
        xl = nxl;
        yl = nyl;
        xh = nxh;
        yh = nyh;
    }
    const double nx = __longlong_as_double((static_cast<long long>(xh) << 32) | xl);
    const double ny = __longlong_as_double((static_cast<long long>(yh) << 32) | yl);
    return nx + ny;
}

// Sums of v[0..7] over the 64 lanes of a wave in one butterfly: three folds
// (lane bits 5, 4, 3) leave in each lane the pair sums of value
// 4 bit3 + 2 bit4 + bit5 of its lane index, then two quad permutes and a row
// shift by 4 add the 8 lanes of each group (bits 0..2).  The total of value i
// ends in lane 4 + 8 (4 (i & 1) + 2 ((i >> 1) & 1) + (i >> 2)) (lanes with
// bits 0..2 = 4); returns it there.
__device__ __forceinline__ double wave_sum8(const double (&v)[8]) {
    const double w0 = fold_pair<5>(v[0], v[1]), w1 = fold_pair<5>(v[2], v[3]);
    const double w2 = fold_pair<5>(v[4], v[5]), w3 = fold_pair<5>(v[6], v[7]);
    const double u0 = fold_pair<4>(w0, w1), u1 = fold_pair<4>(w2, w3);
    double t = fold_pair<3>(u0, u1);
    t += dpp_f64<0xB1, 0xf>(t);  // quad_perm [1,0,3,2]
    t += dpp_f64<0x4E, 0xf>(t);  // quad_perm [2,3,0,1]
    const long long tb = __double_as_longlong(t);
    const int lo = __builtin_amdgcn_update_dpp(0, int(tb & 0xffffffff), 0x114, 0xf, 0xa, false);  // row_shr:4
    const int hi = __builtin_amdgcn_update_dpp(0, int(tb >> 32), 0x114, 0xf, 0xa, false);
    return t + __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned>(lo));
}

// An amplitude as one naturally aligned vector: one LDS access each.
template <typename T>
using amp2 = T __attribute__((ext_vector_type(2)));

// Complex products as FMA chains: u * x is two products and two FMAs, and
// acc + u * x four FMAs (the kernel is compiled with contraction).
template <typename T>
__device__ __forceinline__ T afma(T a, T b, T c) {
    if constexpr (sizeof(T) == 8) return __builtin_fma(a, b, c);
    else return __builtin_fmaf(a, b, c);
}
template <typename T>
__device__ __forceinline__ cplx<T> cfmul(cplx<T> u, cplx<T> x) {
    return cplx<T>{afma(u.re, x.re, -(u.im * x.im)), afma(u.re, x.im, u.im * x.re)};
}
template <typename T>
__device__ __forceinline__ cplx<T> cfmac(cplx<T> acc, cplx<T> u, cplx<T> x) {
    return cplx<T>{afma(-u.im, x.im, afma(u.re, x.re, acc.re)), afma(u.im, x.re, afma(u.re, x.im, acc.im))};
}

template <typename T, int NR, int M>
__device__ __forceinline__ void r_dense1(cplx<T> (&a)[NR], const cplx<T> u0, const cplx<T> u1, const cplx<T> u2,
                                         const cplx<T> u3) {
#pragma unroll
    for (int j = 0; j < NR; ++j) {
        if (j & (1 << M)) continue;
        const int j1 = j | (1 << M);
        const cplx<T> x0 = a[j], x1 = a[j1];
        a[j] = cfmac(cfmul(u1, x1), u0, x0);
        a[j1] = cfmac(cfmul(u3, x1), u2, x0);
    }
}

// 2x2 with real entries: each output is two real-times-complex products.
template <typename T, int NR, int M>
__device__ __forceinline__ void r_dense1r(cplx<T> (&a)[NR], const T u0, const T u1, const T u2, const T u3) {
#pragma unroll
    for (int j = 0; j < NR; ++j) {
        if (j & (1 << M)) continue;
        const int j1 = j | (1 << M);
        const cplx<T> x0 = a[j], x1 = a[j1];
        a[j] = cplx<T>{u0 * x0.re + u1 * x1.re, u0 * x0.im + u1 * x1.im};
        a[j1] = cplx<T>{u2 * x0.re + u3 * x1.re, u2 * x0.im + u3 * x1.im};
    }
}

// <lam| P |psi> summed over this thread's pairs for P = X (Y = 0) or Y (Y = 1)
// on register slot M: with s0 = -1 for Y (the Z factor on the partner of a
// bit-0 amplitude) and +1 for X, the sum is sum over pairs of
// s0 conj(l0) p1 + conj(l1) p0 (the term's i^{#Y} is in its coefficient).
// Returns (Re, Im) of that sum.
template <typename T, int NR, int M, int Y>
__device__ __forceinline__ void r_pair_inner(const cplx<T> (&p)[NR], const cplx<T> (&l)[NR], double &ar, double &ai) {
#pragma unroll
    for (int j = 0; j < NR; ++j) {
        if (j & (1 << M)) continue;
        const int j1 = j | (1 << M);
        const double l0r = double(l[j].re), l0i = double(l[j].im), l1r = double(l[j1].re), l1i = double(l[j1].im);
        const double p0r = double(p[j].re), p0i = double(p[j].im), p1r = double(p[j1].re), p1i = double(p[j1].im);
        ar = afma(l1i, p0i, afma(l1r, p0r, ar));
        ai = afma(-l1i, p0r, afma(l1r, p0i, ai));
        if (Y) {
            ar = afma(-l0i, p1i, afma(-l0r, p1r, ar));
            ai = afma(l0i, p1r, afma(-l0r, p1i, ai));
        } else {
            ar = afma(l0i, p1i, afma(l0r, p1r, ar));
            ai = afma(-l0i, p1r, afma(l0r, p1i, ai));
        }
    }
}

// A rotation on register slot M inside a three-rotation step: the gradient
// Im(c <lam| P |psi>) for P = X or Y (runtime, `z` of the entry) into g, then
// the real 2x2 inverse on both states.  The sum over this thread's pairs is
// split as A1 + s A0, s = -1 for Y, so that the Pauli choice costs no branch.
template <typename T, int NR, int M>
__device__ __forceinline__ void rot_slot(cplx<T> (&p)[NR], cplx<T> (&l)[NR], const AdjRTerm tm,
                                         const cplx<double> *__restrict__ mats, double &g) {
    double a1r = 0.0, a1i = 0.0, a0r = 0.0, a0i = 0.0;
#pragma unroll
    for (int j = 0; j < NR; ++j) {
        if (j & (1 << M)) continue;
        const int j1 = j | (1 << M);
        const double l0r = double(l[j].re), l0i = double(l[j].im), l1r = double(l[j1].re), l1i = double(l[j1].im);
        const double p0r = double(p[j].re), p0i = double(p[j].im), p1r = double(p[j1].re), p1i = double(p[j1].im);
        a1r = afma(l1i, p0i, afma(l1r, p0r, a1r));
        a1i = afma(-l1i, p0r, afma(l1r, p0i, a1i));
        a0r = afma(l0i, p1i, afma(l0r, p1r, a0r));
        a0i = afma(-l0i, p1r, afma(l0r, p1i, a0i));
    }
    const double sg = tm.z ? -1.0 : 1.0;
    const double ar = afma(sg, a0r, a1r), ai = afma(sg, a0i, a1i);
    g = tm.cre * ai + tm.cim * ar;  // Im(c (ar + i ai))
    const T u0 = T(mats[tm.xt].re), u1 = T(mats[tm.xt + 1].re), u2 = T(mats[tm.xt + 2].re), u3 = T(mats[tm.xt + 3].re);
    r_dense1r<T, NR, M>(p, u0, u1, u2, u3);
    r_dense1r<T, NR, M>(l, u0, u1, u2, u3);
}

template <typename T, int NR, int MH, int ML>
__device__ __forceinline__ void r_dense2(cplx<T> (&a)[NR], const cplx<T> (&u)[16]) {
#pragma unroll
    for (int j = 0; j < NR; ++j) {
        if (j & ((1 << MH) | (1 << ML))) continue;
        const int idx[4] = {j, j | (1 << ML), j | (1 << MH), j | (1 << MH) | (1 << ML)};
        cplx<T> x[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) x[e] = a[idx[e]];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            cplx<T> acc = cfmul(u[4 * r + 3], x[3]);
#pragma unroll
            for (int c = 2; c >= 0; --c) acc = cfmac(acc, u[4 * r + c], x[c]);
            a[idx[r]] = acc;
        }
    }
}

// Register permutations (R_PERM2): pure exchanges of amplitudes, no arithmetic.
// r_cswap: where register bit C is set, exchange across register bit X (cx).
template <typename T, int NR, int C, int X>
__device__ __forceinline__ void r_cswap(cplx<T> (&a)[NR]) {
#pragma unroll
    for (int j = 0; j < NR; ++j)
        if (((j >> C) & 1) && !((j >> X) & 1)) {
            const cplx<T> t = a[j];
            a[j] = a[j | (1 << X)];
            a[j | (1 << X)] = t;
        }
}
// r_swap2: exchange register bits A and B (swap).
template <typename T, int NR, int A, int B>
__device__ __forceinline__ void r_swap2(cplx<T> (&a)[NR]) {
#pragma unroll
    for (int j = 0; j < NR; ++j)
        if (((j >> A) & 1) && !((j >> B) & 1)) {
            const int k = j ^ (1 << A) ^ (1 << B);
            const cplx<T> t = a[j];
            a[j] = a[k];
            a[k] = t;
        }
}

// Controlled 2x2 (R_CTRL2): where register bit C is set, the pair across
// register bit X takes u (row-major 2x2); the products run in the order of the
// 4x4's row sums (column 3, then 2), whose other two terms are exact zeros.
template <typename T, int NR, int C, int X>
__device__ __forceinline__ void r_ctrl2(cplx<T> (&a)[NR], const cplx<T> (&u)[4]) {
#pragma unroll
    for (int j = 0; j < NR; ++j)
        if (((j >> C) & 1) && !((j >> X) & 1)) {
            const cplx<T> x = a[j], y = a[j | (1 << X)];
            a[j] = cfmac(cfmul(u[1], y), u[0], x);
            a[j | (1 << X)] = cfmac(cfmul(u[3], y), u[2], x);
        }
}

// sum over this thread's amplitudes of Im(conj(l) * (G p)) for one term set
template <typename T, int NR, int XR>
__device__ __forceinline__ double r_grad_term(const cplx<T> (&p)[NR], const cplx<T> (&l)[NR], const unsigned (&idx)[NR],
                                              const AdjRTerm &tm) {
    double acc = 0.0;
#pragma unroll
    for (int j = 0; j < NR; ++j) {
        const unsigned y = idx[j] ^ tm.xt;
        const cplx<T> v = p[j ^ XR];
        const double sg = (__popc(tm.z & y) & 1) ? -1.0 : 1.0;
        const double vr = sg * double(v.re), vi = sg * double(v.im);
        const double gr = tm.cre * vr - tm.cim * vi, gi = tm.cre * vi + tm.cim * vr;
        acc += double(l[j].re) * gi - double(l[j].im) * gr;
    }
    return acc;
}

template <typename T, int NR>
__device__ __forceinline__ double r_grad_any(const cplx<T> (&p)[NR], const cplx<T> (&l)[NR], const unsigned (&idx)[NR],
                                             const AdjRTerm &tm) {
    switch (tm.xr) {
    case 0: return r_grad_term<T, NR, 0>(p, l, idx, tm);
    case 1: return r_grad_term<T, NR, 1 % NR>(p, l, idx, tm);
    case 2: return r_grad_term<T, NR, 2 % NR>(p, l, idx, tm);
    case 3: return r_grad_term<T, NR, 3 % NR>(p, l, idx, tm);
    case 4: return r_grad_term<T, NR, 4 % NR>(p, l, idx, tm);
    case 5: return r_grad_term<T, NR, 5 % NR>(p, l, idx, tm);
    case 6: return r_grad_term<T, NR, 6 % NR>(p, l, idx, tm);
    default: return r_grad_term<T, NR, 7 % NR>(p, l, idx, tm);
    }
}

template <typename T, int B, int RB>
__global__ __launch_bounds__(1 << (B - RB), sizeof(T) == 4 ? 4 : 2) void k_adjoint_reg(cplx<T> *__restrict__ psi, cplx<T> *__restrict__ lam,
                                                              AdjPass pd, const AdjRStep *__restrict__ steps,
                                                              const cplx<double> *__restrict__ mats,
                                                              const AdjRTerm *__restrict__ terms,
                                                              double *__restrict__ partials, uint64_t ntiles,
                                                              unsigned first_mask) {
    constexpr int NT = 1 << (B - RB), NR = 1 << RB, W = NT / 64;
    constexpr unsigned kTileMask = (1u << B) - 1u;
    extern __shared__ unsigned char smem_raw[];
    cplx<T> *LP = reinterpret_cast<cplx<T> *>(smem_raw);
    cplx<T> *LL = LP + (1 << B);
    const unsigned t = threadIdx.x;
    const unsigned lane = t & 63, wave = t >> 6;
    cplx<T> p[NR], l[NR];
    unsigned tpart = 0, cur_mask = 0;  // this thread's tile bits and the register bits of the layout
    // tile index of amplitude j: the thread's bits | j's register bits (the
    // layout is the same for the whole workgroup, so the latter is scalar)
    auto reg_bits = [&](int j) {  // j's register bits of the current layout (scalar)
        unsigned cm = __builtin_amdgcn_readfirstlane(cur_mask), o = 0;
#pragma unroll
        for (int r = 0; r < RB; ++r) {
            const unsigned low = cm & (~cm + 1u);
            if ((j >> r) & 1) o |= low;
            cm &= cm - 1u;
        }
        return o;
    };
    auto ix = [&](int j) { return tpart | reg_bits(j); };
    // LDS slot of tile index i during a relayout: the low nibble XORed with
    // the higher nibbles, against bank conflicts of strided layouts.  It is
    // linear over XOR, so slot(tpart | o) = slot(tpart) ^ slot(o): one
    // per-thread value per layout and a scalar one per register, kept scalar
    // (readfirstlane) so that each access is a single XOR.
    auto slot = [](unsigned i) {
        unsigned f = 0;
#pragma unroll
        for (int sh = 4; sh < B; sh += 4) f ^= (i >> sh) & 15u;
        return i ^ f;
    };
    // Global offset of a tile: its outer-index bits deposited on the outer
    // qubits (oq ascending).  Tiles are blockIdx.x + k * gridDim.x with
    // gridDim.x a power of two (the host launches it so): the deposit of
    // blockIdx.x is fixed per workgroup, and that of k steps as a counter
    // confined to the outer qubits of the high index bits (hmask).
    uint64_t dep_lo = 0, hmask = 0;
    for (int o = 0; o < pd.n_outer; ++o) {
        const uint64_t qb = uint64_t(1) << pd.oq[o];
        if ((uint64_t(blockIdx.x) >> o) & 1) dep_lo |= qb;
        if (!(((uint64_t(gridDim.x) - 1) >> o) & 1)) hmask |= qb;
    }
    uint64_t high = 0;
    for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x, high = ((high | ~hmask) + 1) & hmask) {
        const uint64_t base = dep_lo | high;
        // global offsets of the current layout: base | the thread's part,
        // and the qubit offset of each register bit (scalar); amplitude j is
        // the first ORed with the offsets of j's register bits
        uint64_t gt = 0, gr[RB];
        auto gmap_layout = [&]() {
            gt = base;
#pragma unroll
            for (int b2 = 0; b2 < B; ++b2)
                if ((tpart >> b2) & 1) gt |= uint64_t(1) << pd.tq[b2];
            unsigned cm = __builtin_amdgcn_readfirstlane(cur_mask);
#pragma unroll
            for (int r = 0; r < RB; ++r) {
                const int bit = __builtin_amdgcn_readfirstlane(__ffs(cm) - 1);
                gr[r] = uint64_t(1) << pd.tq[bit];
                cm &= cm - 1u;
            }
        };
        auto gidx_j = [&](int j) {
            uint64_t g = gt;
#pragma unroll
            for (int r = 0; r < RB; ++r)
                if ((j >> r) & 1) g |= gr[r];
            return g;
        };
        // the thread's tile bits: t deposited on the non-register bits,
        // ascending, i.e. a zero bit inserted into t at each register bit
        auto layout = [&](unsigned mask) {
            unsigned tp = t, cm = __builtin_amdgcn_readfirstlane(mask);
#pragma unroll
            for (int r = 0; r < RB; ++r) {
                const unsigned lo = (cm & (~cm + 1u)) - 1u;
                tp = (tp & lo) | ((tp & ~lo) << 1);
                cm &= cm - 1u;
            }
            tpart = tp;
            cur_mask = mask;
        };
        layout(first_mask);
        gmap_layout();
#pragma unroll
        for (int j = 0; j < NR; ++j) {
            const uint64_t g = gidx_j(j);
            p[j] = psi[g];
            l[j] = lam[g];
        }
        // a product of diagonal gates as one phase table over the tile bits in m
        auto apply_dtable = [&](unsigned m, int mat) __attribute__((always_inline)) {
            unsigned pt = 0;
            {
                unsigned mm = m, pos = 0;
                while (mm) {
                    const unsigned low = mm & (~mm + 1u);
                    if (tpart & low) pt |= 1u << pos;
                    ++pos;
                    mm &= mm - 1u;
                }
            }
            unsigned c[RB];
            {
                unsigned cm = cur_mask;
#pragma unroll
                for (int r = 0; r < RB; ++r) {
                    const unsigned low = cm & (~cm + 1u);
                    c[r] = (m & low) ? (1u << __popc(m & (low - 1u))) : 0u;
                    cm &= cm - 1u;
                }
            }
#pragma unroll
            for (int j = 0; j < NR; ++j) {
                unsigned row = pt;
#pragma unroll
                for (int r = 0; r < RB; ++r)
                    if ((j >> r) & 1) row |= c[r];
                const cplx<T> ph = cvt<T>(mats[mat + int(row)]);
                p[j] = cmul(ph, p[j]);
                l[j] = cmul(ph, l[j]);
            }
        };
        for (int s = 0; s < pd.nsteps; ++s) {
            const AdjRStep st = steps[pd.step0 + s];
            switch (st.kind) {
            case R_RELAYOUT: {
                __syncthreads();  // every thread has read the previous layout's LDS data
                {
                    const unsigned st_t = slot(tpart);
#pragma unroll
                    for (int j = 0; j < NR; ++j) {
                        const unsigned a = st_t ^ __builtin_amdgcn_readfirstlane(slot(reg_bits(j)));
                        reinterpret_cast<amp2<T> *>(LP)[a] = amp2<T>{p[j].re, p[j].im};
                        reinterpret_cast<amp2<T> *>(LL)[a] = amp2<T>{l[j].re, l[j].im};
                    }
                }
                __syncthreads();
                layout(st.mask);
                {
                    const unsigned st_t = slot(tpart);
#pragma unroll
                    for (int j = 0; j < NR; ++j) {
                        const unsigned a = st_t ^ __builtin_amdgcn_readfirstlane(slot(reg_bits(j)));
                        const amp2<T> vp = reinterpret_cast<const amp2<T> *>(LP)[a];
                        const amp2<T> vl = reinterpret_cast<const amp2<T> *>(LL)[a];
                        p[j] = cplx<T>{vp.x, vp.y};
                        l[j] = cplx<T>{vl.x, vl.y};
                    }
                }
                break;
            }
            case R_DENSE1: {
                const cplx<T> u0 = cvt<T>(mats[st.mat]), u1 = cvt<T>(mats[st.mat + 1]);
                const cplx<T> u2 = cvt<T>(mats[st.mat + 2]), u3 = cvt<T>(mats[st.mat + 3]);
                switch (st.a) {
                case 0: r_dense1<T, NR, 0>(p, u0, u1, u2, u3); r_dense1<T, NR, 0>(l, u0, u1, u2, u3); break;
                case 1: r_dense1<T, NR, 1 % RB>(p, u0, u1, u2, u3); r_dense1<T, NR, 1 % RB>(l, u0, u1, u2, u3); break;
                case 2: r_dense1<T, NR, 2 % RB>(p, u0, u1, u2, u3); r_dense1<T, NR, 2 % RB>(l, u0, u1, u2, u3); break;
                default: r_dense1<T, NR, 3 % RB>(p, u0, u1, u2, u3); r_dense1<T, NR, 3 % RB>(l, u0, u1, u2, u3); break;
                }
                break;
            }
            case R_DENSE1R: {
                const T u0 = T(mats[st.mat].re), u1 = T(mats[st.mat + 1].re);
                const T u2 = T(mats[st.mat + 2].re), u3 = T(mats[st.mat + 3].re);
                switch (st.a) {
                case 0: r_dense1r<T, NR, 0>(p, u0, u1, u2, u3); r_dense1r<T, NR, 0>(l, u0, u1, u2, u3); break;
                case 1: r_dense1r<T, NR, 1 % RB>(p, u0, u1, u2, u3); r_dense1r<T, NR, 1 % RB>(l, u0, u1, u2, u3); break;
                default: r_dense1r<T, NR, 2 % RB>(p, u0, u1, u2, u3); r_dense1r<T, NR, 2 % RB>(l, u0, u1, u2, u3); break;
                }
                break;
            }
            case R_GRAD1R: {
                double ar = 0.0, ai = 0.0;
                const T u0 = T(mats[st.mat].re), u1 = T(mats[st.mat + 1].re);
                const T u2 = T(mats[st.mat + 2].re), u3 = T(mats[st.mat + 3].re);
                switch (st.a * 2 + st.b) {
                case 0: r_pair_inner<T, NR, 0, 0>(p, l, ar, ai); r_dense1r<T, NR, 0>(p, u0, u1, u2, u3); r_dense1r<T, NR, 0>(l, u0, u1, u2, u3); break;
                case 1: r_pair_inner<T, NR, 0, 1>(p, l, ar, ai); r_dense1r<T, NR, 0>(p, u0, u1, u2, u3); r_dense1r<T, NR, 0>(l, u0, u1, u2, u3); break;
                case 2: r_pair_inner<T, NR, 1 % RB, 0>(p, l, ar, ai); r_dense1r<T, NR, 1 % RB>(p, u0, u1, u2, u3); r_dense1r<T, NR, 1 % RB>(l, u0, u1, u2, u3); break;
                case 3: r_pair_inner<T, NR, 1 % RB, 1>(p, l, ar, ai); r_dense1r<T, NR, 1 % RB>(p, u0, u1, u2, u3); r_dense1r<T, NR, 1 % RB>(l, u0, u1, u2, u3); break;
                case 4: r_pair_inner<T, NR, 2 % RB, 0>(p, l, ar, ai); r_dense1r<T, NR, 2 % RB>(p, u0, u1, u2, u3); r_dense1r<T, NR, 2 % RB>(l, u0, u1, u2, u3); break;
                default: r_pair_inner<T, NR, 2 % RB, 1>(p, l, ar, ai); r_dense1r<T, NR, 2 % RB>(p, u0, u1, u2, u3); r_dense1r<T, NR, 2 % RB>(l, u0, u1, u2, u3); break;
                }
                const AdjRTerm tm = terms[st.term0];
                double acc = tm.cre * ai + tm.cim * ar;  // Im(c (ar + i ai))
                acc = wave_sum_lane63(acc);
                if (lane == 63) partials[(uint64_t(st.slot) * ntiles + tile) * W + wave] = acc;
                break;
            }
            case R_GRAD1: {
                double ar = 0.0, ai = 0.0;
                const int code = st.a * 2 + st.b;
                switch (code) {
                case 0: r_pair_inner<T, NR, 0, 0>(p, l, ar, ai); break;
                case 1: r_pair_inner<T, NR, 0, 1>(p, l, ar, ai); break;
                case 2: r_pair_inner<T, NR, 1 % RB, 0>(p, l, ar, ai); break;
                case 3: r_pair_inner<T, NR, 1 % RB, 1>(p, l, ar, ai); break;
                case 4: r_pair_inner<T, NR, 2 % RB, 0>(p, l, ar, ai); break;
                default: r_pair_inner<T, NR, 2 % RB, 1>(p, l, ar, ai); break;
                }
                const AdjRTerm tm = terms[st.term0];
                double acc = tm.cre * ai + tm.cim * ar;  // Im(c (ar + i ai))
                acc = wave_sum_lane63(acc);
                if (lane == 63) partials[(uint64_t(st.slot) * ntiles + tile) * W + wave] = acc;
                break;
            }
            case R_DENSE2: {
                cplx<T> u[16];
#pragma unroll
                for (int e = 0; e < 16; ++e) u[e] = cvt<T>(mats[st.mat + e]);
                const int code = st.a * 4 + st.b;
                switch (code) {
                case 1: r_dense2<T, NR, 0, 1 % RB>(p, u); r_dense2<T, NR, 0, 1 % RB>(l, u); break;
                case 2: r_dense2<T, NR, 0, 2 % RB>(p, u); r_dense2<T, NR, 0, 2 % RB>(l, u); break;
                case 4: r_dense2<T, NR, 1 % RB, 0>(p, u); r_dense2<T, NR, 1 % RB, 0>(l, u); break;
                case 6: r_dense2<T, NR, 1 % RB, 2 % RB>(p, u); r_dense2<T, NR, 1 % RB, 2 % RB>(l, u); break;
                case 8: r_dense2<T, NR, 2 % RB, 0>(p, u); r_dense2<T, NR, 2 % RB, 0>(l, u); break;
                default: r_dense2<T, NR, 2 % RB, 1 % RB>(p, u); r_dense2<T, NR, 2 % RB, 1 % RB>(l, u); break;
                }
                break;
            }
            case R_PERM2:
                switch (st.k * 16 + st.a * 4 + st.b) {
                case 1: r_cswap<T, NR, 0, 1 % RB>(p); r_cswap<T, NR, 0, 1 % RB>(l); break;
                case 2: r_cswap<T, NR, 0, 2 % RB>(p); r_cswap<T, NR, 0, 2 % RB>(l); break;
                case 4: r_cswap<T, NR, 1 % RB, 0>(p); r_cswap<T, NR, 1 % RB, 0>(l); break;
                case 6: r_cswap<T, NR, 1 % RB, 2 % RB>(p); r_cswap<T, NR, 1 % RB, 2 % RB>(l); break;
                case 8: r_cswap<T, NR, 2 % RB, 0>(p); r_cswap<T, NR, 2 % RB, 0>(l); break;
                case 9: r_cswap<T, NR, 2 % RB, 1 % RB>(p); r_cswap<T, NR, 2 % RB, 1 % RB>(l); break;
                case 17: r_swap2<T, NR, 0, 1 % RB>(p); r_swap2<T, NR, 0, 1 % RB>(l); break;
                case 18: r_swap2<T, NR, 0, 2 % RB>(p); r_swap2<T, NR, 0, 2 % RB>(l); break;
                default: r_swap2<T, NR, 1 % RB, 2 % RB>(p); r_swap2<T, NR, 1 % RB, 2 % RB>(l); break;  // 22
                }
                break;
            case R_CTRL2: {
                cplx<T> u[4];
#pragma unroll
                for (int e = 0; e < 4; ++e) u[e] = cvt<T>(mats[st.mat + e]);
                switch (st.a * 4 + st.b) {
                case 1: r_ctrl2<T, NR, 0, 1 % RB>(p, u); r_ctrl2<T, NR, 0, 1 % RB>(l, u); break;
                case 2: r_ctrl2<T, NR, 0, 2 % RB>(p, u); r_ctrl2<T, NR, 0, 2 % RB>(l, u); break;
                case 4: r_ctrl2<T, NR, 1 % RB, 0>(p, u); r_ctrl2<T, NR, 1 % RB, 0>(l, u); break;
                case 6: r_ctrl2<T, NR, 1 % RB, 2 % RB>(p, u); r_ctrl2<T, NR, 1 % RB, 2 % RB>(l, u); break;
                case 8: r_ctrl2<T, NR, 2 % RB, 0>(p, u); r_ctrl2<T, NR, 2 % RB, 0>(l, u); break;
                default: r_ctrl2<T, NR, 2 % RB, 1 % RB>(p, u); r_ctrl2<T, NR, 2 % RB, 1 % RB>(l, u); break;
                }
                break;
            }
            case R_DTABLE:
                apply_dtable(st.mask, st.mat);
                break;
            case R_DIAG: {
                const int dim = 1 << st.k;
                const cplx<T> d0 = cvt<T>(mats[st.mat]), d1 = cvt<T>(mats[st.mat + (dim + 1)]);
                const cplx<T> d2 = st.k == 2 ? cvt<T>(mats[st.mat + 2 * (dim + 1)]) : d0;
                const cplx<T> d3 = st.k == 2 ? cvt<T>(mats[st.mat + 3 * (dim + 1)]) : d0;
#pragma unroll
                for (int j = 0; j < NR; ++j) {
                    const unsigned i = ix(j);
                    const int row = st.k == 2 ? int(((i >> st.bits[0]) & 1) << 1 | ((i >> st.bits[1]) & 1))
                                              : int((i >> st.bits[0]) & 1);
                    const cplx<T> ph = row == 0 ? d0 : row == 1 ? d1 : row == 2 ? d2 : d3;
                    p[j] = cmul(ph, p[j]);
                    l[j] = cmul(ph, l[j]);
                }
                break;
            }
            case R_ZGRADS: {
                // conj(l) p once for the group: every gradient of a Z string
                // is s (cre Im(conj(l) p) + cim Re(conj(l) p)) summed, with
                // the sign s split into a per-lane and a per-register factor
                double wr[NR], wi[NR];
#pragma unroll
                for (int j = 0; j < NR; ++j) {
                    const double pr = double(p[j].re), pi = double(p[j].im);
                    const double lr = double(l[j].re), li = double(l[j].im);
                    wr[j] = afma(li, pi, lr * pr);
                    wi[j] = afma(-li, pr, lr * pi);
                }
                double v[8];
#pragma unroll
                for (int g = 0; g < 8; ++g) {
                    v[g] = 0.0;
                    if (g < st.nterms) {
                        const AdjRTerm tm = terms[st.term0 + g];
                        double im_sum = 0.0, re_sum = 0.0;
#pragma unroll
                        for (int j = 0; j < NR; ++j) {
                            const double sj = (__popc(tm.z & reg_bits(j)) & 1) ? -1.0 : 1.0;  // scalar
                            im_sum = afma(sj, wi[j], im_sum);
                            re_sum = afma(sj, wr[j], re_sum);
                        }
                        const double lane_sign = (__popc(tm.z & tpart) & 1) ? -1.0 : 1.0;
                        v[g] = lane_sign * afma(tm.cre, im_sum, tm.cim * re_sum);
                    }
                }
                const double tot = wave_sum8(v);
                if ((lane & 7) == 4) {
                    const int g = 4 * int((lane >> 3) & 1) + 2 * int((lane >> 4) & 1) + int((lane >> 5) & 1);
                    if (g < st.nterms)
                        partials[(uint64_t(terms[st.term0 + g].slot) * ntiles + tile) * W + wave] = tot;
                }
                if (st.b) apply_dtable(st.mask, st.mat);
                break;
            }
            case R_ROT3:
                if constexpr (RB == 3) {
                    // straight-line: the three slots in order, no branches, so
                    // the amplitudes stay in place through the step
                    double v[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
                    rot_slot<T, NR, 0>(p, l, terms[st.term0], mats, v[0]);
                    rot_slot<T, NR, 1>(p, l, terms[st.term0 + 1], mats, v[1]);
                    rot_slot<T, NR, 2>(p, l, terms[st.term0 + 2], mats, v[2]);
                    const double tot = wave_sum8(v);
                    const int i = 4 * int((lane >> 3) & 1) + 2 * int((lane >> 4) & 1) + int((lane >> 5) & 1);
                    if ((lane & 7) == 4 && i < 3) {
                        const int ps = terms[st.term0 + i].slot;
                        if (ps >= 0) partials[(uint64_t(ps) * ntiles + tile) * W + wave] = tot;
                    }
                }
                break;
            default: {  // R_GRAD
                double acc = 0.0;
                if (st.a) {
                    // every term diagonal (Z strings): Im(conj(l) c s p) = s (cre Im(conj(l) p) + cim Re(conj(l) p)),
                    // with the sign s split into a per-lane and a per-register (scalar) factor
                    double wr[NR], wi[NR];
#pragma unroll
                    for (int j = 0; j < NR; ++j) {
                        const double pr = double(p[j].re), pi = double(p[j].im);
                        const double lr = double(l[j].re), li = double(l[j].im);
                        wr[j] = lr * pr + li * pi;
                        wi[j] = lr * pi - li * pr;
                    }
                    for (int q = 0; q < st.nterms; ++q) {
                        const AdjRTerm tm = terms[st.term0 + q];
                        double im_sum = 0.0, re_sum = 0.0;
#pragma unroll
                        for (int j = 0; j < NR; ++j) {
                            const double sj = (__popc(tm.z & reg_bits(j)) & 1) ? -1.0 : 1.0;  // scalar
                            im_sum += sj * wi[j];
                            re_sum += sj * wr[j];
                        }
                        const double lane_sign = (__popc(tm.z & tpart) & 1) ? -1.0 : 1.0;
                        acc += lane_sign * (tm.cre * im_sum + tm.cim * re_sum);
                    }
                } else {
                    unsigned idx[NR];
#pragma unroll
                    for (int j = 0; j < NR; ++j) idx[j] = ix(j);
                    for (int q = 0; q < st.nterms; ++q) acc += r_grad_any<T, NR>(p, l, idx, terms[st.term0 + q]);
                }
                acc = wave_sum_lane63(acc);
                if (lane == 63) partials[(uint64_t(st.slot) * ntiles + tile) * W + wave] = acc;
                break;
            }
            }
        }
        gmap_layout();
#pragma unroll
        for (int j = 0; j < NR; ++j) {
            const uint64_t g = gidx_j(j);
            psi[g] = p[j];
            lam[g] = l[j];
        }
        __syncthreads();  // the next tile's first relayout writes the LDS
    }
}

// First level of the gradient sums: block (s, c) adds chunk c (of kReduceChunks
// contiguous chunks) of slot s's partials, in order.
constexpr int kReduceChunks = 32;
__global__ __launch_bounds__(kThreads) void k_adjoint_reduce_chunks(const double *__restrict__ partials,
                                                                    uint64_t count, double *__restrict__ chunks) {
    __shared__ double scratch[kThreads / 64];
    const int s = blockIdx.x, c = blockIdx.y;
    const uint64_t per = (count + kReduceChunks - 1) / kReduceChunks;
    const uint64_t lo = uint64_t(c) * per, hi = std::min<uint64_t>(count, lo + per);
    double acc = 0.0;
    for (uint64_t i = lo + threadIdx.x; i < hi; i += kThreads) acc += partials[uint64_t(s) * count + i];
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if ((threadIdx.x & 63) == 0) scratch[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = 0.0;
        for (int w = 0; w < kThreads / 64; ++w) r += scratch[w];
        chunks[s * kReduceChunks + c] = r;
    }
}

// grads[slot_base + s] = sum over tiles of partials[s * ntiles + tile], in tile order.
__global__ __launch_bounds__(kThreads) void k_adjoint_reduce(const double *__restrict__ partials, uint64_t ntiles,
                                                             double *__restrict__ grads) {
    __shared__ double scratch[kThreads / 64];
    const int s = blockIdx.x;
    double acc = 0.0;
    for (uint64_t i = threadIdx.x; i < ntiles; i += kThreads) acc += partials[uint64_t(s) * ntiles + i];
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if ((threadIdx.x & 63) == 0) scratch[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = 0.0;
        for (int w = 0; w < kThreads / 64; ++w) r += scratch[w];
        grads[s] = r;
    }
}

int afail(int code, const std::string &msg) { return dq::set_error(code, "adjoint sweep: " + msg); }

#define ADJ_TRY(expr)                                                                                   \
    do {                                                                                                \
        hipError_t _e = (expr);                                                                         \
        if (_e != hipSuccess) return afail(DQ_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

template <typename T, int B>
hipError_t launch_pass(unsigned grid, hipStream_t stream, void *psi, void *lam, const AdjPass &pd,
                       const AdjStep *steps, const void *mats, const AdjTerm *terms, double *partials,
                       uint64_t ntiles) {
    const size_t shmem = 2 * sizeof(cplx<T>) * (size_t(1) << B) + 8 * (kPassThreads / 64);
    if (shmem > 65536) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_adjoint_pass<T, B>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, int(shmem));
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((k_adjoint_pass<T, B>), dim3(grid), dim3(kPassThreads), shmem, stream,
                       static_cast<cplx<T> *>(psi), static_cast<cplx<T> *>(lam), pd, steps,
                       static_cast<const cplx<double> *>(mats), terms, partials, ntiles);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_any(int b, unsigned grid, hipStream_t stream, void *psi, void *lam, const AdjPass &pd,
                      const AdjStep *steps, const void *mats, const AdjTerm *terms, double *partials,
                      uint64_t ntiles) {
    switch (b) {
    case 8: return launch_pass<T, 8>(grid, stream, psi, lam, pd, steps, mats, terms, partials, ntiles);
    case 9: return launch_pass<T, 9>(grid, stream, psi, lam, pd, steps, mats, terms, partials, ntiles);
    case 10: return launch_pass<T, 10>(grid, stream, psi, lam, pd, steps, mats, terms, partials, ntiles);
    case 11: return launch_pass<T, 11>(grid, stream, psi, lam, pd, steps, mats, terms, partials, ntiles);
    case 12: return launch_pass<T, 12>(grid, stream, psi, lam, pd, steps, mats, terms, partials, ntiles);
    default: return hipErrorInvalidValue;
    }
}

template <typename T, int B>
hipError_t launch_reg_pass(unsigned grid, hipStream_t stream, void *psi, void *lam, const AdjPass &pd,
                           const AdjRStep *steps, const void *mats, const AdjRTerm *terms, double *partials,
                           uint64_t ntiles, unsigned first_mask) {
    constexpr int RB = kRegBits;
    const size_t shmem = 2 * sizeof(cplx<T>) * (size_t(1) << B);
    if (shmem > 65536) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_adjoint_reg<T, B, RB>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, int(shmem));
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((k_adjoint_reg<T, B, RB>), dim3(grid), dim3(1 << (B - RB)), shmem, stream,
                       static_cast<cplx<T> *>(psi), static_cast<cplx<T> *>(lam), pd, steps,
                       static_cast<const cplx<double> *>(mats), terms, partials, ntiles, first_mask);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_reg_any(int b, unsigned grid, hipStream_t stream, void *psi, void *lam, const AdjPass &pd,
                          const AdjRStep *steps, const void *mats, const AdjRTerm *terms, double *partials,
                          uint64_t ntiles, unsigned first_mask) {
    switch (b) {
    case 9: return launch_reg_pass<T, 9>(grid, stream, psi, lam, pd, steps, mats, terms, partials, ntiles, first_mask);
    case 10: return launch_reg_pass<T, 10>(grid, stream, psi, lam, pd, steps, mats, terms, partials, ntiles, first_mask);
    case 11: return launch_reg_pass<T, 11>(grid, stream, psi, lam, pd, steps, mats, terms, partials, ntiles, first_mask);
    default: return hipErrorInvalidValue;
    }
}

// Register-resident form of one pass's steps: segments whose dense gates and
// gradient X bits fit kRegBits register bits, joined by relayouts.  Returns
// false when a step does not fit the form (gates on more than 2 qubits).
bool to_register_steps(const AdjStep *st, int nsteps, const std::vector<AdjTerm> &hterms, int b,
                       std::vector<cplx<double>> &hmats, std::vector<AdjRStep> &out, std::vector<AdjRTerm> &rterms,
                       unsigned &first_mask) {
    const size_t out0 = out.size();
    constexpr int RB = kRegBits;
    const unsigned tile_mask = (1u << b) - 1u;
    auto needs = [&](const AdjStep &x) -> unsigned {
        switch (x.kind) {
        case A_DENSE1: return 1u << x.bit[0];
        case A_DENSE2: return (1u << x.bit[0]) | (1u << x.bit[1]);
        case A_GRAD: {
            unsigned m = 0;
            for (int q = 0; q < x.nterms; ++q) m |= hterms[size_t(x.term0 + q)].x;
            return m;
        }
        default: return 0;
        }
    };
    for (int i = 0; i < nsteps; ++i) {
        if (st[i].kind == A_GENERIC) return false;
        if (__builtin_popcount(needs(st[i])) > RB) return false;
    }
    auto fill = [&](unsigned m) {
        for (int bit = b - 1; bit >= 0 && __builtin_popcount(m) < RB; --bit)
            if (!((m >> bit) & 1)) m |= 1u << bit;
        return m & tile_mask;
    };
    // segment masks
    std::vector<std::pair<int, unsigned>> segs;  // (first step, register mask)
    unsigned cur = 0;
    int start = 0;
    for (int i = 0; i < nsteps; ++i) {
        const unsigned q = needs(st[i]);
        if (__builtin_popcount(cur | q) <= RB) {
            cur |= q;
            continue;
        }
        segs.emplace_back(start, fill(cur));
        start = i;
        cur = q;
    }
    segs.emplace_back(start, fill(cur));
    first_mask = segs[0].second;
    auto slot_of = [](unsigned mask, int bit) { return __builtin_popcount(mask & ((1u << bit) - 1u)); };
    size_t si = 0;
    unsigned mask = first_mask;
    for (int i = 0; i < nsteps; ++i) {
        if (si + 1 < segs.size() && segs[si + 1].first == i) {
            ++si;
            mask = segs[si].second;
            AdjRStep r;
            std::memset(&r, 0, sizeof r);
            r.kind = R_RELAYOUT;
            r.mask = mask;
            out.push_back(r);
        }
        const AdjStep &x = st[i];
        AdjRStep r;
        std::memset(&r, 0, sizeof r);
        r.mat = x.mat;
        switch (x.kind) {
        case A_DENSE1: {
            r.kind = R_DENSE1;
            r.a = slot_of(mask, x.bit[0]);
            bool real = true;
            for (int e = 0; e < 4; ++e) real = real && hmats[size_t(x.mat + e)].im == 0.0;
            if (real) r.kind = R_DENSE1R;
            break;
        }
        case A_DENSE2: {
            r.kind = R_DENSE2;
            r.a = slot_of(mask, x.bit[0]);
            r.b = slot_of(mask, x.bit[1]);
            // exact permutations of the cx / swap pattern exchange registers
            // (entries exactly 0 or 1): src[row] = the column of row's 1
            int src[4];
            bool perm = true;
            for (int row = 0; row < 4 && perm; ++row) {
                src[row] = -1;
                for (int col = 0; col < 4; ++col) {
                    const cplx<double> e = hmats[size_t(x.mat + 4 * row + col)];
                    if (e.im != 0.0 || (e.re != 0.0 && e.re != 1.0)) perm = false;
                    if (e.re == 1.0) {
                        if (src[row] >= 0) perm = false;
                        src[row] = col;
                    }
                }
                if (src[row] < 0) perm = false;
            }
            if (perm) {
                const int code = src[0] * 64 + src[1] * 16 + src[2] * 4 + src[3];
                if (code == 0 * 64 + 1 * 16 + 3 * 4 + 2) {  // rows 2, 3 exchanged: cx, control = MSB
                    r.kind = R_PERM2;
                    r.k = 0;
                } else if (code == 0 * 64 + 2 * 16 + 1 * 4 + 3) {  // rows 1, 2 exchanged: swap
                    r.kind = R_PERM2;
                    r.k = 1;
                    const int lo = std::min(r.a, r.b), hi = std::max(r.a, r.b);
                    r.a = lo;
                    r.b = hi;
                }
                break;
            }
            // identity where the MSB is 0 (rows and columns 0, 1), no
            // coupling between the halves: a controlled 2x2 on the LSB
            bool ctrl = true;
            for (int row = 0; row < 4 && ctrl; ++row)
                for (int col = 0; col < 4; ++col) {
                    if (row >= 2 && col >= 2) continue;
                    const cplx<double> e = hmats[size_t(x.mat + 4 * row + col)];
                    const double want = row == col ? 1.0 : 0.0;
                    if (e.re != want || e.im != 0.0) {
                        ctrl = false;
                        break;
                    }
                }
            if (ctrl) {
                r.kind = R_CTRL2;
                const cplx<double> blk[4] = {hmats[size_t(x.mat + 10)], hmats[size_t(x.mat + 11)],
                                             hmats[size_t(x.mat + 14)], hmats[size_t(x.mat + 15)]};
                r.mat = int(hmats.size());
                hmats.insert(hmats.end(), blk, blk + 4);
            }
            break;
        }
        case A_DIAG:
            r.kind = R_DIAG;
            r.k = x.k;
            r.bits[0] = x.bit[0];
            r.bits[1] = x.k == 2 ? x.bit[1] : 0;
            break;
        default: {  // A_GRAD
            r.kind = R_GRAD;
            r.term0 = int(rterms.size());
            r.nterms = x.nterms;
            r.slot = x.slot;
            r.a = 1;  // 1: every term is a Z string
            for (int q = 0; q < x.nterms; ++q)
                if (hterms[size_t(x.term0 + q)].x) r.a = 0;
            if (x.nterms == 1) {
                // one X or Y on a single qubit: the pair form
                const AdjTerm &tm = hterms[size_t(x.term0)];
                if (__builtin_popcount(tm.x) == 1 && (tm.z == 0 || tm.z == tm.x)) {
                    r.kind = R_GRAD1;
                    r.a = slot_of(mask, __builtin_ctz(tm.x));
                    r.b = tm.z ? 1 : 0;
                }
            }
            for (int q = 0; q < x.nterms; ++q) {
                const AdjTerm &tm = hterms[size_t(x.term0 + q)];
                AdjRTerm rt;
                std::memset(&rt, 0, sizeof rt);
                for (int bit = 0; bit < b; ++bit)
                    if ((tm.x >> bit) & 1) rt.xr |= 1u << slot_of(mask, bit);
                rt.xt = tm.x;
                rt.z = tm.z;
                rt.cre = tm.cre;
                rt.cim = tm.cim;
                rterms.push_back(rt);
            }
            break;
        }
        }
        out.push_back(r);
    }
    // Runs of diagonal gates and Z-string gradients: every such gradient is
    // unchanged by diagonal gates (they commute with Z strings and act on
    // both states), so the run becomes its gradients followed by the product
    // of its gates as one phase table.
    std::vector<AdjRStep> res(out.begin(), out.begin() + long(out0));
    size_t i = out0;
    while (i < out.size()) {
        const auto in_run = [&](const AdjRStep &x) { return x.kind == R_DIAG || (x.kind == R_GRAD && x.a == 1); };
        if (!in_run(out[i])) {
            res.push_back(out[i++]);
            continue;
        }
        size_t j = i;
        std::vector<AdjRStep> grads, diags;
        while (j < out.size() && in_run(out[j])) (out[j].kind == R_DIAG ? diags : grads).push_back(out[j++]);
        for (const AdjRStep &g : grads) res.push_back(g);
        if (diags.size() == 1) {
            res.push_back(diags[0]);
        } else if (diags.size() > 1) {
            unsigned m = 0;
            for (const AdjRStep &d : diags) {
                m |= 1u << d.bits[0];
                if (d.k == 2) m |= 1u << d.bits[1];
            }
            const int nb = __builtin_popcount(m);
            std::vector<int> tbits;
            for (int bit = 0; bit < b; ++bit)
                if ((m >> bit) & 1) tbits.push_back(bit);
            AdjRStep tstep;
            std::memset(&tstep, 0, sizeof tstep);
            tstep.kind = R_DTABLE;
            tstep.mask = m;
            tstep.mat = int(hmats.size());
            for (unsigned row = 0; row < (1u << nb); ++row) {
                unsigned tile = 0;  // the row's tile-bit values
                for (int e = 0; e < nb; ++e)
                    if ((row >> e) & 1) tile |= 1u << tbits[size_t(e)];
                double re = 1.0, im = 0.0;
                for (const AdjRStep &d : diags) {
                    const int dim = 1 << d.k;
                    const int dr = d.k == 2 ? int(((tile >> d.bits[0]) & 1) << 1 | ((tile >> d.bits[1]) & 1))
                                            : int((tile >> d.bits[0]) & 1);
                    const cplx<double> ph = hmats[size_t(d.mat + dr * (dim + 1))];
                    const double nr = ph.re * re - ph.im * im, ni = ph.re * im + ph.im * re;
                    re = nr;
                    im = ni;
                }
                hmats.push_back(cplx<double>{re, im});
            }
            res.push_back(tstep);
        }
        i = j;
    }
    // Consecutive one-term Z-string gradients, up to 8 at a time, become one
    // R_ZGRADS step: conj(lam) psi is formed once and the 8 wave sums are one
    // butterfly.  Their terms are copied to consecutive entries.
    std::vector<AdjRStep> grouped(res.begin(), res.begin() + long(out0));
    for (size_t i2 = out0; i2 < res.size();) {
        const auto single_z = [&](const AdjRStep &x) { return x.kind == R_GRAD && x.a == 1 && x.nterms == 1; };
        if (!single_z(res[i2])) {
            grouped.push_back(res[i2++]);
            continue;
        }
        size_t j = i2;
        while (j < res.size() && j - i2 < 8 && single_z(res[j])) ++j;
        if (j - i2 == 1) {
            grouped.push_back(res[i2++]);
            continue;
        }
        AdjRStep gz;
        std::memset(&gz, 0, sizeof gz);
        gz.kind = R_ZGRADS;
        gz.term0 = int(rterms.size());
        gz.nterms = int(j - i2);
        for (size_t q = i2; q < j; ++q) {
            AdjRTerm t = rterms[size_t(res[q].term0)];
            t.slot = res[q].slot;
            rterms.push_back(t);
        }
        grouped.push_back(gz);
        i2 = j;
    }
    // A rotation's pair-form gradient followed by its inverse on the same
    // register slot (real entries) is one step.
    std::vector<AdjRStep> merged(grouped.begin(), grouped.begin() + long(out0));
    for (size_t i3 = out0; i3 < grouped.size(); ++i3) {
        const AdjRStep &g = grouped[i3];
        if (g.kind == R_GRAD1 && i3 + 1 < grouped.size() && grouped[i3 + 1].kind == R_DENSE1R &&
            grouped[i3 + 1].a == g.a) {
            AdjRStep m = g;
            m.kind = R_GRAD1R;
            m.mat = grouped[i3 + 1].mat;
            merged.push_back(m);
            ++i3;
            continue;
        }
        merged.push_back(g);
    }
    // Consecutive rotation steps on distinct register slots (2 or 3 of them)
    // become one R_ROT3 step: rotations on different qubits commute, so they
    // run in slot order, and an absent slot takes the identity with no
    // gradient.  A Z-gradient group absorbs the phase table that follows it.
    int ident = -1;  // mats offset of the real 2x2 identity
    std::vector<AdjRStep> fused(merged.begin(), merged.begin() + long(out0));
    for (size_t i4 = out0; i4 < merged.size();) {
        const AdjRStep &g = merged[i4];
        if (g.kind == R_ZGRADS && i4 + 1 < merged.size() && merged[i4 + 1].kind == R_DTABLE) {
            AdjRStep z = g;
            z.b = 1;
            z.mask = merged[i4 + 1].mask;
            z.mat = merged[i4 + 1].mat;
            fused.push_back(z);
            i4 += 2;
            continue;
        }
        size_t j = i4;
        unsigned used = 0;
        while (j < merged.size() && merged[j].kind == R_GRAD1R && !((used >> merged[j].a) & 1) && RB == 3) {
            used |= 1u << merged[j].a;
            ++j;
        }
        if (j - i4 < 2) {
            fused.push_back(g);
            ++i4;
            continue;
        }
        if (ident < 0) {
            ident = int(hmats.size());
            hmats.push_back(cplx<double>{1.0, 0.0});
            hmats.push_back(cplx<double>{0.0, 0.0});
            hmats.push_back(cplx<double>{0.0, 0.0});
            hmats.push_back(cplx<double>{1.0, 0.0});
        }
        AdjRStep r;
        std::memset(&r, 0, sizeof r);
        r.kind = R_ROT3;
        r.term0 = int(rterms.size());
        r.nterms = 3;
        for (int m = 0; m < 3; ++m) {
            AdjRTerm t;
            std::memset(&t, 0, sizeof t);
            t.xt = unsigned(ident);
            t.slot = -1;
            for (size_t q = i4; q < j; ++q)
                if (merged[q].a == m) {
                    t = rterms[size_t(merged[q].term0)];
                    t.z = unsigned(merged[q].b);
                    t.xt = unsigned(merged[q].mat);
                    t.slot = merged[q].slot;
                }
            rterms.push_back(t);
        }
        fused.push_back(r);
        i4 = j;
    }
    out.swap(fused);
    return true;
}

}  // namespace

extern "C" {

int dq_adjoint_sweep(dq_state *psi, dq_state *lam, int n_ops, const int *k, const int *qubits, const double *mats,
                     const int *n_terms, const uint64_t *xmask, const uint64_t *zmask, const double *coeff,
                     double *grad) {
    if (!psi || !lam || (n_ops > 0 && (!k || !qubits || !mats || !n_terms || !grad)))
        return afail(DQ_ERR_ARG, "null argument");
    if (psi == lam) return afail(DQ_ERR_ARG, "psi and lam must be distinct states");
    if (psi->n != lam->n || psi->dtype != lam->dtype) return afail(DQ_ERR_ARG, "states differ in size or dtype");
    const int n = psi->n;
    if (n < 8) return afail(DQ_ERR_ARG, "the tiled adjoint sweep needs at least 8 qubits");
    if (n_ops < 0) return afail(DQ_ERR_ARG, "n_ops must be >= 0");
    // validate and lay out the ops
    std::vector<size_t> mat_off(size_t(n_ops) + 1, 0), term_off(size_t(n_ops) + 1, 0);
    std::vector<uint64_t> qmask(size_t(n_ops), 0);
    for (int l = 0; l < n_ops; ++l) {
        if (k[l] < 1 || k[l] > 5) return afail(DQ_ERR_ARG, "gate arity must be 1..5");
        for (int j = 0; j < k[l]; ++j) {
            const int q = qubits[5 * l + j];
            if (q < 0 || q >= n) return afail(DQ_ERR_ARG, "qubit index out of range");
            if ((qmask[size_t(l)] >> q) & 1) return afail(DQ_ERR_ARG, "gate qubits must be distinct");
            qmask[size_t(l)] |= uint64_t(1) << q;
        }
        if (n_terms[l] < 0 || n_terms[l] > 64) return afail(DQ_ERR_ARG, "generator term count must be 0..64");
        if (n_terms[l] > 0 && (!xmask || !zmask || !coeff)) return afail(DQ_ERR_ARG, "null argument");
        mat_off[size_t(l) + 1] = mat_off[size_t(l)] + (size_t(1) << (2 * k[l]));
        term_off[size_t(l) + 1] = term_off[size_t(l)] + size_t(n_terms[l]);
        for (size_t t = term_off[size_t(l)]; t < term_off[size_t(l) + 1]; ++t)
            if ((xmask[t] | zmask[t]) & ~qmask[size_t(l)])
                return afail(DQ_ERR_ARG, "a generator term acts outside its gate's qubits");
    }
    for (int l = 0; l < n_ops; ++l) grad[l] = 0.0;
    if (n_ops == 0) return DQ_OK;
    if (int rc = dq_synchronize(psi)) return rc;
    if (int rc = dq_synchronize(lam)) return rc;
    // the sweep rewrites both states in place
    psi->zero_q = 0;
    lam->zero_q = 0;
    const int b = std::min(n, 11);
    const uint64_t low = (uint64_t(1) << kLowBits) - 1;
    // passes: ops in reverse order; an op joins the pass if its qubits fit
    // the tile and miss every qubit of an op left for a later pass.  The
    // tile's qubits are chosen first: starting from the low block, add the
    // qubits of the waiting op whose addition lets the pass take the most
    // ops over the next 256 (ties: fewer qubits, then the earlier op).
    std::vector<std::vector<int>> passes;
    std::vector<uint64_t> pass_S;
    {
        std::vector<int> remaining;
        for (int l = n_ops - 1; l >= 0; --l) remaining.push_back(l);
        auto fitting = [&](uint64_t S, std::vector<uint64_t> *grow) {
            uint64_t blocked = 0;
            int taken = 0;
            const size_t end = std::min<size_t>(remaining.size(), 256);
            for (size_t r = 0; r < end; ++r) {
                const uint64_t q = qmask[size_t(remaining[r])];
                if (q & blocked) {  // held back behind a deferred op, as the pass builder does
                    blocked |= q;
                    continue;
                }
                if ((q & ~S) == 0) {
                    ++taken;
                    continue;
                }
                blocked |= q;
                if (grow && std::find(grow->begin(), grow->end(), q & ~S) == grow->end()) grow->push_back(q & ~S);
            }
            return taken;
        };
        std::vector<uint64_t> grow;
        while (!remaining.empty()) {
            uint64_t S = low;
            for (;;) {
                grow.clear();
                const int base = fitting(S, &grow);
                int best_gain = 0, best_bits = 0;
                uint64_t best = 0;
                for (uint64_t extra : grow) {
                    const int bits = __builtin_popcountll(extra);
                    if (__builtin_popcountll(S | extra) > b) continue;
                    const int gain = fitting(S | extra, nullptr) - base;
                    if (gain > best_gain || (gain == best_gain && gain > 0 && bits < best_bits)) {
                        best_gain = gain;
                        best_bits = bits;
                        best = extra;
                    }
                }
                if (!best) break;
                S |= best;
            }
            std::vector<int> take, rest;
            auto build_pass = [&](uint64_t S0) {
                take.clear();
                rest.clear();
                S = S0;
                uint64_t blocked = 0;
                for (int l : remaining) {
                    const uint64_t q = qmask[size_t(l)];
                    if (!(q & blocked) && __builtin_popcountll(S | q) <= b) {
                        S |= q;
                        take.push_back(l);
                    } else {
                        blocked |= q;
                        rest.push_back(l);
                    }
                }
            };
            build_pass(S);
            if (take.empty()) build_pass(low);  // the chosen tile admits an op by construction; kept as a guard
            if (take.empty()) return afail(DQ_ERR_ARG, "a gate does not fit the tile");
            // fill the tile with the lowest unused qubits
            for (int q = 0; q < n && __builtin_popcountll(S) < b; ++q) S |= uint64_t(1) << q;
            passes.push_back(std::move(take));
            pass_S.push_back(S);
            remaining.swap(rest);
        }
    }
    // device buffers: matrices (complex128), steps, terms, partials, grads
    std::vector<cplx<double>> hmats(mat_off[size_t(n_ops)]);
    std::memcpy(hmats.data(), mats, hmats.size() * sizeof(cplx<double>));
    const uint64_t ntiles = uint64_t(1) << (n - b);
    size_t max_slots = 1;
    std::vector<AdjStep> hsteps;
    std::vector<AdjTerm> hterms;
    std::vector<AdjPass> hpasses;
    std::vector<AdjRStep> rsteps;  // register-resident form of the passes that allow it
    std::vector<AdjRTerm> rterms;
    std::vector<AdjPass> rpasses;
    std::vector<unsigned> first_masks;
    std::vector<bool> use_reg;
    std::vector<std::vector<int>> pass_grad_ops;
    for (size_t p = 0; p < passes.size(); ++p) {
        AdjPass pd;
        std::memset(&pd, 0, sizeof pd);
        pd.b = b;
        int pos[64];
        int tb = 0, ob = 0;
        for (int q = 0; q < n; ++q) {
            if ((pass_S[p] >> q) & 1) {
                pd.tq[tb] = q;
                pos[q] = tb++;
            } else {
                pd.oq[ob++] = q;
                pos[q] = -1;
            }
        }
        pd.n_outer = ob;
        pd.step0 = int(hsteps.size());
        std::vector<int> gops;
        for (int l : passes[p]) {
            if (n_terms[l] > 0) {
                AdjStep g;
                std::memset(&g, 0, sizeof g);
                g.kind = A_GRAD;
                g.term0 = int(hterms.size());
                g.nterms = n_terms[l];
                g.slot = int(gops.size());
                for (size_t t = term_off[size_t(l)]; t < term_off[size_t(l) + 1]; ++t) {
                    AdjTerm tm;
                    tm.x = tm.z = 0;
                    for (int q = 0; q < n; ++q) {
                        if ((xmask[t] >> q) & 1) tm.x |= 1u << pos[q];
                        if ((zmask[t] >> q) & 1) tm.z |= 1u << pos[q];
                    }
                    const int ny = __builtin_popcountll(xmask[t] & zmask[t]) & 3;
                    double re = coeff[2 * t], im = coeff[2 * t + 1];
                    for (int j = 0; j < ny; ++j) {  // multiply by i
                        const double tmp = re;
                        re = -im;
                        im = tmp;
                    }
                    tm.cre = re;
                    tm.cim = im;
                    hterms.push_back(tm);
                }
                hsteps.push_back(g);
                gops.push_back(l);
            }
            AdjStep s;
            std::memset(&s, 0, sizeof s);
            s.k = k[l];
            for (int j = 0; j < k[l]; ++j) s.bit[j] = pos[qubits[5 * l + j]];
            s.mat = int(mat_off[size_t(l)]);
            bool diag = true;
            const int dim = 1 << k[l];
            for (int r = 0; r < dim && diag; ++r)
                for (int c = 0; c < dim; ++c)
                    if (r != c && (hmats[mat_off[size_t(l)] + size_t(r * dim + c)].re != 0.0 ||
                                   hmats[mat_off[size_t(l)] + size_t(r * dim + c)].im != 0.0)) {
                        diag = false;
                        break;
                    }
            s.kind = (diag && k[l] <= 2) ? A_DIAG : k[l] == 1 ? A_DENSE1 : k[l] == 2 ? A_DENSE2 : A_GENERIC;
            hsteps.push_back(s);
        }
        pd.nsteps = int(hsteps.size()) - pd.step0;
        max_slots = std::max(max_slots, gops.size());
        AdjPass rp = pd;
        rp.step0 = int(rsteps.size());
        unsigned fm = 0;
        const bool reg =
            b >= 9 && to_register_steps(hsteps.data() + pd.step0, pd.nsteps, hterms, b, hmats, rsteps, rterms, fm);
        if (!reg) rsteps.resize(size_t(rp.step0));
        rp.nsteps = int(rsteps.size()) - rp.step0;
        use_reg.push_back(reg);
        rpasses.push_back(rp);
        first_masks.push_back(fm);
        hpasses.push_back(pd);
        pass_grad_ops.push_back(std::move(gops));
    }
    const size_t slots_total = std::max<size_t>(1, [&] {
        size_t s = 0;
        for (const auto &g : pass_grad_ops) s += g.size();
        return s;
    }());
    // One device block (grow-only, kept on psi across calls) holds the tables
    // and the partials; the tables are staged in one pinned host block and
    // uploaded with one copy.  Every region starts on a 256-byte boundary.
    auto up = [](size_t v) { return (v + 255) & ~size_t(255); };
    const size_t b_mats = up(std::max<size_t>(1, hmats.size()) * sizeof(cplx<double>));
    const size_t b_steps = up(std::max<size_t>(1, hsteps.size()) * sizeof(AdjStep));
    const size_t b_terms = up(std::max<size_t>(1, hterms.size()) * sizeof(AdjTerm));
    const size_t b_rsteps = up(std::max<size_t>(1, rsteps.size()) * sizeof(AdjRStep));
    const size_t b_rterms = up(std::max<size_t>(1, rterms.size()) * sizeof(AdjRTerm));
    const size_t b_tables = b_mats + b_steps + b_terms + b_rsteps + b_rterms;
    const size_t b_part = up(max_slots * ntiles * 4 * sizeof(double));
    const size_t b_grads = up(slots_total * sizeof(double));
    const size_t b_chunks = up(max_slots * kReduceChunks * sizeof(double));
    const size_t b_total = b_tables + b_part + b_grads + b_chunks;
    if (psi->d_adj_cap < b_total) {
        if (psi->d_adj) (void)hipFree(psi->d_adj);
        psi->d_adj = nullptr;
        psi->d_adj_cap = 0;
        if (hipMalloc(&psi->d_adj, b_total) != hipSuccess) {
            psi->d_adj = nullptr;
            return afail(DQ_ERR_OOM, "hipMalloc of the sweep's buffers failed");
        }
        psi->d_adj_cap = b_total;
    }
    if (psi->h_adj_cap < b_tables + b_grads) {
        if (psi->h_adj) (void)hipHostFree(psi->h_adj);
        psi->h_adj = nullptr;
        psi->h_adj_cap = 0;
        if (hipHostMalloc(reinterpret_cast<void **>(&psi->h_adj), b_tables + b_grads, hipHostMallocDefault) !=
            hipSuccess) {
            psi->h_adj = nullptr;
            return afail(DQ_ERR_OOM, "hipHostMalloc of the sweep's staging buffer failed");
        }
        psi->h_adj_cap = b_tables + b_grads;
    }
    unsigned char *dbase = static_cast<unsigned char *>(psi->d_adj);
    void *d_mats = dbase;
    void *d_steps = dbase + b_mats;
    void *d_terms = dbase + b_mats + b_steps;
    void *d_rsteps = dbase + b_mats + b_steps + b_terms;
    void *d_rterms = dbase + b_mats + b_steps + b_terms + b_rsteps;
    double *d_part = reinterpret_cast<double *>(dbase + b_tables);
    double *d_grads = reinterpret_cast<double *>(dbase + b_tables + b_part);
    double *d_chunks = reinterpret_cast<double *>(dbase + b_tables + b_part + b_grads);
    {
        unsigned char *h = psi->h_adj;
        std::memcpy(h, hmats.data(), hmats.size() * sizeof(cplx<double>));
        std::memcpy(h + b_mats, hsteps.data(), hsteps.size() * sizeof(AdjStep));
        if (!hterms.empty()) std::memcpy(h + b_mats + b_steps, hterms.data(), hterms.size() * sizeof(AdjTerm));
        if (!rsteps.empty())
            std::memcpy(h + b_mats + b_steps + b_terms, rsteps.data(), rsteps.size() * sizeof(AdjRStep));
        if (!rterms.empty())
            std::memcpy(h + b_mats + b_steps + b_terms + b_rsteps, rterms.data(), rterms.size() * sizeof(AdjRTerm));
    }
    hipStream_t stream = psi->stream;
    int rc = DQ_OK;
    auto run = [&]() -> int {
        ADJ_TRY(hipMemcpyAsync(d_mats, psi->h_adj, b_tables, hipMemcpyHostToDevice, stream));
        // a power of two: the register kernel derives tile offsets from it
        const unsigned grid = unsigned(std::min<uint64_t>(ntiles, 8192));
        if (grid & (grid - 1)) return afail(DQ_ERR_ARG, "grid is not a power of two");
        size_t slot_base = 0;
        for (size_t p = 0; p < hpasses.size(); ++p) {
            hipError_t e;
            uint64_t count = ntiles;  // partials per gradient slot
            if (use_reg[p]) {
                count = ntiles * uint64_t((1 << (b - kRegBits)) / 64);
                e = psi->dtype == DQ_C128
                        ? launch_reg_any<double>(b, grid, stream, psi->amps, lam->amps, rpasses[p],
                                                 static_cast<const AdjRStep *>(d_rsteps), d_mats,
                                                 static_cast<const AdjRTerm *>(d_rterms), d_part, ntiles, first_masks[p])
                        : launch_reg_any<float>(b, grid, stream, psi->amps, lam->amps, rpasses[p],
                                                static_cast<const AdjRStep *>(d_rsteps), d_mats,
                                                static_cast<const AdjRTerm *>(d_rterms), d_part, ntiles, first_masks[p]);
            } else {
                e = psi->dtype == DQ_C128
                        ? launch_any<double>(b, grid, stream, psi->amps, lam->amps, hpasses[p],
                                             static_cast<const AdjStep *>(d_steps), d_mats,
                                             static_cast<const AdjTerm *>(d_terms), d_part, ntiles)
                        : launch_any<float>(b, grid, stream, psi->amps, lam->amps, hpasses[p],
                                            static_cast<const AdjStep *>(d_steps), d_mats,
                                            static_cast<const AdjTerm *>(d_terms), d_part, ntiles);
            }
            ADJ_TRY(e);
            const size_t ns = pass_grad_ops[p].size();
            if (ns) {
                hipLaunchKernelGGL(k_adjoint_reduce_chunks, dim3(unsigned(ns), kReduceChunks), dim3(kThreads), 0,
                                   stream, d_part, count, d_chunks);
                ADJ_TRY(hipGetLastError());
                hipLaunchKernelGGL(k_adjoint_reduce, dim3(unsigned(ns)), dim3(kThreads), 0, stream, d_chunks,
                                   uint64_t(kReduceChunks), d_grads + slot_base);
                ADJ_TRY(hipGetLastError());
            }
            slot_base += ns;
            psi->launches++;
            psi->sweeps++;
            psi->swept_amps += uint64_t(1) << n;
        }
        double *hg = reinterpret_cast<double *>(psi->h_adj + b_tables);
        ADJ_TRY(hipMemcpyAsync(hg, d_grads, slots_total * sizeof(double), hipMemcpyDeviceToHost, stream));
        ADJ_TRY(hipStreamSynchronize(stream));
        size_t s = 0;
        for (const auto &g : pass_grad_ops)
            for (int l : g) grad[l] = hg[s++];
        return DQ_OK;
    };
    rc = run();
    (void)hipStreamSynchronize(stream);
    return rc;
}

}  // extern "C"
