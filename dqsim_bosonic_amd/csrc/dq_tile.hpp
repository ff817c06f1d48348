// dq_tile.hpp -- register-resident LDS-tiled sweep kernel for gfx950.
//
// One workgroup owns a tile of 2^B amplitudes (B = tile qubits).  Each of the
// 2^(B-4) threads holds 16 amplitudes in registers.  Which tile qubits index a
// thread's 16 registers (the "register bits", 4 of them) is the *layout*;
// the remaining B-4 tile qubits index the thread.  Gates whose matrix qubits
// are register bits are applied entirely in registers; a layout change is one
// LDS round trip (write all, barrier, read all) with an XOR swizzle that
// spreads strided layouts over the banks (the per-register offsets come from a
// host table).  The host plans, per pass, the layout every op runs in, so most
// gates cost no LDS traffic and no barrier.
//
// Arithmetic: each output is the reference's row sum (src/engine.rs:58-59,
// 146-151) in the same column order.  The host picks a variant per op:
// general, real (only the real parts of the entries multiply), diagonal or
// anti-diagonal (products of zero entries dropped).  Dropped terms are +-0
// and do not change any sum, so the values equal the reference's up to the
// sign of a zero.
#pragma once
#include "dq_device.hpp"

namespace dq {

template <typename T, int R>
struct Layout {
    unsigned tpart;  // tile-index bits contributed by this thread
    int rb;          // the layout (R register bits, 4 bits each, ascending)
    // tile-index bits contributed by register j (uniform), computed where an
    // op needs them rather than held for the whole pass
    __device__ __forceinline__ unsigned jp(int j) const {
        unsigned v = 0;
#pragma unroll
        for (int m = 0; m < R; ++m)
            if ((j >> m) & 1) v |= 1u << ((rb >> (4 * m)) & 15);
        return v;
    }
};

template <int B>
__device__ __forceinline__ unsigned swz(unsigned i) {
    unsigned s = 0;
#pragma unroll
    for (int sh = 4; sh < B; sh += 4) s ^= (i >> sh) & 15u;
    return i ^ s;
}

template <typename T, int B, int R>
__device__ __forceinline__ void make_layout(Layout<T, R> &L, int rb, unsigned t) {
    int rbit[R];
#pragma unroll
    for (int m = 0; m < R; ++m) rbit[m] = (rb >> (4 * m)) & 15;
    L.rb = rb;
    // thread bit c -> the c-th non-register tile bit, ascending: insert a
    // zero bit into t at each register bit, in ascending order (the host
    // packs the register bits of a layout ascending)
    unsigned tp = t;
#pragma unroll
    for (int m = 0; m < R; ++m) {
        const unsigned lo = (1u << rbit[m]) - 1;
        tp = (tp & lo) | ((tp & ~lo) << 1);
    }
    L.tpart = tp;
}

// Global-offset map of tile-index bits: additive over disjoint bit sets.
template <int B>
__device__ __forceinline__ uint64_t gmap(unsigned i, const TilePass &pd) {
    uint64_t g = 0;
#pragma unroll
    for (int bpos = 0; bpos < B; ++bpos)
        if ((i >> bpos) & 1) g |= uint64_t(1) << pd.tq[bpos];
    return g;
}

// The dispatch header of an op (the first 12 words of TileOp), read as one
// block so that the next op's header can be fetched while this op computes.
struct OpHead {
    int kind, relayout, rb, p, q, mat, variant, nsel, lay, sel[3];
};
static_assert(offsetof(TileOp, sel) == 9 * sizeof(int) && offsetof(TileOp, k) == 12 * sizeof(int),
              "TileOp must start with its 12-word dispatch header");

__device__ __forceinline__ OpHead op_head(const TileOp *op) {
    const int *w = reinterpret_cast<const int *>(op);
    OpHead h;
    h.kind = w[0];
    h.relayout = w[1];
    h.rb = w[2];
    h.p = w[3];
    h.q = w[4];
    h.mat = w[5];
    h.variant = w[6];
    h.nsel = w[7];
    h.lay = w[8];
    h.sel[0] = w[9];
    h.sel[1] = w[10];
    h.sel[2] = w[11];
    return h;
}

// ---- register ops ----

template <typename T>
__device__ __forceinline__ cplx<T> rmul(cplx<T> u, cplx<T> v) {  // u real: u.im products dropped
    return cplx<T>{u.re * v.re, u.re * v.im};
}

// Per-lane choice among four uniform values, as two levels of selects on the
// bits of e (a select chain on e's value can be lowered to a stack table).
template <typename T>
__device__ __forceinline__ T sel(bool c, T a, T b) {
    return c ? a : b;
}
template <typename T>
__device__ __forceinline__ cplx<T> pick4(int e, cplx<T> m0, cplx<T> m1, cplx<T> m2, cplx<T> m3) {
    const bool b0 = (e & 1) != 0, b1 = (e & 2) != 0;
    const T lo_re = sel(b0, m1.re, m0.re), lo_im = sel(b0, m1.im, m0.im);
    const T hi_re = sel(b0, m3.re, m2.re), hi_im = sel(b0, m3.im, m2.im);
    return cplx<T>{sel(b1, hi_re, lo_re), sel(b1, hi_im, lo_im)};
}

// Contracted-kernel arithmetic (FAST = 1): explicit FMA chains.  u * x is
// (fma(ur, xr, -ui xi), fma(ur, xi, ui xr)); acc + u * x is four FMAs.  The
// exact kernel (FAST = 0) keeps the reference's separately rounded products
// and sums.
template <typename T>
__device__ __forceinline__ T dq_fma(T a, T b, T c) {
    if constexpr (sizeof(T) == 8) return __builtin_fma(a, b, c);
    else return __builtin_fmaf(a, b, c);
}
template <typename T>
__device__ __forceinline__ cplx<T> fmul(cplx<T> u, cplx<T> x) {
    return cplx<T>{dq_fma(u.re, x.re, -(u.im * x.im)), dq_fma(u.re, x.im, u.im * x.re)};
}
template <typename T>
__device__ __forceinline__ cplx<T> fmac(cplx<T> acc, cplx<T> u, cplx<T> x) {
    return cplx<T>{dq_fma(-u.im, x.im, dq_fma(u.re, x.re, acc.re)), dq_fma(u.im, x.re, dq_fma(u.re, x.im, acc.im))};
}
template <typename T>
__device__ __forceinline__ cplx<T> frmac(cplx<T> acc, cplx<T> u, cplx<T> x) {  // u real
    return cplx<T>{dq_fma(u.re, x.re, acc.re), dq_fma(u.re, x.im, acc.im)};
}

template <int P>
__device__ __forceinline__ constexpr int pair_lo(int jj) {
    return ((jj >> P) << (P + 1)) | (jj & ((1 << P) - 1));
}

// 2x2 on register slot P, pairs (j0, j0 | 1 << P); out = u00*x + u01*y, u10*x + u11*y.
template <typename T, int P, int V, int FAST, int R>
__device__ __forceinline__ void reg_dense1(cplx<T> (&a)[1 << R], const cplx<T> *__restrict__ m) {
    const cplx<T> u00 = m[0], u01 = m[1], u10 = m[2], u11 = m[3];
#pragma unroll
    for (int jj = 0; jj < (1 << (R - 1)); ++jj) {
        const int j0 = pair_lo<P>(jj), j1 = j0 | (1 << P);
        const cplx<T> x = a[j0], y = a[j1];
        if (FAST && V == V_GEN) {
            a[j0] = fmac(fmul(u01, y), u00, x);
            a[j1] = fmac(fmul(u11, y), u10, x);
        } else if (FAST && V == V_REAL) {
            a[j0] = frmac(cplx<T>{u01.re * y.re, u01.re * y.im}, u00, x);
            a[j1] = frmac(cplx<T>{u11.re * y.re, u11.re * y.im}, u10, x);
        } else if (V == V_GEN) {
            a[j0] = cadd(cmul(u00, x), cmul(u01, y));
            a[j1] = cadd(cmul(u10, x), cmul(u11, y));
        } else if (V == V_REAL) {
            a[j0] = cadd(rmul(u00, x), rmul(u01, y));
            a[j1] = cadd(rmul(u10, x), rmul(u11, y));
        } else if (V == V_DIAG) {
            a[j0] = cmul(u00, x);
            a[j1] = cmul(u11, y);
        } else {
            a[j0] = cmul(u01, y);
            a[j1] = cmul(u10, x);
        }
    }
}

// Controlled 2x2 on slot P: the pair changes only where the control pattern holds.
template <typename T, int P, int V, int FAST, int R>
__device__ __forceinline__ void reg_ctrl1(cplx<T> (&a)[1 << R], const cplx<T> *__restrict__ m, const Layout<T, R> &L,
                                          unsigned cmask, unsigned cval) {
    const cplx<T> u00 = m[0], u01 = m[1], u10 = m[2], u11 = m[3];
#pragma unroll
    for (int jj = 0; jj < (1 << (R - 1)); ++jj) {
        const int j0 = pair_lo<P>(jj), j1 = j0 | (1 << P);
        const cplx<T> x = a[j0], y = a[j1];
        cplx<T> n0, n1;
        if (V == V_ANTI) {
            n0 = cmul(u01, y);
            n1 = cmul(u10, x);
        } else if (FAST) {
            n0 = fmac(fmul(u01, y), u00, x);
            n1 = fmac(fmul(u11, y), u10, x);
        } else {
            n0 = cadd(cmul(u00, x), cmul(u01, y));
            n1 = cadd(cmul(u10, x), cmul(u11, y));
        }
        const bool on = ((L.tpart | L.jp(j0)) & cmask) == cval;
        a[j0] = on ? n0 : x;
        a[j1] = on ? n1 : y;
    }
}

// 4x4 on register slots P (matrix MSB) and Q: acc = 0; acc += U[r][c] * v[c].
template <typename T, int P, int Q, int V, int FAST, int R>
__device__ __forceinline__ void reg_dense2(cplx<T> (&a)[1 << R], const cplx<T> *__restrict__ m) {
    constexpr int LO = P < Q ? P : Q, HI = P < Q ? Q : P;
    if constexpr (FAST && V != V_DIAG) {
        // Contracted: rows 0-1 of every group, then rows 2-3, so that only
        // two matrix rows (32 scalar registers for complex128) are live at
        // once; with all four rows live the scalar registers spill and every
        // op reloads them (10-16 v_readlane per op).  Per output the FMA
        // chain is unchanged: column 3 product, then columns 2, 1, 0.
        // V_ROWPH / V_COLPH: real rows (the real parts of the entries), with
        // each output (V_ROWPH) or each input (V_COLPH) times its phase.
        constexpr bool kReal = V == V_REAL || V == V_ROWPH || V == V_COLPH;
        if constexpr (V == V_COLPH) {
#pragma unroll
            for (int g = 0; g < (1 << (R - 2)); ++g) {
                int base = g;
                base = ((base >> LO) << (LO + 1)) | (base & ((1 << LO) - 1));
                base = ((base >> HI) << (HI + 1)) | (base & ((1 << HI) - 1));
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const int j = base | (((c >> 1) & 1) << P) | ((c & 1) << Q);
                    a[j] = fmul(cplx<T>{m[4 * c].im, m[4 * c + 1].im}, a[j]);
                }
            }
        }
        cplx<T> out[1 << R];
#pragma unroll
        for (int half = 0; half < 2; ++half) {
#pragma unroll
            for (int g = 0; g < (1 << (R - 2)); ++g) {
                int base = g;
                base = ((base >> LO) << (LO + 1)) | (base & ((1 << LO) - 1));
                base = ((base >> HI) << (HI + 1)) | (base & ((1 << HI) - 1));
                cplx<T> v[4];
#pragma unroll
                for (int c = 0; c < 4; ++c) v[c] = a[base | (((c >> 1) & 1) << P) | ((c & 1) << Q)];
#pragma unroll
                for (int r = 2 * half; r < 2 * half + 2; ++r) {
                    cplx<T> acc;
                    if (kReal) {
                        acc = cplx<T>{m[r * 4 + 3].re * v[3].re, m[r * 4 + 3].re * v[3].im};
#pragma unroll
                        for (int c = 2; c >= 0; --c) acc = frmac(acc, m[r * 4 + c], v[c]);
                        if (V == V_ROWPH) acc = fmul(cplx<T>{m[4 * r].im, m[4 * r + 1].im}, acc);
                    } else {
                        acc = fmul(m[r * 4 + 3], v[3]);
#pragma unroll
                        for (int c = 2; c >= 0; --c) acc = fmac(acc, m[r * 4 + c], v[c]);
                    }
                    out[base | (((r >> 1) & 1) << P) | ((r & 1) << Q)] = acc;
                }
            }
        }
#pragma unroll
        for (int j = 0; j < (1 << R); ++j) a[j] = out[j];
        return;
    }
#pragma unroll
    for (int g = 0; g < (1 << (R - 2)); ++g) {
        int base = g;
        base = ((base >> LO) << (LO + 1)) | (base & ((1 << LO) - 1));
        base = ((base >> HI) << (HI + 1)) | (base & ((1 << HI) - 1));
        cplx<T> v[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) v[c] = a[base | (((c >> 1) & 1) << P) | ((c & 1) << Q)];
        if (V == V_COLPH) {  // (the host sends the phase forms to contracted kernels only)
#pragma unroll
            for (int c = 0; c < 4; ++c) v[c] = cmul(cplx<T>{m[4 * c].im, m[4 * c + 1].im}, v[c]);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            cplx<T> acc;
            if (V == V_DIAG) {
                acc = cmul(m[r * 5], v[r]);
            } else if (V == V_REAL || V == V_ROWPH || V == V_COLPH) {
                acc = rmul(m[r * 4], v[0]);
#pragma unroll
                for (int c = 1; c < 4; ++c) acc = cadd(acc, rmul(m[r * 4 + c], v[c]));
                if (V == V_ROWPH) acc = cmul(cplx<T>{m[4 * r].im, m[4 * r + 1].im}, acc);
            } else {
                acc = cmul(m[r * 4], v[0]);
#pragma unroll
                for (int c = 1; c < 4; ++c) acc = cadd(acc, cmul(m[r * 4 + c], v[c]));
            }
            a[base | (((r >> 1) & 1) << P) | ((r & 1) << Q)] = acc;
        }
    }
}

template <typename T, int P, int FAST, int R>
__device__ __forceinline__ void dense1_v(cplx<T> (&a)[1 << R], int v, const cplx<T> *m) {
    switch (v) {
    case V_REAL: reg_dense1<T, P, V_REAL, FAST, R>(a, m); break;
    case V_DIAG: reg_dense1<T, P, V_DIAG, FAST, R>(a, m); break;
    case V_ANTI: reg_dense1<T, P, V_ANTI, FAST, R>(a, m); break;
    default: reg_dense1<T, P, V_GEN, FAST, R>(a, m); break;
    }
}

// Register slot p < R (the host packs only register bits of the layout).
template <typename T, int FAST, int R>
__device__ __forceinline__ void dispatch_dense1(cplx<T> (&a)[1 << R], int p, int v, const cplx<T> *m) {
    switch (p) {
    case 0: dense1_v<T, 0, FAST, R>(a, v, m); break;
    case 1: dense1_v<T, 1, FAST, R>(a, v, m); break;
    case 2:
        if constexpr (R > 3) {
            dense1_v<T, 2, FAST, R>(a, v, m);
            break;
        }
        [[fallthrough]];
    default: dense1_v<T, R - 1, FAST, R>(a, v, m); break;
    }
}

template <typename T, int P, int FAST, int R>
__device__ __forceinline__ void ctrl1_v(cplx<T> (&a)[1 << R], int v, const cplx<T> *m, const Layout<T, R> &L,
                                        unsigned cm, unsigned cv) {
    if (v == V_ANTI) reg_ctrl1<T, P, V_ANTI, FAST, R>(a, m, L, cm, cv);
    else reg_ctrl1<T, P, V_GEN, FAST, R>(a, m, L, cm, cv);
}

template <typename T, int FAST, int R>
__device__ __forceinline__ void dispatch_ctrl1(cplx<T> (&a)[1 << R], int p, int v, const cplx<T> *m,
                                               const Layout<T, R> &L, unsigned cmask, unsigned cval) {
    switch (p) {
    case 0: ctrl1_v<T, 0, FAST, R>(a, v, m, L, cmask, cval); break;
    case 1: ctrl1_v<T, 1, FAST, R>(a, v, m, L, cmask, cval); break;
    case 2:
        if constexpr (R > 3) {
            ctrl1_v<T, 2, FAST, R>(a, v, m, L, cmask, cval);
            break;
        }
        [[fallthrough]];
    default: ctrl1_v<T, R - 1, FAST, R>(a, v, m, L, cmask, cval); break;
    }
}

template <typename T, int P, int Q, int FAST, int R>
__device__ __forceinline__ void dense2_v(cplx<T> (&a)[1 << R], int v, const cplx<T> *m) {
    switch (v) {
    case V_REAL: reg_dense2<T, P, Q, V_REAL, FAST, R>(a, m); break;
    case V_DIAG: reg_dense2<T, P, Q, V_DIAG, FAST, R>(a, m); break;
    case V_ROWPH: reg_dense2<T, P, Q, V_ROWPH, FAST, R>(a, m); break;
    case V_COLPH: reg_dense2<T, P, Q, V_COLPH, FAST, R>(a, m); break;
    default: reg_dense2<T, P, Q, V_GEN, FAST, R>(a, m); break;
    }
}

// Slots p != q, both < R.
template <typename T, int FAST, int R>
__device__ __forceinline__ void dispatch_dense2(cplx<T> (&a)[1 << R], int p, int q, int v, const cplx<T> *m) {
    if constexpr (R == 3) {
        switch (p * 4 + q) {
        case 0 * 4 + 1: dense2_v<T, 0, 1, FAST, R>(a, v, m); break;
        case 0 * 4 + 2: dense2_v<T, 0, 2, FAST, R>(a, v, m); break;
        case 1 * 4 + 0: dense2_v<T, 1, 0, FAST, R>(a, v, m); break;
        case 1 * 4 + 2: dense2_v<T, 1, 2, FAST, R>(a, v, m); break;
        case 2 * 4 + 0: dense2_v<T, 2, 0, FAST, R>(a, v, m); break;
        default: dense2_v<T, 2, 1, FAST, R>(a, v, m); break;
        }
    } else {
        switch (p * 4 + q) {
        case 0 * 4 + 1: dense2_v<T, 0, 1, FAST, R>(a, v, m); break;
        case 0 * 4 + 2: dense2_v<T, 0, 2, FAST, R>(a, v, m); break;
        case 0 * 4 + 3: dense2_v<T, 0, 3, FAST, R>(a, v, m); break;
        case 1 * 4 + 0: dense2_v<T, 1, 0, FAST, R>(a, v, m); break;
        case 1 * 4 + 2: dense2_v<T, 1, 2, FAST, R>(a, v, m); break;
        case 1 * 4 + 3: dense2_v<T, 1, 3, FAST, R>(a, v, m); break;
        case 2 * 4 + 0: dense2_v<T, 2, 0, FAST, R>(a, v, m); break;
        case 2 * 4 + 1: dense2_v<T, 2, 1, FAST, R>(a, v, m); break;
        case 2 * 4 + 3: dense2_v<T, 2, 3, FAST, R>(a, v, m); break;
        case 3 * 4 + 0: dense2_v<T, 3, 0, FAST, R>(a, v, m); break;
        case 3 * 4 + 1: dense2_v<T, 3, 1, FAST, R>(a, v, m); break;
        default: dense2_v<T, 3, 2, FAST, R>(a, v, m); break;
        }
    }
}

// The swizzle XORs nibbles together, so it is linear over XOR, and a tile
// index is tpart | jp[j] with disjoint bits:
//   swz(tpart | jp[j]) = swz(tpart) ^ swz(jp[j]),
// one per-thread value per layout and 16 uniform ones.  The uniform ones come
// from the host as byte offsets (the layout table of the pass, one entry of
// 16 per layout), so each LDS access costs one XOR: computed in the kernel
// they were folded back into a full per-lane swizzle of every index.
template <typename T, int B, int R>
__device__ __forceinline__ unsigned lds_base(const Layout<T, R> &L) {
    return swz<B>(L.tpart) * unsigned(sizeof(cplx<T>));
}

// An amplitude as one naturally aligned vector, so that each LDS access is a
// single 2-element access (ds_write_b128 / ds_read_b128 for complex128).
template <typename T>
using amp_vec = T __attribute__((ext_vector_type(2)));

template <typename T, int R>
__device__ __forceinline__ void to_lds(unsigned char *lds, const cplx<T> (&a)[1 << R], unsigned st,
                                       const unsigned *__restrict__ sj) {
#pragma unroll
    for (int j = 0; j < (1 << R); ++j) {
        amp_vec<T> v;
        v.x = a[j].re;
        v.y = a[j].im;
        *reinterpret_cast<amp_vec<T> *>(lds + (st ^ sj[j])) = v;
    }
}

template <typename T, int R>
__device__ __forceinline__ void from_lds(cplx<T> (&a)[1 << R], const unsigned char *lds, unsigned st,
                                         const unsigned *__restrict__ sj) {
#pragma unroll
    for (int j = 0; j < (1 << R); ++j) {
        const amp_vec<T> v = *reinterpret_cast<const amp_vec<T> *>(lds + (st ^ sj[j]));
        a[j] = cplx<T>{v.x, v.y};
    }
}

// Diagonal with one register bit (slot P): registers with that bit set take
// d1, the others d0, chosen at compile time.
template <typename T, int P, int R>
__device__ __forceinline__ void reg_diag_split(cplx<T> (&a)[1 << R], cplx<T> d0, cplx<T> d1) {
#pragma unroll
    for (int j = 0; j < (1 << R); ++j) a[j] = cmul(((j >> P) & 1) ? d1 : d0, a[j]);
}

// Fused 5-qubit unitary on FP32 MFMA (complex64 tiles).  The tile, in LDS,
// is a 32 x 2^(B-5) complex matrix V (rows: the 5 gate bits, bit[0] = MSB;
// columns: the other tile bits).  W = U V is computed as the real product
//   [Wr; Wi] = [[Ur, -Ui], [Ui, Ur]] [Vr; Vi]
// with v_mfma_f32_32x32x2f32, the K index ordered (gate index s, re/im h):
// lane l supplies B = (h ? Im : Re) V[s][c] with h = l / 32, c = its column,
// and A = U[i][s] recombined per row block, i = l % 32.  Wave w owns columns
// [32 w, 32 w + 32): it reads and writes only those, so no barrier is needed
// between its reads and writes.  Output layout of the 32x32 MFMA (checked on
// the device): D[row][col] with col = lane % 32 and, for accumulator v,
// row = 8 (v / 4) + 4 (lane / 32) + v % 4.
typedef float dq_floatx16 __attribute__((ext_vector_type(16)));

template <int B>
__device__ __forceinline__ void fused5_mfma(cplx<float> *lds, const cplx<float> *__restrict__ u, const TileOp &op,
                                            unsigned t) {
    const unsigned lane = t & 63, wave = t >> 6;
    const unsigned i = lane & 31, h = lane >> 5;
    unsigned fmask = 0;
#pragma unroll
    for (int b = 0; b < 5; ++b) fmask |= 1u << op.bit[b];
    // column c = 32 * wave + i, scattered onto the non-gate tile bits
    const unsigned c = 32 * wave + i;
    unsigned cpart = 0;
    {
        int cb = 0;
#pragma unroll
        for (int bpos = 0; bpos < B; ++bpos) {
            if ((fmask >> bpos) & 1) continue;
            if ((c >> cb) & 1) cpart |= 1u << bpos;
            ++cb;
        }
    }
    const unsigned sc = swz<B>(cpart);
    dq_floatx16 d0 = {0}, d1 = {0};
#pragma unroll 4
    for (int s = 0; s < 32; ++s) {
        unsigned gpart = 0;
#pragma unroll
        for (int b = 0; b < 5; ++b)
            if ((s >> (4 - b)) & 1) gpart |= 1u << op.bit[b];
        const cplx<float> v = lds[sc ^ swz<B>(gpart)];
        const cplx<float> us = u[i * 32 + s];
        const float bv = h ? v.im : v.re;
        const float a0 = h ? -us.im : us.re;  // rows of Wr
        const float a1 = h ? us.re : us.im;   // rows of Wi
        d0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a0, bv, d0, 0, 0, 0);
        d1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a1, bv, d1, 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        const int row = 8 * (v / 4) + 4 * int(h) + (v % 4);
        unsigned gpart = 0;
#pragma unroll
        for (int b = 0; b < 5; ++b)
            if ((row >> (4 - b)) & 1) gpart |= 1u << op.bit[b];
        lds[sc ^ swz<B>(gpart)] = cplx<float>{d0[v], d1[v]};
    }
}

// Streaming accesses of the sweep: every amplitude is read once and written
// once per sweep, so its loads and stores are marked non-temporal (measured:
// 2% less time per 30-qubit sweep than cached accesses).
template <typename T>
__device__ __forceinline__ cplx<T> load_nt(const cplx<T> *p) {
    return cplx<T>{__builtin_nontemporal_load(&p->re), __builtin_nontemporal_load(&p->im)};
}
template <typename T>
__device__ __forceinline__ void store_nt(cplx<T> *p, cplx<T> v) {
    __builtin_nontemporal_store(v.re, &p->re);
    __builtin_nontemporal_store(v.im, &p->im);
}

// One HBM sweep with register-resident amplitudes (see the file header).
// The next tile's amplitudes are loaded into registers before the current
// tile's ops run, so each tile's HBM latency overlaps the previous tile's work.
// FAST distinguishes the instantiation compiled with contraction (FMA, 1) from
// the exact one (0); FAST = 2 is the complex64 contracted instantiation that
// also carries the MFMA block op (T_FUSED5), kept apart because its
// accumulators cost registers (occupancy) in sweeps that have no blocks.
//
// The sweep walks tiles vblock, vblock + vgrid, ... (vgrid a power of two):
// a launch passes blockIdx.x / gridDim.x; a workgroup that owns a whole state
// passes 0 / 1 and walks every tile in order.  smem_raw: the tile's LDS, 2^B
// amplitudes (the sweep kernels declare it statically, so its address is a
// constant folded into each LDS access).
template <typename T, int B, int FAST, int R>
__device__ __forceinline__ void tile_sweep(cplx<T> *__restrict__ amps, const TilePass &pd,
                                           const TileOp *__restrict__ ops, const cplx<T> *__restrict__ mats,
                                           uint64_t ntiles, double *__restrict__ partials, uint64_t vblock,
                                           uint64_t vgrid, unsigned char *smem_raw) {
    double red = 0.0;  // this thread's share of the epilogue reduction
    cplx<T> *lds = reinterpret_cast<cplx<T> *>(smem_raw);
    // layout tables of this pass: 16 LDS byte offsets per layout
    const unsigned *laytab =
        reinterpret_cast<const unsigned *>(reinterpret_cast<const unsigned char *>(mats) + pd.lay_off);
    const unsigned t = threadIdx.x;
    constexpr int NA = 1 << R;  // amplitudes per thread
    Layout<T, R> L;
    // global offsets of the load/store layouts: per-thread part + 4 uniform
    // per-register-bit offsets (additive over disjoint bits)
    uint64_t g_t, gs_t, g_r[R], gs_r[R];
    {
        Layout<T, R> L0;
        make_layout<T, B, R>(L0, pd.load_rb, t);
        g_t = gmap<B>(L0.tpart, pd);
        make_layout<T, B, R>(L0, pd.store_rb, t);
        gs_t = gmap<B>(L0.tpart, pd);
#pragma unroll
        for (int mm = 0; mm < R; ++mm) {
            g_r[mm] = uint64_t(1) << pd.tq[(pd.load_rb >> (4 * mm)) & 15];
            gs_r[mm] = uint64_t(1) << pd.tq[(pd.store_rb >> (4 * mm)) & 15];
        }
    }
    // Global offset of a tile: its outer-index bits deposited on the outer
    // qubits (oq ascending).  Tiles are vblock + k * vgrid with vgrid a
    // power of two (the host launches it so), so the offset is the deposit
    // of vblock, fixed per workgroup, ORed with the deposit
    // of k on the outer qubits of the high index bits (hmask); that one steps
    // from k to k + 1 as a counter confined to the bits of hmask.
    uint64_t dep_lo = 0, hmask = 0;
    for (int j = 0; j < pd.n_outer; ++j) {
        const uint64_t qb = uint64_t(1) << pd.oq[j];
        if ((vblock >> j) & 1) dep_lo |= qb;
        if (!(((vgrid - 1) >> j) & 1)) hmask |= qb;
    }
    auto next_high = [&](uint64_t h) { return ((h | ~hmask) + 1) & hmask; };
    cplx<T> nxt[NA];
    uint64_t tile = vblock;
    uint64_t high = 0;  // deposit of the current tile's high index bits
    if (tile < ntiles) {
        const uint64_t base = dep_lo + g_t;
#pragma unroll
        for (int j = 0; j < NA; ++j) {
            uint64_t off = base;
#pragma unroll
            for (int mm = 0; mm < R; ++mm)
                if ((j >> mm) & 1) off += g_r[mm];
            nxt[j] = load_nt(amps + off);
        }
    }
    // One tile: take the prefetched amplitudes, prefetch the next tile, run
    // the pass's ops, store.  The first tile runs outside the loop so that
    // the loop head is entered with the same memory operations pending from
    // both sides (the next tile's loads, then this tile's stores); the wait
    // for the loads then leaves the stores in flight.
    auto process = [&](uint64_t tile) __attribute__((always_inline)) {
        const uint64_t base = dep_lo | high;
        cplx<T> a[NA];
#pragma unroll
        for (int j = 0; j < NA; ++j) a[j] = nxt[j];
        const uint64_t next = tile + vgrid;
        if (next < ntiles) {
            const uint64_t nb = (dep_lo | next_high(high)) + g_t;
#pragma unroll
            for (int j = 0; j < NA; ++j) {
                uint64_t off = nb;
#pragma unroll
                for (int mm = 0; mm < R; ++mm)
                    if ((j >> mm) & 1) off += g_r[mm];
                nxt[j] = load_nt(amps + off);
            }
        }
        make_layout<T, B, R>(L, pd.load_rb, t);
        unsigned st = lds_base<T, B, R>(L);
        const unsigned *cs = laytab + NA * pd.load_lay;
        int cur = pd.load_rb;
        OpHead h = op_head(ops + pd.op_begin);
        for (int o = 0; o < pd.op_count; ++o) {
            const TileOp &op = ops[pd.op_begin + o];
            // the next op's header is fetched now, used next iteration
            const OpHead hn = op_head(ops + pd.op_begin + (o + 1 < pd.op_count ? o + 1 : o));
            if (h.relayout) {
                __syncthreads();
                to_lds<T, R>(smem_raw, a, st, cs);
                __syncthreads();
                make_layout<T, B, R>(L, h.rb, t);
                st = lds_base<T, B, R>(L);
                cs = laytab + NA * h.lay;
                from_lds<T, R>(a, smem_raw, st, cs);
                cur = h.rb;
            }
            const cplx<T> *m = mats + h.mat;
            if (h.nsel > 0) {  // one matrix per value of the selector bits (constant over the tile)
                unsigned e = 0;
#pragma unroll
                for (int j = 0; j < 3; ++j)
                    if (j < h.nsel) e |= unsigned((tile >> h.sel[j]) & 1) << j;
                m += e * (h.kind == T_DENSE2 ? 16u : 4u);
            }
            switch (h.kind) {
            case T_DENSE1:
                dispatch_dense1<T, FAST, R>(a, h.p, h.variant, m);
                break;
            case T_CTRL1:
                if ((unsigned(tile) & op.ocmask) == op.ocval)
                    dispatch_ctrl1<T, FAST, R>(a, h.p, h.variant, m, L, op.cmask, op.cval);
                break;
            case T_DENSE2:
                dispatch_dense2<T, FAST, R>(a, h.p, h.q, h.variant, m);
                break;
            case T_DIAGT: {
                // e = thread part | register part; the thread part picks the
                // candidate entries once per op (uniform values, no gathers).
                int et = 0;
#pragma unroll
                for (int b2 = 0; b2 < 2; ++b2) {
                    if (b2 >= op.k) break;
                    const int bp = op.bit[b2];
                    if (bp >= kOuterBit) et |= int((tile >> (bp - kOuterBit)) & 1) << (op.k - 1 - b2);
                    else if (b2 != h.q || h.p < 0) et |= int((L.tpart >> bp) & 1) << (op.k - 1 - b2);
                }
                const cplx<T> m0 = m[0], m1 = m[1];
                const cplx<T> m2 = op.k == 2 ? m[2] : m0, m3 = op.k == 2 ? m[3] : m1;
                if (h.p < 0) {
                    const cplx<T> d = pick4(et, m0, m1, m2, m3);
#pragma unroll
                    for (int j = 0; j < NA; ++j) a[j] = cmul(d, a[j]);
                } else {
                    const int rw = 1 << (op.k - 1 - h.q);  // weight of the register bit in e
                    const cplx<T> d0 = pick4(et, m0, m1, m2, m3), d1 = pick4(et | rw, m0, m1, m2, m3);
                    switch (h.p) {
                    case 0: reg_diag_split<T, 0, R>(a, d0, d1); break;
                    case 1: reg_diag_split<T, 1, R>(a, d0, d1); break;
                    case 2:
                        if constexpr (R > 3) {
                            reg_diag_split<T, 2, R>(a, d0, d1);
                            break;
                        }
                        [[fallthrough]];
                    default: reg_diag_split<T, R - 1, R>(a, d0, d1); break;
                    }
                }
                break;
            }
            case T_DIAG:
#pragma unroll
                for (int j = 0; j < NA; ++j) {
                    const unsigned i = L.tpart | L.jp(j);
                    int e = 0;
                    for (int b2 = 0; b2 < op.k; ++b2) {
                        const int bp = op.bit[b2];
                        const int v = bp >= kOuterBit ? int((tile >> (bp - kOuterBit)) & 1) : int((i >> bp) & 1);
                        e = (e << 1) | v;
                    }
                    a[j] = cmul(m[e], a[j]);
                }
                break;
            case T_COLLAPSE: {
                const double nrm = op.norm;  // renormalise in FP64 (exact division for T = double)
                const int bp = op.bit[0];
                const bool outer = bp >= kOuterBit;
                const int ov = outer ? int((tile >> (bp - kOuterBit)) & 1) : 0;
#pragma unroll
                for (int j = 0; j < NA; ++j) {
                    const unsigned i = L.tpart | L.jp(j);
                    const int v = outer ? ov : int((i >> (bp & 31)) & 1);
                    if (v == op.outcome) {
                        a[j].re = T(double(a[j].re) / nrm);
                        a[j].im = T(double(a[j].im) / nrm);
                    } else {
                        a[j] = cplx<T>{T(0), T(0)};
                    }
                }
                break;
            }
            case T_FUSED5:
                if constexpr (sizeof(T) == 4 && B >= 12 && FAST == 2 && R == 4) {  // only the MFMA instantiation
                    __syncthreads();
                    to_lds<T, R>(smem_raw, a, st, cs);
                    __syncthreads();
                    fused5_mfma<B>(reinterpret_cast<cplx<float> *>(lds), reinterpret_cast<const cplx<float> *>(m),
                                   op, t);
                    __syncthreads();
                    from_lds<T, R>(a, smem_raw, st, cs);
                }
                break;
            default: {  // T_LDSOP
                __syncthreads();
                to_lds<T, R>(smem_raw, a, st, cs);
                __syncthreads();
                const int k = op.k;
                const int dim = 1 << k;
                unsigned gm = 0;
                for (int b2 = 0; b2 < k; ++b2) gm |= 1u << op.bit[b2];
#pragma unroll
                for (int j = 0; j < NA; ++j) {
                    const unsigned i = L.tpart | L.jp(j);
                    int r = 0;
                    for (int b2 = 0; b2 < k; ++b2) r = (r << 1) | int((i >> op.bit[b2]) & 1);
                    const unsigned base_i = i & ~gm;
                    cplx<T> acc{T(0), T(0)};
                    for (int c = 0; c < dim; ++c) {
                        unsigned off = 0;
                        for (int b2 = 0; b2 < k; ++b2)
                            if ((c >> (k - 1 - b2)) & 1) off |= 1u << op.bit[b2];
                        const cplx<T> pr = cmul(m[r * dim + c], lds[swz<B>(base_i | off)]);
                        acc = (c == 0) ? pr : cadd(acc, pr);
                    }
                    a[j] = acc;
                }
                break;
            }
            }
            h = hn;
        }
        if (cur != pd.store_rb) {
            __syncthreads();
            to_lds<T, R>(smem_raw, a, st, cs);
            __syncthreads();
            make_layout<T, B, R>(L, pd.store_rb, t);
            from_lds<T, R>(a, smem_raw, lds_base<T, B, R>(L), laytab + NA * pd.store_lay);
        }
        if (pd.red_bit >= 0) {
            const int rb2 = pd.red_bit;
            const bool outer = rb2 >= kOuterBit;
            const bool ov = outer && ((tile >> (rb2 - kOuterBit)) & 1);
#pragma unroll
            for (int j = 0; j < NA; ++j) {
                const bool on = outer ? ov : (((L.tpart | L.jp(j)) >> (rb2 & 31)) & 1);
                const double re = double(a[j].re), im = double(a[j].im);
                if (on) red += re * re + im * im;
            }
        }
        const uint64_t sb = base + gs_t;
#pragma unroll
        for (int j = 0; j < NA; ++j) {
            uint64_t off = sb;
#pragma unroll
            for (int mm = 0; mm < R; ++mm)
                if ((j >> mm) & 1) off += gs_r[mm];
            store_nt(amps + off, a[j]);
        }
        // The next tile's first LDS write is preceded by a barrier.
    };
    if (tile < ntiles) {
        process(tile);
        for (tile += vgrid; tile < ntiles; tile += vgrid) {
            high = next_high(high);
            process(tile);
        }
    }
    if (pd.red_bit >= 0) {
        // fixed-shape block sum: butterflies within each group of kGroup
        // lanes (a wave, or the whole block when it is smaller than a wave:
        // 32 threads at B = 9), then the group totals in order
        constexpr int kThr = 1 << (B - R);
        constexpr int kGroup = kThr < 64 ? kThr : 64;
        for (int off = kGroup / 2; off > 0; off >>= 1) red += __shfl_xor(red, off, kGroup);
        __syncthreads();
        double *scratch = reinterpret_cast<double *>(smem_raw);
        if ((t % kGroup) == 0) scratch[t / kGroup] = red;
        __syncthreads();
        if (t == 0) {
            double r = 0.0;
            for (int w = 0; w < kThr / kGroup; ++w) r += scratch[w];
            partials[vblock] = r;
        }
    }
}

template <typename T, int B, int FAST>
__global__ __launch_bounds__(1 << (B - 4)) void k_tile_reg(cplx<T> *__restrict__ amps, TilePass pd,
                                                           const TileOp *__restrict__ ops,
                                                           const cplx<T> *__restrict__ mats, uint64_t ntiles,
                                                           double *__restrict__ partials) {
    __shared__ __align__(16) unsigned char tile_lds[sizeof(cplx<T>) << B];
    tile_sweep<T, B, FAST, 4>(amps, pd, ops, mats, ntiles, partials, blockIdx.x, gridDim.x, tile_lds);
}

// Eight amplitudes per thread (3 register bits): twice the threads of
// k_tile_reg for the same tile and LDS, half the registers per thread, so a
// 12-qubit complex128 tile runs 8 waves per workgroup and two workgroups per
// CU give 4 waves per SIMD (k_tile_reg: 2) to hide FP64, LDS and barrier
// latencies.  Ops on 2 register bits (T_DENSE2) need 2 of the 3 bits.
template <typename T, int B, int FAST>
__global__ __launch_bounds__(1 << (B - 3), 4) void k_tile_reg8(cplx<T> *__restrict__ amps, TilePass pd,
                                                               const TileOp *__restrict__ ops,
                                                               const cplx<T> *__restrict__ mats, uint64_t ntiles,
                                                               double *__restrict__ partials) {
    __shared__ __align__(16) unsigned char tile_lds[sizeof(cplx<T>) << B];
    tile_sweep<T, B, FAST, 3>(amps, pd, ops, mats, ntiles, partials, blockIdx.x, gridDim.x, tile_lds);
}

// complex64 without the MFMA block op (FAST = 1): the sweep fits 128 VGPRs,
// and a 12-qubit tile takes 32 KiB of LDS, so 4 waves per SIMD (one more than
// the register allocation reaches unasked).  The MFMA instantiation would
// spill at this bound and keeps k_tile_reg.
template <int B, int FAST>
__global__ __launch_bounds__(1 << (B - 4), 4) void k_tile_reg_c64(cplx<float> *__restrict__ amps, TilePass pd,
                                                                  const TileOp *__restrict__ ops,
                                                                  const cplx<float> *__restrict__ mats,
                                                                  uint64_t ntiles, double *__restrict__ partials) {
    __shared__ __align__(16) unsigned char tile_lds[sizeof(cplx<float>) << B];
    tile_sweep<float, B, FAST, 4>(amps, pd, ops, mats, ntiles, partials, blockIdx.x, gridDim.x, tile_lds);
}

}  // namespace dq
