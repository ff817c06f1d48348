// dq_tile.hpp -- register-resident LDS-tiled sweep kernel for gfx950.
//
// One workgroup owns a tile of 2^B amplitudes (B = tile qubits).  Each of the
// 2^(B-4) threads holds 16 amplitudes in registers.  Which tile qubits index a
// thread's 16 registers (the "register bits", 4 of them) is the *layout*;
// the remaining B-4 tile qubits index the thread.  Gates whose matrix qubits
// are register bits are applied entirely in registers; a layout change is one
// LDS round trip (write all, barrier, read all) with an XOR swizzle so that
// strided layouts stay free of bank conflicts.  The host plans, per pass, the
// layout every op runs in, so most gates cost no LDS traffic and no barrier.
//
// Arithmetic: each output is the reference's row sum (src/engine.rs:58-59,
// 146-151) in the same column order.  Products with an exactly zero entry are
// skipped and products with entries 1, -1, real or imaginary use only the
// nonzero parts; the skipped terms are +-0 and do not change any sum, so the
// values equal the reference's up to the sign of a zero.
#pragma once
#include "dq_device.hpp"

namespace dq {

enum TileKind : int {
    T_DENSE1 = 0,    // 2x2 on register slot p (diagonal and anti-diagonal included)
    T_CTRL1 = 1,     // 2x2 on register slot p where (tile index & cmask) == cval
    T_DENSE2 = 2,    // 4x4 on register slots p (matrix MSB) and q
    T_DIAG = 3,      // diagonal 2^k on any tile bits, elementwise
    T_COLLAPSE = 4,  // measurement collapse on tile bit bit[0], elementwise
    T_LDSOP = 5      // any 2^k x 2^k matrix through LDS, one output per amplitude
};

// Product shape of a matrix entry, 3 bits per entry in TileOp::modes.
enum EntryMode : int { M_ZERO = 0, M_ONE = 1, M_MONE = 2, M_REAL = 3, M_IMAG = 4, M_GEN = 5 };

struct TileOp {
    int kind;
    int k;
    int relayout;  // 1: change to layout `rb` before this op
    int rb;        // layout active from this op: 4 tile bits, 4 bits each, ascending
    int p, q;      // register slots (see TileKind)
    int bit[5];    // tile bits, bit[0] = MSB of the gate index
    int sbit[5];   // ascending
    unsigned cmask, cval;
    int mat;       // offset (complex entries) into the matrix pool
    int outcome;
    unsigned long long modes;
    double norm;
};

struct TilePass {
    int n_outer;
    int op_begin;
    int op_count;
    int load_rb;   // layout of the global load
    int store_rb;  // layout of the global store
    int tq[16];    // tile bit -> qubit
    int oq[48];    // outer bit -> qubit
};

template <typename T>
struct Layout {
    unsigned tpart;    // tile-index bits contributed by this thread
    unsigned jp[16];   // tile-index bits contributed by register j (uniform)
};

template <int B>
__device__ __forceinline__ unsigned swz(unsigned i) {
    unsigned s = 0;
#pragma unroll
    for (int sh = 4; sh < B; sh += 4) s ^= (i >> sh) & 15u;
    return i ^ s;
}

template <typename T, int B>
__device__ __forceinline__ void make_layout(Layout<T> &L, int rb, unsigned t) {
    int rbit[4];
    unsigned rmask = 0;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        rbit[m] = (rb >> (4 * m)) & 15;
        rmask |= 1u << rbit[m];
    }
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        unsigned v = 0;
#pragma unroll
        for (int m = 0; m < 4; ++m)
            if ((j >> m) & 1) v |= 1u << rbit[m];
        L.jp[j] = v;
    }
    unsigned tp = 0;
    int c = 0;
#pragma unroll
    for (int bpos = 0; bpos < B; ++bpos) {
        if ((rmask >> bpos) & 1) continue;
        if ((t >> c) & 1) tp |= 1u << bpos;
        ++c;
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

// ---- exact products by entry shape ----

template <typename T, int N>
__device__ __forceinline__ void prod_n(cplx<T> (&out)[N], const cplx<T> (&v)[N], cplx<T> u, int mode) {
    switch (mode) {
    case M_ONE:
#pragma unroll
        for (int i = 0; i < N; ++i) out[i] = v[i];
        break;
    case M_MONE:
#pragma unroll
        for (int i = 0; i < N; ++i) out[i] = cplx<T>{-v[i].re, -v[i].im};
        break;
    case M_REAL:
#pragma unroll
        for (int i = 0; i < N; ++i) out[i] = cplx<T>{u.re * v[i].re, u.re * v[i].im};
        break;
    case M_IMAG:
#pragma unroll
        for (int i = 0; i < N; ++i) out[i] = cplx<T>{-(u.im * v[i].im), u.im * v[i].re};
        break;
    default:
#pragma unroll
        for (int i = 0; i < N; ++i) out[i] = cmul(u, v[i]);
        break;
    }
}

// acc (+)= u * v with the reference's order; `first` means acc holds 0.
template <typename T, int N>
__device__ __forceinline__ void mac_n(cplx<T> (&acc)[N], bool &first, const cplx<T> (&v)[N], cplx<T> u, int mode) {
    if (mode == M_ZERO) return;
    if (first) {
        prod_n<T, N>(acc, v, u, mode);
        first = false;
        return;
    }
    cplx<T> p[N];
    prod_n<T, N>(p, v, u, mode);
#pragma unroll
    for (int i = 0; i < N; ++i) acc[i] = cadd(acc[i], p[i]);
}

template <typename T, int N>
__device__ __forceinline__ void zero_if_first(cplx<T> (&acc)[N], bool first) {
    if (first) {
#pragma unroll
        for (int i = 0; i < N; ++i) acc[i] = cplx<T>{T(0), T(0)};
    }
}

__device__ __forceinline__ int emode(unsigned long long modes, int e) { return int((modes >> (3 * e)) & 7); }

// 2x2 on register slot P: pairs (j0, j0 | 1 << P).
template <typename T, int P>
__device__ __forceinline__ void reg_dense1(cplx<T> (&a)[16], const cplx<T> *__restrict__ m, unsigned long long modes) {
    cplx<T> x[8], y[8];
#pragma unroll
    for (int jj = 0; jj < 8; ++jj) {
        const int j0 = ((jj >> P) << (P + 1)) | (jj & ((1 << P) - 1));
        x[jj] = a[j0];
        y[jj] = a[j0 | (1 << P)];
    }
    cplx<T> r0[8], r1[8];
    bool f0 = true, f1 = true;
    mac_n<T, 8>(r0, f0, x, m[0], emode(modes, 0));
    mac_n<T, 8>(r0, f0, y, m[1], emode(modes, 1));
    zero_if_first<T, 8>(r0, f0);
    mac_n<T, 8>(r1, f1, x, m[2], emode(modes, 2));
    mac_n<T, 8>(r1, f1, y, m[3], emode(modes, 3));
    zero_if_first<T, 8>(r1, f1);
#pragma unroll
    for (int jj = 0; jj < 8; ++jj) {
        const int j0 = ((jj >> P) << (P + 1)) | (jj & ((1 << P) - 1));
        a[j0] = r0[jj];
        a[j0 | (1 << P)] = r1[jj];
    }
}

// Controlled 2x2 on slot P: pair applied only where the control pattern holds.
template <typename T, int P>
__device__ __forceinline__ void reg_ctrl1(cplx<T> (&a)[16], const cplx<T> *__restrict__ m, unsigned long long modes,
                                          const Layout<T> &L, unsigned cmask, unsigned cval) {
    cplx<T> x[8], y[8];
#pragma unroll
    for (int jj = 0; jj < 8; ++jj) {
        const int j0 = ((jj >> P) << (P + 1)) | (jj & ((1 << P) - 1));
        x[jj] = a[j0];
        y[jj] = a[j0 | (1 << P)];
    }
    cplx<T> r0[8], r1[8];
    bool f0 = true, f1 = true;
    mac_n<T, 8>(r0, f0, x, m[0], emode(modes, 0));
    mac_n<T, 8>(r0, f0, y, m[1], emode(modes, 1));
    zero_if_first<T, 8>(r0, f0);
    mac_n<T, 8>(r1, f1, x, m[2], emode(modes, 2));
    mac_n<T, 8>(r1, f1, y, m[3], emode(modes, 3));
    zero_if_first<T, 8>(r1, f1);
#pragma unroll
    for (int jj = 0; jj < 8; ++jj) {
        const int j0 = ((jj >> P) << (P + 1)) | (jj & ((1 << P) - 1));
        const bool on = ((L.tpart | L.jp[j0]) & cmask) == cval;
        if (on) {
            a[j0] = r0[jj];
            a[j0 | (1 << P)] = r1[jj];
        }
    }
}

// 4x4 on register slots P (matrix MSB) and Q.
template <typename T, int P, int Q>
__device__ __forceinline__ void reg_dense2(cplx<T> (&a)[16], const cplx<T> *__restrict__ m, unsigned long long modes) {
    constexpr int LO = P < Q ? P : Q, HI = P < Q ? Q : P;
    int base[4];
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        int v = g;
        v = ((v >> LO) << (LO + 1)) | (v & ((1 << LO) - 1));
        v = ((v >> HI) << (HI + 1)) | (v & ((1 << HI) - 1));
        base[g] = v;
    }
    cplx<T> v[4][4];  // v[c][g]
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int g = 0; g < 4; ++g) v[c][g] = a[base[g] | (((c >> 1) & 1) << P) | ((c & 1) << Q)];
    cplx<T> out[4][4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        bool f = true;
#pragma unroll
        for (int c = 0; c < 4; ++c) mac_n<T, 4>(out[r], f, v[c], m[r * 4 + c], emode(modes, r * 4 + c));
        zero_if_first<T, 4>(out[r], f);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int g = 0; g < 4; ++g) a[base[g] | (((r >> 1) & 1) << P) | ((r & 1) << Q)] = out[r][g];
}

template <typename T>
__device__ __forceinline__ void dispatch_dense1(cplx<T> (&a)[16], int p, const cplx<T> *m, unsigned long long modes) {
    switch (p) {
    case 0: reg_dense1<T, 0>(a, m, modes); break;
    case 1: reg_dense1<T, 1>(a, m, modes); break;
    case 2: reg_dense1<T, 2>(a, m, modes); break;
    default: reg_dense1<T, 3>(a, m, modes); break;
    }
}

template <typename T>
__device__ __forceinline__ void dispatch_ctrl1(cplx<T> (&a)[16], int p, const cplx<T> *m, unsigned long long modes,
                                               const Layout<T> &L, unsigned cmask, unsigned cval) {
    switch (p) {
    case 0: reg_ctrl1<T, 0>(a, m, modes, L, cmask, cval); break;
    case 1: reg_ctrl1<T, 1>(a, m, modes, L, cmask, cval); break;
    case 2: reg_ctrl1<T, 2>(a, m, modes, L, cmask, cval); break;
    default: reg_ctrl1<T, 3>(a, m, modes, L, cmask, cval); break;
    }
}

template <typename T>
__device__ __forceinline__ void dispatch_dense2(cplx<T> (&a)[16], int p, int q, const cplx<T> *m,
                                                unsigned long long modes) {
    switch (p * 4 + q) {
    case 0 * 4 + 1: reg_dense2<T, 0, 1>(a, m, modes); break;
    case 0 * 4 + 2: reg_dense2<T, 0, 2>(a, m, modes); break;
    case 0 * 4 + 3: reg_dense2<T, 0, 3>(a, m, modes); break;
    case 1 * 4 + 0: reg_dense2<T, 1, 0>(a, m, modes); break;
    case 1 * 4 + 2: reg_dense2<T, 1, 2>(a, m, modes); break;
    case 1 * 4 + 3: reg_dense2<T, 1, 3>(a, m, modes); break;
    case 2 * 4 + 0: reg_dense2<T, 2, 0>(a, m, modes); break;
    case 2 * 4 + 1: reg_dense2<T, 2, 1>(a, m, modes); break;
    case 2 * 4 + 3: reg_dense2<T, 2, 3>(a, m, modes); break;
    case 3 * 4 + 0: reg_dense2<T, 3, 0>(a, m, modes); break;
    case 3 * 4 + 1: reg_dense2<T, 3, 1>(a, m, modes); break;
    default: reg_dense2<T, 3, 2>(a, m, modes); break;
    }
}

template <typename T, int B>
__device__ __forceinline__ void to_lds(cplx<T> *lds, const cplx<T> (&a)[16], const Layout<T> &L) {
#pragma unroll
    for (int j = 0; j < 16; ++j) lds[swz<B>(L.tpart | L.jp[j])] = a[j];
}

template <typename T, int B>
__device__ __forceinline__ void from_lds(cplx<T> (&a)[16], const cplx<T> *lds, const Layout<T> &L) {
#pragma unroll
    for (int j = 0; j < 16; ++j) a[j] = lds[swz<B>(L.tpart | L.jp[j])];
}

// One HBM sweep with register-resident amplitudes (see the file header).
template <typename T, int B>
__global__ __launch_bounds__(1 << (B - 4)) void k_tile_reg(cplx<T> *__restrict__ amps, TilePass pd,
                                                           const TileOp *__restrict__ ops,
                                                           const cplx<T> *__restrict__ mats, uint64_t ntiles) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cplx<T> *lds = reinterpret_cast<cplx<T> *>(smem_raw);
    const unsigned t = threadIdx.x;
    Layout<T> L0, L;
    make_layout<T, B>(L0, pd.load_rb, t);
    const uint64_t g_t = gmap<B>(L0.tpart, pd);
    uint64_t g_j[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) g_j[j] = gmap<B>(L0.jp[j], pd);
    Layout<T> LS;
    make_layout<T, B>(LS, pd.store_rb, t);
    const uint64_t gs_t = gmap<B>(LS.tpart, pd);
    uint64_t gs_j[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) gs_j[j] = gmap<B>(LS.jp[j], pd);

    for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        uint64_t base = 0;
        for (int j = 0; j < pd.n_outer; ++j)
            if ((tile >> j) & 1) base |= uint64_t(1) << pd.oq[j];
        cplx<T> a[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) a[j] = amps[base + g_t + g_j[j]];
        L = L0;
        int cur = pd.load_rb;
        for (int o = 0; o < pd.op_count; ++o) {
            const TileOp &op = ops[pd.op_begin + o];
            if (op.relayout) {
                __syncthreads();
                to_lds<T, B>(lds, a, L);
                __syncthreads();
                make_layout<T, B>(L, op.rb, t);
                from_lds<T, B>(a, lds, L);
                cur = op.rb;
            }
            const cplx<T> *m = mats + op.mat;
            switch (op.kind) {
            case T_DENSE1:
                dispatch_dense1<T>(a, op.p, m, op.modes);
                break;
            case T_CTRL1:
                dispatch_ctrl1<T>(a, op.p, m, op.modes, L, op.cmask, op.cval);
                break;
            case T_DENSE2:
                dispatch_dense2<T>(a, op.p, op.q, m, op.modes);
                break;
            case T_DIAG:
#pragma unroll
                for (int j = 0; j < 16; ++j) {
                    const unsigned i = L.tpart | L.jp[j];
                    int e = 0;
                    for (int b2 = 0; b2 < op.k; ++b2) e = (e << 1) | int((i >> op.bit[b2]) & 1);
                    a[j] = cmul(m[e], a[j]);
                }
                break;
            case T_COLLAPSE: {
                const T nrm = T(op.norm);
#pragma unroll
                for (int j = 0; j < 16; ++j) {
                    const unsigned i = L.tpart | L.jp[j];
                    if (int((i >> op.bit[0]) & 1) == op.outcome) {
                        a[j].re = a[j].re / nrm;
                        a[j].im = a[j].im / nrm;
                    } else {
                        a[j] = cplx<T>{T(0), T(0)};
                    }
                }
                break;
            }
            default: {  // T_LDSOP
                __syncthreads();
                to_lds<T, B>(lds, a, L);
                __syncthreads();
                const int k = op.k;
                const int dim = 1 << k;
                unsigned gm = 0;
                for (int b2 = 0; b2 < k; ++b2) gm |= 1u << op.bit[b2];
#pragma unroll
                for (int j = 0; j < 16; ++j) {
                    const unsigned i = L.tpart | L.jp[j];
                    int r = 0;
                    for (int b2 = 0; b2 < k; ++b2) r = (r << 1) | int((i >> op.bit[b2]) & 1);
                    const unsigned base_i = i & ~gm;
                    cplx<T> acc{T(0), T(0)};
                    bool first = true;
                    for (int c = 0; c < dim; ++c) {
                        const cplx<T> u = m[r * dim + c];
                        if (u.re == T(0) && u.im == T(0)) continue;
                        unsigned off = 0;
                        for (int b2 = 0; b2 < k; ++b2)
                            if ((c >> (k - 1 - b2)) & 1) off |= 1u << op.bit[b2];
                        const cplx<T> pr = cmul(u, lds[swz<B>(base_i | off)]);
                        acc = first ? pr : cadd(acc, pr);
                        first = false;
                    }
                    a[j] = acc;
                }
                break;
            }
            }
        }
        if (cur != pd.store_rb) {
            __syncthreads();
            to_lds<T, B>(lds, a, L);
            __syncthreads();
            from_lds<T, B>(a, lds, LS);
        }
#pragma unroll
        for (int j = 0; j < 16; ++j) amps[base + gs_t + gs_j[j]] = a[j];
        // The next tile's first LDS write is preceded by a barrier.
    }
}

}  // namespace dq
