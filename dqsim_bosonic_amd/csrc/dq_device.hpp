// dq_device.hpp -- device-side building blocks of the gfx950 statevector engine.
//
// Arithmetic contract: every gate update uses the reference's complex op order
// (num-complex 0.4.6 as called from src/engine.rs:58-59,146-151):
//   mul (ar*br - ai*bi, ar*bi + ai*br); add componentwise; no contraction.
// The library is compiled with -ffp-contract=off, so each product and sum is
// rounded exactly as on the reference.  Specialised op kinds (diagonal,
// monomial, controlled-2x2) skip only terms whose matrix entry is exactly
// zero, whose products are +-0 and leave every sum unchanged, so they give
// the dense formula's values (up to the sign of a zero).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dq_types.hpp"

namespace dq {

template <typename T>
struct cplx {
    T re, im;
};

template <typename T>
__device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename T>
__device__ __forceinline__ cplx<T> cadd(cplx<T> a, cplx<T> b) {
    return {a.re + b.re, a.im + b.im};
}

// Index of the g-th group base: insert a zero at each ascending position.
template <typename I>
__device__ __forceinline__ I insert_zeros(I g, const int *sbit, int k) {
    for (int j = 0; j < k; ++j) {
        const int p = sbit[j];
        const I lo = g & ((I(1) << p) - 1);
        g = ((g >> p) << (p + 1)) | lo;
    }
    return g;
}

template <typename I>
__device__ __forceinline__ I group_offset(int j, const int *bit, int k) {
    I o = 0;
    for (int b = 0; b < k; ++b)
        if ((j >> (k - 1 - b)) & 1) o |= I(1) << bit[b];
    return o;
}

template <typename I>
__device__ __forceinline__ int gather_bits(I i, const int *bit, int k) {
    int j = 0;
    for (int b = 0; b < k; ++b) j = (j << 1) | int((i >> bit[b]) & 1);
    return j;
}

// Selects v[idx] with compile-time indexing only, so v stays in registers.
template <typename T, int DIM>
__device__ __forceinline__ cplx<T> pick(const cplx<T> (&v)[DIM], int idx) {
    cplx<T> r = v[0];
#pragma unroll
    for (int j = 1; j < DIM; ++j)
        if (idx == j) r = v[j];
    return r;
}

// Monomial gate: out[r] = ph[r] * in[perm[r]] on every group of 2^K.
// Up to 3 qubits the offsets, permutation and phases stay in registers for
// the whole sweep; for 4-5 qubits they are re-read per row (a hoisted copy of
// 32 rows would spill to scratch memory).
template <typename T, typename I, int K>
__device__ __forceinline__ void mono_k(cplx<T> *__restrict__ a, int nbits, const DevOp &op,
                                       const cplx<T> *__restrict__ mats,
                                       const int *__restrict__ perms, I tid, I nthr) {
    constexpr int DIM = 1 << K;
    const I ngroups = I(1) << (nbits - K);
    if constexpr (K <= 3) {
        I off[DIM];
        int pm[DIM];
        cplx<T> ph[DIM];
#pragma unroll
        for (int r = 0; r < DIM; ++r) {
            off[r] = group_offset<I>(r, op.bit, K);
            pm[r] = perms[op.aux + r];
            ph[r] = mats[op.mat + r];
        }
        for (I g = tid; g < ngroups; g += nthr) {
            const I base = insert_zeros<I>(g, op.sbit, K);
            cplx<T> v[DIM];
#pragma unroll
            for (int j = 0; j < DIM; ++j) v[j] = a[base | off[j]];
#pragma unroll
            for (int r = 0; r < DIM; ++r) a[base | off[r]] = cmul(ph[r], pick<T, DIM>(v, pm[r]));
        }
    } else {
        for (I g = tid; g < ngroups; g += nthr) {
            const I base = insert_zeros<I>(g, op.sbit, K);
            cplx<T> v[DIM];
#pragma unroll
            for (int j = 0; j < DIM; ++j) v[j] = a[base | group_offset<I>(j, op.bit, K)];
#pragma unroll 1
            for (int r = 0; r < DIM; ++r)
                a[base | group_offset<I>(r, op.bit, K)] = cmul(mats[op.mat + r], pick<T, DIM>(v, perms[op.aux + r]));
        }
    }
}

// Dense gate, the reference's row loop: acc = 0; acc += U[r][c] * v[c].
// Rows are fully unrolled up to 3 qubits; for 4-5 qubits the row loop stays
// a loop so the 2^2K matrix entries are not hoisted out of the sweep.
template <typename T, typename I, int K>
__device__ __forceinline__ void dense_k(cplx<T> *__restrict__ a, int nbits, const DevOp &op,
                                        const cplx<T> *__restrict__ mats, I tid, I nthr) {
    constexpr int DIM = 1 << K;
    const cplx<T> *__restrict__ U = mats + op.mat;
    const I ngroups = I(1) << (nbits - K);
    if constexpr (K <= 3) {
        I off[DIM];
#pragma unroll
        for (int j = 0; j < DIM; ++j) off[j] = group_offset<I>(j, op.bit, K);
        for (I g = tid; g < ngroups; g += nthr) {
            const I base = insert_zeros<I>(g, op.sbit, K);
            cplx<T> v[DIM];
#pragma unroll
            for (int j = 0; j < DIM; ++j) v[j] = a[base | off[j]];
#pragma unroll
            for (int r = 0; r < DIM; ++r) {
                cplx<T> acc{T(0), T(0)};
#pragma unroll
                for (int c = 0; c < DIM; ++c) acc = cadd(acc, cmul(U[r * DIM + c], v[c]));
                a[base | off[r]] = acc;
            }
        }
    } else {
        for (I g = tid; g < ngroups; g += nthr) {
            const I base = insert_zeros<I>(g, op.sbit, K);
            cplx<T> v[DIM];
#pragma unroll
            for (int j = 0; j < DIM; ++j) v[j] = a[base | group_offset<I>(j, op.bit, K)];
#pragma unroll 1
            for (int r = 0; r < DIM; ++r) {
                cplx<T> acc{T(0), T(0)};
#pragma unroll
                for (int c = 0; c < DIM; ++c) acc = cadd(acc, cmul(U[r * DIM + c], v[c]));
                a[base | group_offset<I>(r, op.bit, K)] = acc;
            }
        }
    }
}

// Applies one op to the 2^nbits amplitudes at `a` (LDS tile or global
// state), work items distributed over `nthr` threads starting at `tid`.
// Monomial and dense ops on more than KMAX qubits are skipped: a caller
// with KMAX < 5 applies those itself (so their register arrays are not part
// of its kernel).
template <typename T, typename I, int KMAX = 5>
__device__ __forceinline__ void apply_op(cplx<T> *__restrict__ a, int nbits, const DevOp &op,
                                         const cplx<T> *__restrict__ mats,
                                         const int *__restrict__ perms, I tid, I nthr) {
    const int k = op.k;
    switch (op.kind) {
    case K_DENSE1: {
        const int p = op.bit[0];
        const I half = I(1) << p;
        const I npairs = I(1) << (nbits - 1);
        const cplx<T> u00 = mats[op.mat], u01 = mats[op.mat + 1];
        const cplx<T> u10 = mats[op.mat + 2], u11 = mats[op.mat + 3];
        for (I g = tid; g < npairs; g += nthr) {
            const I i0 = ((g >> p) << (p + 1)) | (g & (half - 1));
            const I i1 = i0 | half;
            const cplx<T> x = a[i0], y = a[i1];
            a[i0] = cadd(cmul(u00, x), cmul(u01, y));
            a[i1] = cadd(cmul(u10, x), cmul(u11, y));
        }
        break;
    }
    case K_CTRL1: {
        const int p = op.bit[k - 1];
        const I half = I(1) << p;
        const I npairs = I(1) << (nbits - 1);
        const I cm = I(op.cmask), cv = I(op.cval);
        const cplx<T> u00 = mats[op.mat], u01 = mats[op.mat + 1];
        const cplx<T> u10 = mats[op.mat + 2], u11 = mats[op.mat + 3];
        for (I g = tid; g < npairs; g += nthr) {
            const I i0 = ((g >> p) << (p + 1)) | (g & (half - 1));
            if ((i0 & cm) != cv) continue;
            const I i1 = i0 | half;
            const cplx<T> x = a[i0], y = a[i1];
            a[i0] = cadd(cmul(u00, x), cmul(u01, y));
            a[i1] = cadd(cmul(u10, x), cmul(u11, y));
        }
        break;
    }
    case K_DIAG: {
        const I total = I(1) << nbits;
        for (I i = tid; i < total; i += nthr) {
            const int j = gather_bits<I>(i, op.bit, k);
            a[i] = cmul(mats[op.mat + j], a[i]);
        }
        break;
    }
    case K_COLLAPSE: {
        const I total = I(1) << nbits;
        const int p = op.bit[0];
        const double nrm = op.norm;  // renormalise in FP64 (exact division for T = double)
        for (I i = tid; i < total; i += nthr) {
            if (int((i >> p) & 1) == op.outcome) {
                cplx<T> v = a[i];
                v.re = T(double(v.re) / nrm);
                v.im = T(double(v.im) / nrm);
                a[i] = v;
            } else {
                a[i] = cplx<T>{T(0), T(0)};
            }
        }
        break;
    }
    case K_MONO:
        switch (k) {
        case 1: mono_k<T, I, 1>(a, nbits, op, mats, perms, tid, nthr); break;
        case 2: mono_k<T, I, 2>(a, nbits, op, mats, perms, tid, nthr); break;
        case 3: mono_k<T, I, 3>(a, nbits, op, mats, perms, tid, nthr); break;
        case 4: if constexpr (KMAX >= 4) mono_k<T, I, 4>(a, nbits, op, mats, perms, tid, nthr); break;
        default: if constexpr (KMAX >= 5) mono_k<T, I, 5>(a, nbits, op, mats, perms, tid, nthr); break;
        }
        break;
    case K_DENSEK:
        switch (k) {
        case 1: dense_k<T, I, 1>(a, nbits, op, mats, tid, nthr); break;
        case 2: dense_k<T, I, 2>(a, nbits, op, mats, tid, nthr); break;
        case 3: dense_k<T, I, 3>(a, nbits, op, mats, tid, nthr); break;
        case 4: if constexpr (KMAX >= 4) dense_k<T, I, 4>(a, nbits, op, mats, tid, nthr); break;
        default: if constexpr (KMAX >= 5) dense_k<T, I, 5>(a, nbits, op, mats, tid, nthr); break;
        }
        break;
    default:
        break;
    }
}

}  // namespace dq
