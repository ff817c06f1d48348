// dq_types.hpp -- plain data shared by the host planner and the kernels.
#pragma once
#include <stdint.h>

namespace dq {

enum OpKind : int {
    K_DENSE1 = 0,   // 2x2 on bit[0]
    K_DIAG = 1,     // diagonal 2^k, entries at mat
    K_MONO = 2,     // monomial: out[r] = ph[r] * in[perm[r]], ph at mat, perm at aux
    K_CTRL1 = 3,    // 2x2 on bit[k-1] where (i & cmask) == cval
    K_DENSEK = 4,   // dense 2^k x 2^k, acc = 0; acc += U[r][c] * v[c]
    K_COLLAPSE = 5,  // keep bit[0] == outcome (divide by norm), zero the rest
    K_FUSED = 6      // host only: a fused block of gates on 5 qubits (dense 32 x 32)
};

// One queued operation, positions relative to the index space it runs on
// (global qubits for direct launches, tile-local bits inside a tile pass).
struct DevOp {
    int kind;
    int k;
    int mat;          // offset (complex entries) into the matrix pool
    int aux;          // offset (ints) into the perm pool (K_MONO)
    int bit[5];       // bit[0] = MSB of the gate index
    int sbit[5];      // the same positions, ascending
    int outcome;
    int pad;
    uint64_t cmask;   // K_CTRL1
    uint64_t cval;
    double norm;      // K_COLLAPSE
};

enum TileKind : int {
    T_DENSE1 = 0,    // 2x2 on register slot p; variant: V_GEN / V_REAL / V_DIAG / V_ANTI
    T_CTRL1 = 1,     // 2x2 on register slot p where (tile index & cmask) == cval; V_GEN / V_ANTI
    T_DENSE2 = 2,    // 4x4 on register slots p (matrix MSB) and q; V_GEN / V_REAL / V_DIAG
    T_DIAG = 3,      // diagonal 2^k on any tile bits, elementwise
    T_COLLAPSE = 4,  // measurement collapse on tile bit bit[0], elementwise
    T_LDSOP = 5,     // any 2^k x 2^k matrix through LDS, one output per amplitude
    T_DIAGT = 6,     // diagonal, k <= 2: outer bits, thread bits and at most one register bit
                     // (slot p, gate-index bit q)
    T_FUSED5 = 7     // complex64 only: a fused 5-qubit unitary (32x32) on tile bits, as a
                     // real 64x64 GEMM on FP32 MFMA over the tile's 32-amplitude columns
};

// Matrix shape of a register op, chosen on the host.  V_GEN .. V_ANTI
// compute the reference's row sums term by term; V_REAL drops only the
// products of zero imaginary parts, V_DIAG / V_ANTI only the products of
// zero entries.  V_ROWPH and V_COLPH (4x4 only, contracted kernels only,
// i.e. fused ops of fusion levels 2-3) factor U = D R (each row a real row
// times one phase) or U = R D (each column): the op is a real 4x4 plus one
// complex product per amplitude, 12 FP64 instructions per amplitude instead
// of 16.  Their matrix slot holds R in the real parts of the 16 entries and
// phase k in the imaginary parts of entries 4k and 4k + 1.
enum OpVariant : int { V_GEN = 0, V_REAL = 1, V_DIAG = 2, V_ANTI = 3, V_ROWPH = 4, V_COLPH = 5 };

struct TileOp {
    // dispatch header: the first 12 words, read as one block by the kernel
    int kind;
    int relayout;  // 1: change to layout `rb` before this op
    int rb;        // layout active from this op: 4 tile bits, 4 bits each, ascending
    int p, q;      // register slots (see TileKind)
    int mat;       // offset (complex entries) into the matrix pool
    int variant;
    int nsel;      // T_DENSE1/T_DENSE2: > 0 selects matrix number sum_j bit(sel[j]) << j
    int lay;       // entry of layout `rb` in the pass's layout table
    int sel[3];    // outer-index bits of the selector
    // operands
    int k;
    int bit[5];    // tile bits, bit[0] = MSB of the gate index
    int sbit[5];   // ascending
    unsigned cmask, cval;
    int outcome;
    unsigned ocmask, ocval;  // T_CTRL1 controls on outer qubits (outer-index bits)
    int pad;
    double norm;
};

// A TileOp bit >= kOuterBit names outer-index bit (bit - kOuterBit) of the
// tile: the qubit is not in the tile and its value is constant over the tile.
constexpr int kOuterBit = 32;

struct TilePass {
    int red_bit;   // >= 0: sum |amp|^2 over the qubit at this tile/outer bit being 1 (per-block partials)
    int n_outer;
    int op_begin;
    int op_count;
    int load_rb;   // layout of the global load
    int store_rb;  // layout of the global store
    int lay_off;   // byte offset, from the matrix pool, of the pass's layout table:
                   // per layout, the LDS byte offsets swz(register bits of j) * amplitude size, j < 16
    int load_lay;  // table entries of load_rb and store_rb
    int store_lay;
    int tq[16];    // tile bit -> qubit
    int oq[48];    // outer bit -> qubit
};

}  // namespace dq
