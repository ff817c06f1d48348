// dq_plan.hpp -- host-side gate queue, classification and sweep planning.
//
// Pure C++ (no HIP calls): the engine uses it to turn a queue of gate calls
// into LDS-tiled sweeps, and the CPU test-suite exercises it directly.
#pragma once
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <deque>
#include <string>
#include <vector>

#include "dq_types.hpp"

namespace dq {

// Sets the library's error message (dq_last_error) and returns code.
int set_error(int code, const std::string &msg);

struct PlanConfig {
    int n = 0;
    int fuse = 2;        // 0: one sweep per op; 1: program-order passes; 2: passes with lookahead
    int tile_bits = 12;  // qubits per tile
    int low_bits = 4;    // qubits 0..low_bits-1 are in every tile (coalesced runs)
    int reg_tiles = 1;   // register-resident tiles for tiles of >= 9 qubits
    double fuse_min_cost = 128.0;  // fusion level 3: least cost of a fused block's gates
    int pair_fusion = 1;           // fusion levels 2-3: merge neighbouring gates on <= 2 qubits
    int reg_bits = 4;              // register bits of a register-tile layout (4: 16, 3: 8 amplitudes per thread)
    int phase_forms = 1;           // fusion levels 2-3: 4x4 ops factored as U = D R or R D (V_ROWPH / V_COLPH)
};

inline bool dq_tile_bits_supported(int b) { return b == 9 || b == 12 || b == 13; }

struct HostOp {
    int kind = K_DENSE1;
    int k = 1;
    int q[5] = {0, 0, 0, 0, 0};  // global qubits, q[0] = MSB of the gate index
    uint64_t cmask = 0, cval = 0;  // global masks (K_CTRL1)
    int outcome = 0;
    double norm = 1.0;
    std::vector<double> mat;    // (re, im) pairs: 4 (2x2), 2^k (diag/mono), 4^k (dense)
    std::vector<int> perm;      // K_MONO
    std::vector<double> dense;  // the full 2^k x 2^k matrix as given (register/LDS paths)
    int nsel = 0;               // > 0: one matrix per value of these outer qubits (pair fusion)
    int selq[3] = {0, 0, 0};    // selector qubits; selector bit j = qubit selq[j]
    uint64_t qmask() const {
        uint64_t m = 0;
        for (int j = 0; j < k; ++j) m |= uint64_t(1) << q[j];
        return m;
    }
};

struct Pass {
    uint64_t S = 0;  // tile qubits (mask); 0 = direct launch of a single op
    std::vector<int> ops;
};

inline int popcount64(uint64_t x) { return __builtin_popcountll(x); }


inline bool is_zero(const double *c) { return c[0] == 0.0 && c[1] == 0.0; }

// Classifies a dense 2^k matrix into the cheapest op kind that reproduces the
// dense formula exactly (see dq_device.hpp): diagonal (every off-diagonal
// entry is exactly zero), monomial (one nonzero per row, a permutation),
// controlled-2x2 (identity outside the bottom-right 2x2 block), else dense.
inline void classify(HostOp &op, const double *u) {
    const int dim = 1 << op.k;
    auto at = [&](int r, int c) { return u + 2 * (size_t(r) * dim + c); };
    op.dense.assign(u, u + 2 * size_t(dim) * dim);
    bool diag = true, mono = true;
    std::vector<int> perm(dim, -1), colcount(dim, 0);
    for (int r = 0; r < dim; ++r) {
        int nz = 0;
        for (int c = 0; c < dim; ++c) {
            if (is_zero(at(r, c))) continue;
            ++nz;
            perm[r] = c;
            if (c != r) diag = false;
        }
        if (nz != 1) mono = false;
        else colcount[perm[r]]++;
    }
    for (int c = 0; c < dim && mono; ++c)
        if (colcount[c] != 1) mono = false;
    if (diag) {
        op.kind = K_DIAG;
        op.mat.resize(2 * dim);
        for (int r = 0; r < dim; ++r) {
            op.mat[2 * r] = at(r, r)[0];
            op.mat[2 * r + 1] = at(r, r)[1];
        }
        return;
    }
    if (op.k == 1) {
        op.kind = K_DENSE1;
        op.mat.assign(u, u + 8);
        return;
    }
    // controlled 2x2: identity outside the bottom-right 2x2 block
    bool ctrl = true;
    for (int r = 0; r < dim && ctrl; ++r)
        for (int c = 0; c < dim; ++c) {
            if (r >= dim - 2 && c >= dim - 2) continue;
            const double *e = at(r, c);
            const bool want_one = (r == c);
            if (want_one ? !(e[0] == 1.0 && e[1] == 0.0) : !is_zero(e)) { ctrl = false; break; }
        }
    if (ctrl) {
        op.kind = K_CTRL1;
        op.mat.resize(8);
        for (int r = 0; r < 2; ++r)
            for (int c = 0; c < 2; ++c) {
                op.mat[2 * (2 * r + c)] = at(dim - 2 + r, dim - 2 + c)[0];
                op.mat[2 * (2 * r + c) + 1] = at(dim - 2 + r, dim - 2 + c)[1];
            }
        op.cmask = op.cval = 0;
        for (int j = 0; j < op.k - 1; ++j) {
            op.cmask |= uint64_t(1) << op.q[j];
            op.cval |= uint64_t(1) << op.q[j];
        }
        return;
    }
    if (mono) {
        op.kind = K_MONO;
        op.mat.resize(2 * dim);
        op.perm = perm;
        for (int r = 0; r < dim; ++r) {
            op.mat[2 * r] = at(r, perm[r])[0];
            op.mat[2 * r + 1] = at(r, perm[r])[1];
        }
        return;
    }
    op.kind = K_DENSEK;
    op.mat.assign(u, u + 2 * size_t(dim) * dim);
}


// Tile size of the passes: the configured size, or the whole state if it is
// smaller; register-resident tiles exist for 9, 12 and 13 qubits.
inline int effective_tile_bits(const PlanConfig *s) {
    const int b = std::min(s->tile_bits, s->n);
    if (s->fuse && s->reg_tiles && b >= 9 && !dq_tile_bits_supported(b)) return b < 12 ? 9 : 12;
    return b;
}

// Qubits an op needs inside the tile.  With `anywhere`, diagonal ops and
// collapses need none (an outer qubit's bit is constant over a tile) and a
// controlled 2x2 needs only its target.
inline uint64_t tile_need(const HostOp &h, bool anywhere) {
    if (!anywhere) return h.qmask();
    if (h.kind == K_DIAG || h.kind == K_COLLAPSE) return 0;
    if (h.kind == K_CTRL1) return uint64_t(1) << h.q[h.k - 1];
    return h.qmask();
}

inline bool is_diagonal(const HostOp &h) { return h.kind == K_DIAG || h.kind == K_COLLAPSE; }

// Ops among the first `window` of `remaining` that a pass with tile qubits S
// takes, by the pass builder's rules with S fixed.  `grow` receives, in order
// of first occurrence, the qubits each op that does not fit would add to S
// (ops already held back by deferred ops excluded).
inline int count_fitting(const PlanConfig *s, const std::vector<HostOp> &queue, const std::vector<int> &remaining,
                         size_t window, uint64_t S, bool anywhere, std::vector<uint64_t> *grow) {
    uint64_t nd_block = 0, d_block = 0;
    int taken = 0;
    const size_t end = std::min(window, remaining.size());
    for (size_t r = 0; r < end; ++r) {
        const HostOp &h = queue[remaining[r]];
        const uint64_t qm = h.qmask();
        const bool diag = anywhere && is_diagonal(h);
        const uint64_t block = (diag && s->fuse >= 2) ? nd_block : (nd_block | d_block);
        if (qm & block) {
            (diag ? d_block : nd_block) |= qm;
            continue;
        }
        const uint64_t extra = tile_need(h, anywhere) & ~S;
        if (!extra) {
            ++taken;
            continue;
        }
        nd_block |= qm;
        if (grow && std::find(grow->begin(), grow->end(), extra) == grow->end()) grow->push_back(extra);
    }
    return taken;
}

// Tile qubits of the next pass at fusion level 2 (register tiles): starting
// from the low block, repeatedly add the qubits of the waiting op whose
// addition lets the pass take the most ops (ties: fewer qubits, then the
// earliest op), judged over the next kPlanWindow ops, while the tile has
// room.  Taking qubits in order of first use instead lets a qubit that one
// early op touches crowd out qubits that many later ops need; on layered
// random circuits this choice cuts the passes by about a tenth.
constexpr size_t kPlanWindow = 256;

inline uint64_t choose_tile(const PlanConfig *s, const std::vector<HostOp> &queue, const std::vector<int> &remaining,
                            uint64_t S, int b, bool anywhere) {
    std::vector<uint64_t> grow;
    for (;;) {
        grow.clear();
        const int base = count_fitting(s, queue, remaining, kPlanWindow, S, anywhere, &grow);
        int best_gain = 0, best_bits = 0;
        uint64_t best = 0;
        for (uint64_t extra : grow) {
            const int bits = popcount64(extra);
            if (popcount64(S | extra) > b) continue;
            const int gain = count_fitting(s, queue, remaining, kPlanWindow, S | extra, anywhere, nullptr) - base;
            if (gain > best_gain || (gain == best_gain && gain > 0 && bits < best_bits)) {
                best_gain = gain;
                best_bits = bits;
                best = extra;
            }
        }
        if (!best) return S;
        S |= best;
    }
}

// Pass builder (see dq_engine.hip).  `anywhere` (register tiles only):
// diagonal ops may name outer qubits, and at fusion level 2 they may move
// past other diagonal ops (diagonal matrices commute).  At fusion level 2 with
// `anywhere` the tile qubits are chosen first (choose_tile); otherwise a
// pass takes qubits in order of first use.
inline std::vector<Pass> schedule(const PlanConfig *s, const std::vector<HostOp> &queue, bool anywhere = false) {
    std::vector<Pass> passes;
    const int n = s->n;
    const int b = effective_tile_bits(s);
    // leave room for a 5-qubit gate on qubits outside the low block
    const int low = b < n ? std::min(s->low_bits, b - 5) : b;
    const uint64_t low_mask = (low >= 64) ? ~uint64_t(0) : ((uint64_t(1) << low) - 1);
    std::vector<int> remaining(queue.size());
    for (size_t i = 0; i < remaining.size(); ++i) remaining[i] = int(i);
    while (!remaining.empty()) {
        Pass p;
        p.S = (anywhere && s->fuse >= 2 && b < n) ? choose_tile(s, queue, remaining, low_mask, b, true) : low_mask;
        uint64_t nd_block = 0, d_block = 0;  // qubits of deferred non-diagonal / diagonal ops
        std::vector<int> rest;
        for (int idx : remaining) {
            const HostOp &h = queue[idx];
            const uint64_t qm = h.qmask();
            const bool diag = anywhere && is_diagonal(h);
            const uint64_t block = (diag && s->fuse >= 2) ? nd_block : (nd_block | d_block);
            if (qm & block) {
                (diag ? d_block : nd_block) |= qm;
                rest.push_back(idx);
                continue;
            }
            const uint64_t need = tile_need(h, anywhere);
            if (popcount64(p.S | need) <= b) {
                p.S |= need;
                p.ops.push_back(idx);
            } else {
                nd_block |= qm;
                rest.push_back(idx);
                // program-order mode: the pass is a prefix of the queue
                if (s->fuse == 1) nd_block = d_block = ~uint64_t(0);
            }
        }
        // fill the tile up to b qubits with the lowest unused qubits
        for (int q = 0; q < n && popcount64(p.S) < b; ++q) p.S |= uint64_t(1) << q;
        passes.push_back(std::move(p));
        remaining.swap(rest);
    }
    return passes;
}

// ---- gate fusion into 5-qubit blocks (fusion level 3, complex64 tiles) ----

using cd = std::complex<double>;

// Full 2^k x 2^k matrix of a queued op (qubits h.q, q[0] = MSB).
inline std::vector<cd> op_full_matrix(const HostOp &h) {
    const int dim = 1 << h.k;
    std::vector<cd> m(size_t(dim) * dim, cd(0, 0));
    if (h.kind == K_CTRL1 && h.dense.size() == 8 && h.k > 1) {  // controls as given to dq_apply_1q
        for (int r = 0; r < dim; ++r) m[size_t(r) * dim + r] = cd(1, 0);
        for (int base = 0; base < dim; base += 2) {
            bool on = true;
            for (int j = 0; j < h.k - 1; ++j) {
                const int bitv = (base >> (h.k - 1 - j)) & 1;
                const int want = (h.cval >> h.q[j]) & 1;
                if (bitv != want) on = false;
            }
            if (!on) continue;
            for (int r = 0; r < 2; ++r)
                for (int c = 0; c < 2; ++c)
                    m[size_t(base + r) * dim + base + c] = cd(h.dense[2 * (2 * r + c)], h.dense[2 * (2 * r + c) + 1]);
        }
        return m;
    }
    for (size_t e = 0; e < m.size(); ++e) m[e] = cd(h.dense[2 * e], h.dense[2 * e + 1]);
    return m;
}

// M <- E(G) M, with E(G) the op's matrix embedded on the ordered qubit list qs
// (qs[0] = MSB of the 2^m index).
inline void embed_apply(std::vector<cd> &M, const std::vector<int> &qs, const HostOp &h) {
    const int m = int(qs.size()), dim = 1 << m, gk = h.k, gdim = 1 << gk;
    const std::vector<cd> G = op_full_matrix(h);
    int gpos[5];
    for (int j = 0; j < gk; ++j)
        for (int t = 0; t < m; ++t)
            if (qs[t] == h.q[j]) gpos[j] = m - 1 - t;  // bit position in the 2^m index
    std::vector<cd> out(M.size(), cd(0, 0));
    for (int r = 0; r < dim; ++r) {
        int gr = 0, rest = r;
        for (int j = 0; j < gk; ++j) {
            gr = (gr << 1) | ((r >> gpos[j]) & 1);
            rest &= ~(1 << gpos[j]);
        }
        for (int gc = 0; gc < gdim; ++gc) {
            const cd g = G[size_t(gr) * gdim + gc];
            if (g == cd(0, 0)) continue;
            int k = rest;
            for (int j = 0; j < gk; ++j)
                if ((gc >> (gk - 1 - j)) & 1) k |= 1 << gpos[j];
            for (int c = 0; c < dim; ++c) out[size_t(r) * dim + c] += g * M[size_t(k) * dim + c];
        }
    }
    M.swap(out);
}

constexpr int kFuseMinQubits = 3;  // fused blocks narrower than this stay per-gate
constexpr int kFuseMinOps = 4;     // ... and so do blocks with fewer gates

// Cost model of fusion level 3, in FP32 flops per amplitude.  A fused block
// is a 32x32 complex GEMM: 32 complex MACs = 256 flops per amplitude on MFMA,
// which runs at twice the rate of scalar FP32 VALU code, so it costs about
// 128 VALU flops per amplitude.  The gates it replaces cost (VALU, per
// amplitude): a dense 2x2 = 2 complex mul + 1 add = 14; a diagonal or
// monomial = 1 complex mul = 6; a dense 2^k gate = 2^k complex MACs = 8 * 2^k;
// a controlled 2x2 with c controls = 14 / 2^c.  A block is fused only when
// the gates it replaces cost at least PlanConfig::fuse_min_cost (default 128).

inline double gate_cost(const HostOp &h) {
    switch (h.kind) {
    case K_DENSE1: return 14.0;
    case K_DIAG:
    case K_MONO: return 6.0;
    case K_CTRL1: return 14.0 / double(1 << (h.k - 1));
    case K_DENSEK: return 8.0 * double(1 << h.k);
    default: return 0.0;
    }
}

// Rewrites a pass's op list into 5-qubit blocks (fusion level 3).  Greedy, in
// rounds: a round walks the remaining ops in order and takes every tile op
// that fits the open block (union of qubits <= 5) and touches no "frozen"
// qubit; an op that is not taken freezes its qubits, so any later op that
// would have to follow it stays behind too.  Every op taken therefore
// commutes with every op it overtakes (disjoint qubits).  Blocks of at least
// kFuseMinQubits qubits and kFuseMinOps gates, whose gates cost at least
// min_cost, become one K_FUSED op (their FP64 product, stored in
// `store`); other blocks keep their gates.
inline std::vector<const HostOp *> fuse_blocks(const std::vector<const HostOp *> &in, const int *pos, const int *tq,
                                               int b, double min_cost, std::deque<HostOp> &store) {
    std::vector<const HostOp *> out, rest(in), next;
    auto tile_op = [&](const HostOp &h) {
        // ops holding one matrix per outer-qubit value (pair fusion) act on
        // qubits outside the tile: they stay out of blocks
        if (h.kind == K_COLLAPSE || h.k > 5 || h.nsel > 0) return false;
        for (int j = 0; j < h.k; ++j)
            if (pos[h.q[j]] >= kOuterBit) return false;
        return true;
    };
    while (!rest.empty()) {
        size_t lead = 0;  // ops outside the tile at the front go out as they are
        while (lead < rest.size() && !tile_op(*rest[lead])) out.push_back(rest[lead++]);
        if (lead == rest.size()) break;
        std::vector<const HostOp *> blk;
        uint64_t bq = 0, frozen = 0;
        next.clear();
        for (size_t i = lead; i < rest.size(); ++i) {
            const HostOp &h = *rest[i];
            const uint64_t qm = h.qmask();
            if (tile_op(h) && !(qm & frozen) && __builtin_popcountll(bq | qm) <= 5) {
                blk.push_back(&h);
                bq |= qm;
            } else {
                next.push_back(&h);
                frozen |= qm;
            }
        }
        double cost = 0.0;
        for (const HostOp *h : blk) cost += gate_cost(*h);
        if (__builtin_popcountll(bq) >= kFuseMinQubits && int(blk.size()) >= kFuseMinOps && cost >= min_cost) {
            std::vector<int> qs;
            for (const HostOp *h : blk)
                for (int j = 0; j < h->k; ++j)
                    if (std::find(qs.begin(), qs.end(), h->q[j]) == qs.end()) qs.push_back(h->q[j]);
            for (int t = b - 1; t >= 0 && qs.size() < 5; --t)  // pad with other tile qubits
                if (std::find(qs.begin(), qs.end(), tq[t]) == qs.end()) qs.push_back(tq[t]);
            std::vector<cd> M(32 * 32, cd(0, 0));
            for (int r = 0; r < 32; ++r) M[size_t(r) * 32 + r] = cd(1, 0);
            for (const HostOp *h : blk) embed_apply(M, qs, *h);
            HostOp f;
            f.kind = K_FUSED;
            f.k = 5;
            f.outcome = int(blk.size());  // gates in the block (statistics only)
            for (int j = 0; j < 5; ++j) f.q[j] = qs[j];
            f.dense.resize(2 * 32 * 32);
            for (size_t e = 0; e < M.size(); ++e) {
                f.dense[2 * e] = M[e].real();
                f.dense[2 * e + 1] = M[e].imag();
            }
            store.push_back(std::move(f));
            out.push_back(&store.back());
        } else {
            out.insert(out.end(), blk.begin(), blk.end());
        }
        rest.swap(next);
    }
    return out;
}

// Pair fusion (fusion levels 2 and 3): gates on at most 2 tile qubits that
// are neighbours on every tile qubit's timeline merge into one op, their FP64
// product, classified again (a product of diagonals stays diagonal).  Walk
// the ops in order keeping, per tile qubit, the op that last touched it (if
// it is a mergeable op):
//   * a 1-qubit gate on a joins the last op on a;
//   * a 2-qubit gate on (a, b) joins the last op on a if that op is the last
//     op on b too (same pair); else it absorbs the last ops on a and b when
//     those act on a (resp. b) alone;
//   * any other op (3+ tile qubits, collapse) is a barrier on its qubits.
// A merged gate only moves past ops that touch none of its tile qubits (the
// op it joins is the last one on each of them), so the product is the same
// operator as the sequence.
// Gates may also act on qubits outside the tile if they are diagonal in them
// (diagonal gates, controls of a controlled 2x2): every op of a tile pass
// touches outer qubits only diagonally, so such qubits never order ops.  The
// merged op then holds one matrix per value of its (at most kMaxSel) outer
// qubits, which are constant over a tile: the sweep picks the matrix per tile.
constexpr int kMaxSel = 3;

inline std::vector<const HostOp *> fuse_pairs(const std::vector<const HostOp *> &in, const int *pos,
                                              std::deque<HostOp> &store) {
    struct FOp {
        std::vector<const HostOp *> gates;
        uint64_t qm = 0;  // tile qubits
        uint64_t om = 0;  // outer (selector) qubits
        bool dead = false;
    };
    std::vector<FOp> ops;
    ops.reserve(in.size());
    int last[64];
    for (int q = 0; q < 64; ++q) last[q] = -1;
    auto split = [&](const HostOp &h, uint64_t &tm, uint64_t &om) {
        tm = om = 0;
        for (int j = 0; j < h.k; ++j) (pos[h.q[j]] < kOuterBit ? tm : om) |= uint64_t(1) << h.q[j];
    };
    auto mergeable = [&](const HostOp &h, uint64_t tm, uint64_t om) {
        if (h.kind == K_COLLAPSE) return false;
        const int nt = __builtin_popcountll(tm);
        if (nt < 1 || nt > 2 || __builtin_popcountll(om) > kMaxSel) return false;
        if (!om) return true;
        if (h.kind == K_DIAG) return true;
        // a controlled 2x2 with every outer qubit among its controls
        return h.kind == K_CTRL1 && pos[h.q[h.k - 1]] < kOuterBit;
    };
    for (const HostOp *hp : in) {
        const HostOp &h = *hp;
        uint64_t tm, om;
        split(h, tm, om);
        if (!mergeable(h, tm, om)) {
            FOp f;
            f.gates.push_back(hp);
            f.qm = tm;
            f.om = om;
            ops.push_back(f);
            for (int j = 0; j < h.k; ++j) last[h.q[j]] = -1;
            continue;
        }
        const int a = __builtin_ctzll(tm);
        const int b = __builtin_popcountll(tm) == 2 ? 63 - __builtin_clzll(tm) : -1;
        auto fits = [&](const FOp &f) { return __builtin_popcountll(f.om | om) <= kMaxSel; };
        if (b < 0) {
            if (last[a] >= 0 && fits(ops[size_t(last[a])])) {
                FOp &F = ops[size_t(last[a])];
                F.gates.push_back(hp);
                F.om |= om;
                continue;
            }
            FOp f;
            f.gates.push_back(hp);
            f.qm = tm;
            f.om = om;
            ops.push_back(f);
            last[a] = int(ops.size() - 1);
            continue;
        }
        const int la = last[a], lb = last[b];
        if (la >= 0 && la == lb && fits(ops[size_t(la)])) {  // the last op on both is on exactly this pair
            ops[size_t(la)].gates.push_back(hp);
            ops[size_t(la)].om |= om;
            continue;
        }
        // a new op at this position absorbs the single-qubit predecessors on a
        // and b (each is the last op on its qubit, so it moves forward past
        // ops that do not touch that qubit)
        FOp f;
        f.om = om;
        for (const int l : {la, lb}) {
            if (l < 0) continue;
            FOp &P = ops[size_t(l)];
            if (P.dead || (P.qm != (uint64_t(1) << a) && P.qm != (uint64_t(1) << b))) continue;
            if (__builtin_popcountll(f.om | P.om) > kMaxSel) continue;
            f.gates.insert(f.gates.end(), P.gates.begin(), P.gates.end());
            f.om |= P.om;
            P.dead = true;
        }
        f.gates.push_back(hp);
        f.qm = tm;
        ops.push_back(f);
        last[a] = last[b] = int(ops.size() - 1);
    }
    std::vector<const HostOp *> out;
    out.reserve(ops.size());
    for (const FOp &f : ops) {
        if (f.dead) continue;
        if (f.gates.size() == 1) {
            out.push_back(f.gates[0]);
            continue;
        }
        // tile qubits in order of first appearance, then the outer qubits
        std::vector<int> tqs, oqs;
        for (const HostOp *g : f.gates)
            for (int j = 0; j < g->k; ++j) {
                const int q = g->q[j];
                auto &v = pos[q] < kOuterBit ? tqs : oqs;
                if (std::find(v.begin(), v.end(), q) == v.end()) v.push_back(q);
            }
        std::vector<int> qs(tqs);
        qs.insert(qs.end(), oqs.begin(), oqs.end());
        const int m = int(qs.size()), dim = 1 << m;
        std::vector<cd> M(size_t(dim) * dim, cd(0, 0));
        for (int r = 0; r < dim; ++r) M[size_t(r) * dim + r] = cd(1, 0);
        for (const HostOp *g : f.gates) embed_apply(M, qs, *g);
        std::vector<double> flat(2 * M.size());
        for (size_t e = 0; e < M.size(); ++e) {
            flat[2 * e] = M[e].real();
            flat[2 * e + 1] = M[e].imag();
        }
        HostOp fused;
        fused.k = m;
        for (int j = 0; j < m; ++j) fused.q[j] = qs[size_t(j)];
        classify(fused, flat.data());
        if (!oqs.empty() && !(fused.kind == K_DIAG && m <= 2)) {
            // one 2^kt x 2^kt matrix per value of the outer qubits (the
            // product is block-diagonal in them): rows and columns of M whose
            // outer bits (the low ones, after the tile qubits) equal s
            const int kt = int(tqs.size()), no = int(oqs.size()), sd = 1 << kt;
            HostOp sel;
            sel.kind = kt == 1 ? K_DENSE1 : K_DENSEK;
            sel.k = kt;
            for (int j = 0; j < kt; ++j) sel.q[j] = tqs[size_t(j)];
            sel.nsel = no;
            for (int j = 0; j < no; ++j) sel.selq[j] = oqs[size_t(j)];
            sel.dense.resize(size_t(2) * (size_t(1) << no) * sd * sd);
            size_t w = 0;
            for (int sv = 0; sv < (1 << no); ++sv) {
                // selector value bit j = outer qubit oqs[j], which is index bit no-1-j of M's low part
                int low = 0;
                for (int j = 0; j < no; ++j)
                    if ((sv >> j) & 1) low |= 1 << (no - 1 - j);
                for (int r = 0; r < sd; ++r)
                    for (int c = 0; c < sd; ++c) {
                        const cd v = M[size_t((r << no) | low) * dim + size_t((c << no) | low)];
                        sel.dense[w++] = v.real();
                        sel.dense[w++] = v.imag();
                    }
            }
            sel.mat = sel.dense;
            store.push_back(std::move(sel));
        } else {
            store.push_back(std::move(fused));
        }
        out.push_back(&store.back());
    }
    return out;
}

// ---- register-resident tile passes (k_tile_reg) ----


// A layout of r register bits is packed as r 4-bit tile-bit numbers, ascending.
inline int pack_rb(unsigned mask, int r = 4) {
    int packed = 0, m = 0;
    for (int bpos = 0; bpos < 16 && m < r; ++bpos)
        if ((mask >> bpos) & 1) packed |= bpos << (4 * m++);
    return packed;
}

inline unsigned unpack_rb(int packed, int r = 4) {
    unsigned m = 0;
    for (int j = 0; j < r; ++j) m |= 1u << ((packed >> (4 * j)) & 15);
    return m;
}

inline int rb_slot(int packed, int tile_bit, int r = 4) {
    for (int j = 0; j < r; ++j)
        if (((packed >> (4 * j)) & 15) == tile_bit) return j;
    return -1;
}

// Variant of a 2x2 or 4x4 register op (see OpVariant).
inline int op_variant(const std::vector<double> &m, int dim) {
    const size_t cnt = m.size() / 2;
    bool real = true, diag = true, anti = true;
    for (size_t e = 0; e < cnt; ++e) {
        const double re = m[2 * e], im = m[2 * e + 1];
        const int r = int(e % size_t(dim * dim)) / dim, c = int(e) % dim;  // several matrices: per matrix
        const bool zero = re == 0.0 && im == 0.0;
        if (im != 0.0) real = false;
        if (r != c && !zero) diag = false;
        if (r + c != dim - 1 && !zero) anti = false;
    }
    if (diag) return V_DIAG;
    if (anti && dim == 2) return V_ANTI;
    if (real) return V_REAL;
    return V_GEN;
}

// Phase-factored forms of 4x4 matrices (V_ROWPH / V_COLPH; see OpVariant).
// `m` holds one or more 4x4 matrices (the selector blocks of an op).  A row
// (column) qualifies when every entry, turned by the phase of its largest
// entry, has an imaginary part within 4e-15 of that entry's magnitude, the
// rounding of a product of a few gate matrices; the dropped imaginary parts
// are of that order.  Returns the variant and rewrites `m` into its encoding,
// or returns V_GEN and leaves `m` alone.
inline int phase_factor4(std::vector<double> &m) {
    const size_t nblk = m.size() / 32;
    if (nblk == 0 || m.size() % 32) return V_GEN;
    for (const double x : m)
        if (!std::isfinite(x)) return V_GEN;  // NaN / inf entries stay general
    for (const int form : {V_ROWPH, V_COLPH}) {
        std::vector<double> enc(m.size(), 0.0);
        bool ok = true;
        for (size_t b = 0; b < nblk && ok; ++b) {
            const double *u = m.data() + 32 * b;
            double *e = enc.data() + 32 * b;
            for (int k = 0; k < 4 && ok; ++k) {
                // entry j of row k (V_ROWPH) or of column k (V_COLPH)
                auto at = [&](int j) { return form == V_ROWPH ? 4 * k + j : 4 * j + k; };
                int jm = 0;
                double am = -1.0;
                for (int j = 0; j < 4; ++j) {
                    const double a = std::hypot(u[2 * at(j)], u[2 * at(j) + 1]);
                    if (a > am) {
                        am = a;
                        jm = j;
                    }
                }
                if (!(am > 0.0)) {
                    ok = false;
                    break;
                }
                const double pr = u[2 * at(jm)] / am, pi = u[2 * at(jm) + 1] / am;
                for (int j = 0; j < 4; ++j) {
                    const double ur = u[2 * at(j)], ui = u[2 * at(j) + 1];
                    const double wr = ur * pr + ui * pi, wi = ui * pr - ur * pi;  // u * conj(phase)
                    if (!(std::fabs(wi) <= 4e-15 * am)) {
                        ok = false;
                        break;
                    }
                    e[2 * at(j)] = wr;
                }
                e[2 * (4 * k) + 1] = pr;
                e[2 * (4 * k + 1) + 1] = pi;
            }
        }
        if (ok) {
            m.swap(enc);
            return form;
        }
    }
    return V_GEN;
}

inline std::vector<double> diag_to_dense(const std::vector<double> &d) {
    const size_t dim = d.size() / 2;
    std::vector<double> m(2 * dim * dim, 0.0);
    for (size_t r = 0; r < dim; ++r) {
        m[2 * (r * dim + r)] = d[2 * r];
        m[2 * (r * dim + r) + 1] = d[2 * r + 1];
    }
    return m;
}

struct RegPlan {
    TilePass pd;
    int b;
    int reg_bits = 4;
    std::vector<TileOp> ops;
    std::vector<int> layouts;  // distinct layouts of the pass (entries of its layout table)
    int fused_blocks = 0;  // K_FUSED ops formed in this pass
    int fused_gates = 0;   // gates they replace
    int relayouts = 0;     // LDS layout changes before ops, the store's included
};

// Register-layout order of a pass's ops (fusion level 2 and up).  A layout
// holds 4 tile bits; every change of layout is an LDS round trip of the whole
// tile with two barriers.  Ops commute when their tile bits are disjoint or
// both are diagonal (ops touch outer qubits only diagonally), so they may be
// reordered within those limits: repeatedly run every ready op whose register
// bits fit the current layout, then pick the next layout as the one (built
// from a ready op and the needs of the ops after it, as the layout planner
// builds them) that lets the most ops run.  Passes of more than kOrderMaxOps
// ops keep program order (the search is quadratic in the op count).
constexpr size_t kOrderMaxOps = 64;

inline std::vector<const HostOp *> order_for_layouts(const std::vector<const HostOp *> &in, const int *pos,
                                                     int reg_bits = 4) {
    const size_t n = in.size();
    if (n < 3 || n > kOrderMaxOps) return in;
    std::vector<unsigned> need(n, 0), soft(n, 0), qm(n, 0);
    std::vector<char> diag(n, 0);
    for (size_t i = 0; i < n; ++i) {
        const HostOp &h = *in[i];
        bool all_inner = true;
        for (int j = 0; j < h.k; ++j) {
            if (pos[h.q[j]] < kOuterBit) qm[i] |= 1u << pos[h.q[j]];
            else all_inner = false;
        }
        diag[i] = (h.kind == K_DIAG || h.kind == K_COLLAPSE) ? 1 : 0;
        if (h.kind == K_DENSE1 || h.kind == K_MONO || h.kind == K_DENSEK) {
            if (h.k <= 2) need[i] = qm[i];
        } else if (h.kind == K_CTRL1) {
            need[i] = 1u << pos[h.q[h.k - 1]];
        } else if (h.kind == K_DIAG && h.k <= 2 && all_inner) {
            soft[i] = qm[i];
        }
    }
    // i must precede j: shared tile bits, not both diagonal
    std::vector<std::vector<int>> succ(n);
    std::vector<int> indeg(n, 0);
    for (size_t j = 0; j < n; ++j)
        for (size_t i = 0; i < j; ++i)
            if ((qm[i] & qm[j]) && !(diag[i] && diag[j])) {
                succ[i].push_back(int(j));
                ++indeg[j];
            }
    std::vector<char> done(n, 0);
    std::vector<int> deg = indeg;
    // runs every op that becomes ready and fits layout S, in program order;
    // returns how many ran (the state is updated in place)
    auto run = [&](unsigned S, std::vector<char> &dn, std::vector<int> &dg, std::vector<int> *order) {
        int cnt = 0;
        for (bool progress = true; progress;) {
            progress = false;
            for (size_t k = 0; k < n; ++k) {
                if (dn[k] || dg[k] || (need[k] & ~S)) continue;
                dn[k] = 1;
                for (int j : succ[k]) --dg[j];
                if (order) order->push_back(int(k));
                ++cnt;
                progress = true;
            }
        }
        return cnt;
    };
    std::vector<int> order;
    run(0u, done, deg, &order);
    std::vector<unsigned> cands;
    while (order.size() < n) {
        cands.clear();
        for (size_t k = 0; k < n; ++k) {
            if (done[k] || deg[k]) continue;
            for (int mode = 0; mode < 2; ++mode) {
                unsigned S = need[k];
                for (size_t j = 0; j < n; ++j) {
                    if (done[j] || j == k) continue;
                    const unsigned nd = need[j] ? need[j] : (mode ? soft[j] : 0u);
                    if (!nd || (nd & ~S) == 0) continue;
                    if (__builtin_popcount(S | nd) <= reg_bits) S |= nd;
                    else if (need[j] && mode == 0) break;
                }
                if (std::find(cands.begin(), cands.end(), S) == cands.end()) cands.push_back(S);
            }
        }
        int best = -1;
        unsigned bestS = 0;
        for (unsigned S : cands) {
            std::vector<char> dn = done;
            std::vector<int> dg = deg;
            const int c = run(S, dn, dg, nullptr);
            if (c > best) {
                best = c;
                bestS = S;
            }
        }
        if (best <= 0) break;  // cannot happen: a ready op always fits a layout built from it
        run(bestS, done, deg, &order);
    }
    if (order.size() != n) return in;
    std::vector<const HostOp *> out(n);
    for (size_t i = 0; i < n; ++i) out[i] = in[size_t(order[i])];
    return out;
}

// The kernel's LDS swizzle of a tile index (dq_tile.hpp: swz): the low nibble
// XORed with every higher nibble of the b tile bits.
inline unsigned tile_swizzle(unsigned i, int b) {
    unsigned s = 0;
    for (int sh = 4; sh < b; sh += 4) s ^= (i >> sh) & 15u;
    return i ^ s;
}

// One layout-table entry: for register j < 2^r of layout rb, the LDS byte
// offset of its register bits, swizzled, for amplitudes of elem bytes.
inline void append_layout_table(std::vector<uint32_t> &out, int rb, int b, unsigned elem, int r = 4) {
    for (int j = 0; j < (1 << r); ++j) {
        unsigned v = 0;
        for (int m = 0; m < r; ++m)
            if ((j >> m) & 1) v |= 1u << ((rb >> (4 * m)) & 15);
        out.push_back(tile_swizzle(v, b) * elem);
    }
}

// Qubits on which a queued op can turn an amplitude that is zero at 0 into a
// nonzero one: every qubit it acts on, except those on which it acts
// diagonally (a diagonal op or a collapse scales each amplitude by a factor;
// a controlled op acts on its target only where its controls match).
inline uint64_t op_support(const HostOp &h) {
    if (h.kind == K_DIAG || h.kind == K_COLLAPSE) return 0;
    if (h.kind == K_CTRL1) return uint64_t(1) << h.q[h.k - 1];
    return h.qmask();
}

// Qubits known to be 0 (zero) or 1 (one) in every amplitude.
struct KnownBits {
    uint64_t zero = 0, one = 0;
};

// Whether a queued op is a 2x2 with exact zeros on its diagonal (x, y, any
// anti-diagonal) on one qubit: it maps a qubit that is 1 everywhere to one
// that is 0 everywhere and back, each new amplitude an exact product with
// one old amplitude plus an exact zero product.
inline bool is_antidiag1(const HostOp &h) {
    if (h.k != 1 || h.mat.size() != 8) return false;
    if (!(h.kind == K_DENSE1 || (h.kind == K_CTRL1 && h.cmask == 0))) return false;
    return is_zero(h.mat.data()) && is_zero(h.mat.data() + 6);
}

// Advances the known bits over one register pass, its ops in queue order
// (program order; a pass keeps the order of ops that do not commute): a
// collapse makes its qubit known (it writes exact zeros over the other
// half), an anti-diagonal 2x2 swaps 0 and 1, diagonal ops and controls keep
// what is known, and any other op forgets its support.
inline void advance_known(const std::vector<HostOp> &queue, const Pass &p, KnownBits &kb) {
    std::vector<int> idx(p.ops);
    std::sort(idx.begin(), idx.end());
    for (int i : idx) {
        const HostOp &h = queue[size_t(i)];
        const uint64_t b = uint64_t(1) << h.q[h.k - 1];
        if (h.kind == K_COLLAPSE) {
            (h.outcome ? kb.one : kb.zero) |= b;
            (h.outcome ? kb.zero : kb.one) &= ~b;
        } else if (is_antidiag1(h)) {
            const bool z = (kb.zero & b) != 0, o = (kb.one & b) != 0;
            kb.zero = (kb.zero & ~b) | (o ? b : 0);
            kb.one = (kb.one & ~b) | (z ? b : 0);
        } else {
            kb.zero &= ~op_support(h);
            kb.one &= ~op_support(h);
        }
    }
}

// Plans the layout of every op of one pass and encodes the ops.  zero_q:
// qubits that are 0 in every amplitude before the pass.  Outer qubits among
// them take the highest outer-index bits and are left out of pd.n_outer, so a
// launch over 2^n_outer tiles visits exactly the tiles where they are 0 (the
// others hold zeros, which the pass would leave zero), and ops that read such
// a qubit's outer bit (controls, selectors, diagonals, collapses) read its
// value there, 0.
inline void plan_pass(const PlanConfig *s, const std::vector<HostOp> &queue, const Pass &p, std::vector<double> &mats,
               RegPlan &out, uint64_t zero_q = 0) {
    const int n = s->n;
    TilePass &pd = out.pd;
    std::memset(&pd, 0, sizeof pd);
    int pos[64];  // tile bit of a tile qubit; kOuterBit + outer-index bit otherwise
    int tb = 0, ob = 0;
    for (int q = 0; q < n; ++q) {
        if ((p.S >> q) & 1) {
            pd.tq[tb] = q;
            pos[q] = tb++;
        } else if (!((zero_q >> q) & 1)) {
            pd.oq[ob] = q;
            pos[q] = kOuterBit + ob++;
        }
    }
    pd.n_outer = ob;
    for (int q = 0; q < n; ++q)
        if (!((p.S >> q) & 1) && ((zero_q >> q) & 1)) {
            pd.oq[ob] = q;
            pos[q] = kOuterBit + ob++;
        }
    const int b = tb;
    out.b = b;
    const int R = s->reg_bits;  // register bits per layout
    out.reg_bits = R;
    auto inner = [&](int bitpos) { return bitpos < kOuterBit; };
    // the low qubits stay thread bits at load/store (coalesced runs)
    const unsigned lowmask = (1u << std::max(0, std::min(s->low_bits, b - R))) - 1;
    const unsigned dflt = ((1u << b) - 1) & ~((1u << (b - R)) - 1);
    std::deque<HostOp> fused_store, pair_store;
    std::vector<const HostOp *> eff;
    for (int idx : p.ops) eff.push_back(&queue[idx]);
    if (s->fuse >= 2 && s->pair_fusion) eff = fuse_pairs(eff, pos, pair_store);
    if (s->fuse == 3 && b >= 12) {
        eff = fuse_blocks(eff, pos, pd.tq, b, s->fuse_min_cost, fused_store);
        for (const HostOp &f : fused_store) {
            ++out.fused_blocks;
            out.fused_gates += f.outcome;
        }
    }
    if (s->fuse >= 2) eff = order_for_layouts(eff, pos, R);
    const size_t nops = eff.size();
    std::vector<unsigned> hard(nops, 0), soft(nops, 0);
    for (size_t i = 0; i < nops; ++i) {
        const HostOp &h = *eff[i];
        unsigned qm = 0;
        bool all_inner = true;
        for (int j = 0; j < h.k; ++j) {
            if (inner(pos[h.q[j]])) qm |= 1u << pos[h.q[j]];
            else all_inner = false;
        }
        switch (h.kind) {
        case K_DENSE1:
        case K_MONO:
        case K_DENSEK:
            if (h.k <= 2) hard[i] = qm;
            break;
        case K_CTRL1:
            hard[i] = 1u << pos[h.q[h.k - 1]];
            break;
        case K_DIAG:
            if (h.k <= 2 && all_inner) soft[i] = qm;
            break;
        default:
            break;
        }
    }
    auto build = [&](size_t i, unsigned curmask) {
        unsigned S = hard[i];
        for (size_t j = i + 1; j < nops && j < i + 64; ++j) {
            const unsigned need = hard[j] ? hard[j] : soft[j];
            if (!need || (need & ~S) == 0) continue;
            if (__builtin_popcount(S | need) <= R) S |= need;
            else if (hard[j]) break;
        }
        // fill to R bits: keep current register bits, then high tile bits, then the rest
        for (int bpos = b - 1; bpos >= 0 && __builtin_popcount(S) < R; --bpos)
            if (((curmask >> bpos) & 1) && !((lowmask >> bpos) & 1)) S |= 1u << bpos;
        for (int bpos = b - 1; bpos >= 0 && __builtin_popcount(S) < R; --bpos)
            if (!((lowmask >> bpos) & 1)) S |= 1u << bpos;
        for (int bpos = b - 1; bpos >= 0 && __builtin_popcount(S) < R; --bpos) S |= 1u << bpos;
        return S;
    };
    std::vector<int> lay(nops, -1), relay(nops, 0);
    int cur = -1;
    for (size_t i = 0; i < nops; ++i) {
        if (hard[i] && (cur < 0 || (hard[i] & ~unpack_rb(cur, R)))) {
            cur = pack_rb(build(i, cur < 0 ? 0u : unpack_rb(cur, R)), R);
            relay[i] = 1;
        }
        lay[i] = cur;
    }
    // load layout: the first planned layout if it keeps the low qubits as thread bits
    size_t first = nops;
    for (size_t i = 0; i < nops; ++i)
        if (relay[i]) { first = i; break; }
    if (first < nops && !(unpack_rb(lay[first], R) & lowmask)) {
        pd.load_rb = lay[first];
        relay[first] = 0;
    } else {
        pd.load_rb = pack_rb(dflt, R);
    }
    for (size_t i = 0; i < first && i < nops; ++i) lay[i] = pd.load_rb;
    const int last = nops ? lay[nops - 1] : pd.load_rb;
    pd.store_rb = (unpack_rb(last, R) & lowmask) ? pack_rb(dflt, R) : last;
    auto lay_index = [&](int rb) {
        for (size_t e = 0; e < out.layouts.size(); ++e)
            if (out.layouts[e] == rb) return int(e);
        out.layouts.push_back(rb);
        return int(out.layouts.size() - 1);
    };
    pd.load_lay = lay_index(pd.load_rb);
    pd.store_lay = lay_index(pd.store_rb);
    out.relayouts = int(std::count(relay.begin(), relay.end(), 1)) + (last != pd.store_rb ? 1 : 0);

    for (size_t i = 0; i < nops; ++i) {
        const HostOp &h = *eff[i];
        TileOp t;
        std::memset(&t, 0, sizeof t);
        t.k = h.k;
        t.relayout = relay[i];
        t.rb = lay[i];
        t.lay = lay_index(lay[i]);
        for (int j = 0; j < h.k; ++j) t.bit[j] = pos[h.q[j]];
        for (int j = 0; j < h.k; ++j) t.sbit[j] = t.bit[j];
        std::sort(t.sbit, t.sbit + h.k);
        const unsigned regmask = unpack_rb(lay[i], R);
        unsigned qm = 0;
        bool all_inner = true;
        for (int j = 0; j < h.k; ++j) {
            if (inner(t.bit[j])) qm |= 1u << t.bit[j];
            else all_inner = false;
        }
        const std::vector<double> *m = &h.mat;
        std::vector<double> tmp;
        switch (h.kind) {
        case K_DENSE1:
        case K_MONO:
        case K_DENSEK:
            if (h.k == 1) {
                t.kind = T_DENSE1;
                t.p = rb_slot(lay[i], t.bit[0], R);
                m = &h.dense;
            } else if (h.k == 2) {
                t.kind = T_DENSE2;
                t.p = rb_slot(lay[i], t.bit[0], R);
                t.q = rb_slot(lay[i], t.bit[1], R);
                m = &h.dense;
            } else {
                t.kind = T_LDSOP;
                m = &h.dense;
            }
            break;
        case K_CTRL1:
            t.kind = T_CTRL1;
            t.p = rb_slot(lay[i], t.bit[h.k - 1], R);
            for (int j = 0; j < h.k - 1; ++j) {
                const bool one = (h.cval & (uint64_t(1) << h.q[j])) != 0;
                if (inner(t.bit[j])) {
                    t.cmask |= 1u << t.bit[j];
                    if (one) t.cval |= 1u << t.bit[j];
                } else {
                    t.ocmask |= 1u << (t.bit[j] - kOuterBit);
                    if (one) t.ocval |= 1u << (t.bit[j] - kOuterBit);
                }
            }
            break;
        case K_DIAG:
            if (h.k <= 2 && all_inner && (qm & ~regmask) == 0) {
                tmp = diag_to_dense(h.mat);
                m = &tmp;
                t.kind = h.k == 1 ? T_DENSE1 : T_DENSE2;
                t.p = rb_slot(lay[i], t.bit[0], R);
                if (h.k == 2) t.q = rb_slot(lay[i], t.bit[1], R);
            } else if (h.k <= 2 && __builtin_popcount(qm & regmask) <= 1) {
                // thread bits (and at most one register bit): T_DIAGT
                t.kind = T_DIAGT;
                t.p = -1;
                t.q = -1;
                for (int j = 0; j < h.k; ++j)
                    if (inner(t.bit[j]) && ((regmask >> t.bit[j]) & 1)) {
                        t.p = rb_slot(lay[i], t.bit[j], R);
                        t.q = j;
                    }
            } else {
                t.kind = T_DIAG;
            }
            break;
        case K_COLLAPSE:
            t.kind = T_COLLAPSE;
            t.outcome = h.outcome;
            t.norm = h.norm;
            break;
        case K_FUSED:
            t.kind = T_FUSED5;
            m = &h.dense;
            break;
        default:
            break;
        }
        t.mat = int(mats.size() / 2);
        if (t.kind == T_DENSE1 || t.kind == T_DENSE2) t.variant = op_variant(*m, t.kind == T_DENSE1 ? 2 : 4);
        if (t.kind == T_DENSE2 && t.variant == V_GEN && s->fuse >= 2 && s->phase_forms) {
            // contracted kernels only: U = D R or R D as a real op and phases
            if (m != &tmp) tmp = *m;
            const int v = phase_factor4(tmp);
            if (v != V_GEN) {
                t.variant = v;
                m = &tmp;
            }
        }
        mats.insert(mats.end(), m->begin(), m->end());
        if (t.kind == T_CTRL1) t.variant = op_variant(*m, 2) == V_ANTI ? V_ANTI : V_GEN;
        if ((t.kind == T_DENSE1 || t.kind == T_DENSE2) && h.nsel > 0) {
            t.nsel = h.nsel;
            for (int j = 0; j < h.nsel; ++j) t.sel[j] = pos[h.selq[j]] - kOuterBit;
        }
        out.ops.push_back(t);
    }
}


}  // namespace dq
