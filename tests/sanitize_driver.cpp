// sanitize_driver.cpp -- runs the host-side C++ of the engine (JSON decoder,
// gate library, ChaCha8 stream, sweep planner) on circuits given as files,
// built with AddressSanitizer and UndefinedBehaviorSanitizer by
// tests/test_sanitize.py.  No GPU code is involved: the planner headers are
// pure C++.
//
// For every circuit and every planner configuration (fusion level, tile
// qubits, low qubits, register bits) it decodes the circuit, builds each
// unitary's matrix, classifies it, splits the op stream at measurements,
// schedules the passes, plans every pass (pair fusion, 5-qubit blocks,
// layout order, layouts) and encodes the layout tables,
// then checks the plan's invariants: every op lands in exactly one pass,
// register slots are below the layout's register bits and distinct, a tile
// holds the configured number of qubits, and an emulation of the tile's
// amplitudes through the pass (emulate_pass) agrees with the kernel's view of
// every layout.  Exit status 0 means no sanitizer report and no broken
// invariant.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../dqsim_bosonic_amd/csrc/dq_plan.hpp"
#include "../dqsim_bosonic_amd/csrc/host/circuit.hpp"
#include "../dqsim_bosonic_amd/csrc/host/gates.hpp"
#include "../dqsim_bosonic_amd/csrc/host/rng.hpp"

namespace dq {
int set_error(int code, const std::string &) { return code; }
}  // namespace dq

namespace {

int failures = 0;
#define EXPECT(cond)                                                            \
    do {                                                                        \
        if (!(cond)) {                                                          \
            std::fprintf(stderr, "invariant failed: %s (line %d)\n", #cond, __LINE__); \
            ++failures;                                                         \
        }                                                                       \
    } while (0)

// Generic gate names of the statevector dispatcher that have a matrix.
const char *generic_kind(const std::string &name) {
    if (name == "nonlocal_cz" || name == "remote_cz") return "cz";
    if (name == "remote_cx") return "cx";
    if (name == "remote_epr") return "remote_link_phi_plus";
    if (name == "remote_link_phi_plus" || name == "remote_link_psi_minus" || name == "remote_link_psi_plus")
        return name.c_str();
    return nullptr;
}

bool unitary(const dqh::Instr &in, dqh::Mat &m) {
    if (in.kind == "gate") {
        const char *k = generic_kind(in.name);
        return k && dqh::gate_matrix(k, {}, m);
    }
    return dqh::gate_matrix(in.kind, in.params, m);
}

// The kernel's layout formulas (dq_tile.hpp: make_layout and Layout::jp),
// restated for the emulation below: register j of thread t holds tile index
// tpart(t) | jp(j); the thread bits go, ascending, to the non-register tile
// bits.
unsigned tpart_of(int word, unsigned t, int R) {
    unsigned tp = t;
    for (int m = 0; m < R; ++m) {
        const unsigned lo = (1u << ((word >> (4 * m)) & 15)) - 1;
        tp = (tp & lo) | ((tp & ~lo) << 1);
    }
    return tp;
}

unsigned jp_of(int word, int j, int R) {
    unsigned v = 0;
    for (int m = 0; m < R; ++m)
        if ((j >> m) & 1) v |= 1u << ((word >> (4 * m)) & 15);
    return v;
}

// Emulates where every amplitude of one tile sits (thread, register) through
// a planned pass -- the load, the LDS relayouts (each amplitude written at
// and read back from its tile index), the store -- as the kernel runs it
// (dq_tile.hpp: tile_sweep), and checks: the register bits of a layout are
// packed ascending (make_layout inserts them in that order) and distinct;
// every layout puts each tile index in exactly one (thread, register); the
// register ops find their tile bits in the slots they name; the load and
// store layouts keep the low tile bits on thread bits 0..low-1 (coalesced
// runs).
void emulate_pass(const dq::RegPlan &rp, unsigned lowmask) {
    const int R = rp.reg_bits, b = rp.b;
    const unsigned nt = 1u << (b - R), na = 1u << R;
    std::vector<unsigned> held(size_t(nt) * na);
    auto at = [&](unsigned t, unsigned j) -> unsigned & { return held[size_t(t) * na + j]; };
    auto ascending = [&](int word) {
        for (int m = 1; m < R; ++m)
            if (((word >> (4 * m)) & 15) <= ((word >> (4 * (m - 1))) & 15)) return false;
        return ((word >> (4 * (R - 1))) & 15) < b;
    };
    auto set_layout = [&](int word) {
        for (unsigned t = 0; t < nt; ++t)
            for (unsigned j = 0; j < na; ++j) at(t, j) = tpart_of(word, t, R) | jp_of(word, int(j), R);
    };
    auto bijective = [&]() {
        std::vector<char> seen(size_t(1) << b, 0);
        for (unsigned i : held) {
            if (i >= (1u << b) || seen[i]) return false;
            seen[i] = 1;
        }
        return true;
    };
    auto coalesced = [&](int word) {
        for (unsigned t = 0; t < nt; ++t)
            if ((tpart_of(word, t, R) & lowmask) != (t & lowmask)) return false;
        for (int j = 0; j < int(na); ++j)
            if (jp_of(word, j, R) & lowmask) return false;
        return true;
    };
    auto slot_holds = [&](int slot, int bit) {
        if (slot < 0 || slot >= R || bit < 0 || bit >= b) return false;
        for (unsigned t = 0; t < nt; ++t)
            for (unsigned j = 0; j < na; ++j)
                if (((at(t, j) >> bit) & 1) != ((j >> slot) & 1)) return false;
        return true;
    };
    EXPECT(ascending(rp.pd.load_rb) && coalesced(rp.pd.load_rb));
    set_layout(rp.pd.load_rb);
    EXPECT(bijective());
    int cur = rp.pd.load_rb;
    for (const dq::TileOp &op : rp.ops) {
        if (op.relayout) {
            EXPECT(ascending(op.rb));
            set_layout(op.rb);
            EXPECT(bijective());
            cur = op.rb;
        }
        EXPECT(op.rb == cur);
        if (op.kind == dq::T_DENSE1) EXPECT(slot_holds(op.p, op.bit[0]));
        if (op.kind == dq::T_DENSE2) EXPECT(slot_holds(op.p, op.bit[0]) && slot_holds(op.q, op.bit[1]));
        if (op.kind == dq::T_CTRL1) EXPECT(slot_holds(op.p, op.bit[op.k - 1]));
        if (op.kind == dq::T_DIAGT && op.p >= 0) EXPECT(slot_holds(op.p, op.bit[op.q]));
    }
    // the kernel moves to the store layout through LDS when the last layout differs
    EXPECT(ascending(rp.pd.store_rb) && coalesced(rp.pd.store_rb));
    EXPECT(rp.relayouts == int(std::count_if(rp.ops.begin(), rp.ops.end(),
                                             [](const dq::TileOp &o) { return o.relayout != 0; })) +
                               (cur != rp.pd.store_rb ? 1 : 0));
}

void check_plan(const dq::PlanConfig &cfg, const std::vector<dq::HostOp> &queue) {
    const bool anywhere = cfg.fuse && cfg.reg_tiles && dq::effective_tile_bits(&cfg) >= 9;
    const std::vector<dq::Pass> passes = dq::schedule(&cfg, queue, anywhere);
    std::vector<int> seen(queue.size(), 0);
    std::vector<double> mats;
    // qubits 0 everywhere before each pass, as a flush from |0...0> tracks them
    dq::KnownBits kb;
    kb.zero = cfg.n ? ~uint64_t(0) >> (64 - cfg.n) : 0;
    const uint64_t &zero_q = kb.zero;
    for (const dq::Pass &p : passes) {
        for (int i : p.ops) {
            EXPECT(i >= 0 && size_t(i) < queue.size());
            if (i >= 0 && size_t(i) < queue.size()) seen[size_t(i)]++;
        }
        if (!p.S || !anywhere) continue;
        EXPECT(__builtin_popcountll(p.S) == dq::effective_tile_bits(&cfg));
        dq::RegPlan rp;
        dq::plan_pass(&cfg, queue, p, mats, rp, kb.zero);
        {
            // the outer index lists every outer qubit once, the ones not known
            // to be 0 first; n_outer counts those
            uint64_t seen_o = 0;
            int live = 0;
            for (int j = 0; j < cfg.n - rp.b; ++j) {
                const int q = rp.pd.oq[j];
                EXPECT(q >= 0 && q < cfg.n && !((p.S >> q) & 1) && !((seen_o >> q) & 1));
                if (q < 0 || q >= cfg.n) break;
                seen_o |= uint64_t(1) << q;
                EXPECT(((zero_q >> q) & 1) == (j >= rp.pd.n_outer));
                live += !((zero_q >> q) & 1);
            }
            EXPECT(live == rp.pd.n_outer);
        }
        dq::advance_known(queue, p, kb);
        const int R = rp.reg_bits;
        for (const dq::TileOp &t : rp.ops) {
            if (t.kind == dq::T_DENSE1 || t.kind == dq::T_CTRL1) EXPECT(t.p >= 0 && t.p < R);
            if (t.kind == dq::T_DENSE2) EXPECT(t.p >= 0 && t.p < R && t.q >= 0 && t.q < R && t.p != t.q);
            EXPECT(t.mat >= 0 && size_t(t.mat) * 2 <= mats.size());
            EXPECT(t.lay >= 0 && size_t(t.lay) < rp.layouts.size());
        }
        const int low = std::max(0, std::min(cfg.low_bits, rp.b - R));
        emulate_pass(rp, (1u << low) - 1);
        std::vector<uint32_t> table;
        for (int rb : rp.layouts) dq::append_layout_table(table, rb, rp.b, 16u, R);
        EXPECT(table.size() == rp.layouts.size() * (size_t(1) << R));
        for (uint32_t off : table) EXPECT(off < (16u << rp.b));
    }
    for (int c : seen) EXPECT(c == 1);
}

int run_file(const std::string &path) {
    std::ifstream f(path);
    std::stringstream ss;
    ss << f.rdbuf();
    dqh::Circuit c = dqh::Decoder().decode(ss.str());
    const int n = int(c.num_qubits());
    std::vector<std::vector<dq::HostOp>> segments(1);
    for (const dqh::Instr &in : c.instructions) {
        if (in.kind == "measure" || in.kind == "reset" || in.kind == "conditional") {
            if (!segments.back().empty()) segments.emplace_back();
            continue;
        }
        dqh::Mat m;
        if (!unitary(in, m) || m.k == 0) continue;
        dq::HostOp op;
        op.k = m.k;
        for (int j = 0; j < m.k; ++j) op.q[j] = int(in.qubits[size_t(j)]);
        dq::classify(op, m.v.data());
        segments.back().push_back(std::move(op));
    }
    for (int fuse : {1, 2, 3})
        for (int tb : {9, 12, 13})
            for (int lb : {3, 4})
                for (int rbits : {3, 4}) {
                    dq::PlanConfig cfg;
                    cfg.n = n;
                    cfg.fuse = fuse;
                    cfg.tile_bits = tb;
                    cfg.low_bits = lb;
                    cfg.reg_bits = fuse >= 2 ? rbits : 4;
                    cfg.fuse_min_cost = fuse == 3 ? 0.0 : 128.0;
                    for (const auto &seg : segments)
                        if (!seg.empty()) check_plan(cfg, seg);
                }
    // the measurement stream: seeds around the u64 wrap
    for (uint64_t seed : {uint64_t(0), uint64_t(42), ~uint64_t(0)}) {
        dqh::ChaCha8Rng rng = dqh::ChaCha8Rng::seed_from_u64(seed);
        for (int i = 0; i < 1000; ++i) {
            const double r = rng.gen_f64();
            EXPECT(r >= 0.0 && r < 1.0);
        }
    }
    return n;
}

}  // namespace

int main(int argc, char **argv) {
    for (int i = 1; i < argc; ++i) {
        const int n = run_file(argv[i]);
        std::printf("%s: %d qubits\n", argv[i], n);
    }
    if (failures) std::fprintf(stderr, "%d invariant failures\n", failures);
    return failures ? 1 : 0;
}
