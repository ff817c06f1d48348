// pblock.hpp -- The P-block simulator (src/distributed/pblock).
//
// Part of the _core extension module: included once, by core.cpp, after its
// library includes and namespace aliases; the definitions are internal to
// that translation unit.
#pragma once

namespace {

// ---- P-block simulator (src/distributed/pblock) ----

// Block, model.rs:7-28: a statevector over a sorted set of physical qubits
// (local bit i = qubits[i]), here resident on the device.
struct PBlock {
    std::vector<uint64_t> qubits;
    std::shared_ptr<DeviceState> st;
    int local(uint64_t phys) const {
        for (size_t i = 0; i < qubits.size(); ++i)
            if (qubits[i] == phys) return int(i);
        throw py::value_error("qubit " + std::to_string(phys) + " not in block");
    }
};

PBlock new_block(std::vector<uint64_t> qubits) {
    std::sort(qubits.begin(), qubits.end());
    qubits.erase(std::unique(qubits.begin(), qubits.end()), qubits.end());
    if (qubits.size() > 34) throw py::value_error("a block of more than 34 qubits is not supported");
    PBlock b;
    b.st = std::make_shared<DeviceState>(int(qubits.size()));
    b.qubits = std::move(qubits);
    return b;
}

// merge_blocks, model.rs:30-58: the tensor product over the sorted union of
// the qubits, on the device (dq_state_tensor).
PBlock merge_blocks(const PBlock &a, const PBlock &b) {
    std::vector<uint64_t> qs(a.qubits);
    qs.insert(qs.end(), b.qubits.begin(), b.qubits.end());
    std::sort(qs.begin(), qs.end());
    qs.erase(std::unique(qs.begin(), qs.end()), qs.end());
    if (qs.size() > 34) throw py::value_error("merged block of " + std::to_string(qs.size()) + " qubits exceeds 34");
    auto pos = [&](uint64_t q) { return int(std::lower_bound(qs.begin(), qs.end(), q) - qs.begin()); };
    std::vector<int> ap, bp;
    for (uint64_t q : a.qubits) ap.push_back(pos(q));
    for (uint64_t q : b.qubits) bp.push_back(pos(q));
    PBlock m;
    m.st = std::make_shared<DeviceState>(int(qs.size()));
    check(dq_state_tensor(m.st->s, a.st->s, ap.data(), b.st->s, bp.data()), "block merge");
    m.qubits = std::move(qs);
    return m;
}

// BlockPool, model.rs:60-132: one block per node (nodes ascending), merged
// on demand when an instruction spans blocks.
class BlockPool {
   public:
    explicit BlockPool(const std::map<uint64_t, std::vector<uint64_t>> &qpn, bool device = true) {
        for (const auto &kv : qpn) {
            const size_t idx = blocks.size();
            for (uint64_t q : kv.second) qubit_to_block[q] = idx;
            if (device) {
                blocks.push_back(std::make_unique<PBlock>(new_block(kv.second)));
            } else {
                auto b = std::make_unique<PBlock>();
                b->qubits = kv.second;
                std::sort(b->qubits.begin(), b->qubits.end());
                b->qubits.erase(std::unique(b->qubits.begin(), b->qubits.end()), b->qubits.end());
                blocks.push_back(std::move(b));
            }
        }
        device_ = device;
    }

    std::vector<std::unique_ptr<PBlock>> blocks;  // null: merged away
    std::map<uint64_t, size_t> qubit_to_block;
    uint64_t merge_count = 0;

    size_t block_of(uint64_t q) const {
        auto it = qubit_to_block.find(q);
        if (it == qubit_to_block.end())
            throw py::value_error("qubit " + std::to_string(q) + " is not assigned to any block");
        return it->second;
    }

    size_t ensure_single_block(const std::vector<uint64_t> &qubits) {
        if (qubits.empty()) return 0;
        const size_t target = block_of(qubits[0]);
        for (size_t i = 1; i < qubits.size(); ++i) {
            const size_t bq = block_of(qubits[i]);
            if (bq != target) merge(target, bq);
        }
        return target;
    }

    PBlock merge_all() {
        std::unique_ptr<PBlock> acc;
        for (auto &b : blocks) {
            if (!b) continue;
            if (!acc)
                acc = std::move(b);
            else
                acc = std::make_unique<PBlock>(merge_blocks(*acc, *b));
        }
        return acc ? std::move(*acc) : new_block({});
    }

   private:
    bool device_ = true;
    void merge(size_t keep, size_t drop) {
        if (keep == drop) return;
        std::unique_ptr<PBlock> a = std::move(blocks[keep]), b = std::move(blocks[drop]);
        for (uint64_t q : b->qubits) qubit_to_block[q] = keep;
        if (device_) {
            blocks[keep] = std::make_unique<PBlock>(merge_blocks(*a, *b));
        } else {  // bookkeeping only (the merge count of the shots profile)
            a->qubits.insert(a->qubits.end(), b->qubits.begin(), b->qubits.end());
            std::sort(a->qubits.begin(), a->qubits.end());
            blocks[keep] = std::move(a);
        }
        merge_count++;
    }
};

// The distributed circuit as the P-block simulator reads it
// (pblock/simulator.rs:286-348): entries (global order, node, local index)
// from `_instruction_index` keyed by the instruction objects' identities,
// shared instructions counted once, unknown ones after all known ones in
// encounter order; node circuits decoded from their JSON.
struct Distributed {
    std::map<uint64_t, std::vector<uint64_t>> qpn;
    std::map<uint64_t, Circuit> circuits;
    struct Entry {
        int64_t order;
        uint64_t node;
        size_t local;
    };
    std::vector<Entry> entries;
    uint64_t num_cbits = 0;
    const Instr &instr(const Entry &e) const { return circuits.at(e.node).instructions.at(e.local); }
};

Distributed read_distributed(py::object d) {
    Distributed out;
    std::unordered_map<uint64_t, int64_t> index;
    py::dict idx = d.attr("_instruction_index").cast<py::dict>();
    for (auto kv : idx) index[kv.first.cast<uint64_t>()] = kv.second.cast<int64_t>();
    for (auto kv : d.attr("qubits_per_node").cast<py::dict>())
        out.qpn[kv.first.cast<uint64_t>()] = kv.second.cast<std::vector<uint64_t>>();
    py::dict circuits = d.attr("circuits").cast<py::dict>();
    std::vector<uint64_t> nodes;
    for (auto kv : circuits) nodes.push_back(kv.first.cast<uint64_t>());
    std::sort(nodes.begin(), nodes.end());
    std::unordered_set<uintptr_t> seen;
    int64_t fallback = std::numeric_limits<int64_t>::max() / 2;
    for (uint64_t node : nodes) {
        py::object c = circuits[py::int_(node)];
        const std::string text = py::str(c.attr("model_dump_json")());
        py::list ins = c.attr("instructions").cast<py::list>();
        for (size_t i = 0; i < ins.size(); ++i) {
            const uintptr_t id = reinterpret_cast<uintptr_t>(ins[i].ptr());
            if (!seen.insert(id).second) continue;
            auto it = index.find(uint64_t(id));
            out.entries.push_back({it != index.end() ? it->second : fallback++, node, i});
        }
        try {
            out.circuits[node] = Decoder().decode(text);
        } catch (const JsonError &e) {
            throw py::value_error("Circuit JSON parse error for node " + std::to_string(node) + ": " + e.what());
        }
        if (out.circuits[node].instructions.size() != ins.size())
            throw py::value_error("node " + std::to_string(node) + ": instruction list and JSON differ in length");
        out.num_cbits = std::max(out.num_cbits, out.circuits[node].num_cbits());
    }
    std::stable_sort(out.entries.begin(), out.entries.end(),
                     [](const Distributed::Entry &a, const Distributed::Entry &b) { return a.order < b.order; });
    return out;
}

// dispatch / dispatch_par, pblock/simulator.rs:463-710 and :746-985, on a
// device block.  The two differ only in the reduction order of p1.
void pblock_dispatch(PBlock &blk, const Instr &in, std::map<uint64_t, int> &cbits, ChaCha8Rng &rng) {
    const std::string &k = in.kind;
    dq_state *s = blk.st->s;
    if (k == "id" || k == "u0" || k == "barrier" || k == "classical") return;
    auto measure = [&](uint64_t q) {
        const int lq = blk.local(q);
        double p1 = 0.0;
        check(dq_prob_one(s, lq, &p1), "prob_one");
        const int outcome = rng.gen_f64() < p1 ? 1 : 0;
        const double norm =
            outcome ? std::max(std::sqrt(p1), 1e-15) : std::max(std::sqrt(std::max(1.0 - p1, 0.0)), 1e-15);
        check(dq_collapse(s, lq, outcome, norm), "collapse");
        return outcome;
    };
    if (k == "measure") {
        cbits[in.cbit] = measure(in.qubits[0]);
        return;
    }
    if (k == "reset") {
        if (measure(in.qubits[0]) == 1) {
            Mat x;
            gate_matrix("x", {}, x);
            check(dq_apply_1q(s, x.v.data(), blk.local(in.qubits[0]), nullptr, nullptr, 0), "apply_1q");
        }
        return;
    }
    if (k == "conditional") {
        uint64_t actual = 0;
        for (uint64_t b = 0; b < in.creg_size; ++b) {
            auto it = cbits.find(in.creg_base + b);
            const uint64_t v = it == cbits.end() ? 0 : uint64_t(int64_t(it->second));
            actual |= v << (b & 63);
        }
        if (actual == in.creg_value && in.op) pblock_dispatch(blk, *in.op, cbits, rng);
        return;
    }
    std::vector<int> lq;
    for (uint64_t q : in.qubits) lq.push_back(blk.local(q));
    Mat m;
    if (k == "gate") {
        try {
            if (!generic_shot_gate(in, ShotFlavor::PBlock, m)) return;
        } catch (const std::runtime_error &e) {
            const std::string msg = e.what();
            if (msg.rfind("remote_rzz", 0) == 0) throw py::value_error(msg);
            throw NotImplemented(msg);
        }
    } else if (!gate_matrix(k, in.params, m)) {
        throw NotImplemented("unsupported instruction kind " + rust_debug(k));
    }
    if (int(lq.size()) != m.k)
        throw py::value_error("gate expects " + std::to_string(m.k) + " qubits, got " + std::to_string(lq.size()));
    if (m.k == 1)
        check(dq_apply_1q(s, m.v.data(), lq[0], nullptr, nullptr, 0), "apply_1q");
    else
        check(dq_apply_nq(s, m.v.data(), lq.data(), m.k), "apply_nq");
}

// PBlockResult, pblock/simulator.rs:38-104
class PBlockResult {
   public:
    PBlockResult(PBlock blk, std::map<uint64_t, int> cbits) : blk_(std::move(blk)), cbits_(std::move(cbits)) {}

    py::array_t<std::complex<double>> statevector() const {
        py::array_t<std::complex<double>> out(static_cast<py::ssize_t>(size_t(1) << blk_.qubits.size()));
        auto *p = reinterpret_cast<double *>(out.mutable_data());
        int rc;
        {
            py::gil_scoped_release nogil;
            rc = dq_copy_to_host(blk_.st->s, p);
        }
        check(rc, "statevector copy");
        return out;
    }
    size_t num_qubits() const { return blk_.qubits.size(); }
    std::vector<uint64_t> physical_qubits() const { return blk_.qubits; }
    py::dict classical_bits() const {
        py::dict d;
        for (const auto &kv : cbits_) d[py::int_(kv.first)] = py::int_(kv.second);
        return d;
    }
    // qubits: physical indices, qubits[0] = MSB of the key; default all in
    // descending order (bit 0 of the key = the lowest physical qubit).
    py::dict probabilities(const std::optional<std::vector<uint64_t>> &qubits) const {
        std::vector<uint64_t> phys;
        if (qubits)
            phys = *qubits;
        else
            phys.assign(blk_.qubits.rbegin(), blk_.qubits.rend());
        std::vector<int> local;
        for (uint64_t q : phys) {
            auto it = std::find(blk_.qubits.begin(), blk_.qubits.end(), q);
            if (it == blk_.qubits.end()) throw py::value_error("qubit " + std::to_string(q) + " not in result");
            local.push_back(int(it - blk_.qubits.begin()));
        }
        std::vector<double> probs(size_t(1) << local.size());
        check(dq_marginal_probs(blk_.st->s, local.data(), int(local.size()), probs.data()), "marginal probabilities");
        py::dict d;
        for (size_t j = 0; j < probs.size(); ++j)
            if (probs[j] > 0.0) d[py::int_(j)] = py::float_(probs[j]);
        return d;
    }

   private:
    PBlock blk_;
    std::map<uint64_t, int> cbits_;
};

// PBlockSimulator, pblock/simulator.rs:111-460
class PBlockSimulator {
   public:
    explicit PBlockSimulator(std::optional<uint64_t> seed) : seed_(seed) {}

    // simulate, :285-460: blocks merged on demand; runs of consecutive
    // instructions on distinct single blocks form an epoch whose members each
    // draw from their own ChaCha8Rng seeded by one u64 of the main stream.
    PBlockResult simulate(py::object distributed) const {
        const Distributed d = read_distributed(distributed);
        BlockPool pool(d.qpn);
        std::map<uint64_t, int> cbits;
        ChaCha8Rng rng = ChaCha8Rng::seed_from_u64(seed_ ? *seed_ : ChaCha8Rng::entropy_u64());
        guarded([&] {
            size_t cursor = 0;
            const size_t ne = d.entries.size();
            while (cursor < ne) {
                std::vector<std::pair<size_t, size_t>> epoch;  // (entry, block)
                std::set<size_t> used;
                size_t j = cursor;
                for (; j < ne; ++j) {
                    const std::vector<uint64_t> qs = instr_qubits(d.instr(d.entries[j]));
                    if (qs.empty()) continue;
                    std::set<size_t> touched;
                    for (uint64_t q : qs) {
                        auto it = pool.qubit_to_block.find(q);
                        if (it != pool.qubit_to_block.end()) touched.insert(it->second);
                    }
                    if (touched.size() != 1) break;
                    const size_t blk = *touched.begin();
                    if (used.count(blk)) break;
                    epoch.emplace_back(j, blk);
                    used.insert(blk);
                }
                if (epoch.size() > 1) {
                    std::vector<uint64_t> seeds;
                    for (size_t e = 0; e < epoch.size(); ++e) seeds.push_back(rng.next_u64());
                    std::vector<std::map<uint64_t, int>> local(epoch.size());
                    for (size_t e = 0; e < epoch.size(); ++e) {
                        ChaCha8Rng grng = ChaCha8Rng::seed_from_u64(seeds[e]);
                        pblock_dispatch(*pool.blocks[epoch[e].second], d.instr(d.entries[epoch[e].first]), local[e],
                                        grng);
                    }
                    for (const auto &l : local)
                        for (const auto &kv : l) cbits[kv.first] = kv.second;
                    cursor = j;
                } else {
                    const Instr &in = d.instr(d.entries[cursor]);
                    const std::vector<uint64_t> qs = instr_qubits(in);
                    if (!qs.empty()) {
                        const size_t blk = pool.ensure_single_block(qs);
                        pblock_dispatch(*pool.blocks[blk], in, cbits, rng);
                    }
                    ++cursor;
                }
            }
        });
        PBlock fin = pool.merge_all();
        check(dq_synchronize(fin.st->s), "synchronize");
        return PBlockResult(std::move(fin), std::move(cbits));
    }

    // simulate_shots, :128-283: every shot from fresh blocks with
    // ChaCha8Rng::seed_from_u64(base_seed + i) over the globally ordered,
    // fused entry stream (fuse_pblock_entries, src/types.rs:306-345).  The
    // block structure of a shot follows from the instruction stream alone, and
    // the blocks of a shot are the factors of one product state, so the shots
    // run as the batched statevector shots of the globally ordered stream.
    py::dict simulate_shots(py::object distributed, uint64_t shots, bool profile) const {
        const auto total_t0 = Clock::now();
        const Distributed d = read_distributed(distributed);
        const uint64_t base_seed = seed_ ? *seed_ : ChaCha8Rng::entropy_u64();
        const double pre = std::chrono::duration<double>(Clock::now() - total_t0).count();
        const auto fuse_t0 = Clock::now();
        Circuit global;
        uint64_t n = 0;
        for (const auto &kv : d.qpn)
            for (uint64_t q : kv.second) n = std::max(n, q + 1);
        for (const auto &e : d.entries) global.instructions.push_back(d.instr(e));
        for (const Instr &in : global.instructions)
            for (uint64_t q : instr_qubits(in))
                if (!d.qpn.empty() && q >= n) throw py::value_error("qubit " + std::to_string(q) + " is not assigned to any block");
        // merges per shot (the profile's merge_calls): the block bookkeeping alone
        BlockPool shape(d.qpn, false);
        for (const Instr &in : global.instructions) {
            const std::vector<uint64_t> qs = instr_qubits(in);
            if (!qs.empty()) shape.ensure_single_block(qs);
        }
        if (n < 1) n = 1;
        if (n > 34) throw py::value_error("distributed circuit has " + std::to_string(n) + " qubits; at most 34");
        dq_shots *S = nullptr;
        check(dq_shots_create(int(n), int(d.num_cbits), &S), "shots");
        std::unique_ptr<dq_shots, void (*)(dq_shots *)> guard(S, dq_shots_destroy);
        try {
            build_shots_program(S, global, n, ShotFlavor::PBlock);
        } catch (const std::runtime_error &e) {
            PyErr_SetString(PyExc_RuntimeError, e.what());
            throw py::error_already_set();
        }
        const double fusion = std::chrono::duration<double>(Clock::now() - fuse_t0).count();
        const int words = dq_shots_words(S);
        std::vector<uint64_t> bits(size_t(shots) * size_t(words));
        std::vector<double> times(profile ? size_t(shots) : 0);
        const auto shots_t0 = Clock::now();
        int rc;
        {
            py::gil_scoped_release nogil;
            rc = dq_shots_run(S, base_seed, shots, bits.data(), profile ? times.data() : nullptr);
        }
        check(rc, "simulate_shots");
        std::map<std::string, uint64_t> counts;
        for (uint64_t i = 0; i < shots; ++i)
            counts[format_shot(bits.data() + i * uint64_t(words), int(d.num_cbits))]++;
        const double shots_total = std::chrono::duration<double>(Clock::now() - shots_t0).count();
        py::dict out;
        for (auto &kv : counts) out[py::str(kv.first)] = py::int_(kv.second);
        if (profile) {
            py::dict p;
            p["preprocessing_time"] = pre;
            p["fusion_time"] = fusion;
            p["shots_total_time"] = shots_total;
            p["total_time"] = std::chrono::duration<double>(Clock::now() - total_t0).count();
            p["num_shots"] = py::int_(shots);
            p["shot_times"] = py::cast(times);
            p["merge_calls"] = py::int_(shape.merge_count * shots);
            write_profile_json("pblock", p);
        }
        return out;
    }

   private:
    std::optional<uint64_t> seed_;

    template <typename F>
    static void guarded(F &&f) {
        try {
            f();
        } catch (const NotImplemented &e) {
            PyErr_SetString(PyExc_NotImplementedError, e.what());
            throw py::error_already_set();
        }
    }
};

// codecs.rs:30-37
void parse_distributed_mode(const std::string &mode) {
    std::string m;
    size_t a = 0, b = mode.size();
    while (a < b && std::isspace(static_cast<unsigned char>(mode[a]))) ++a;
    while (b > a && std::isspace(static_cast<unsigned char>(mode[b - 1]))) --b;
    for (size_t i = a; i < b; ++i) {
        char c = char(std::tolower(static_cast<unsigned char>(mode[i])));
        m += (c == '-') ? '_' : c;
    }
    if (m != "p_block")
        throw py::value_error("Unsupported distributed simulation mode " + rust_debug(m) + "; expected 'p_block'");
}

}  // namespace
