// mps_sim.hpp -- The matrix-product-state simulator (src/monolithic/mps/simulator.rs), its
// option parsing (src/codecs.rs:39-66) and host-only hooks for tests.
//
// Part of the _core extension module: included once, by core.cpp, after its
// library includes and namespace aliases; the definitions are internal to
// that translation unit.
#pragma once

namespace {

// ---- matrix-product-state simulator (src/monolithic/mps/simulator.rs) ----

// RAII owner of the device contraction context of an MPS (dq_mps_*).
struct MpsContext {
    dq_mps *m = nullptr;
    explicit MpsContext(int n) { check(dq_mps_create(n, &m), "mps context"); }
    ~MpsContext() { dq_mps_destroy(m); }
    MpsContext(const MpsContext &) = delete;
    MpsContext &operator=(const MpsContext &) = delete;
    void load(const MpsNetwork &net) {
        net.pack(bond_, flat_);
        check(dq_mps_load(m, bond_.data(), flat_.data()), "mps load");
    }

   private:
    std::vector<uint64_t> bond_;
    std::vector<double> flat_;
};

// run_instruction of the MPS simulator, mps/simulator.rs:423-563.
class MpsRunner {
   public:
    MpsRunner(int n, std::optional<uint64_t> max_bond, double thr, uint64_t seed, MpsContext *ctx)
        : net(n, max_bond, thr), n_(n), rng_(ChaCha8Rng::seed_from_u64(seed)), ctx_(ctx) {}

    MpsNetwork net;
    std::map<uint64_t, int> cbits;

    void run(const Instr &in) {
        const std::string &k = in.kind;
        if (k == "id" || k == "u0" || k == "barrier" || k == "classical") return;
        if (k == "measure") {
            cbits[in.cbit] = measure(in.qubits[0]);
            return;
        }
        if (k == "reset") {
            const int q = qubit(in.qubits[0]);
            if (measure(uint64_t(q)) == 1) net.apply_1q(q, named("x"));
            return;
        }
        if (k == "conditional") {
            uint64_t actual = 0;
            for (uint64_t b = 0; b < in.creg_size; ++b) {
                auto it = cbits.find(in.creg_base + b);
                const uint64_t v = it == cbits.end() ? 0 : uint64_t(int64_t(it->second));
                actual |= v << (b & 63);
            }
            if (actual == in.creg_value) run(*in.op);
            return;
        }
        if (k == "gate") {
            std::string name = in.name;
            for (auto &c : name) c = char(std::tolower(static_cast<unsigned char>(c)));
            const char *mat = nullptr;
            if (name == "remote_link_phi_plus" || name == "remote_epr" || name == "epr") mat = "remote_link_phi_plus";
            else if (name == "remote_link_psi_minus") mat = "remote_link_psi_minus";
            else if (name == "remote_link_psi_plus") mat = "remote_link_psi_plus";
            else if (name == "nonlocal_cz" || name == "remote_cz") mat = "cz";
            else if (name == "remote_cx") mat = "cx";
            else if (name == "remote_barrier" || name == "remote_cu1" || name.rfind("circuit-", 0) == 0) return;
            else throw NotImplemented("MPS simulator does not support generic gate " + rust_debug(name));
            if (in.qubits.size() < 2)
                throw py::value_error("gate " + rust_debug(name) + " expects 2 qubits, got " +
                                      std::to_string(in.qubits.size()));
            net.apply_2q(qubit(in.qubits[0]), qubit(in.qubits[1]), named(mat));
            return;
        }
        Mat m;
        if (!gate_matrix(k, in.params, m) || m.k > 2)
            throw NotImplemented(
                "MPS simulator only supports one- and two-qubit unitary gates plus measurement, reset, and "
                "conditionals");
        if (m.k == 1)
            net.apply_1q(qubit(in.qubits[0]), m);
        else
            net.apply_2q(qubit(in.qubits[0]), qubit(in.qubits[1]), m);
    }

    // Mps::to_statevector (mps/simulator.rs:246-264) on the device.
    void to_state(dq_state *dst) {
        ctx_->load(net);
        check(dq_mps_to_state(ctx_->m, dst), "to_statevector");
    }

   private:
    int n_;
    ChaCha8Rng rng_;
    MpsContext *ctx_;

    static Mat named(const char *name) {
        Mat m;
        gate_matrix(name, {}, m);
        return m;
    }
    int qubit(uint64_t q) const {
        if (q >= uint64_t(n_))
            throw py::value_error("qubit index " + std::to_string(q) + " out of range for " + std::to_string(n_) +
                                  "-qubit circuit");
        return int(q);
    }
    // Mps::measure, mps/simulator.rs:266-279: p1 of the dense state (here
    // summed on the device without writing it), clamp, one draw, project.
    int measure(uint64_t q64) {
        const int q = qubit(q64);
        ctx_->load(net);
        double p1 = 0.0;
        check(dq_mps_prob_one(ctx_->m, q, &p1), "measure");
        p1 = std::min(std::max(p1, 0.0), 1.0);
        const int outcome = rng_.gen_f64() < p1 ? 1 : 0;
        net.project(q, outcome, outcome ? p1 : 1.0 - p1);
        return outcome;
    }
};

// format_cbits (types.rs:43-51) over a cbit map: MSB first, missing bits 0.
std::string format_cbit_map(const std::map<uint64_t, int> &cbits, uint64_t num_cbits) {
    std::string s;
    for (uint64_t i = num_cbits; i-- > 0;) {
        auto it = cbits.find(i);
        s += std::to_string(it == cbits.end() ? 0 : it->second);
    }
    return s;
}

Circuit decode_circuit(py::object circuit) {
    const std::string text = py::str(circuit.attr("model_dump_json")());
    try {
        return Decoder().decode(text);
    } catch (const JsonError &e) {
        throw py::value_error(std::string("Circuit JSON parse error: ") + e.what());
    }
}

// MpsSimulator, mps/simulator.rs:300-421
class MpsSimulator {
   public:
    MpsSimulator(std::optional<uint64_t> seed, std::optional<uint64_t> max_bond, double threshold)
        : seed_(seed), max_bond_(max_bond), thr_(threshold) {}

    // simulate, mps/simulator.rs:317-345
    SimulationResult simulate(py::object circuit) const {
        const Circuit c = decode_circuit(circuit);
        const int n = checked_width(c);
        const uint64_t seed = seed_ ? *seed_ : ChaCha8Rng::entropy_u64();
        std::unique_ptr<MpsContext> ctx = n > 0 ? std::make_unique<MpsContext>(n) : nullptr;
        MpsRunner run(n, max_bond_, thr_, seed, ctx.get());
        auto st = std::make_shared<DeviceState>(n);
        guarded([&] {
            for (const Instr &in : c.instructions) run.run(in);
            if (n > 0) run.to_state(st->s);
        });
        return SimulationResult(st, n, run.cbits, py::none());
    }

    // simulate_shots, mps/simulator.rs:347-421: every shot from |0...0> with
    // ChaCha8Rng::seed_from_u64(base_seed + shot), counts of format_cbits.
    py::dict simulate_shots(py::object circuit, uint64_t shots, bool profile) const {
        const auto total_t0 = Clock::now();
        const auto pre_t0 = Clock::now();
        const Circuit c = decode_circuit(circuit);
        const uint64_t num_cbits = c.num_cbits();
        const uint64_t base_seed = seed_ ? *seed_ : ChaCha8Rng::entropy_u64();
        const double pre = std::chrono::duration<double>(Clock::now() - pre_t0).count();
        const int n = checked_width(c);
        const auto shots_t0 = Clock::now();
        std::unique_ptr<MpsContext> ctx = n > 0 ? std::make_unique<MpsContext>(n) : nullptr;
        std::map<std::string, uint64_t> counts;
        std::vector<double> shot_times;
        uint64_t svd_calls = 0;
        double svd_time = 0.0;
        guarded([&] {
            for (uint64_t shot = 0; shot < shots; ++shot) {
                const auto t0 = Clock::now();
                MpsRunner run(n, max_bond_, thr_, base_seed + shot, ctx.get());
                for (const Instr &in : c.instructions) run.run(in);
                counts[format_cbit_map(run.cbits, num_cbits)]++;
                if (profile) {
                    shot_times.push_back(std::chrono::duration<double>(Clock::now() - t0).count());
                    svd_calls += run.net.svd_calls;
                    svd_time += run.net.svd_time;
                }
            }
        });
        const double shots_total = std::chrono::duration<double>(Clock::now() - shots_t0).count();
        py::dict d;
        for (auto &kv : counts) d[py::str(kv.first)] = py::int_(kv.second);
        if (profile) {
            py::dict p;
            p["preprocessing_time"] = pre;
            p["shots_total_time"] = shots_total;
            p["total_time"] = std::chrono::duration<double>(Clock::now() - total_t0).count();
            p["num_shots"] = py::int_(shots);
            p["shot_times"] = py::cast(shot_times);
            p["svd_calls"] = py::int_(svd_calls);
            p["svd_time"] = svd_time;
            write_profile_json("mps", p);
        }
        return d;
    }

   private:
    std::optional<uint64_t> seed_;
    std::optional<uint64_t> max_bond_;
    double thr_;

    static int checked_width(const Circuit &c) {
        const uint64_t n = c.num_qubits();
        if (n > 34) throw py::value_error("circuit has " + std::to_string(n) + " qubits; at most 34 are supported");
        return int(n);
    }
    template <typename F>
    static void guarded(F &&f) {
        try {
            f();
        } catch (const NotImplemented &e) {
            PyErr_SetString(PyExc_NotImplementedError, e.what());
            throw py::error_already_set();
        } catch (const py::error_already_set &) {
            throw;
        } catch (const py::builtin_exception &) {
            throw;
        } catch (const std::runtime_error &e) {
            PyErr_SetString(PyExc_RuntimeError, e.what());
            throw py::error_already_set();
        }
    }
};

// Host-only hooks for tests: the Jacobi SVD, and the tensor network after the
// unitary instructions of a circuit (no measurement: that needs the device).
py::tuple svd_jacobi_hook(py::array_t<std::complex<double>, py::array::c_style | py::array::forcecast> a) {
    if (a.ndim() != 2) throw py::value_error("expected a matrix");
    const int m = int(a.shape(0)), n = int(a.shape(1));
    std::vector<Cx> in(size_t(m) * n);
    std::memcpy(in.data(), a.data(), in.size() * sizeof(Cx));
    std::vector<Cx> U, Vt;
    std::vector<double> sv;
    svd_jacobi(m, n, in, U, sv, Vt);
    const int k = int(sv.size());
    py::array_t<std::complex<double>> u({m, k}), vt({k, n});
    std::memcpy(u.mutable_data(), U.data(), U.size() * sizeof(Cx));
    std::memcpy(vt.mutable_data(), Vt.data(), Vt.size() * sizeof(Cx));
    return py::make_tuple(u, py::cast(sv), vt);
}

py::list mps_network_hook(const std::string &text, std::optional<uint64_t> max_bond, double threshold) {
    Circuit c;
    try {
        c = Decoder().decode(text);
    } catch (const JsonError &e) {
        throw py::value_error(std::string("Circuit JSON parse error: ") + e.what());
    }
    const int n = int(c.num_qubits());
    MpsRunner run(n, max_bond, threshold, 0, nullptr);
    for (const Instr &in : c.instructions) {
        if (in.kind == "measure" || in.kind == "reset" || in.kind == "conditional")
            throw py::value_error("the host-only network hook takes unitary instructions only");
        run.run(in);
    }
    py::list out;
    for (const MpsTensor &t : run.net.tensors) {
        py::array_t<std::complex<double>> a({py::ssize_t(t.left), py::ssize_t(2), py::ssize_t(t.right)});
        std::memcpy(a.mutable_data(), t.data.data(), t.data.size() * sizeof(Cx));
        out.append(a);
    }
    return out;
}

// parse_mps_options, codecs.rs:39-66 (+ :86-107)
MpsSimulator mps_from_options(std::optional<uint64_t> seed, const py::kwargs &options) {
    std::optional<uint64_t> max_bond;
    double thr = 1e-12;
    for (auto item : options) {
        const std::string key = py::str(item.first);
        py::handle v = item.second;
        if (key == "max_bond_dimension") {
            if (v.is_none()) {
                max_bond.reset();
                continue;
            }
            const uint64_t b = v.cast<uint64_t>();
            if (b == 0) throw py::value_error("max_bond_dimension must be positive or None");
            max_bond = b;
        } else if (key == "truncation_threshold") {
            const double t = v.cast<double>();
            if (!std::isfinite(t) || t < 0.0)
                throw py::value_error("truncation_threshold must be a finite non-negative float");
            thr = t;
        } else {
            throw py::type_error("Unsupported MPS option " + rust_debug(key));
        }
    }
    return MpsSimulator(seed, max_bond, thr);
}

}  // namespace
