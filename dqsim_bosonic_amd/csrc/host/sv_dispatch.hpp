// sv_dispatch.hpp -- The statevector instruction dispatcher: generic gate names, unitary
// matrices of instructions and the Runner that drives a device state
// (src/monolithic/statevector/simulator.rs:386-632).
//
// Part of the _core extension module: included once, by core.cpp, after its
// library includes and namespace aliases; the definitions are internal to
// that translation unit.
#pragma once

namespace {

// Generic `gate` names accepted by the statevector dispatcher
// (simulator.rs:571-600) -> the gate library's matrix name.
const std::unordered_map<std::string, std::string> &generic_gates() {
    static const std::unordered_map<std::string, std::string> m = {
        {"remote_link_phi_plus", "remote_link_phi_plus"}, {"remote_link_psi_minus", "remote_link_psi_minus"},
        {"remote_link_psi_plus", "remote_link_psi_plus"}, {"nonlocal_cz", "cz"},
        {"remote_cz", "cz"},                              {"remote_cx", "cx"},
        {"remote_epr", "remote_link_phi_plus"}};
    return m;
}

struct NotImplemented : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// Matrix of a unitary instruction (simulator.rs:386-632, the gate arms):
// false for the kinds that are no-ops (id, u0, barrier, classical and the
// generic remote_barrier / remote_cu1); throws for unknown kinds.
bool unitary_of(const Instr &in, Mat &m) {
    const std::string &k = in.kind;
    if (k == "id" || k == "u0" || k == "barrier" || k == "classical") return false;
    if (k == "gate") {
        std::string name = in.name;
        for (auto &c : name) c = char(std::tolower(static_cast<unsigned char>(c)));
        if (name == "remote_barrier" || name == "remote_cu1") return false;
        auto it = generic_gates().find(name);
        if (it == generic_gates().end())
            throw NotImplemented("Unsupported generic gate: " + rust_debug(name) + ". Decompose it before simulating.");
        gate_matrix(it->second, {}, m);
        return true;
    }
    if (!gate_matrix(k, in.params, m)) throw NotImplemented("unsupported instruction kind " + rust_debug(k));
    return true;
}

class Runner {
   public:
    Runner(dq_state *s, int n, uint64_t seed, bool profile)
        : s_(s), n_(n), rng_(ChaCha8Rng::seed_from_u64(seed)), profile_(profile) {}

    std::map<uint64_t, int> cbits;
    SimulationProfile acc;
    std::vector<double> p1s;  // the p1 of every measurement and reset, in program order

    // run_instruction, simulator.rs:386-632
    void run(const Instr &in) {
        const std::string &k = in.kind;
        if (k == "id" || k == "u0" || k == "barrier" || k == "classical") return;
        if (k == "measure") {
            cbits[in.cbit] = measure(in.qubits[0]);
            return;
        }
        if (k == "reset") {
            if (measure(in.qubits[0]) == 1) apply_x(int(qubit(in.qubits[0])));
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
        Mat m;
        if (!unitary_of(in, m)) return;
        if (k == "gate") {
            apply_n(m, in.qubits);
            return;
        }
        if (in.qubits.size() == 1) {
            apply_1(m, int(qubit(in.qubits[0])));
        } else {
            apply_n(m, in.qubits);
        }
    }

   private:
    dq_state *s_;
    int n_;
    ChaCha8Rng rng_;
    bool profile_;

    uint64_t qubit(uint64_t q) const {
        if (q >= uint64_t(n_))
            throw py::value_error("qubit index " + std::to_string(q) + " out of range for " + std::to_string(n_) +
                                  "-qubit circuit");
        return q;
    }

    template <typename F>
    void timed(uint64_t &calls, double &secs, F &&f) {
        if (!profile_) {
            f();
            return;
        }
        const auto t0 = Clock::now();
        f();
        check(dq_synchronize(s_), "synchronize");
        secs += std::chrono::duration<double>(Clock::now() - t0).count();
        calls++;
    }

    // do_oq, simulator.rs:890-907
    void apply_1(const Mat &m, int q) {
        timed(acc.oq_calls, acc.oq_time, [&] { check(dq_apply_1q(s_, m.v.data(), q, nullptr, nullptr, 0), "apply_1q"); });
    }
    void apply_x(int q) {
        Mat x;
        gate_matrix("x", {}, x);
        apply_1(x, q);
    }
    // do_nq, simulator.rs:909-926
    void apply_n(const Mat &m, const std::vector<uint64_t> &qubits) {
        if (int(qubits.size()) != m.k)
            throw py::value_error("gate expects " + std::to_string(m.k) + " qubits, got " +
                                  std::to_string(qubits.size()));
        int qs[5];
        for (int j = 0; j < m.k; ++j) qs[j] = int(qubit(qubits[j]));
        timed(acc.nq_calls, acc.nq_time, [&] { check(dq_apply_nq(s_, m.v.data(), qs, m.k), "apply_nq"); });
    }
    // do_mq + engine::measure_qubit, simulator.rs:928-946, src/engine.rs:193-257
    int measure(uint64_t q64) {
        const int q = int(qubit(q64));
        int outcome = 0;
        timed(acc.mq_calls, acc.mq_time, [&] {
            double p1 = 0.0;
            check(dq_prob_one(s_, q, &p1), "prob_one");
            p1s.push_back(p1);
            outcome = rng_.gen_f64() < p1 ? 1 : 0;
            const double norm =
                outcome ? std::max(std::sqrt(p1), 1e-15) : std::max(std::sqrt(std::max(1.0 - p1, 0.0)), 1e-15);
            check(dq_collapse(s_, q, outcome, norm), "collapse");
        });
        return outcome;
    }
};

}  // namespace
