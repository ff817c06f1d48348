// statevector.hpp -- StatevectorSimulator (src/monolithic/statevector/simulator.rs:195-379) and
// the resident Engine used for repeated runs of one decoded circuit.
//
// Part of the _core extension module: included once, by core.cpp, after its
// library includes and namespace aliases; the definitions are internal to
// that translation unit.
#pragma once

namespace {

class StatevectorSimulator {
   public:
    StatevectorSimulator(std::optional<uint64_t> seed, bool profile, int fusion = 2, int tile_bits = 0,
                         int low_bits = 0)
        : seed_(seed), profile_(profile), fusion_(fusion), tile_bits_(tile_bits), low_bits_(low_bits) {}

    // simulator.rs:325-379
    SimulationResult simulate(py::object circuit) const {
        const std::string text = py::str(circuit.attr("model_dump_json")());
        Circuit c;
        try {
            c = Decoder().decode(text);
        } catch (const JsonError &e) {
            throw py::value_error(std::string("Circuit JSON parse error: ") + e.what());
        }
        return run_circuit(c);
    }

    // The same simulation from the opcode-stream wire format
    // (include/dqsim_opstream.h): no JSON on either side.
    SimulationResult simulate_opstream(py::array records, uint64_t num_qubits, uint64_t num_cbits,
                                       const std::vector<std::string> &names) const {
        return run_circuit(opstream_circuit(records, num_qubits, num_cbits, names));
    }

    py::dict simulate_shots_opstream(py::array records, uint64_t num_qubits, uint64_t num_cbits,
                                     const std::vector<std::string> &names, uint64_t shots, bool profile) const {
        return shots_circuit(opstream_circuit(records, num_qubits, num_cbits, names), shots, profile,
                             Clock::now());
    }

    static Circuit opstream_circuit(py::array records, uint64_t num_qubits, uint64_t num_cbits,
                                    const std::vector<std::string> &names) {
        if (records.itemsize() != py::ssize_t(sizeof(dq_op_record)) || !(records.flags() & py::array::c_style))
            throw py::value_error("records must be a C-contiguous array of 64-byte dq_op_record");
        try {
            return decode_opstream(static_cast<const dq_op_record *>(records.data()), size_t(records.size()),
                                   num_qubits, num_cbits, names);
        } catch (const OpstreamError &e) {
            throw py::value_error(std::string("opcode stream error: ") + e.what());
        }
    }

    SimulationResult run_circuit(const Circuit &c) const {
        const uint64_t n64 = c.num_qubits();
        if (n64 > 40) throw py::value_error("circuit has " + std::to_string(n64) + " qubits; at most 40 are supported");
        const int n = int(n64);
        auto st = std::make_shared<DeviceState>(n);
        check(dq_set_options(st->s, fusion_, tile_bits_, low_bits_), "options");
        const uint64_t seed = seed_ ? *seed_ : ChaCha8Rng::entropy_u64();
        Runner run(st->s, n, seed, profile_);
        const auto t0 = Clock::now();
        try {
            for (const Instr &in : c.instructions) run.run(in);
            check(dq_synchronize(st->s), "synchronize");
        } catch (const NotImplemented &e) {
            PyErr_SetString(PyExc_NotImplementedError, e.what());
            throw py::error_already_set();
        }
        run.acc.total_time = std::chrono::duration<double>(Clock::now() - t0).count();
        py::object prof = py::none();
        if (profile_) prof = py::cast(run.acc);
        return SimulationResult(st, n, run.cbits, prof);
    }

    // simulate_shots, simulator.rs:215-322: dict[str, int] of formatted
    // classical registers over `shots` independent runs.
    py::dict simulate_shots(py::object circuit, uint64_t shots, bool profile) const {
        const auto total_t0 = Clock::now();
        const std::string text = py::str(circuit.attr("model_dump_json")());
        Circuit c;
        try {
            c = Decoder().decode(text);
        } catch (const JsonError &e) {
            throw py::value_error(std::string("Circuit JSON parse error: ") + e.what());
        }
        return shots_circuit(c, shots, profile, total_t0);
    }

    py::dict shots_circuit(const Circuit &c, uint64_t shots, bool profile, Clock::time_point total_t0) const {
        const uint64_t n = c.num_qubits();
        if (n < 1 || n > 34) throw py::value_error("circuit has " + std::to_string(n) + " qubits; 1..34 are supported");
        const int num_cbits = int(c.num_cbits());
        const uint64_t base_seed = seed_ ? *seed_ : ChaCha8Rng::entropy_u64();
        const double preprocessing = std::chrono::duration<double>(Clock::now() - total_t0).count();
        const auto fuse_t0 = Clock::now();
        dq_shots *S = nullptr;
        check(dq_shots_create(int(n), num_cbits, &S), "shots");
        std::unique_ptr<dq_shots, void (*)(dq_shots *)> guard(S, dq_shots_destroy);
        try {
            build_shots_program(S, c, n);
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
        for (uint64_t i = 0; i < shots; ++i) counts[format_shot(bits.data() + i * uint64_t(words), num_cbits)]++;
        const double shots_total = std::chrono::duration<double>(Clock::now() - shots_t0).count();
        py::dict d;
        for (auto &kv : counts) d[py::str(kv.first)] = py::int_(kv.second);
        if (profile) write_shots_profile(preprocessing, fusion, shots_total,
                                         std::chrono::duration<double>(Clock::now() - total_t0).count(), times);
        return d;
    }

    // write_shots_profile (profiling.rs:8-18): ./dqsim_profiles/statevector_shots_profile_<unix ns>.json
    static void write_shots_profile(double pre, double fusion, double shots_total, double total,
                                    const std::vector<double> &times) {
        py::dict p;
        p["preprocessing_time"] = pre;
        p["fusion_time"] = fusion;
        p["shots_total_time"] = shots_total;
        p["total_time"] = total;
        p["num_shots"] = py::int_(times.size());
        p["shot_times"] = py::cast(times);
        write_profile_json("statevector", p);
    }

    std::optional<uint64_t> seed_;
    bool profile_;
    int fusion_;
    int tile_bits_;
    int low_bits_;
};

// Runs one decoded circuit repeatedly on a resident device state: the same
// dispatcher as simulate(), without per-call allocation or JSON decoding.
// Used by the benchmark (one step = reset to |0...0> + the whole circuit).
class Engine {
   public:
    Engine(int n, int fusion, int tile_bits, int low_bits, int dtype)
        : n_(n), st_(std::make_shared<DeviceState>(n, dtype)) {
        check(dq_set_options(st_->s, fusion, tile_bits, low_bits), "options");
    }
    void load(const std::string &text) {
        try {
            circ_ = Decoder().decode(text);
        } catch (const JsonError &e) {
            throw py::value_error(std::string("Circuit JSON parse error: ") + e.what());
        }
        if (circ_.num_qubits() != uint64_t(n_)) throw py::value_error("circuit width differs from the engine's");
    }
    py::dict step(uint64_t seed) {
        check(dq_state_reset(st_->s), "reset");
        Runner run(st_->s, n_, seed, false);
        for (const Instr &in : circ_.instructions) run.run(in);
        check(dq_synchronize(st_->s), "synchronize");
        p1s_ = run.p1s;
        py::dict d;
        for (const auto &kv : run.cbits) d[py::int_(kv.first)] = py::int_(kv.second);
        return d;
    }
    void set_timing(bool on) { check(dq_set_timing(st_->s, on ? 1 : 0), "timing"); }
    py::tuple timing() {
        double ms = 0;
        uint64_t cnt = 0;
        check(dq_get_timing(st_->s, &ms, &cnt), "timing");
        return py::make_tuple(ms, cnt);
    }
    uint64_t timed_amplitudes() {
        uint64_t a = 0;
        check(dq_get_timed_amplitudes(st_->s, &a), "timed amplitudes");
        return a;
    }
    void set_zero_tiles(bool on) { check(dq_set_zero_tiles(st_->s, on ? 1 : 0), "zero tiles"); }
    py::tuple stats() {
        uint64_t a = 0, b = 0, c = 0;
        check(dq_get_stats(st_->s, &a, &b, &c), "stats");
        return py::make_tuple(a, b, c);
    }
    void set_fuse_min_cost(double c) { check(dq_set_fuse_min_cost(st_->s, c), "fusion cost threshold"); }
    void set_pair_fusion(bool on) { check(dq_set_pair_fusion(st_->s, on ? 1 : 0), "pair fusion"); }
    void set_reg_bits(int r) { check(dq_set_reg_bits(st_->s, r), "register bits"); }
    py::tuple fused_stats() {
        uint64_t a = 0, b = 0;
        check(dq_get_fused_stats(st_->s, &a, &b), "fused stats");
        return py::make_tuple(a, b);
    }
    // Distance between this engine's state and another's (any precisions).
    py::dict compare(Engine &other) {
        if (other.n_ != n_) throw py::value_error("engines differ in width");
        double m = 0, r = 0, i = 0, t = 0;
        check(dq_compare_states(st_->s, other.st_->s, &m, &r, &i, &t), "compare");
        py::dict d;
        d["max_abs"] = m;
        d["inner"] = std::complex<double>(r, i);
        d["fidelity"] = r * r + i * i;
        d["tvd"] = t;
        return d;
    }
    SimulationResult result() const { return SimulationResult(st_, n_, {}, py::none()); }
    // p1 of each measurement and reset of the last step, in program order
    const std::vector<double> &p1_record() const { return p1s_; }

   private:
    int n_;
    std::shared_ptr<DeviceState> st_;
    Circuit circ_;
    std::vector<double> p1s_;
};

}  // namespace
