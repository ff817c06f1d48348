// core.cpp -- the Python module `dqsim_bosonic_amd._core`.
//
// Host side of the statevector path, in C++ because the reference's host is
// compiled code (Rust/PyO3).  It keeps the reference's Python surface:
//   StatevectorSimulator(seed=None, profile=False).simulate(circuit)
//       src/monolithic/statevector/simulator.rs:206-210, :325-379
//   SimulationResult (statevector, num_qubits, classical_bits, profile,
//       probabilities, counts, fidelity)          simulator.rs:94-192
//   SimulationProfile (+ __repr__)                simulator.rs:52-88
//   simulate_monolithic / simulate_monolithic_shots (state-vector mode)
//       src/simulator.rs:13-45, :63-95; src/codecs.rs:19-28, :68-84
// and the reference's instruction dispatcher (run_instruction,
// simulator.rs:386-632).  All state mutation and reductions go through the
// C ABI of include/dqsim_gpu.h (the HIP engine); measurement draws stay on
// the host in the reference's ChaCha8 order.
#include <pybind11/complex.h>
#include <pybind11/functional.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <algorithm>
#include <cctype>
#include <chrono>
#include <array>
#include <cmath>
#include <complex>
#include <cstring>
#include <functional>
#include <cstdio>
#include <map>
#include <memory>
#include <optional>
#include <sstream>
#include <string>
#include <limits>
#include <set>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "../../../include/dqsim_gpu.h"
#include "circuit.hpp"
#include "gates.hpp"
#include "mps.hpp"
#include "opstream.hpp"
#include "rng.hpp"
#include "../dq_plan.hpp"

namespace py = pybind11;
using namespace dqh;

namespace {

[[noreturn]] void raise_dq(int rc, const char *what) {
    std::string msg = std::string(what) + ": " + dq_last_error();
    if (rc == DQ_ERR_ARG) throw py::value_error(msg);
    if (rc == DQ_ERR_OOM) {
        PyErr_SetString(PyExc_MemoryError, msg.c_str());
        throw py::error_already_set();
    }
    throw std::runtime_error(msg);
}

inline void check(int rc, const char *what) {
    if (rc != DQ_OK) raise_dq(rc, what);
}

// Rust's `{:?}` of a string: quoted with escapes.
std::string rust_debug(const std::string &s) {
    std::string o = "\"";
    for (unsigned char c : s) {
        switch (c) {
        case '"': o += "\\\""; break;
        case '\\': o += "\\\\"; break;
        case '\n': o += "\\n"; break;
        case '\r': o += "\\r"; break;
        case '\t': o += "\\t"; break;
        default:
            if (c < 0x20 || c == 0x7f) {
                char buf[16];
                std::snprintf(buf, sizeof buf, "\\u{%x}", c);
                o += buf;
            } else {
                o += char(c);
            }
        }
    }
    return o + "\"";
}

// RAII owner of a device state.
struct DeviceState {
    dq_state *s = nullptr;
    explicit DeviceState(int n, int dtype = DQ_C128) { check(dq_state_create(n, dtype, &s), "statevector allocation"); }
    ~DeviceState() { dq_state_destroy(s); }
    DeviceState(const DeviceState &) = delete;
    DeviceState &operator=(const DeviceState &) = delete;
};

struct SimulationProfile {
    uint64_t oq_calls = 0, nq_calls = 0, mq_calls = 0;
    double oq_time = 0, nq_time = 0, mq_time = 0, total_time = 0;

    // simulator.rs:72-87
    std::string repr() const {
        const double total = std::max(total_time, 1e-9);
        char buf[512];
        std::snprintf(buf, sizeof buf,
                      "SimulationProfile (total: %.2f ms)\n"
                      "  apply_one_qubit : %4llu calls  %8.2f ms  (%.1f%%)\n"
                      "  apply_n_qubit   : %4llu calls  %8.2f ms  (%.1f%%)\n"
                      "  measure_qubit   : %4llu calls  %8.2f ms  (%.1f%%)",
                      total_time * 1000.0, (unsigned long long)oq_calls, oq_time * 1000.0,
                      100.0 * oq_time / total, (unsigned long long)nq_calls, nq_time * 1000.0,
                      100.0 * nq_time / total, (unsigned long long)mq_calls, mq_time * 1000.0,
                      100.0 * mq_time / total);
        return buf;
    }
};

using Clock = std::chrono::steady_clock;

std::vector<int> qubits_list(const std::optional<std::vector<int64_t>> &qubits, int n) {
    std::vector<int> qs;
    if (!qubits) {
        for (int q = n - 1; q >= 0; --q) qs.push_back(q);
        return qs;
    }
    for (int64_t q : *qubits) {
        if (q < 0 || q >= n) throw py::value_error("qubit index " + std::to_string(q) + " out of range");
        qs.push_back(int(q));
    }
    return qs;
}

class SimulationResult {
   public:
    SimulationResult(std::shared_ptr<DeviceState> st, int n, std::map<uint64_t, int> cbits, py::object prof)
        : st_(std::move(st)), n_(n), cbits_(std::move(cbits)), prof_(std::move(prof)) {}

    int num_qubits() const { return n_; }

    // simulator.rs:122-125: a fresh complex128 copy on every access.
    py::array_t<std::complex<double>> statevector() const {
        py::array_t<std::complex<double>> out(size_t(1) << n_);
        auto *p = reinterpret_cast<double *>(out.mutable_data());
        {
            py::gil_scoped_release nogil;
            check(dq_copy_to_host(st_->s, p), "statevector copy");
        }
        return out;
    }

    py::dict classical_bits() const {
        py::dict d;
        for (const auto &kv : cbits_) d[py::int_(kv.first)] = py::int_(kv.second);
        return d;
    }

    py::object profile() const { return prof_; }

    std::vector<double> marginal(const std::vector<int> &qs) const {
        std::vector<double> probs(size_t(1) << qs.size());
        check(dq_marginal_probs(st_->s, qs.data(), int(qs.size()), probs.data()), "marginal probabilities");
        return probs;
    }

    // Lazy results (no reference counterpart): the distribution as an array
    // (index j as in probabilities(), no dict) ...
    py::array_t<double> probabilities_array(const std::optional<std::vector<int64_t>> &qubits) const {
        const std::vector<int> qs = qubits_list(qubits, n_);
        if (qs.size() > 34) throw py::value_error("at most 34 qubits");
        py::array_t<double> out(static_cast<py::ssize_t>(size_t(1) << qs.size()));
        double *p = out.mutable_data();
        {
            py::gil_scoped_release nogil;
            check(dq_marginal_probs(st_->s, qs.data(), int(qs.size()), p), "marginal probabilities");
        }
        return out;
    }

    // ... and the k most probable basis states over all qubits, selected on
    // the device: [(index, probability)], decreasing, ties by index.
    py::list top_k(uint64_t k) const {
        std::vector<uint64_t> idx(static_cast<size_t>(k));
        std::vector<double> p(static_cast<size_t>(k));
        int rc;
        {
            py::gil_scoped_release nogil;
            rc = dq_top_k(st_->s, k, idx.data(), p.data());
        }
        check(rc, "top_k");
        py::list out;
        for (size_t j = 0; j < idx.size(); ++j) out.append(py::make_tuple(py::int_(idx[j]), py::float_(p[j])));
        return out;
    }

    // simulator.rs:148-159
    py::dict probabilities(const std::optional<std::vector<int64_t>> &qubits) const {
        const std::vector<int> qs = qubits_list(qubits, n_);
        const std::vector<double> probs = marginal(qs);
        py::dict d;
        for (size_t j = 0; j < probs.size(); ++j)
            if (probs[j] > 0.0) d[py::int_(j)] = py::float_(probs[j]);
        return d;
    }

    // simulator.rs:162-181 over engine::sample_counts (src/engine.rs:398-433)
    py::dict counts(uint64_t shots, const std::optional<std::vector<int64_t>> &qubits,
                    const std::optional<uint64_t> &seed) const {
        const std::vector<int> qs = qubits_list(qubits, n_);
        ChaCha8Rng rng = seed ? ChaCha8Rng::seed_from_u64(*seed) : ChaCha8Rng::from_entropy();
        const std::vector<double> probs = marginal(qs);
        const size_t dim = probs.size();
        const int k = int(qs.size());
        std::vector<double> cdf(dim);
        double acc = 0.0;
        for (size_t i = 0; i < dim; ++i) {
            acc += probs[i];
            cdf[i] = acc;
        }
        cdf[dim - 1] = 1.0;
        std::unordered_map<uint64_t, uint64_t> hist;
        for (uint64_t s = 0; s < shots; ++s) {
            const double r = rng.gen_f64();
            size_t idx = size_t(std::lower_bound(cdf.begin(), cdf.end(), r) - cdf.begin());
            if (idx > dim - 1) idx = dim - 1;
            hist[idx]++;
        }
        py::dict d;
        for (const auto &kv : hist) {
            std::string bits;
            uint64_t v = kv.first;
            do {
                bits.insert(bits.begin(), char('0' + (v & 1)));
                v >>= 1;
            } while (v);
            if (int(bits.size()) < k) bits.insert(bits.begin(), size_t(k) - bits.size(), '0');
            d[py::str(bits)] = py::int_(kv.second);
        }
        return d;
    }

    // simulator.rs:184-191
    double fidelity(py::array_t<std::complex<double>, py::array::c_style> other) const {
        if (other.ndim() != 1) throw py::value_error("expected a one-dimensional array");
        if (uint64_t(other.shape(0)) != (uint64_t(1) << n_)) return 0.0;
        double re = 0, im = 0;
        check(dq_inner(st_->s, reinterpret_cast<const double *>(other.data()), &re, &im), "fidelity");
        return re * re + im * im;
    }

    dq_state *handle() const { return st_->s; }

   private:
    std::shared_ptr<DeviceState> st_;
    int n_;
    std::map<uint64_t, int> cbits_;
    py::object prof_;
};

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
            outcome = rng_.gen_f64() < p1 ? 1 : 0;
            const double norm =
                outcome ? std::max(std::sqrt(p1), 1e-15) : std::max(std::sqrt(std::max(1.0 - p1, 0.0)), 1e-15);
            check(dq_collapse(s_, q, outcome, norm), "collapse");
        });
        return outcome;
    }
};

// ---- shots (simulator.rs:215-322) ----

struct SCx {
    double re, im;
};
inline SCx cmulx(SCx a, SCx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
inline SCx caddx(SCx a, SCx b) { return {a.re + b.re, a.im + b.im}; }
using M2 = std::array<SCx, 4>;  // row-major 2x2

// gate_matrix_1q (src/types.rs:164-230): the fusable single-qubit kinds with
// the matrices the shots path multiplies (written with its own formulas,
// distinct from gates.rs for u3/u2/rx/ry/rz).
bool fused_1q_matrix(const Instr &in, M2 &m) {
    const std::string &k = in.kind;
    auto r = [](double x) { return SCx{x, 0.0}; };
    auto c = [](double a, double b) { return SCx{a, b}; };
    auto neg = [](SCx a) { return SCx{-a.re, -a.im}; };
    const double s2 = 0.70710678118654752440;  // FRAC_1_SQRT_2
    const auto &p = in.params;
    if (k == "id" || k == "u0") m = {r(1.0), r(0.0), r(0.0), r(1.0)};
    else if (k == "x") m = {r(0.0), r(1.0), r(1.0), r(0.0)};
    else if (k == "y") m = {r(0.0), c(0.0, -1.0), c(0.0, 1.0), r(0.0)};
    else if (k == "z") m = {r(1.0), r(0.0), r(0.0), r(-1.0)};
    else if (k == "h") m = {r(s2), r(s2), r(s2), r(-s2)};
    else if (k == "s") m = {r(1.0), r(0.0), r(0.0), c(0.0, 1.0)};
    else if (k == "sdg") m = {r(1.0), r(0.0), r(0.0), c(0.0, -1.0)};
    else if (k == "t") m = {r(1.0), r(0.0), r(0.0), c(s2, s2)};
    else if (k == "tdg") m = {r(1.0), r(0.0), r(0.0), c(s2, -s2)};
    else if (k == "sx") m = {c(0.5, 0.5), c(0.5, -0.5), c(0.5, -0.5), c(0.5, 0.5)};
    else if (k == "sxdg") m = {c(0.5, -0.5), c(0.5, 0.5), c(0.5, 0.5), c(0.5, -0.5)};
    else if (k == "u3" || k == "u" || k == "u2") {
        double theta, phi, lam;
        if (k == "u2") {
            theta = 3.14159265358979323846 / 2.0;
            phi = p[0];
            lam = p[1];
        } else {
            theta = p[0];
            phi = p[1];
            lam = p[2];
        }
        const double ct = std::cos(theta / 2.0), st = std::sin(theta / 2.0);
        m = {r(ct), neg(cmulx(c(std::cos(lam), std::sin(lam)), r(st))), cmulx(c(std::cos(phi), std::sin(phi)), r(st)),
             cmulx(c(std::cos(phi + lam), std::sin(phi + lam)), r(ct))};
    } else if (k == "u1" || k == "p") {
        m = {r(1.0), r(0.0), r(0.0), c(std::cos(p[0]), std::sin(p[0]))};
    } else if (k == "rx") {
        const double cv = std::cos(p[0] / 2.0), sv = std::sin(p[0] / 2.0);
        m = {r(cv), c(0.0, -sv), c(0.0, -sv), r(cv)};
    } else if (k == "ry") {
        const double cv = std::cos(p[0] / 2.0), sv = std::sin(p[0] / 2.0);
        m = {r(cv), r(-sv), r(sv), r(cv)};
    } else if (k == "rz") {
        const double ph = p[0];
        m = {c(std::cos(-ph / 2.0), std::sin(-ph / 2.0)), r(0.0), r(0.0), c(std::cos(ph / 2.0), std::sin(ph / 2.0))};
    } else {
        return false;
    }
    return true;
}

// matmul2x2 (types.rs:149-160): a * b.
M2 matmul2(const M2 &a, const M2 &b) {
    return {caddx(cmulx(a[0], b[0]), cmulx(a[1], b[2])), caddx(cmulx(a[0], b[1]), cmulx(a[1], b[3])),
            caddx(cmulx(a[2], b[0]), cmulx(a[3], b[2])), caddx(cmulx(a[2], b[1]), cmulx(a[3], b[3]))};
}

// is_identity_2x2 (types.rs:239-244), num-complex norm = hypot.
bool near_identity(const M2 &m) {
    return std::hypot(m[0].re - 1.0, m[0].im - 0.0) < 1e-10 && std::hypot(m[1].re, m[1].im) < 1e-10 &&
           std::hypot(m[2].re, m[2].im) < 1e-10 && std::hypot(m[3].re - 1.0, m[3].im - 0.0) < 1e-10;
}

std::vector<uint64_t> instr_qubits(const Instr &in) {  // Instruction::qubits, types.rs:348-418
    if (in.kind == "conditional") return in.op ? instr_qubits(*in.op) : std::vector<uint64_t>{};
    if (in.kind == "barrier" || in.kind == "classical") return {};
    return in.qubits;
}

// Generic `gate` names of the two shot dispatchers: the statevector one
// (run_instruction_par, simulator.rs:802-830) and the P-block one
// (dispatch_par, src/distributed/pblock/simulator.rs:912-951).  Returns false
// for the no-op names; throws (RuntimeError in the shots path) otherwise.
enum class ShotFlavor { Statevector, PBlock };

bool generic_shot_gate(const Instr &in, ShotFlavor flavor, Mat &m) {
    std::string name = in.name;
    for (auto &ch : name) ch = char(std::tolower(static_cast<unsigned char>(ch)));
    if (name == "remote_barrier" || name == "remote_cu1" || name.rfind("circuit-", 0) == 0) return false;
    if (flavor == ShotFlavor::Statevector) {
        auto it = generic_gates().find(name);
        if (it == generic_gates().end())
            throw std::runtime_error("Unsupported generic gate: " + rust_debug(name) + ". Decompose it before simulating.");
        gate_matrix(it->second, {}, m);
        return true;
    }
    if (name == "remote_rzz") {
        if (in.params.empty()) throw std::runtime_error("remote_rzz requires one theta parameter");
        gate_matrix("rzz", {in.params[0]}, m);
        return true;
    }
    if (name == "teleport")
        throw std::runtime_error(
            "Symbolic 'teleport' gate cannot be simulated natively. Distribute with lowered=True to get a decomposed "
            "circuit.");
    static const std::unordered_map<std::string, std::string> pb = {
        {"remote_link_phi_plus", "remote_link_phi_plus"}, {"remote_link_psi_minus", "remote_link_psi_minus"},
        {"remote_link_psi_plus", "remote_link_psi_plus"}, {"nonlocal_cz", "cz"}, {"remote_cz", "cz"},
        {"remote_cx", "cx"}, {"remote_epr", "remote_link_phi_plus"}};
    auto it = pb.find(name);
    if (it == pb.end()) throw std::runtime_error("Unsupported gate in pblock simulator: " + rust_debug(name));
    gate_matrix(it->second, {}, m);
    return true;
}

// Program ops one unfused instruction becomes (see emit_shot_instr).
int count_shot_ops(const Instr &in) {
    const std::string &k = in.kind;
    if (k == "barrier" || k == "classical" || k == "id" || k == "u0") return 0;
    if (k == "conditional") return 1 + (in.op ? count_shot_ops(*in.op) : 0);
    if (k == "gate") {
        std::string name = in.name;
        for (auto &ch : name) ch = char(std::tolower(static_cast<unsigned char>(ch)));
        if (name == "remote_barrier" || name == "remote_cu1" || name.rfind("circuit-", 0) == 0) return 0;
    }
    return 1;
}

// Appends one unfused instruction (run_instruction_par, simulator.rs:638-866)
// to a shots program; returns the number of program ops appended.
int emit_shot_instr(dq_shots *S, const Instr &in, uint64_t n, ShotFlavor flavor) {
    const std::string &k = in.kind;
    auto qcheck = [&](uint64_t q) {
        if (q >= n)
            throw py::value_error("qubit index " + std::to_string(q) + " out of range for " + std::to_string(n) +
                                  "-qubit circuit");
        return int(q);
    };
    if (k == "barrier" || k == "classical" || k == "id" || k == "u0") return 0;
    if (k == "measure") {
        check(dq_shots_add_measure(S, qcheck(in.qubits[0]), int(in.cbit)), "shots measure");
        return 1;
    }
    if (k == "reset") {
        check(dq_shots_add_reset(S, qcheck(in.qubits[0])), "shots reset");
        return 1;
    }
    if (k == "conditional") {
        const int inner = in.op ? count_shot_ops(*in.op) : 0;
        check(dq_shots_add_conditional(S, int(in.creg_base), int(in.creg_size), in.creg_value, inner),
              "shots conditional");
        if (in.op) emit_shot_instr(S, *in.op, n, flavor);
        return 1 + inner;
    }
    Mat m;
    if (k == "gate") {
        if (!generic_shot_gate(in, flavor, m)) return 0;
    } else if (!gate_matrix(k, in.params, m)) {
        throw std::runtime_error("unsupported instruction kind " + rust_debug(k));
    }
    if (int(in.qubits.size()) != m.k)
        throw py::value_error("gate expects " + std::to_string(m.k) + " qubits, got " + std::to_string(in.qubits.size()));
    int qs[5];
    for (int j = 0; j < m.k; ++j) qs[j] = qcheck(in.qubits[size_t(j)]);
    check(dq_shots_add_gate(S, m.v.data(), qs, m.k), "shots gate");
    return 1;
}

// fuse_circuit (types.rs:254-293): a fused single-qubit gate (instr < 0)
// or an instruction of the circuit kept as it is.
struct FusedItem {
    long instr = -1;
    uint64_t qubit = 0;
    M2 m{};
};

std::vector<FusedItem> fuse_instrs(const Circuit &c, uint64_t n) {
    std::vector<FusedItem> out;
    std::map<uint64_t, M2> pending;  // the final pending gates are flushed in ascending qubit order
    auto flush = [&](uint64_t q) {
        auto it = pending.find(q);
        if (it == pending.end()) return;
        const M2 m = it->second;
        pending.erase(it);
        if (near_identity(m)) return;
        FusedItem f;
        f.qubit = q;
        f.m = m;
        out.push_back(f);
    };
    for (size_t i = 0; i < c.instructions.size(); ++i) {
        const Instr &in = c.instructions[i];
        M2 mat;
        if (in.qubits.size() == 1 && fused_1q_matrix(in, mat)) {
            const uint64_t q = in.qubits[0];
            if (q >= n)
                throw py::value_error("qubit index " + std::to_string(q) + " out of range for " + std::to_string(n) +
                                      "-qubit circuit");
            auto it = pending.find(q);
            const M2 acc = it == pending.end() ? M2{SCx{1.0, 0.0}, SCx{0.0, 0.0}, SCx{0.0, 0.0}, SCx{1.0, 0.0}} : it->second;
            pending[q] = matmul2(mat, acc);
            continue;
        }
        for (uint64_t q : instr_qubits(in)) flush(q);
        FusedItem f;
        f.instr = long(i);
        out.push_back(f);
    }
    std::vector<uint64_t> rest;
    for (auto &kv : pending) rest.push_back(kv.first);
    for (uint64_t q : rest) flush(q);
    return out;
}

// The program of the shot loop (run_fused_par, simulator.rs:870-884).
void build_shots_program(dq_shots *S, const Circuit &c, uint64_t n, ShotFlavor flavor = ShotFlavor::Statevector) {
    for (const FusedItem &f : fuse_instrs(c, n)) {
        if (f.instr >= 0) {
            emit_shot_instr(S, c.instructions[size_t(f.instr)], n, flavor);
            continue;
        }
        double u[8];
        for (int e = 0; e < 4; ++e) {
            u[2 * e] = f.m[size_t(e)].re;
            u[2 * e + 1] = f.m[size_t(e)].im;
        }
        const int q32 = int(f.qubit);
        check(dq_shots_add_gate(S, u, &q32, 1), "shots gate");
    }
}

// write_shots_profile (profiling.rs:8-18): ./dqsim_profiles/<prefix>_shots_profile_<unix ns>.json,
// pretty-printed.
void write_profile_json(const char *prefix, const py::dict &p) {
    py::module_ json = py::module_::import("json");
    py::module_ os = py::module_::import("os");
    py::module_ time = py::module_::import("time");
    os.attr("makedirs")("dqsim_profiles", py::arg("exist_ok") = true);
    const std::string path = std::string("dqsim_profiles/") + prefix + "_shots_profile_" +
                             std::string(py::str(time.attr("time_ns")())) + ".json";
    py::object f = py::module_::import("builtins").attr("open")(path, "w");
    f.attr("write")(json.attr("dumps")(p, py::arg("indent") = 2));
    f.attr("close")();
}

// format_cbits (types.rs:43-51): MSB first, missing bits 0.
std::string format_shot(const uint64_t *w, int num_cbits) {
    std::string s(size_t(num_cbits), '0');
    for (int i = 0; i < num_cbits; ++i)
        if ((w[i / 64] >> (i % 64)) & 1) s[size_t(num_cbits - 1 - i)] = '1';
    return s;
}

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
        {
            py::gil_scoped_release nogil;
            check(dq_copy_to_host(blk_.st->s, p), "statevector copy");
        }
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
    py::tuple stats() {
        uint64_t a = 0, b = 0, c = 0;
        check(dq_get_stats(st_->s, &a, &b, &c), "stats");
        return py::make_tuple(a, b, c);
    }
    void set_fuse_min_cost(double c) { check(dq_set_fuse_min_cost(st_->s, c), "fusion cost threshold"); }
    void set_pair_fusion(bool on) { check(dq_set_pair_fusion(st_->s, on ? 1 : 0), "pair fusion"); }
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

   private:
    int n_;
    std::shared_ptr<DeviceState> st_;
    Circuit circ_;
};

// ---- batching (no reference counterpart) ----
//
// simulate_batch(circuits, observable=None, dtype="c128", workers=8): <H> of
// every circuit from |0...0>, H = sum_t c_t Z(zmask_t) (default: the open
// chain sum_{i=0}^{n-2} Z_i Z_{i+1} of each circuit's width).  Circuits must
// be unitary (mid-circuit measure / reset / conditional are rejected).
py::array_t<double> simulate_batch(py::list circuits, py::object observable, const std::string &dtype,
                                   int workers) {
    int dt;
    if (dtype == "c128" || dtype == "complex128") dt = DQ_C128;
    else if (dtype == "c64" || dtype == "complex64") dt = DQ_C64;
    else throw py::value_error("dtype must be 'c128' or 'c64'");
    std::vector<uint64_t> zm;
    std::vector<double> zc;
    const bool custom = !observable.is_none();
    if (custom) {
        for (py::handle t : observable) {
            auto tup = py::reinterpret_borrow<py::tuple>(t);
            if (tup.size() != 2) throw py::value_error("observable terms are (coeff, zmask) pairs");
            zc.push_back(tup[0].cast<double>());
            zm.push_back(tup[1].cast<uint64_t>());
        }
    }
    dq_batch *b = nullptr;
    check(dq_batch_create(dt, &b), "batch");
    std::unique_ptr<dq_batch, void (*)(dq_batch *)> guard(b, dq_batch_destroy);
    check(dq_batch_set_workers(b, workers), "batch workers");
    const size_t count = circuits.size();
    for (size_t i = 0; i < count; ++i) {
        py::handle obj = circuits[i];
        const std::string text = py::isinstance<py::str>(obj) ? obj.cast<std::string>()
                                                               : std::string(py::str(obj.attr("model_dump_json")()));
        Circuit c;
        try {
            c = Decoder().decode(text);
        } catch (const JsonError &e) {
            throw py::value_error(std::string("Circuit JSON parse error: ") + e.what());
        }
        const uint64_t n = c.num_qubits();
        if (n < 1 || n > 34) throw py::value_error("simulate_batch: circuit " + std::to_string(i) + " has " +
                                                   std::to_string(n) + " qubits (1..34 supported)");
        int idx = 0;
        check(dq_batch_add_instance(b, int(n), &idx), "batch instance");
        for (const Instr &in : c.instructions) {
            if (in.kind == "measure" || in.kind == "reset" || in.kind == "conditional")
                throw py::value_error("simulate_batch: circuit " + std::to_string(i) + " has a '" + in.kind +
                                      "' instruction; batch instances must be unitary");
            Mat m;
            if (!unitary_of(in, m)) continue;
            if (int(in.qubits.size()) != m.k)
                throw py::value_error("gate expects " + std::to_string(m.k) + " qubits, got " +
                                      std::to_string(in.qubits.size()));
            int qs[5];
            for (int j = 0; j < m.k; ++j) {
                if (in.qubits[size_t(j)] >= n)
                    throw py::value_error("qubit index " + std::to_string(in.qubits[size_t(j)]) + " out of range for " +
                                          std::to_string(n) + "-qubit circuit");
                qs[j] = int(in.qubits[size_t(j)]);
            }
            check(dq_batch_apply(b, idx, m.v.data(), qs, m.k), "batch gate");
        }
        if (custom) check(dq_batch_set_observable(b, idx, int(zm.size()), zm.data(), zc.data()), "observable");
    }
    py::array_t<double> out(static_cast<py::ssize_t>(count));
    if (count == 0) return out;
    double *o = out.mutable_data();
    int rc;
    {
        py::gil_scoped_release nogil;
        rc = dq_batch_run(b, o);
    }
    check(rc, "simulate_batch");
    return out;
}

// The hardware-efficient-ansatz family built directly from arrays (no JSON):
// instance i has ns[i] qubits and ds[i] layers; per layer ry(q) then rz(q)
// on every qubit with angles (ry, rz) taken in order from `angles`, then
// cx(q, q+1) for q = 0..n-2 and cx(n-1, 0).  Same gates as the JSON circuits
// of that family; the observable is the open ZZ chain.
py::tuple simulate_hea_batch(py::array_t<int64_t, py::array::c_style | py::array::forcecast> ns,
                             py::array_t<int64_t, py::array::c_style | py::array::forcecast> ds,
                             py::array_t<double, py::array::c_style | py::array::forcecast> angles,
                             const std::string &dtype, int workers) {
    const int dt = (dtype == "c64" || dtype == "complex64") ? DQ_C64 : DQ_C128;
    if (dt == DQ_C128 && dtype != "c128" && dtype != "complex128") throw py::value_error("dtype must be 'c128' or 'c64'");
    const size_t count = size_t(ns.size());
    if (size_t(ds.size()) != count) throw py::value_error("ns and ds differ in length");
    size_t need = 0;
    for (size_t i = 0; i < count; ++i) {
        if (ns.data()[i] < 2 || ns.data()[i] > 34 || ds.data()[i] < 0)
            throw py::value_error("widths must be in [2, 34] and depths >= 0");
        need += size_t(2 * ns.data()[i] * ds.data()[i]);
    }
    if (size_t(angles.size()) != need) throw py::value_error("angles must hold 2 * n * d values per instance");
    dq_batch *b = nullptr;
    check(dq_batch_create(dt, &b), "batch");
    std::unique_ptr<dq_batch, void (*)(dq_batch *)> guard(b, dq_batch_destroy);
    check(dq_batch_set_workers(b, workers), "batch workers");
    Mat cx;
    gate_matrix("cx", {}, cx);
    const double *a = angles.data();
    for (size_t i = 0; i < count; ++i) {
        const int n = int(ns.data()[i]), d = int(ds.data()[i]);
        int idx = 0;
        check(dq_batch_add_instance(b, n, &idx), "batch instance");
        for (int l = 0; l < d; ++l) {
            for (int q = 0; q < n; ++q) {
                Mat m;
                gate_matrix("ry", {a[2 * q]}, m);
                check(dq_batch_apply(b, idx, m.v.data(), &q, 1), "ry");
            }
            for (int q = 0; q < n; ++q) {
                Mat m;
                gate_matrix("rz", {a[2 * q + 1]}, m);
                check(dq_batch_apply(b, idx, m.v.data(), &q, 1), "rz");
            }
            for (int q = 0; q < n; ++q) {
                const int qs[2] = {q, (q + 1) % n};
                check(dq_batch_apply(b, idx, cx.v.data(), qs, 2), "cx");
            }
            a += 2 * n;
        }
    }
    py::array_t<double> out(static_cast<py::ssize_t>(std::max<size_t>(count, 1)));
    int rc;
    {
        py::gil_scoped_release nogil;
        rc = dq_batch_run(b, out.mutable_data());
    }
    check(rc, "simulate_hea_batch");
    uint64_t lds = 0, eng = 0;
    double ms = 0;
    check(dq_batch_stats(b, &lds, &eng, &ms), "batch stats");
    py::dict st;
    st["lds_instances"] = lds;
    st["engine_instances"] = eng;
    st["lds_gpu_ms"] = ms;
    return py::make_tuple(out[py::slice(0, py::ssize_t(count), 1)], st);
}

py::tuple batch_stats_probe(py::list circuits, const std::string &dtype) {
    // paths the batch would take for these widths (no device work)
    const int lim = (dtype == "c64" || dtype == "complex64") ? 14 : 13;
    uint64_t lds = 0, eng = 0;
    for (py::handle obj : circuits) {
        const std::string text = py::isinstance<py::str>(obj) ? obj.cast<std::string>()
                                                               : std::string(py::str(obj.attr("model_dump_json")()));
        Circuit c = Decoder().decode(text);
        (c.num_qubits() <= uint64_t(lim) ? lds : eng)++;
    }
    return py::make_tuple(lds, eng);
}

// ---- adjoint differentiation (no reference counterpart) ----
//
// E(theta) = <psi(theta)| H |psi(theta)>, psi = U_L ... U_1 |0>, H a Pauli sum.
// Trainable gates have the form U = exp(-i theta G / 2) with G a Pauli sum
// (rx/ry/rz/rxx/rzz: one Pauli string; crx/cry/crz, p/u1: two; cp/cu1: four).
// Reverse sweep (Jones & Gacon, arXiv:2009.02823): lambda = H psi; for each
// gate from the last: dE/dtheta += Im <lambda| G |psi>; then psi <- U^dag psi,
// lambda <- U^dag lambda, applied in place (no stored per-layer states).

struct PauliSumH {
    std::vector<uint64_t> x, z;
    std::vector<double> c;  // (re, im) pairs
    void add(double cr, double ci, uint64_t xm, uint64_t zm) {
        x.push_back(xm);
        z.push_back(zm);
        c.push_back(cr);
        c.push_back(ci);
    }
    int size() const { return int(x.size()); }
};

struct POp {
    std::string name;
    std::vector<int> qubits;
    int param = -1;     // index into the parameter vector, or -1 (fixed)
    double value = 0;   // fixed angle when param < 0
};

// Generator of a trainable gate, U = exp(-i theta G / 2).
PauliSumH generator(const POp &op) {
    PauliSumH g;
    const auto bit = [&](int j) { return uint64_t(1) << op.qubits[size_t(j)]; };
    const std::string &k = op.name;
    if (k == "rx") g.add(1, 0, bit(0), 0);
    else if (k == "ry") g.add(1, 0, bit(0), bit(0));
    else if (k == "rz") g.add(1, 0, 0, bit(0));
    else if (k == "rxx") g.add(1, 0, bit(0) | bit(1), 0);
    else if (k == "rzz") g.add(1, 0, 0, bit(0) | bit(1));
    else if (k == "p" || k == "u1") {  // diag(1, e^{i t}) = e^{i t/2} exp(-i t Z / 2): G = Z - I
        g.add(1, 0, 0, bit(0));
        g.add(-1, 0, 0, 0);
    } else if (k == "cp" || k == "cu1") {  // G = -2 |11><11| = -(I - Z_a - Z_b + Z_a Z_b) / 2
        g.add(-0.5, 0, 0, 0);
        g.add(0.5, 0, 0, bit(0));
        g.add(0.5, 0, 0, bit(1));
        g.add(-0.5, 0, 0, bit(0) | bit(1));
    } else if (k == "crx" || k == "cry" || k == "crz") {  // G = |1><1|_c (x) P_t = (I - Z_c)/2 P_t
        const uint64_t c = bit(0), t = bit(1);
        const uint64_t px = (k == "crz") ? 0 : t, pz = (k == "crx") ? 0 : t;
        g.add(0.5, 0, px, pz);
        g.add(-0.5, 0, px, pz | c);
    } else {
        throw py::value_error("gate '" + k + "' cannot carry a trainable parameter");
    }
    return g;
}

Mat dagger(const Mat &m) {
    Mat d = m;
    for (int r = 0; r < m.dim(); ++r)
        for (int c = 0; c < m.dim(); ++c) {
            const Cx v = m.get(c, r);
            d.set(r, c, {v.re, -v.im});
        }
    return d;
}

class Adjoint {
   public:
    Adjoint(int n, int dtype, int fusion) : n_(n), dtype_(dtype), psi_(n, dtype), lam_(n, dtype) {
        check(dq_set_options(psi_.s, fusion, 0, 0), "options");
        check(dq_set_options(lam_.s, fusion, 0, 0), "options");
    }

    void set_circuit(const std::vector<std::tuple<std::string, std::vector<int>, int, double>> &ops) {
        ops_.clear();
        n_params_ = 0;
        for (const auto &t : ops) {
            POp op;
            op.name = std::get<0>(t);
            op.qubits = std::get<1>(t);
            op.param = std::get<2>(t);
            op.value = std::get<3>(t);
            for (int q : op.qubits)
                if (q < 0 || q >= n_) throw py::value_error("qubit index out of range");
            static const char *multi[] = {"u3", "u", "u2", "cu3", "cu"};
            for (const char *mname : multi)
                if (op.name == mname)
                    throw py::value_error("multi-parameter gate '" + op.name + "' is not supported in a parametric circuit");
            Mat m;
            if (!gate_matrix(op.name, {0.0, 0.0, 0.0, 0.0}, m)) throw py::value_error("unknown gate '" + op.name + "'");
            if (int(op.qubits.size()) != m.k) throw py::value_error("wrong number of qubits for '" + op.name + "'");
            if (op.param >= 0) {
                generator(op);  // validates
                n_params_ = std::max(n_params_, op.param + 1);
            }
            ops_.push_back(op);
        }
    }

    void set_observable(const std::vector<uint64_t> &x, const std::vector<uint64_t> &z,
                        const std::vector<std::complex<double>> &c) {
        if (x.size() != z.size() || x.size() != c.size()) throw py::value_error("observable arrays differ in length");
        h_ = PauliSumH();
        for (size_t k = 0; k < x.size(); ++k) h_.add(c[k].real(), c[k].imag(), x[k], z[k]);
    }

    int num_params() const { return n_params_; }

    // true (default): the reverse sweep runs as tile passes over both states
    // (dq_adjoint_sweep, n >= 8); false: one inner product per parameter.
    void set_tiled_sweep(bool on) { sweep_ = on; }

    double expectation(const std::vector<double> &params) {
        forward(params);
        return energy();
    }

    py::tuple gradient(const std::vector<double> &params) {
        forward(params);
        check(dq_apply_pauli_sum(lam_.s, psi_.s, h_.size(), h_.x.data(), h_.z.data(), h_.c.data()), "H psi");
        const double e = energy_with_lambda();
        std::vector<double> grad(size_t(n_params_), 0.0);
        if (n_ >= 8 && sweep_) {
            // the whole reverse sweep as tile passes over both states
            const size_t L = ops_.size();
            std::vector<int> ks(L), qs(5 * L, 0), nt(L, 0);
            std::vector<double> mats, coeff, g(L, 0.0);
            std::vector<uint64_t> xm, zm;
            for (size_t l = 0; l < L; ++l) {
                const POp &op = ops_[l];
                const Mat ud = dagger(matrix(op, params));
                ks[l] = ud.k;
                for (int j = 0; j < ud.k; ++j) qs[5 * l + size_t(j)] = op.qubits[size_t(j)];
                mats.insert(mats.end(), ud.v.begin(), ud.v.end());
                if (op.param >= 0) {
                    const PauliSumH gen = generator(op);
                    nt[l] = gen.size();
                    xm.insert(xm.end(), gen.x.begin(), gen.x.end());
                    zm.insert(zm.end(), gen.z.begin(), gen.z.end());
                    coeff.insert(coeff.end(), gen.c.begin(), gen.c.end());
                }
            }
            check(dq_adjoint_sweep(psi_.s, lam_.s, int(L), ks.data(), qs.data(), mats.data(), nt.data(), xm.data(),
                                   zm.data(), coeff.data(), g.data()),
                  "adjoint sweep");
            for (size_t l = 0; l < L; ++l)
                if (ops_[l].param >= 0) grad[size_t(ops_[l].param)] += g[l];
            return py::make_tuple(e, grad);
        }
        for (size_t l = ops_.size(); l-- > 0;) {
            const POp &op = ops_[l];
            if (op.param >= 0) {
                const PauliSumH g = generator(op);
                double re = 0, im = 0;
                check(dq_pauli_inner(lam_.s, psi_.s, g.size(), g.x.data(), g.z.data(), g.c.data(), &re, &im),
                      "generator inner product");
                grad[size_t(op.param)] += im;
            }
            const Mat ud = dagger(matrix(op, params));
            apply(psi_.s, ud, op.qubits);
            apply(lam_.s, ud, op.qubits);
        }
        check(dq_synchronize(psi_.s), "synchronize");
        check(dq_synchronize(lam_.s), "synchronize");
        return py::make_tuple(e, grad);
    }

    SimulationResult state() { return SimulationResult(psi_share(), n_, {}, py::none()); }

   private:
    int n_;
    int dtype_;
    DeviceState psi_, lam_;
    std::vector<POp> ops_;
    PauliSumH h_;
    int n_params_ = 0;
    bool sweep_ = true;

    std::shared_ptr<DeviceState> psi_share() {
        // a fresh state holding a copy of psi, owned by the result
        auto st = std::make_shared<DeviceState>(n_, dtype_);
        check(dq_state_copy(st->s, psi_.s), "copy");
        return st;
    }

    Mat matrix(const POp &op, const std::vector<double> &params) const {
        double th = op.value;
        if (op.param >= 0) {
            if (size_t(op.param) >= params.size()) throw py::value_error("parameter vector too short");
            th = params[size_t(op.param)];
        }
        Mat m;
        gate_matrix(op.name, {th, th, th, th}, m);
        return m;
    }

    static void apply(dq_state *s, const Mat &m, const std::vector<int> &qs) {
        if (m.k == 1) check(dq_apply_1q(s, m.v.data(), qs[0], nullptr, nullptr, 0), "apply");
        else check(dq_apply_nq(s, m.v.data(), qs.data(), m.k), "apply");
    }

    void forward(const std::vector<double> &params) {
        check(dq_state_reset(psi_.s), "reset");
        for (const POp &op : ops_) apply(psi_.s, matrix(op, params), op.qubits);
    }

    double energy() {
        double re = 0, im = 0;
        check(dq_pauli_inner(psi_.s, psi_.s, h_.size(), h_.x.data(), h_.z.data(), h_.c.data(), &re, &im), "energy");
        return re;
    }

    double energy_with_lambda() {
        const uint64_t zero = 0;
        const double one[2] = {1.0, 0.0};
        double re = 0, im = 0;
        check(dq_pauli_inner(psi_.s, lam_.s, 1, &zero, &zero, one, &re, &im), "energy");
        return re;
    }
};

enum class Mode { StateVector, Mps, Stabilizer };

// codecs.rs:19-28
Mode parse_mode(const std::string &mode) {
    std::string m;
    size_t a = 0, b = mode.size();
    while (a < b && std::isspace(static_cast<unsigned char>(mode[a]))) ++a;
    while (b > a && std::isspace(static_cast<unsigned char>(mode[b - 1]))) --b;
    for (size_t i = a; i < b; ++i) {
        char c = char(std::tolower(static_cast<unsigned char>(mode[i])));
        m += (c == '-') ? '_' : c;
    }
    if (m == "state_vector" || m == "statevector" || m == "sv") return Mode::StateVector;
    if (m == "mps" || m == "matrix_product_state") return Mode::Mps;
    if (m == "stabilizer" || m == "stab") return Mode::Stabilizer;
    throw py::value_error("Unsupported monolithic simulation mode " + rust_debug(m) +
                          "; expected 'state_vector' or 'mps'");
}

// codecs.rs:68-84
void reject_options(const py::kwargs &options, const std::string &mode) {
    if (options.size() == 0) return;
    std::string keys = "[";
    bool first = true;
    for (auto item : options) {
        if (!first) keys += ", ";
        first = false;
        keys += rust_debug(py::str(item.first));
    }
    keys += "]";
    throw py::type_error("Options " + keys + " are not supported for monolithic mode " + rust_debug(mode));
}

[[noreturn]] void other_engine(const char *what) {
    PyErr_SetString(PyExc_NotImplementedError, what);
    throw py::error_already_set();
}

}  // namespace

// Plain op list for the planner bindings: (qubits, flat complex matrix,
// control values or empty); with control values the matrix is the 2x2
// applied to qubits[-1] under controls qubits[:-1] (as dq_apply_1q).
static std::vector<dq::HostOp> host_ops(py::list ops) {
    std::vector<dq::HostOp> queue;
    for (py::handle item : ops) {
        py::tuple t = py::reinterpret_borrow<py::tuple>(item);
        auto qs = t[0].cast<std::vector<int>>();
        auto u = t[1].cast<std::vector<std::complex<double>>>();
        auto cv = t[2].cast<std::vector<int>>();
        dq::HostOp op;
        op.k = int(qs.size());
        for (int j = 0; j < op.k; ++j) op.q[j] = qs[size_t(j)];
        std::vector<double> flat(2 * u.size());
        for (size_t e = 0; e < u.size(); ++e) {
            flat[2 * e] = u[e].real();
            flat[2 * e + 1] = u[e].imag();
        }
        if (!cv.empty()) {
            op.kind = dq::K_CTRL1;
            for (int j = 0; j + 1 < op.k; ++j) {
                op.cmask |= uint64_t(1) << qs[size_t(j)];
                if (cv[size_t(j)]) op.cval |= uint64_t(1) << qs[size_t(j)];
            }
            op.mat = flat;
            op.dense = flat;
        } else {
            dq::classify(op, flat.data());
        }
        queue.push_back(std::move(op));
    }
    return queue;
}

PYBIND11_MODULE(_core, m) {
    m.doc() = "MI355X statevector engine: the reference's dqsim statevector API over HIP kernels";

    py::class_<SimulationProfile>(m, "SimulationProfile")
        .def_readonly("apply_one_qubit_calls", &SimulationProfile::oq_calls)
        .def_readonly("apply_one_qubit_time", &SimulationProfile::oq_time)
        .def_readonly("apply_n_qubit_calls", &SimulationProfile::nq_calls)
        .def_readonly("apply_n_qubit_time", &SimulationProfile::nq_time)
        .def_readonly("measure_qubit_calls", &SimulationProfile::mq_calls)
        .def_readonly("measure_qubit_time", &SimulationProfile::mq_time)
        .def_readonly("total_time", &SimulationProfile::total_time)
        .def("__repr__", &SimulationProfile::repr);

    py::class_<SimulationResult>(m, "SimulationResult")
        .def_property_readonly("statevector", &SimulationResult::statevector)
        .def_property_readonly("num_qubits", &SimulationResult::num_qubits)
        .def_property_readonly("classical_bits", &SimulationResult::classical_bits)
        .def_property_readonly("profile", &SimulationResult::profile)
        .def("probabilities", &SimulationResult::probabilities, py::arg("qubits") = py::none())
        .def("probabilities_array", &SimulationResult::probabilities_array, py::arg("qubits") = py::none())
        .def("top_k", &SimulationResult::top_k, py::arg("k"))
        .def("counts", &SimulationResult::counts, py::arg("shots") = 1000, py::arg("qubits") = py::none(),
             py::arg("seed") = py::none())
        .def("fidelity", &SimulationResult::fidelity, py::arg("other"));

    py::class_<StatevectorSimulator>(m, "StatevectorSimulator")
        .def(py::init<std::optional<uint64_t>, bool, int, int, int>(), py::arg("seed") = py::none(),
             py::arg("profile") = false, py::kw_only(), py::arg("fusion") = 2, py::arg("tile_bits") = 0,
             py::arg("low_bits") = 0)
        .def("simulate", &StatevectorSimulator::simulate, py::arg("circuit"))
        .def("simulate_shots", &StatevectorSimulator::simulate_shots, py::arg("circuit"), py::arg("shots") = 1000,
             py::arg("profile") = false)
        .def("simulate_opstream", &StatevectorSimulator::simulate_opstream, py::arg("records"), py::arg("num_qubits"),
             py::arg("num_cbits") = 0, py::arg("names") = std::vector<std::string>{})
        .def("simulate_shots_opstream", &StatevectorSimulator::simulate_shots_opstream, py::arg("records"),
             py::arg("num_qubits"), py::arg("num_cbits") = 0, py::arg("names") = std::vector<std::string>{},
             py::arg("shots") = 1000, py::arg("profile") = false);

    // mps/simulator.rs:300-421
    py::class_<MpsSimulator>(m, "MpsSimulator")
        .def(py::init<std::optional<uint64_t>, std::optional<uint64_t>, double>(), py::arg("seed") = py::none(),
             py::arg("max_bond_dimension") = py::none(), py::arg("truncation_threshold") = 1e-12)
        .def("simulate", &MpsSimulator::simulate, py::arg("circuit"))
        .def("simulate_shots", &MpsSimulator::simulate_shots, py::arg("circuit"), py::arg("shots") = 1000,
             py::arg("profile") = false);

    m.def("_svd_jacobi", &svd_jacobi_hook, py::arg("a"));
    m.def("_mps_network", &mps_network_hook, py::arg("circuit_json"), py::arg("max_bond_dimension") = py::none(),
          py::arg("truncation_threshold") = 1e-12);

    // pblock/simulator.rs:38-460 and simulator.rs (crate root) :47-61, :97-114
    py::class_<PBlockResult>(m, "PBlockResult")
        .def_property_readonly("statevector", &PBlockResult::statevector)
        .def_property_readonly("num_qubits", &PBlockResult::num_qubits)
        .def_property_readonly("physical_qubits", &PBlockResult::physical_qubits)
        .def_property_readonly("classical_bits", &PBlockResult::classical_bits)
        .def("probabilities", &PBlockResult::probabilities, py::arg("qubits") = py::none());
    py::class_<PBlockSimulator>(m, "PBlockSimulator")
        .def(py::init<std::optional<uint64_t>>(), py::arg("seed") = py::none())
        .def("simulate", &PBlockSimulator::simulate, py::arg("distributed"))
        .def("simulate_shots", &PBlockSimulator::simulate_shots, py::arg("distributed"), py::arg("shots") = 1000,
             py::arg("profile") = false);
    m.def(
        "simulate_distributed",
        [](py::object distributed, const std::string &mode, std::optional<uint64_t> seed) {
            parse_distributed_mode(mode);
            return PBlockSimulator(seed).simulate(distributed);
        },
        py::arg("distributed"), py::arg("mode") = "p_block", py::arg("seed") = py::none());
    m.def(
        "simulate_distributed_shots",
        [](py::object distributed, const std::string &mode, uint64_t shots, std::optional<uint64_t> seed,
           bool profile) {
            parse_distributed_mode(mode);
            return PBlockSimulator(seed).simulate_shots(distributed, shots, profile);
        },
        py::arg("distributed"), py::arg("mode") = "p_block", py::arg("shots") = 1000, py::arg("seed") = py::none(),
        py::arg("profile") = false);

    // simulator.rs (crate root) simulate_monolithic_shots
    m.def(
        "simulate_monolithic_shots",
        [](py::object circuit, const std::string &mode, uint64_t shots, std::optional<uint64_t> seed, bool profile,
           py::kwargs options) -> py::object {
            switch (parse_mode(mode)) {
            case Mode::StateVector:
                reject_options(options, mode);
                return StatevectorSimulator(seed, false).simulate_shots(circuit, shots, profile);
            case Mode::Mps:
                return mps_from_options(seed, options).simulate_shots(circuit, shots, profile);
            case Mode::Stabilizer:
                other_engine("mode 'stabilizer' is a different algorithm and is not provided by this statevector engine");
            }
            return py::none();
        },
        py::arg("circuit"), py::arg("mode") = "state_vector", py::arg("shots") = 1000, py::arg("seed") = py::none(),
        py::arg("profile") = false);

    m.def(
        "simulate_monolithic",
        [](py::object circuit, const std::string &mode, std::optional<uint64_t> seed, bool profile,
           py::kwargs options) -> py::object {
            switch (parse_mode(mode)) {
            case Mode::StateVector:
                reject_options(options, mode);
                return py::cast(StatevectorSimulator(seed, profile).simulate(circuit));
            case Mode::Mps:
                return py::cast(mps_from_options(seed, options).simulate(circuit));
            case Mode::Stabilizer:
                other_engine(
                    "Full statevector extraction ('simulate_monolithic') is not supported for stabilizer "
                    "simulation. Please use shot-based sampling via 'simulate_monolithic_shots' instead.");
            }
            return py::none();
        },
        py::arg("circuit"), py::arg("mode") = "state_vector", py::arg("seed") = py::none(),
        py::arg("profile") = false);

    py::class_<Engine>(m, "Engine")
        .def(py::init<int, int, int, int, int>(), py::arg("n"), py::arg("fusion") = 2, py::arg("tile_bits") = 0,
             py::arg("low_bits") = 0, py::arg("dtype") = 0)
        .def("load", &Engine::load)
        .def("step", &Engine::step, py::arg("seed"))
        .def("set_timing", &Engine::set_timing)
        .def("timing", &Engine::timing)
        .def("stats", &Engine::stats)
        .def("fused_stats", &Engine::fused_stats)
        .def("set_fuse_min_cost", &Engine::set_fuse_min_cost)
        .def("set_pair_fusion", &Engine::set_pair_fusion)
        .def("compare", &Engine::compare)
        .def("result", &Engine::result);

    m.def("simulate_batch", &simulate_batch, py::arg("circuits"), py::arg("observable") = py::none(),
          py::arg("dtype") = "c128", py::arg("workers") = 8);
    m.def("_batch_paths", &batch_stats_probe, py::arg("circuits"), py::arg("dtype") = "c128");
    m.def("simulate_hea_batch", &simulate_hea_batch, py::arg("ns"), py::arg("ds"), py::arg("angles"),
          py::arg("dtype") = "c128", py::arg("workers") = 8);

    py::class_<Adjoint>(m, "Adjoint")
        .def(py::init<int, int, int>(), py::arg("n"), py::arg("dtype") = 0, py::arg("fusion") = 2)
        .def("set_circuit", &Adjoint::set_circuit)
        .def("set_observable", &Adjoint::set_observable)
        .def("num_params", &Adjoint::num_params)
        .def("set_tiled_sweep", &Adjoint::set_tiled_sweep)
        .def("expectation", &Adjoint::expectation)
        .def("gradient", &Adjoint::gradient)
        .def("state", &Adjoint::state);

    // Host-logic hooks used by the CPU test-suite (no device needed).
    m.def("_gate_matrix", [](const std::string &name, std::vector<double> params) {
        Mat mm;
        if (!gate_matrix(name, params, mm)) throw py::key_error(name);
        const size_t d = size_t(1) << mm.k;
        py::array_t<std::complex<double>> out({d, d});
        std::memcpy(out.mutable_data(), mm.v.data(), mm.v.size() * sizeof(double));
        return out;
    });
    m.def("_rng_stream", [](uint64_t seed, size_t count) {
        ChaCha8Rng r = ChaCha8Rng::seed_from_u64(seed);
        std::vector<uint64_t> v(count);
        for (auto &x : v) x = r.next_u64();
        return v;
    });
    auto listing = [](const Circuit &c) {
        py::list out;
        std::function<py::dict(const Instr &)> conv = [&](const Instr &in) {
            py::dict d;
            d["kind"] = in.kind;
            d["qubits"] = in.qubits;
            d["params"] = in.params;
            d["name"] = in.name;
            d["cbit"] = in.cbit;
            d["condition"] = py::make_tuple(in.creg_base, in.creg_size, in.creg_value);
            if (in.op) d["op"] = conv(*in.op);
            return d;
        };
        for (const Instr &in : c.instructions) out.append(conv(in));
        return py::make_tuple(c.num_qubits(), c.num_cbits(), out);
    };
    m.def("_decode_opstream", [listing](py::array records, uint64_t nq, uint64_t ncb, std::vector<std::string> names) {
        return listing(StatevectorSimulator::opstream_circuit(records, nq, ncb, names));
    });
    m.def("_decode", [](const std::string &text) {
        Circuit c;
        try {
            c = Decoder().decode(text);
        } catch (const JsonError &e) {
            throw py::value_error(std::string("Circuit JSON parse error: ") + e.what());
        }
        py::list out;
        std::function<py::dict(const Instr &)> conv = [&](const Instr &in) {
            py::dict d;
            d["kind"] = in.kind;
            d["qubits"] = in.qubits;
            d["params"] = in.params;
            d["name"] = in.name;
            d["cbit"] = in.cbit;
            d["condition"] = py::make_tuple(in.creg_base, in.creg_size, in.creg_value);
            if (in.op) d["op"] = conv(*in.op);
            return d;
        };
        for (const Instr &in : c.instructions) out.append(conv(in));
        return py::make_tuple(c.num_qubits(), c.num_cbits(), out);
    });
    // Fusion level 3's block former on a plain op list (qubits not in
    // tile_qubits are outer).  ops: (qubits, flat complex matrix, control
    // values or empty); with control values the matrix is the 2x2 applied to
    // qubits[-1] under controls qubits[:-1] (as dq_apply_1q).  Returns, in
    // order, (qubits, 32x32 matrix, gate count) for a fused block or
    // (original index,) for an op kept as it is.
    m.def("_fuse_blocks", [](std::vector<int> tile_qubits, py::list ops, double min_cost, int max_qubits) {
        std::vector<dq::HostOp> queue = host_ops(ops);
        int pos[64], tq[16] = {0};
        for (int q = 0; q < 64; ++q) pos[q] = dq::kOuterBit + q;
        for (size_t t = 0; t < tile_qubits.size() && t < 16; ++t) {
            pos[tile_qubits[t]] = int(t);
            tq[t] = tile_qubits[t];
        }
        std::vector<const dq::HostOp *> in;
        for (const dq::HostOp &h : queue) in.push_back(&h);
        std::deque<dq::HostOp> store, pstore;
        // max_qubits 2: pair fusion; 5: blocks; 7: pairs, then blocks (as the planner at fusion level 3)
        auto eff = max_qubits == 5 ? dq::fuse_blocks(in, pos, tq, int(tile_qubits.size()), min_cost, store)
                                   : dq::fuse_pairs(in, pos, pstore);
        if (max_qubits == 7) eff = dq::fuse_blocks(eff, pos, tq, int(tile_qubits.size()), min_cost, store);
        py::list out;
        for (const dq::HostOp *h : eff) {
            if (h >= queue.data() && h < queue.data() + queue.size()) {
                out.append(py::make_tuple(int(h - queue.data())));
                continue;
            }
            const size_t d = size_t(1) << h->k, ns = size_t(1) << h->nsel;
            py::array_t<std::complex<double>> m({ns, d, d});
            std::memcpy(m.mutable_data(), h->dense.data(), h->dense.size() * sizeof(double));
            std::vector<int> qs(h->q, h->q + h->k), sq(h->selq, h->selq + h->nsel);
            out.append(py::make_tuple(qs, m, h->kind == dq::K_FUSED ? h->outcome : -1, sq));
        }
        return out;
    }, py::arg("tile_qubits"), py::arg("ops"), py::arg("min_cost") = 0.0, py::arg("max_qubits") = 5);
    // The pass builder on a plain op list (as _fuse_blocks); returns, per
    // pass, (tile qubit mask, op indices in the order the pass applies them).
    m.def("_schedule", [](int n, py::list ops, int fuse, int tile_bits, int low_bits, bool reg_tiles) {
        std::vector<dq::HostOp> queue = host_ops(ops);
        dq::PlanConfig cfg;
        cfg.n = n;
        cfg.fuse = fuse;
        cfg.tile_bits = tile_bits;
        cfg.low_bits = low_bits;
        cfg.reg_tiles = reg_tiles ? 1 : 0;
        const bool anywhere = fuse && reg_tiles && dq::effective_tile_bits(&cfg) >= 9;
        py::list out;
        for (const dq::Pass &p : dq::schedule(&cfg, queue, anywhere)) out.append(py::make_tuple(p.S, p.ops));
        return out;
    }, py::arg("n"), py::arg("ops"), py::arg("fuse") = 2, py::arg("tile_bits") = 12, py::arg("low_bits") = 4,
       py::arg("reg_tiles") = true);
    m.def("_fuse_circuit", [](const std::string &text) {
        Circuit c = Decoder().decode(text);
        py::list out;
        for (const FusedItem &f : fuse_instrs(c, c.num_qubits())) {
            if (f.instr >= 0) {
                out.append(py::make_tuple("orig", f.instr));
            } else {
                py::list m;
                for (const SCx &e : f.m) m.append(py::make_tuple(e.re, e.im));
                out.append(py::make_tuple("fused", f.qubit, py::tuple(m)));
            }
        }
        return out;
    });
    m.def("_parse_mode", [](const std::string &mode) { return int(parse_mode(mode)); });
}
