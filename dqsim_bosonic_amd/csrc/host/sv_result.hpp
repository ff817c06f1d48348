// sv_result.hpp -- Errors, the resident device state, SimulationProfile and SimulationResult
// (src/monolithic/statevector/simulator.rs:38-192).
//
// Part of the _core extension module: included once, by core.cpp, after its
// library includes and namespace aliases; the definitions are internal to
// that translation unit.
#pragma once

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

// The qubit list of probabilities() / counts(): the default is [n-1, ..., 0]
// (simulator.rs:150-152).  As in the reference, which takes Vec<usize>
// (simulator.rs:148), a negative index is an OverflowError; indices at or
// above n are accepted (see SimulationResult::marginal).
std::vector<int> qubits_list(const std::optional<std::vector<int64_t>> &qubits, int n) {
    std::vector<int> qs;
    if (!qubits) {
        for (int q = n - 1; q >= 0; --q) qs.push_back(q);
        return qs;
    }
    for (int64_t q : *qubits) {
        if (q < 0) {
            PyErr_SetString(PyExc_OverflowError, "can't convert negative int to unsigned");
            throw py::error_already_set();
        }
        // the reference's mask (1 << q) needs q < 64
        if (q >= 64) throw py::value_error("qubit index " + std::to_string(q) + " out of range");
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
        int rc;
        {
            py::gil_scoped_release nogil;
            rc = dq_copy_to_host(st_->s, p);
        }
        check(rc, "statevector copy");
        return out;
    }

    py::dict classical_bits() const {
        py::dict d;
        for (const auto &kv : cbits_) d[py::int_(kv.first)] = py::int_(kv.second);
        return d;
    }

    py::object profile() const { return prof_; }

    // engine::marginal_probs (src/engine.rs:368-390): probs[j] sums |psi_i|^2
    // over the i with (i & mask) == desired_j, mask and desired_j built from
    // the listed qubits (qubits[0] = MSB of j).  The device computes the
    // marginal over the distinct listed qubits below n; each bin j then takes
    // the bin of its desired pattern, which is the same set of terms.  A
    // qubit at or above n makes every bin that wants it set empty (0.0); a
    // repeated qubit makes every j that sets it in any position want it set.
    std::vector<double> marginal(const std::vector<int> &qs) const {
        if (qs.size() > 34) throw py::value_error("at most 34 qubits");
        std::vector<double> probs(size_t(1) << qs.size());
        marginal_to(qs, probs.data());
        return probs;
    }
    void marginal_to(const std::vector<int> &qs, double *out) const {
        const int k = int(qs.size());
        std::vector<int> distinct;
        bool plain = true;
        uint64_t seen = 0;
        for (int q : qs) {
            if (q >= n_ || ((seen >> q) & 1)) {
                plain = false;
                continue;
            }
            seen |= uint64_t(1) << q;
            distinct.push_back(q);
        }
        std::vector<double> sub(plain ? 0 : size_t(1) << distinct.size());
        int rc;
        {
            py::gil_scoped_release nogil;
            rc = dq_marginal_probs(st_->s, distinct.data(), int(distinct.size()), plain ? out : sub.data());
        }
        check(rc, "marginal probabilities");
        if (plain) return;
        const int kd = int(distinct.size());
        std::vector<int> pos(64, -1);  // qubit -> bit of the distinct marginal's index (MSB first)
        for (int j = 0; j < kd; ++j) pos[size_t(distinct[size_t(j)])] = kd - 1 - j;
        for (size_t j = 0; j < (size_t(1) << k); ++j) {
            bool empty = false;
            size_t jd = 0;
            for (int b = 0; b < k; ++b) {
                if (!((j >> (k - 1 - b)) & 1)) continue;
                const int q = qs[size_t(b)];
                if (q >= n_) {
                    empty = true;
                    break;
                }
                jd |= size_t(1) << pos[size_t(q)];
            }
            out[j] = empty ? 0.0 : sub[jd];
        }
    }

    // Lazy results (no reference counterpart): the distribution as an array
    // (index j as in probabilities(), no dict) ...
    py::array_t<double> probabilities_array(const std::optional<std::vector<int64_t>> &qubits) const {
        const std::vector<int> qs = qubits_list(qubits, n_);
        if (qs.size() > 34) throw py::value_error("at most 34 qubits");
        py::array_t<double> out(static_cast<py::ssize_t>(size_t(1) << qs.size()));
        marginal_to(qs, out.mutable_data());
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

}  // namespace
