// adjoint.hpp -- Adjoint differentiation (no reference counterpart).
//
// Part of the _core extension module: included once, by core.cpp, after its
// library includes and namespace aliases; the definitions are internal to
// that translation unit.
#pragma once

namespace {

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

}  // namespace
