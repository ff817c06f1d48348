// shots.hpp -- Batched shots: the fused shot program, profile JSON and shot keys
// (src/monolithic/statevector/simulator.rs:215-322, src/types.rs:43-51,248-293).
//
// Part of the _core extension module: included once, by core.cpp, after its
// library includes and namespace aliases; the definitions are internal to
// that translation unit.
#pragma once

namespace {

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

}  // namespace
