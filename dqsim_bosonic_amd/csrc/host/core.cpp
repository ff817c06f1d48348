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
//
// The subsystems sit in headers included below, one each: results
// (sv_result.hpp), the dispatcher (sv_dispatch.hpp), shots (shots.hpp), the
// statevector simulator (statevector.hpp), the MPS simulator (mps_sim.hpp),
// the P-block simulator (pblock.hpp), batching (batch.hpp) and adjoint
// differentiation (adjoint.hpp).  This file holds the mode routing and the
// Python bindings.
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

#include "sv_result.hpp"
#include "sv_dispatch.hpp"
#include "shots.hpp"
#include "statevector.hpp"
#include "mps_sim.hpp"
#include "pblock.hpp"
#include "batch.hpp"
#include "adjoint.hpp"

namespace {

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
        if (u.empty()) {  // ([qubit], [], [outcome]): a collapse
            if (op.k != 1 || cv.size() != 1) throw std::invalid_argument("a collapse is ([qubit], [], [outcome])");
            op.kind = dq::K_COLLAPSE;
            op.outcome = cv[0];
            queue.push_back(std::move(op));
            continue;
        }
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
    // A profile with given fields, for checking the __repr__ format on fixed values.
    m.def("_profile_repr", [](uint64_t oq, double oqt, uint64_t nq, double nqt, uint64_t mq, double mqt,
                              double total) {
        SimulationProfile p;
        p.oq_calls = oq, p.oq_time = oqt, p.nq_calls = nq, p.nq_time = nqt, p.mq_calls = mq, p.mq_time = mqt;
        p.total_time = total;
        return p.repr();
    });

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
        .def("timed_amplitudes", &Engine::timed_amplitudes)
        .def("set_zero_tiles", &Engine::set_zero_tiles)
        .def("stats", &Engine::stats)
        .def("fused_stats", &Engine::fused_stats)
        .def("set_fuse_min_cost", &Engine::set_fuse_min_cost)
        .def("set_pair_fusion", &Engine::set_pair_fusion)
        .def("set_reg_bits", &Engine::set_reg_bits)
        .def("compare", &Engine::compare)
        .def("p1_record", &Engine::p1_record)
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
    // The register-tile plan of an op list (schedule, then plan_pass per
    // pass): totals of passes, tile ops and LDS layout changes (the stores'
    // included).
    m.def("_plan_stats", [](int n, py::list ops, int fuse, int tile_bits, int low_bits, int reg_bits) {
        std::vector<dq::HostOp> queue = host_ops(ops);
        dq::PlanConfig cfg;
        cfg.n = n;
        cfg.fuse = fuse;
        cfg.tile_bits = tile_bits;
        cfg.low_bits = low_bits;
        cfg.reg_bits = reg_bits;
        const bool anywhere = fuse && dq::effective_tile_bits(&cfg) >= 9;
        py::dict d;
        long passes = 0, tops = 0, relay = 0;
        std::vector<double> mats;
        for (const dq::Pass &p : dq::schedule(&cfg, queue, anywhere)) {
            ++passes;
            if (!p.S || !anywhere) continue;
            dq::RegPlan rp;
            dq::plan_pass(&cfg, queue, p, mats, rp);
            tops += long(rp.ops.size());
            relay += rp.relayouts;
        }
        d["passes"] = passes;
        d["tile_ops"] = tops;
        d["relayouts"] = relay;
        return d;
    }, py::arg("n"), py::arg("ops"), py::arg("fuse") = 2, py::arg("tile_bits") = 12, py::arg("low_bits") = 4,
       py::arg("reg_bits") = 4);
    // The phase-factored form of 4x4 matrices (one or more, as (re, im)
    // pairs, row-major): (variant, encoded entries), the entries unchanged
    // when the variant is V_GEN.
    m.def("_phase_factor4", [](std::vector<double> mat) {
        const int v = dq::phase_factor4(mat);
        return py::make_tuple(v, mat);
    });
    // The register-tile plan of an op list, pass by pass: tile bits, load and
    // store layouts, and each tile op as (kind, variant, relayout, layout,
    // register slots p and q, tile bits).
    m.def("_plan_passes", [](int n, py::list ops, int fuse, int tile_bits, int low_bits, int reg_bits) {
        std::vector<dq::HostOp> queue = host_ops(ops);
        dq::PlanConfig cfg;
        cfg.n = n;
        cfg.fuse = fuse;
        cfg.tile_bits = tile_bits;
        cfg.low_bits = low_bits;
        cfg.reg_bits = reg_bits;
        const bool anywhere = fuse && dq::effective_tile_bits(&cfg) >= 9;
        py::list out;
        std::vector<double> mats;
        for (const dq::Pass &p : dq::schedule(&cfg, queue, anywhere)) {
            if (!p.S || !anywhere) continue;
            dq::RegPlan rp;
            dq::plan_pass(&cfg, queue, p, mats, rp);
            py::list tops;
            for (const dq::TileOp &t : rp.ops) {
                py::list bits;
                for (int j = 0; j < t.k; ++j) bits.append(t.bit[j]);
                tops.append(py::make_tuple(t.kind, t.variant, t.relayout, t.rb, t.p, t.q, bits));
            }
            py::dict d;
            d["b"] = rp.b;
            d["load_rb"] = rp.pd.load_rb;
            d["store_rb"] = rp.pd.store_rb;
            d["relayouts"] = rp.relayouts;
            d["ops"] = tops;
            out.append(d);
        }
        return out;
    }, py::arg("n"), py::arg("ops"), py::arg("fuse") = 2, py::arg("tile_bits") = 12, py::arg("low_bits") = 4,
       py::arg("reg_bits") = 4);
    // The register passes of an op list applied to |0...0>, as a flush plans
    // them: per pass (tile mask, op indices, qubits 0 everywhere before it,
    // n_outer, outer qubits in outer-index order).
    m.def("_zero_passes", [](int n, py::list ops, int fuse, int tile_bits, int low_bits, int reg_bits) {
        std::vector<dq::HostOp> queue = host_ops(ops);
        dq::PlanConfig cfg;
        cfg.n = n;
        cfg.fuse = fuse;
        cfg.tile_bits = tile_bits;
        cfg.low_bits = low_bits;
        cfg.reg_bits = reg_bits;
        if (!(fuse && dq::effective_tile_bits(&cfg) >= 9)) throw std::invalid_argument("not a register-tile plan");
        dq::KnownBits kb;
        kb.zero = n ? ~uint64_t(0) >> (64 - n) : 0;
        py::list out;
        std::vector<double> mats;
        for (const dq::Pass &p : dq::schedule(&cfg, queue, true)) {
            dq::RegPlan rp;
            dq::plan_pass(&cfg, queue, p, mats, rp, kb.zero);
            std::vector<int> oq(rp.pd.oq, rp.pd.oq + (n - rp.b));
            out.append(py::make_tuple(p.S, p.ops, kb.zero, rp.pd.n_outer, oq));
            dq::advance_known(queue, p, kb);
        }
        return out;
    }, py::arg("n"), py::arg("ops"), py::arg("fuse") = 2, py::arg("tile_bits") = 12, py::arg("low_bits") = 4,
       py::arg("reg_bits") = 4);
    // The register-layout order of one pass's ops (dq_plan.hpp:
    // order_for_layouts); qubits not in tile_qubits are outer.  Returns the
    // op indices in the order the pass would apply them.
    m.def("_order_for_layouts", [](std::vector<int> tile_qubits, py::list ops) {
        std::vector<dq::HostOp> queue = host_ops(ops);
        int pos[64];
        for (int q = 0; q < 64; ++q) pos[q] = dq::kOuterBit + q;
        for (size_t t = 0; t < tile_qubits.size() && t < 16; ++t) pos[tile_qubits[t]] = int(t);
        std::vector<const dq::HostOp *> in;
        for (const dq::HostOp &h : queue) in.push_back(&h);
        std::vector<int> out;
        for (const dq::HostOp *h : dq::order_for_layouts(in, pos)) out.push_back(int(h - queue.data()));
        return out;
    }, py::arg("tile_qubits"), py::arg("ops"));
    // One layout-table entry of the register sweep (dq_plan.hpp): the 16 LDS
    // byte offsets of layout rb for tiles of b qubits and amplitudes of elem bytes.
    m.def("_layout_table", [](int rb, int b, unsigned elem) {
        std::vector<uint32_t> out;
        dq::append_layout_table(out, rb, b, elem);
        return out;
    }, py::arg("rb"), py::arg("b"), py::arg("elem"));
    m.def("_tile_swizzle", &dq::tile_swizzle, py::arg("i"), py::arg("b"));
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
