// batch.hpp -- Batched independent circuits (no reference counterpart).
//
// Part of the _core extension module: included once, by core.cpp, after its
// library includes and namespace aliases; the definitions are internal to
// that translation unit.
#pragma once

namespace {

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
    uint64_t cnt[3] = {0, 0, 0};
    double graph_ms = 0, engine_s = 0, bytes[3] = {0, 0, 0};
    check(dq_batch_path_stats(b, cnt, &graph_ms, &engine_s, bytes), "batch stats");
    py::dict st;
    st["lds_instances"] = cnt[0];
    st["persistent_instances"] = cnt[1];
    st["engine_instances"] = cnt[2];
    st["graph_gpu_ms"] = graph_ms;
    st["engine_wall_s"] = engine_s;
    st["persistent_bytes"] = bytes[1];
    st["engine_bytes"] = bytes[2];
    // the autotuner's choices for the widths of this batch
    py::dict tuned;
    std::set<int> widths;
    for (size_t i = 0; i < count; ++i) widths.insert(int(ns.data()[i]));
    for (int n : widths) {
        int path = 0, fu = 0, tb = 0, rb = 0;
        if (dq_batch_tuned(n, dt, &path, &fu, &tb, &rb) != DQ_OK) continue;
        py::dict c;
        c["path"] = path == 1 ? "persistent" : "engine";
        c["fusion"] = fu;
        c["tile_bits"] = tb;
        c["reg_bits"] = rb;
        tuned[py::int_(n)] = c;
    }
    st["tuned"] = tuned;
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

}  // namespace
