#!/usr/bin/env python3
"""Benchmark of the statevector hot path on one MI355X.

Workload: the at-scale plain FP64 random circuit -- 30 qubits, complex128
(16 GiB state), 40 layers of 61 gates (a random gate from {h,s,t,sx,rx,ry,rz}
on every qubit, {cz,rzz,cp} on a random perfect matching, 8
remote_link_psi_plus + 8 nonlocal_cz across the two 15-qubit halves), with
2 measurements + 2 resets + 1 conditional x after every 10th layer.

One step = reset the resident device state to |0...0> and run the whole
circuit through the dispatcher (gates, mid-circuit measurements with host
ChaCha8 draws, resets, conditional); the circuit is decoded once, before
timing.  value = gates applied per second over the K timed steps (a
conditional gate counts when its condition held).

draw_margins: the distance of each measurement's ChaCha8 draw from its p1.
p1 is a reduction whose summation order differs from the reference's (whose
own order is not fixed), so an outcome could differ from the reference's only
for a draw within rounding of p1; near_ties counts draws within 1e-12.

roofline: the dominant kernel is the LDS-tiled sweep (k_tile_reg); its
launches are bracketed by HIP events on the engine's stream.  Algorithmic
bytes per launch = one read + one write of every amplitude the launch visits,
2 * 16 * (amplitudes visited) / launches: 2^n per sweep, except the first
sweeps after each reset, which visit only the tiles where the qubits no op has
touched yet are 0 (the other tiles hold zeros and stay zero).

cpu_baseline: the oracle's OpenMP restatement of the reference's Rayon
kernels (src/engine.rs) on the host cores, on a bounded sample: the circuit's
gates in program order from |0...0> until ~15 s have elapsed.

adjoint (second half of the metric): gradient of E = <psi|H|psi> for the
26-qubit, 40-layer ansatz (2080 trainable ry/rz parameters, rzz brickwork and
crx on every 4th pair fixed; H = sum Z_i Z_{i+1} + 0.7 sum X_i), one parameter
set drawn from the seed.  grads/s = full gradient vectors per second from the
adjoint sweep; parameter shift is timed on a stated subset of parameters on the
same GPU and extrapolated linearly to all of them; the subset also gives the
adjoint's error against parameter shift (FP64) and the complex64 gradient's
error against the FP64 one.

precision (complex64 form): the same circuit family at 32 qubits, depth 60,
blocks of 2 x 16 qubits, run as complex64 with fusion level 3 (runs of gates
on <= 5 qubits multiplied into 32x32 blocks in FP64 and applied on FP32 MFMA
inside the LDS-tiled sweep) and as complex128 (fusion level 2) on the same GPU
(32 GiB + 64 GiB).  Reported: gates/s of each, and on device max |d amp|,
1 - |<psi64|psi32>|^2 and the total variation distance of the probabilities.
--precision-modes f2,f3,f3all,f3nopair,f3nopair_all also times the complex64
step in other fusion modes (f2: VALU register ops only; f3: blocks that pass
the cost threshold; f3all: every block; nopair: without pair fusion), which
shows what moving gates from the VALU into MFMA blocks costs.

batch (batching form): 4096 hardware-efficient-ansatz circuits, n ~ U{10..24},
depth ~ U{8..64} (ry, rz on every qubit then a cx ring per layer), <sum_i
Z_i Z_{i+1}> of each, complex128 and complex64: circuits/s, the instances
each path took (LDS-resident, persistent, engine), the widths' tuned
configurations, and the state bytes the paths moved per second.

mps (matrix-product states): Mps::to_statevector of a random 30-site MPS with
bond <= 16 (upload, growth of both halves, tiled complex GEMM join writing the
16 GiB statevector) and the p1 of Mps::measure (the same join over half of
the index space, summed without writing); CPU baseline: the reference's
per-basis loop restated in C, single-threaded, on a bounded sample.

pblock (distributed circuits): PBlockSimulator.simulate of a 4-node x 7-qubit
layered circuit (u3 layers, cx chains inside nodes, remote_cx and nonlocal_cz
across nodes, measure + conditional every 5 layers); instructions/s, CPU
baseline the oracle's restatement on a bounded prefix.

shots (batched shots): simulate_shots of a 12-qubit random-block circuit with
mid-circuit measure / reset / conditional and a final measurement of every
qubit; shots/s on the GPU, and the oracle's sequential restatement on a
bounded sample as its CPU baseline.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "30q FP64 gates/s + % HBM peak; adjoint grads/s vs plain param-shift"
HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def traffic_from_profiles(n: int):
    """HBM bytes per tile-pass launch from the committed rocprofv3 counter
    summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel factor)."""
    path = os.path.join(ROOT, "profiles", f"tile_pass_{n}q_traffic.txt")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    return None


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(circ, n: int, budget_s: float) -> dict:
    import numpy as np

    from oracle import oracle as orc

    sim = orc.OracleSimulator(seed=42, parallel=True)
    state = np.zeros(1 << n, dtype=np.complex128)
    state[0] = 1.0
    insts = json.loads(circ.model_dump_json())["instructions"]
    cbits, rng, p1s = {}, orc.Rng(42), []
    done = 0
    t0 = time.perf_counter()
    for inst in insts:
        if inst["kind"] in ("measure", "reset", "conditional"):
            break
        sim._run(state, n, inst, cbits, rng, p1s)
        done += 1
        if time.perf_counter() - t0 >= budget_s:
            break
    dt = time.perf_counter() - t0
    return {"value": done / dt, "unit": "gates/s", "cores": int(orc.lib().or_num_threads()),
            "kind": "port", "cpu_model": cpu_model(),
            "sample": f"first {done} gates of the same {n}-qubit circuit from |0...0> "
                      f"(OpenMP restatement of src/engine.rs, {dt:.1f} s)"}


def adjoint_leg(n: int, layers: int, seed: int, ps_subset: int, sets: int = 64) -> dict:
    """Adjoint gradients of E(theta) = <psi(theta)|H|psi(theta)> for `sets`
    parameter vectors drawn U[0, 2 pi) from the seed, in complex128 and
    complex64; on the first set, parameter shift and central differences
    (h = 1e-6) on a random subset of parameters."""
    import numpy as np

    from dqsim_bosonic_amd import adjoint as A

    pc, H = A.hea_benchmark(n, layers, seed)
    thetas = np.random.default_rng(seed).uniform(0.0, 2 * np.pi, (sets, pc.num_params))
    est = A.Estimator(pc, H, dtype="c128")
    est.gradient(thetas[0])  # warm-up
    grads = np.empty_like(thetas)
    energies = np.empty(sets)
    t0 = time.perf_counter()
    for k in range(sets):
        energies[k], grads[k] = est.gradient(thetas[k])
    t_adj = (time.perf_counter() - t0) / sets
    theta, g = thetas[0], grads[0]
    idx = np.random.default_rng(seed + 1).choice(pc.num_params, size=ps_subset, replace=False)
    t0 = time.perf_counter()
    ps = est.parameter_shift(theta, indices=idx)
    t_ps_sub = time.perf_counter() - t0
    t_ps_full = t_ps_sub * pc.num_params / ps_subset
    fd = est.finite_difference(theta, indices=idx, h=1e-6)
    err_ps = float(np.max(np.abs(g[idx] - ps)))
    err_fd = float(np.max(np.abs(g[idx] - fd)))
    est32 = A.Estimator(pc, H, dtype="c64")
    est32.gradient(thetas[0])
    rel32 = []
    t0 = time.perf_counter()
    for k in range(sets):
        _, g32 = est32.gradient(thetas[k])
        rel32.append(float(np.linalg.norm(g32 - grads[k]) / np.linalg.norm(grads[k])))
    t_adj32 = (time.perf_counter() - t0) / sets
    return {"workload": f"hea {n}q {layers} layers, {pc.num_params} params, H = ZZ chain + 0.7 X sum, "
                        f"{sets} parameter sets U[0, 2pi) from seed {seed}",
            "parameter_sets": sets, "energy_set0": float(energies[0]),
            "grads_per_s": round(1.0 / t_adj, 4), "adjoint_s_per_gradient": round(t_adj, 4),
            "param_shift_grads_per_s": round(1.0 / t_ps_full, 6),
            "param_shift_s_extrapolated": round(t_ps_full, 1),
            "param_shift_sample": f"{ps_subset} of {pc.num_params} parameters of set 0 ({t_ps_sub:.2f} s), "
                                  "extrapolated linearly",
            "speedup_vs_param_shift": round(t_ps_full / t_adj, 1),
            "max_abs_err_vs_param_shift_f64": err_ps,
            "max_abs_err_vs_central_fd_f64": err_fd, "fd_step": 1e-6,
            "c64_grads_per_s": round(1.0 / t_adj32, 4),
            "c64_max_rel_err_vs_f64": max(rel32)}


def batch_leg(count: int, seed: int, workers: int) -> dict:
    """4096 HEA circuits (n ~ U{10..24}, d ~ U{8..64}) -> <sum Z_i Z_i+1> each.

    The warm-up runs one instance of every width, which lets the autotuner
    pick each width's path and sweep configuration (cached for the timed
    runs).  `hbm_GBps` is the algorithmic HBM bytes of the persistent and
    engine paths (one read and one write of a state per sweep, plus the
    initial write and the final read) over the wall time of the whole run;
    the LDS-resident instances move no state through HBM."""
    import numpy as np

    from dqsim_bosonic_amd import _core
    from dqsim_bosonic_amd import workloads as W

    ns, ds, angles = W.hea_batch_arrays(count, seed)
    flat = np.concatenate([a.ravel() for a in angles])
    gates = ds * 3 * ns
    out = {"workload": f"{count} HEA circuits, n ~ U{{10..24}}, depth ~ U{{8..64}}, seed {seed}; "
                       "<sum_i Z_i Z_i+1> each", "gates": int(np.sum(gates))}
    first = [int(np.flatnonzero(ns == n)[0]) for n in np.unique(ns)]
    res = {}
    for dt in ("c128", "c64"):
        _core.simulate_hea_batch(ns[first], ds[first], np.concatenate([angles[i].ravel() for i in first]),
                                 dtype=dt, workers=workers)  # warm-up and autotuning
        t0 = time.perf_counter()
        e, st = _core.simulate_hea_batch(ns, ds, flat, dtype=dt, workers=workers)
        el = time.perf_counter() - t0
        res[dt] = e
        hbm = float(st["persistent_bytes"]) + float(st["engine_bytes"])
        out[dt] = {"circuits_per_s": round(count / el, 2), "seconds": round(el, 3),
                   "gates_per_s": round(float(np.sum(gates)) / el, 1),
                   "hbm_bytes": hbm, "hbm_GBps": round(hbm / el / 1e9, 1),
                   "hbm_frac_of_peak": round(hbm / el / 1e9 / HBM_PEAK_GBS, 4),
                   "lds_instances": int(st["lds_instances"]),
                   "persistent_instances": int(st["persistent_instances"]),
                   "engine_instances": int(st["engine_instances"]),
                   "graph_gpu_ms": round(float(st["graph_gpu_ms"]), 3),
                   "engine_wall_s": round(float(st["engine_wall_s"]), 3),
                   "tuned": {int(k): v for k, v in st["tuned"].items()}}
        log(f"[bench] batch {dt}: {el:.3f} s {out[dt]}")
    out["c64_max_abs_diff_vs_c128"] = float(np.max(np.abs(res["c64"] - res["c128"])))
    return out


def shots_leg(shots: int, seed: int, cpu_seconds: float) -> dict:
    """simulate_shots of a 12-qubit random-block circuit with mid-circuit
    measure / reset / conditional and a final measurement of every qubit."""
    from dqsim_bosonic_amd import StatevectorSimulator
    from dqsim_bosonic_amd import workloads as W

    circ = W.shots_blocks(12, 20, seed)
    sim = StatevectorSimulator(seed=seed)
    sim.simulate_shots(circ, shots=256)  # warm-up
    t0 = time.perf_counter()
    counts = sim.simulate_shots(circ, shots=shots)
    el = time.perf_counter() - t0
    out = {"workload": f"shots_blocks 12q depth 20 ({W.gate_count(circ)} gates, "
                       f"{sum(1 for i in circ.instructions if i.kind == 'measure')} measurements per shot)",
           "shots": shots, "shots_per_s": round(shots / el, 1), "seconds": round(el, 3),
           "distinct_outcomes": len(counts)}
    if cpu_seconds > 0:
        from oracle import oracle as orc

        text = circ.model_dump_json()
        done, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < cpu_seconds:
            orc.simulate_shots_json(text, 8, seed + done)
            done += 8
        cpu_el = time.perf_counter() - t0
        out["cpu_baseline"] = {"shots_per_s": round(done / cpu_el, 2), "cores": 1, "kind": "port",
                               "sample": f"{done} shots through the oracle's sequential restatement "
                                         f"({cpu_el:.1f} s)"}
    return out


def c64_mode_engine(n: int, mode: str):
    """A complex64 engine in one of the fusion modes of --precision-modes."""
    from dqsim_bosonic_amd import _core

    if mode not in ("f2", "f3", "f3all", "f3nopair", "f3nopair_all"):
        raise ValueError(f"unknown complex64 fusion mode {mode!r}")
    e = _core.Engine(n, 2 if mode == "f2" else 3, 0, 0, 1)
    if mode.endswith("all"):
        e.set_fuse_min_cost(0.0)
    if "nopair" in mode:
        e.set_pair_fusion(False)
    return e


def c64_modes(n: int, text: str, gates: int, seed: int, modes: list[str]) -> dict:
    out = {}
    for m in modes:
        e = c64_mode_engine(n, m)
        e.load(text)
        e.step(seed)
        e.set_timing(True)
        f0, s0 = e.fused_stats(), e.stats()[0]
        t0 = time.perf_counter()
        e.step(seed)
        dt = time.perf_counter() - t0
        ms, cnt = e.timing()
        f1, s1 = e.fused_stats(), e.stats()[0]
        out[m] = {"s_per_step": round(dt, 3), "gates_per_s": round(gates / dt, 2), "sweeps": s1 - s0,
                  "avg_sweep_ms": round(ms / max(cnt, 1), 3), "blocks": f1[0] - f0[0],
                  "gates_in_blocks": f1[1] - f0[1]}
        log(f"[bench] precision mode {m}: {out[m]}")
        del e
    return out


def precision_leg(n: int, depth: int, seed: int, steps: int, modes: list[str] | None = None) -> dict:
    from dqsim_bosonic_amd import _core
    from dqsim_bosonic_amd import workloads as W

    circ = W.random_blocks(n, depth, seed=seed)
    gates = W.gate_count(circ)
    text = circ.model_dump_json()
    out = {"workload": f"random_blocks {n}q depth {depth} ({gates} gates), blocks 2x{n // 2}"}
    if modes:
        out["c64_modes"] = c64_modes(n, text, gates, seed, modes)
    e32 = _core.Engine(n, 3, 0, 0, 1)
    e32.load(text)
    rec32 = e32.step(seed)  # warm-up
    e32.set_timing(True)
    f0 = e32.fused_stats()
    s0 = e32.stats()[0]
    t0 = time.perf_counter()
    for _ in range(steps):
        rec32 = e32.step(seed)
    dt32 = (time.perf_counter() - t0) / steps
    ms32, cnt32 = e32.timing()
    swept32 = e32.timed_amplitudes()
    f1 = e32.fused_stats()
    s1 = e32.stats()[0]
    e32.set_timing(False)
    log(f"[bench] precision c64: {dt32:.3f} s/step")
    e64 = _core.Engine(n, 2, 0, 0, 0)
    e64.load(text)
    t0 = time.perf_counter()
    rec64 = e64.step(seed)
    dt64 = time.perf_counter() - t0
    log(f"[bench] precision c128: {dt64:.3f} s/step")
    cmp = e32.compare(e64)
    avg32 = ms32 / max(cnt32, 1)
    blocks = (f1[0] - f0[0]) / steps
    alg_bytes = 2.0 * 8.0 * swept32 / max(cnt32, 1)
    # complex 32x32 GEMM per 32-amplitude group: 32*32 complex MACs = 8192 flops
    gemm_flops_per_sweep = blocks / max((s1 - s0) / steps, 1) * 8192.0 * float(1 << (n - 5))
    out.update({
        "c64_gates_per_s": round(gates / dt32, 2), "c128_gates_per_s": round(gates / dt64, 2),
        "c64_s_per_step": round(dt32, 3), "c128_s_per_step": round(dt64, 3),
        "measurement_records_equal": rec32 == rec64,
        "max_abs_err": cmp["max_abs"], "fidelity_deficit": 1.0 - cmp["fidelity"], "tvd": cmp["tvd"],
        "fused_blocks_per_step": blocks, "fused_gates_per_step": (f1[1] - f0[1]) / steps,
        "sweeps_per_step": (s1 - s0) / steps,
        "c64_avg_sweep_ms": round(avg32, 4),
        "c64_sweep_hbm_frac": round(alg_bytes / (avg32 / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
        "c64_sweep_mfma_frac": round(gemm_flops_per_sweep / (avg32 / 1e3) / 157.3e12, 4),
    })
    return out


def mps_leg(n: int, chi: int, seed: int, reps: int, cpu_seconds: float) -> dict:
    """Matrix-product-state contraction (Mps::to_statevector and the p1 of
    Mps::measure) of a random n-site MPS with bond min(2^q, 2^(n-q), chi)."""
    from dqsim_bosonic_amd import capi
    from dqsim_bosonic_amd import workloads as W

    ts = W.random_mps(n, chi, seed)
    m = capi.Mps(n)
    st = capi.State(n)
    m.load(ts)
    m.to_state(st)  # warm-up
    t0 = time.perf_counter()
    for _ in range(reps):
        m.load(ts)  # a fresh upload re-grows both halves: the whole call
        m.to_state(st)
    to_state_s = (time.perf_counter() - t0) / reps
    qs = [0, n // 2, n - 1]
    t0 = time.perf_counter()
    for q in qs:
        m.load(ts)
        m.prob_one(q)
    p1_s = (time.perf_counter() - t0) / len(qs)
    amps = float(1 << n)
    out = {"workload": f"random MPS {n} sites, bond <= {chi}, cut at bond {m.cut()}",
           "to_statevector_ms": round(1e3 * to_state_s, 3),
           "amplitudes_per_s": round(amps / to_state_s, 1),
           "write_GBps": round(16.0 * amps / to_state_s / 1e9, 1),
           "fp64_TFLOPs": round(8.0 * chi * amps / to_state_s / 1e12, 2),
           "measure_p1_ms": round(1e3 * p1_s, 3)}
    del st, m
    if cpu_seconds > 0:
        from oracle import oracle as orc

        done, t0, chunk = 0, time.perf_counter(), 1 << 12
        while time.perf_counter() - t0 < cpu_seconds:
            orc.mps_amplitudes_c(ts, done, chunk)
            done += chunk
        cpu_el = time.perf_counter() - t0
        out["cpu_baseline"] = {"amplitudes_per_s": round(done / cpu_el, 1), "cores": 1, "kind": "port",
                               "sample": f"first {done} basis states through the C restatement of "
                                         f"Mps::to_statevector's loop ({cpu_el:.1f} s)"}
    return out


def pblock_leg(nodes: int, per_node: int, depth: int, seed: int, cpu_seconds: float) -> dict:
    """PBlockSimulator.simulate of a distributed layered circuit: blocks per
    node merged on demand on the device (dq_state_tensor), epochs as the
    reference forms them."""
    from dqsim_bosonic_amd import PBlockSimulator
    from dqsim_bosonic_amd import workloads as W

    circ, dist = W.distributed_layers(nodes, per_node, depth, seed)
    _, small = W.distributed_layers(2, 3, 2, seed)
    PBlockSimulator(seed=seed).simulate(small)  # warm-up
    t0 = time.perf_counter()
    res = PBlockSimulator(seed=seed).simulate(dist)
    el = time.perf_counter() - t0
    count = len(circ.instructions)
    out = {"workload": f"distributed_layers {nodes} nodes x {per_node} qubits, depth {depth} "
                       f"({count} instructions, merged to {res.num_qubits} qubits)",
           "seconds": round(el, 4), "instructions_per_s": round(count / el, 1)}
    if cpu_seconds > 0:
        from oracle import oracle as orc

        # the oracle's 28-qubit merges alone take tens of seconds (numpy index
        # arithmetic over 2^28 entries), so its sample is the same family at
        # 4 x 5 qubits, timed on a bounded prefix, and the GPU is timed on it too
        c20, d20 = W.distributed_layers(nodes, 5, depth, seed)
        t0 = time.perf_counter()
        PBlockSimulator(seed=seed).simulate(d20)
        gpu20 = time.perf_counter() - t0
        t0 = time.perf_counter()
        _, _, done = orc._pblock_run(d20, seed, budget_s=cpu_seconds)
        cpu_el = time.perf_counter() - t0
        out["gpu_20q_instructions_per_s"] = round(len(c20.instructions) / gpu20, 1)
        out["cpu_baseline"] = {"instructions_per_s": round(done / cpu_el, 2), "cores": 1, "kind": "port",
                               "sample": f"distributed_layers {nodes} x 5 qubits (20 qubits): first {done} of "
                                         f"{len(c20.instructions)} instructions through the oracle's restatement "
                                         f"of PBlockSimulator::simulate ({cpu_el:.1f} s); compare "
                                         f"gpu_20q_instructions_per_s"}
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--qubits", type=int, default=30)
    ap.add_argument("--depth", type=int, default=40)
    ap.add_argument("--seed", type=int, default=42)
    ap.add_argument("--fusion", type=int, default=2)
    ap.add_argument("--tile-bits", type=int, default=0)
    ap.add_argument("--low-bits", type=int, default=0)
    ap.add_argument("--reg-bits", type=int, default=4, help="register bits of the sweep layouts (3 or 4)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--adjoint-qubits", type=int, default=26)
    ap.add_argument("--adjoint-layers", type=int, default=40)
    ap.add_argument("--ps-subset", type=int, default=8)
    ap.add_argument("--adjoint-sets", type=int, default=64)
    ap.add_argument("--no-precision", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-shots", action="store_true")
    ap.add_argument("--no-mps", action="store_true")
    ap.add_argument("--no-pblock", action="store_true")
    ap.add_argument("--mps-qubits", type=int, default=30)
    ap.add_argument("--mps-bond", type=int, default=16)
    ap.add_argument("--shots", type=int, default=100000)
    ap.add_argument("--batch-count", type=int, default=4096)
    ap.add_argument("--batch-workers", type=int, default=8)
    ap.add_argument("--precision-qubits", type=int, default=32)
    ap.add_argument("--precision-depth", type=int, default=60)
    ap.add_argument("--precision-modes", default="",
                    help="comma list of complex64 fusion modes to time as well: f2,f3,f3all,f3nopair,f3nopair_all")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the engine runs a state on one GPU; sharding across GPUs is not implemented)", file=sys.stderr)
        return 2

    import dqsim_bosonic_amd as dq  # noqa: F401  (loads the HIP engine)
    from dqsim_bosonic_amd import _core
    from dqsim_bosonic_amd import workloads as W

    n = args.qubits
    circ = W.random_blocks(n, args.depth, seed=args.seed)
    gates = W.gate_count(circ)
    text = circ.model_dump_json()
    log(f"[bench] workload random_blocks n={n} depth={args.depth} gates={gates}")

    eng = _core.Engine(n, args.fusion, args.tile_bits, args.low_bits)
    eng.set_reg_bits(args.reg_bits)
    eng.load(text)
    for w in range(args.warmup):
        t = time.perf_counter()
        rec = eng.step(args.seed)
        log(f"[bench] warmup {w}: {time.perf_counter() - t:.3f} s record={rec}")

    eng.set_timing(True)
    sweeps0, launches0, _ = eng.stats()
    t0 = time.perf_counter()
    records = []
    for k in range(args.steps):
        records.append(eng.step(args.seed))
        log(f"[bench] step {k} done")
    t1 = time.perf_counter()
    kernel_ms, timed = eng.timing()
    swept = eng.timed_amplitudes()
    sweeps1, launches1, _ = eng.stats()
    eng.set_timing(False)
    assert all(r == records[0] for r in records), "measurement record changed between steps"
    # Distance of each measurement's draw from its p1 (one ChaCha8 draw per
    # measure and reset, in program order: (next_u64 >> 11) 2^-53).  p1 is a
    # sum whose order differs from the reference's (its own order is not
    # fixed either); an outcome could differ only for a draw within rounding
    # of p1, so the record is exact when no draw lies within 1e-12 of it.
    p1s = list(eng.p1_record())
    draws = [(u >> 11) * 2.0 ** -53 for u in _core._rng_stream(args.seed, len(p1s))]
    gaps = [abs(r - p) for r, p in zip(draws, p1s)]
    margins = {"measurements_per_step": len(p1s), "min_abs_draw_minus_p1": min(gaps) if gaps else None,
               "near_ties_below_1e-12": sum(1 for g in gaps if g < 1e-12)}

    elapsed = t1 - t0
    ms_per_step = 1000.0 * elapsed / args.steps
    # gates applied per step: a conditional's gate counts only when its
    # condition held (the record is the same in every step)
    skipped = 0
    for ins in circ.instructions:
        if ins.kind != "conditional":
            continue
        c = ins.model_dump()["condition"]
        val = sum(int(records[0].get(c["creg_base"] + b, 0)) << b for b in range(c["creg_size"]))
        if val != c["creg_value"]:
            skipped += 1
    applied = gates - skipped
    value = applied * args.steps / elapsed
    sweeps_per_step = (sweeps1 - sweeps0) / args.steps
    avg_launch_ms = kernel_ms / max(timed, 1)
    # one read and one write of each amplitude a sweep visits: 2^n per sweep,
    # less in the first sweeps after the reset, which visit only the tiles
    # where qubits no op has touched yet are 0
    alg_bytes = 2.0 * 16.0 * swept / max(timed, 1)
    achieved = alg_bytes / (avg_launch_ms / 1e3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_from_profiles(n),
                "traffic_source": f"profiles/tile_pass_{n}q_traffic.txt: rocprofv3 FETCH_SIZE + WRITE_SIZE of this "
                                  "kernel from a separate counter run (not measured in this run)",
                "kernel": f"k_tile_reg{'8' if args.reg_bits == 3 else ''}<double,{args.tile_bits or 12}>",
                "avg_launch_ms": round(avg_launch_ms, 4),
                "launches_timed": int(timed), "sweeps_per_step": sweeps_per_step,
                "alg_bytes_per_launch": round(alg_bytes), "amplitudes_per_launch_frac_of_state":
                    round(swept / max(timed, 1) / float(1 << n), 4),
                "kernel_time_frac_of_step": round(kernel_ms / 1e3 / elapsed, 4)}

    adj = None
    if not args.no_adjoint:
        log("[bench] adjoint gradient ...")
        del eng  # free the 30-qubit state first
        adj = adjoint_leg(args.adjoint_qubits, args.adjoint_layers, args.seed, args.ps_subset, args.adjoint_sets)
        log(f"[bench] adjoint {adj}")

    prec = None
    if not args.no_precision:
        log("[bench] precision (complex64 + MFMA blocks vs complex128) ...")
        modes = [m for m in args.precision_modes.split(",") if m]
        prec = precision_leg(args.precision_qubits, args.precision_depth, args.seed, 1, modes)
        log(f"[bench] precision {prec}")

    batch = None
    if not args.no_batch:
        log("[bench] batch (HEA circuits, <ZZ>) ...")
        batch = batch_leg(args.batch_count, args.seed, args.batch_workers)
        log(f"[bench] batch {batch}")

    shots = None
    if not args.no_shots:
        log("[bench] shots ...")
        shots = shots_leg(args.shots, args.seed, 0.0 if args.no_cpu_baseline else 5.0)
        log(f"[bench] shots {shots}")

    mps = None
    if not args.no_mps:
        log("[bench] mps contraction ...")
        mps = mps_leg(args.mps_qubits, args.mps_bond, args.seed, 5, 0.0 if args.no_cpu_baseline else 5.0)
        log(f"[bench] mps {mps}")

    pblock = None
    if not args.no_pblock:
        log("[bench] pblock ...")
        pblock = pblock_leg(4, 7, 20, args.seed, 0.0 if args.no_cpu_baseline else 5.0)
        log(f"[bench] pblock {pblock}")

    cpu = None
    if not args.no_cpu_baseline:
        log("[bench] cpu baseline ...")
        cpu = cpu_baseline(circ, n, args.cpu_seconds)

    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "gates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"random_blocks {n}q complex128 depth {args.depth} "
                               "(61 gates/layer; measure+reset+conditional every 10 layers)",
                   "qubits": n, "depth": args.depth, "gates": gates, "gates_applied_per_step": applied,
                   "seed": args.seed,
                   "fusion": args.fusion, "reg_bits": args.reg_bits,
                   "measurement_record": records[0]},
        "draw_margins": margins,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "adjoint": adj,
        "precision": prec,
        "batch": batch,
        "shots": shots,
        "mps": mps,
        "pblock": pblock,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
