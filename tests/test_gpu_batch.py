"""GPU tests of the batching form (dq_batch_* / simulate_batch).

Oracle: each instance's statevector from the oracle's restatement of
src/engine.rs (gate by gate from |0...0>), and <H> = sum_x |psi_x|^2 h(x)
computed from it in numpy; the open ZZ chain is also cross-checked against
probabilities() of the per-circuit simulator (a closed form over the
distribution).  complex128 to 1e-12, complex64 to 1e-5 (FP32 rounding over
a few hundred gates).  Instances cover the three device paths: LDS-resident
(n <= 13 complex128, n <= 14 complex64), the persistent launch (up to 20
qubits) and the engine workers, the last two as the autotuner chooses, or
forced with autotuning off.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

capi = pytest.importorskip("dqsim_bosonic_amd.capi")
dq = pytest.importorskip("dqsim_bosonic_amd")
from dqsim_bosonic_amd import workloads as W  # noqa: E402

GATES_1Q = ["h", "s", "t", "sx", "u3", "rx", "ry", "rz", "x", "y", "z"]
GATES_NQ = ["cx", "cz", "cp", "rzz", "rxx", "swap", "cry", "ccx", "cswap", "rccx", "c3x", "c4x", "cu3"]
NPARAM = {"u3": 3, "rx": 1, "ry": 1, "rz": 1, "cp": 1, "rzz": 1, "rxx": 1, "cry": 1, "cu3": 3}


def zz_chain(n):
    return [(1.0, 3 << i) for i in range(n - 1)]


def expect_z(psi, terms):
    idx = np.arange(psi.size, dtype=np.uint64)
    p = np.abs(psi) ** 2
    tot = 0.0
    for c, zm in terms:
        par = np.zeros(psi.size, dtype=np.uint64)
        m = np.uint64(zm)
        x = idx & m
        while np.any(x):
            par ^= x & np.uint64(1)
            x >>= np.uint64(1)
        tot += c * float(np.sum(p * (1.0 - 2.0 * par)))
    return tot


def random_gates(n, count, rng):
    out = []
    for _ in range(count):
        if rng.random() < 0.5:
            g = GATES_1Q[int(rng.integers(len(GATES_1Q)))]
            out.append((orc.gate_matrix(g, rng.uniform(-4, 4, NPARAM.get(g, 0))), [int(rng.integers(n))]))
        else:
            g = GATES_NQ[int(rng.integers(len(GATES_NQ)))]
            u = orc.gate_matrix(g, rng.uniform(-4, 4, NPARAM.get(g, 0)))
            k = int(np.log2(u.shape[0]))
            if k > n:
                continue
            out.append((u, [int(x) for x in rng.choice(n, size=k, replace=False)]))
    return out


def oracle_state(n, gates):
    psi = np.zeros(1 << n, dtype=complex)
    psi[0] = 1.0
    for u, qs in gates:
        if len(qs) == 1:
            orc.apply_one_qubit(psi, n, u, qs[0])
        else:
            orc.apply_n_qubit(psi, n, u, qs)
    return psi


@pytest.mark.parametrize("dtype,tol", [(capi.C128, 1e-12), (capi.C64, 1e-5)])
def test_random_instances_both_paths(dtype, tol):
    rng = np.random.default_rng(5)
    widths = [1, 2, 5, 10, 12, 13, 14, 15, 16, 11, 14, 9]
    b = capi.Batch(dtype)
    b.set_workers(3)
    want = []
    for i, n in enumerate(widths):
        gates = random_gates(n, 120, rng)
        idx = b.add(n)
        for u, qs in gates:
            b.apply(idx, u, qs)
        psi = oracle_state(n, gates)
        if i % 3 == 2:  # a custom observable on some instances
            terms = [(float(rng.normal()), int(rng.integers(1, 1 << n))) for _ in range(5)]
            b.set_observable(idx, terms)
        else:
            terms = zz_chain(n)
        want.append(expect_z(psi, terms))
    got = b.run()
    lds, eng, _ = b.stats()
    lim = 13 if dtype == capi.C128 else 14
    ps = b.path_stats()
    assert lds == sum(1 for n in widths if n <= lim) and eng == ps["engine"]
    assert ps["persistent"] + ps["engine"] == len(widths) - lds
    assert np.max(np.abs(got - np.array(want))) <= tol, np.abs(got - np.array(want))


def test_permutation_runs_and_big_gates():
    """Long runs of cx / swap / x (composed into one permutation sweep)
    between superposing layers, and 4-5 qubit dense and monomial gates."""
    rng = np.random.default_rng(8)
    for n in (10, 13):
        gates = [(orc.gate_matrix("h", []), [q]) for q in range(n)]
        for _ in range(4):
            for _ in range(25):
                g = ["cx", "swap", "x", "ccx"][int(rng.integers(4))]
                u = orc.gate_matrix(g, [])
                k = int(np.log2(u.shape[0]))
                gates.append((u, [int(x) for x in rng.choice(n, size=k, replace=False)]))
            for q in range(n):
                gates.append((orc.gate_matrix("u3", rng.uniform(-4, 4, 3)), [q]))
            a = rng.normal(size=(16, 16)) + 1j * rng.normal(size=(16, 16))
            q4, _ = np.linalg.qr(a)
            gates.append((q4, [int(x) for x in rng.choice(n, size=4, replace=False)]))
            gates.append((orc.gate_matrix("c4x", []), [int(x) for x in rng.choice(n, size=5, replace=False)]))
        b = capi.Batch(capi.C128)
        idx = b.add(n)
        for u, qs in gates:
            b.apply(idx, u, qs)
        terms = [(0.5, 0b1011), (-1.25, (1 << n) - 1), (2.0, 1 << (n - 1))]
        b.set_observable(idx, terms)
        got = b.run()[0]
        assert abs(got - expect_z(oracle_state(n, gates), terms)) <= 1e-12


def test_simulate_batch_hea_against_probabilities():
    circuits = W.hea_batch(24, seed=3, n_range=(10, 15), d_range=(2, 6))
    got = dq.simulate_batch(circuits)
    for c, e in zip(circuits, got):
        res = dq.StatevectorSimulator(seed=1).simulate(c)
        n = int(np.log2(len(res.statevector)))
        # {index: p} over all qubits; the open chain is symmetric under the
        # reversal of the index bits, so either bit order gives the same value
        pr = res.probabilities()
        idx = np.array(list(pr.keys()), dtype=np.int64)
        p = np.array(list(pr.values()))
        zz = sum(1.0 - 2.0 * (((idx >> i) ^ (idx >> (i + 1))) & 1) for i in range(n - 1))
        assert abs(e - float(np.dot(p, zz))) <= 1e-11
    got64 = dq.simulate_batch(circuits, dtype="c64")
    assert np.max(np.abs(got64 - got)) <= 1e-5


def test_simulate_batch_rejects_measurement_and_bad_args():
    c = W.random_blocks(6, 10, seed=1)  # has measure / reset / conditional
    with pytest.raises(ValueError):
        dq.simulate_batch([c])
    with pytest.raises(ValueError):
        dq.simulate_batch([], dtype="f16")
    b = capi.Batch()
    i = b.add(4)
    with pytest.raises(capi.DqError):
        b.apply(i, np.eye(2), [4])
    with pytest.raises(capi.DqError):
        b.set_observable(i, [(1.0, 1 << 5)])
    with pytest.raises(capi.DqError):
        b.add(0)
    assert dq.simulate_batch([]).shape == (0,)


def test_hea_arrays_match_circuit_objects():
    from dqsim_bosonic_amd import _core

    ns, ds, angles = W.hea_batch_arrays(40, seed=9, n_range=(10, 16), d_range=(1, 5))
    circuits = W.hea_batch(40, seed=9, n_range=(10, 16), d_range=(1, 5))
    flat = np.concatenate([a.ravel() for a in angles])
    for dtype in ("c128", "c64"):
        got, st = _core.simulate_hea_batch(ns, ds, flat, dtype=dtype, workers=4)
        want = dq.simulate_batch(circuits, dtype=dtype)
        assert np.max(np.abs(got - want)) <= 1e-14
        lim = 13 if dtype == "c128" else 14
        assert st["lds_instances"] == int(np.sum(ns <= lim))


def hea_gates(n, d, angles):
    ry = lambda t: orc.gate_matrix("ry", [t])  # noqa: E731
    rz = lambda t: orc.gate_matrix("rz", [t])  # noqa: E731
    cx = orc.gate_matrix("cx", [])
    out = []
    for layer in range(d):
        for q in range(n):
            out.append((ry(angles[layer][q][0]), [q]))
        for q in range(n):
            out.append((rz(angles[layer][q][1]), [q]))
        for q in range(n):
            out.append((cx, [q, (q + 1) % n]))
    return out


@pytest.mark.parametrize("dtype,tol", [(capi.C128, 1e-12), (capi.C64, 1e-5)])
def test_persistent_path_random_instances(dtype, tol):
    """Autotuning off: every instance of 14..20 qubits takes the persistent
    launch (one workgroup per instance, tile passes from the engine's
    planner, all in one captured graph).  Random gates of 1 to 5 qubits."""
    rng = np.random.default_rng(21)
    widths = [14, 15, 16, 17, 14, 15, 16, 14, 18, 15, 14, 16]
    b = capi.Batch(dtype)
    b.set_autotune(False)
    insts = []
    for n in widths:
        gates = random_gates(n, 60, rng)
        i = b.add(n)
        for u, qs in gates:
            b.apply(i, u, qs)
        terms = zz_chain(n) if n % 2 else [(0.7, 1 << 3), (-1.3, (1 << 2) | (1 << (n - 1)))]
        if n % 2 == 0:
            b.set_observable(i, terms)
        insts.append((n, gates, terms))
    got = b.run()
    st = b.path_stats()
    lim = 13 if dtype == capi.C128 else 14  # complex64 14-qubit states still fit LDS
    assert st["persistent"] == sum(1 for n in widths if n > lim) and st["engine"] == 0
    assert st["lds"] == sum(1 for n in widths if n <= lim)
    assert st["bytes"][1] > 0 and st["graph_ms"] > 0
    for e, (n, gates, terms) in zip(got, insts):
        assert abs(e - expect_z(oracle_state(n, gates), terms)) <= tol * max(1.0, abs(e))


def test_persistent_path_288_mixed_hea_instances():
    """288 hardware-efficient-ansatz instances of 10..17 qubits and 2..6
    layers (LDS-resident up to 13 qubits, persistent above) in one run:
    every <ZZ chain> against the oracle's statevector, 1e-12."""
    ns, ds, angles = W.hea_batch_arrays(288, seed=8, n_range=(10, 17), d_range=(2, 6))
    b = capi.Batch(capi.C128)
    b.set_autotune(False)
    ref = []
    for n, d, a in zip(ns, ds, angles):
        n, d = int(n), int(d)
        gates = hea_gates(n, d, a)
        i = b.add(n)
        for u, qs in gates:
            b.apply(i, u, qs)
        ref.append((n, gates))
    got = b.run()
    st = b.path_stats()
    assert st["persistent"] == int(np.sum(ns >= 14)) and st["lds"] == int(np.sum(ns <= 13))
    err = max(abs(e - expect_z(oracle_state(n, g), zz_chain(n))) for e, (n, g) in zip(got, ref))
    assert err <= 1e-12, err


def test_autotuner_choice_and_results():
    """With autotuning on, each width above the LDS limit gets a cached
    choice (path, fusion level, tile qubits, register bits); the results
    are the same whichever path it picks."""
    rng = np.random.default_rng(3)
    widths = [14, 15, 21]
    b = capi.Batch(capi.C128)
    insts = []
    for n in widths:
        gates = random_gates(n, 40, rng)
        i = b.add(n)
        for u, qs in gates:
            b.apply(i, u, qs)
        insts.append((n, gates))
    got = b.run()
    for n in widths:
        t = capi.batch_tuned(n, capi.C128)
        assert t is not None and t["path"] in ("persistent", "engine")
        assert t["tile_bits"] in (12, 13) and t["reg_bits"] in (3, 4)
    assert capi.batch_tuned(21, capi.C128)["path"] == "engine"
    st = b.path_stats()
    assert st["persistent"] + st["engine"] == 3
    for e, (n, gates) in zip(got, insts):
        assert abs(e - expect_z(oracle_state(n, gates), zz_chain(n))) <= 1e-12 * max(1.0, abs(e))


def test_batch_workload_4096_at_stated_size():
    """The batching workload at its stated size: the 4096 HEA circuits (n ~
    U{10..24}, depth ~ U{8..64}, seed 42) through the autotuned paths (one
    captured graph for the LDS-resident and persistent instances, engine
    workers for the rest), complex128 and complex64.  Every complex64 <ZZ>
    within 1e-5 of the complex128 one; complex128 against the oracle's
    statevector to 1e-12 for the shallowest instance of each width 10..16."""
    from dqsim_bosonic_amd import _core

    ns, ds, angles = W.hea_batch_arrays(4096, 42)
    flat = np.concatenate([a.ravel() for a in angles])
    e128, st = _core.simulate_hea_batch(ns, ds, flat, dtype="c128", workers=8)
    e64, _ = _core.simulate_hea_batch(ns, ds, flat, dtype="c64", workers=8)
    assert int(st["lds_instances"]) + int(st["persistent_instances"]) + int(st["engine_instances"]) == 4096
    assert int(st["engine_instances"]) > 0 and int(st["persistent_instances"]) > 0
    assert np.max(np.abs(np.asarray(e64) - np.asarray(e128))) <= 1e-5
    for n in range(10, 17):
        idx = np.flatnonzero(ns == n)
        i = int(idx[np.argmin(ds[idx])])
        want = expect_z(oracle_state(n, hea_gates(n, int(ds[i]), angles[i])), zz_chain(n))
        assert abs(e128[i] - want) <= 1e-12 * max(1.0, abs(want)), (n, i)
