"""GPU tests of the matrix-product-state contraction (dq_mps_*) and of
MpsSimulator / simulate_monolithic(mode="mps") on top of it.

Oracle: ``mps_contract`` (pinned to the literal loop of
src/monolithic/mps/simulator.rs:246-264 in test_mps_host.py), the
transfer-matrix norm / p1 and single-amplitude chain products for sizes the
dense oracle does not reach, the oracle statevector simulator for exact
(untruncated) MPS runs, and the oracle MPS (``OracleMps``) for truncated
runs without degenerate cuts.  Tolerances: 1e-12 relative for contractions
(FP64, a different association than the reference's left-to-right chain),
1e-8 absolute for simulator states (the reference's own TOL).
"""
from __future__ import annotations

import json
import math

import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

capi = pytest.importorskip("dqsim_bosonic_amd.capi")
dq = pytest.importorskip("dqsim_bosonic_amd")
from dqsim_bosonic_amd import circuit as C  # noqa: E402
from dqsim_bosonic_amd import workloads as W  # noqa: E402


def rand_tensors(rng, bonds):
    return [rng.normal(size=(bonds[q], 2, bonds[q + 1])) + 1j * rng.normal(size=(bonds[q], 2, bonds[q + 1]))
            for q in range(len(bonds) - 1)]


def chain_bonds(n, chi, rng=None):
    b = [min(1 << min(q, n - q), chi) for q in range(n + 1)]
    if rng is not None:  # ragged: shrink some bonds at random
        b = [1] + [max(1, int(x) - int(rng.integers(0, max(1, x // 2 + 1)))) for x in b[1:-1]] + [1]
    return b


@pytest.mark.parametrize("n,chi", [(1, 1), (2, 2), (3, 2), (5, 3), (7, 4), (11, 5), (12, 8), (13, 6), (14, 4),
                                   (16, 7), (18, 17)])
def test_contraction_every_cut(n, chi):
    rng = np.random.default_rng(n * 100 + chi)
    ts = rand_tensors(rng, chain_bonds(n, chi, rng))
    want = orc.mps_contract(ts)
    scale = np.max(np.abs(want))
    idx = np.arange(1 << n)
    m = capi.Mps(n)
    m.load(ts)
    st = capi.State(n)
    cuts = range(n + 1) if n <= 13 else (0, 1, n // 2, n - 7, n)
    for cut in cuts:
        m.set_cut(cut)
        assert m.cut() == cut
        m.to_state(st)
        got = st.get()
        assert np.max(np.abs(got - want)) <= 1e-12 * scale, cut
        for q in {0, n // 2, n - 1}:
            p1 = float(np.sum(np.abs(want[(idx >> q) & 1 == 1]) ** 2))
            assert abs(m.prob_one(q) - p1) <= 1e-12 * float(np.sum(np.abs(want) ** 2)), (cut, q)
    m.set_cut(-1)
    m.to_state(st)
    assert np.max(np.abs(st.get() - want)) <= 1e-12 * scale


def test_prob_one_every_qubit_tiled_both_halves():
    n = 16
    rng = np.random.default_rng(3)
    ts = rand_tensors(rng, chain_bonds(n, 9))
    want = orc.mps_contract(ts)
    idx = np.arange(1 << n)
    tot = float(np.sum(np.abs(want) ** 2))
    m = capi.Mps(n)
    m.load(ts)
    for cut in (7, 8, 9):  # >= 64 rows and columns on either side: tiled join
        m.set_cut(cut)
        for q in range(n):
            p1 = float(np.sum(np.abs(want[(idx >> q) & 1 == 1]) ** 2))
            assert abs(m.prob_one(q) - p1) <= 1e-12 * tot, (cut, q)


def test_contraction_into_complex64_state():
    n = 15
    rng = np.random.default_rng(5)
    ts = W.random_mps(n, 8, 5)
    want = orc.mps_contract(ts)
    st = capi.State(n, capi.C64)
    m = capi.Mps(n)
    m.load(ts)
    m.to_state(st)
    assert np.max(np.abs(st.get() - want)) <= 2e-7
    del rng


def test_large_contraction_against_closed_forms():
    """26 sites, bond 16: the norm and p1 of several qubits against the
    transfer-matrix contraction, 256 random amplitudes against chain products."""
    n = 26
    ts = W.random_mps(n, 16, 26)
    m = capi.Mps(n)
    m.load(ts)
    st = capi.State(n)
    m.to_state(st)
    psi = st.get()
    assert abs(float(np.sum(np.abs(psi) ** 2)) - orc.mps_norm2_transfer(ts)) <= 1e-11
    rng = np.random.default_rng(0)
    xs = [int(x) for x in rng.integers(0, 1 << n, size=256)] + [0, (1 << n) - 1]
    for x in xs:
        a = orc.mps_amplitude(ts, x)
        assert abs(psi[x] - a) <= 1e-12 * max(1e-6, abs(a)) + 1e-15
    for q in (0, 5, 12, 13, 20, 25):
        assert abs(m.prob_one(q) - orc.mps_norm2_transfer(ts, q)) <= 1e-11


def test_load_and_argument_errors():
    m = capi.Mps(3)
    with pytest.raises(capi.DqError):
        m.prob_one(0)  # nothing loaded
    ts = [np.ones((1, 2, 2), complex), np.ones((2, 2, 2), complex), np.ones((2, 2, 1), complex)]
    m.load(ts)
    with pytest.raises(capi.DqError):
        m.prob_one(3)
    with pytest.raises(capi.DqError):
        m.to_state(capi.State(4))
    with pytest.raises(capi.DqError):
        m.set_cut(4)
    bad = [np.ones((1, 2, 2), complex), np.ones((2, 2, 2), complex), np.ones((2, 2, 2), complex)]
    with pytest.raises(Exception):
        m.load(bad)  # last bond must be 1
    with pytest.raises(capi.DqError):
        capi.Mps(0)


# ---- MpsSimulator (mps/simulator.rs:300-421) ----

SEED = 42
TOL = 1e-8


def assert_matches_statevector(c):
    sv = dq.simulate_monolithic(c, mode="state_vector", seed=SEED)
    mp = dq.simulate_monolithic(c, mode="mps", seed=SEED)
    assert np.max(np.abs(sv.statevector - mp.statevector)) < TOL
    a, b = sv.probabilities(), mp.probabilities()
    for k in set(a) | set(b):
        assert abs(a.get(k, 0.0) - b.get(k, 0.0)) < TOL
    assert abs(sum(b.values()) - 1.0) < TOL


def test_reference_simulation_cases():
    # tests/test_mps_simulation.py of the reference
    assert_matches_statevector(C.circuit(3, [C.instruction("h", qubit=0),
                                             C.instruction("ry", qubit=1, theta=math.pi / 3),
                                             C.instruction("ry", qubit=2, theta=math.pi / 5)]))
    assert_matches_statevector(C.circuit(4, [C.instruction("h", qubit=0), C.instruction("h", qubit=3),
                                             C.instruction("cx", control=3, target=1),
                                             C.instruction("cz", control=2, target=0),
                                             C.instruction("swap", a=3, b=0),
                                             C.instruction("cx", control=0, target=2)]))
    ins = [C.instruction("h", qubit=0), C.instruction("h", qubit=2), C.instruction("ry", qubit=0, theta=0.123),
           C.instruction("cx", control=2, target=1), C.instruction("cx", control=1, target=0),
           C.instruction("swap", a=1, b=0), C.instruction("ry", qubit=2, theta=0.456),
           C.instruction("cx", control=2, target=1), C.instruction("swap", a=2, b=1),
           C.instruction("ry", qubit=1, theta=0.789), C.instruction("cz", control=1, target=0),
           C.instruction("cx", control=0, target=1)]
    assert_matches_statevector(C.circuit(3, ins))
    c = C.circuit(2, [C.instruction("x", qubit=0), C.instruction("measure", qubit=0, cbit=0),
                      C.instruction("conditional", condition=C.Condition(0, 1, 1), op=C.instruction("x", qubit=1))],
                  n_cbits=1)
    r = dq.simulate_monolithic(c, mode="mps", seed=SEED)
    assert r.classical_bits == {0: 1} and r.probabilities() == {3: 1.0}
    r = dq.simulate_monolithic(C.circuit(1, [C.instruction("x", qubit=0), C.instruction("reset", qubit=0)]),
                               mode="mps", seed=SEED)
    assert r.probabilities() == {0: 1.0}
    ins = [C.instruction("h", qubit=0), C.instruction("h", qubit=1), C.instruction("h", qubit=2),
           C.instruction("cx", control=0, target=3), C.instruction("cx", control=1, target=2),
           C.instruction("ry", qubit=3, theta=0.77), C.instruction("cx", control=3, target=1)]
    probs = dq.MpsSimulator(seed=SEED, max_bond_dimension=1).simulate(C.circuit(4, ins)).probabilities()
    assert all(p >= -1e-12 for p in probs.values()) and abs(sum(probs.values()) - 1.0) < 1e-8
    ins = [C.instruction("h", qubit=0), C.instruction("h", qubit=1), C.instruction("cx", control=0, target=2),
           C.instruction("cx", control=1, target=2)]
    probs = dq.simulate_monolithic(C.circuit(3, ins), mode="mps", seed=SEED, max_bond_dimension=1,
                                   truncation_threshold=1e-12).probabilities()
    assert all(p >= -1e-12 for p in probs.values()) and abs(sum(probs.values()) - 1.0) < 1e-8
    with pytest.raises(NotImplementedError, match="one- and two-qubit"):
        dq.simulate_monolithic(C.circuit(3, [C.instruction("h", qubit=0),
                                             C.instruction("ccx", control1=0, control2=1, target=2)]),
                               mode="mps", seed=SEED)


def test_reference_convergence_cases():
    # tests/test_mps_converges.py of the reference
    ins = [C.instruction("h", qubit=q) for q in range(6)]
    for layer in range(3):
        for ctl, tgt in [(0, 5), (1, 4), (2, 3), (0, 2), (3, 5), (1, 3)]:
            ins.append(C.instruction("cx", control=ctl, target=tgt))
        for q in range(6):
            ins.append(C.instruction("ry", qubit=q, theta=0.2 * (layer + 1) * (q + 1)))
    c = C.circuit(6, ins)
    sv = dq.simulate_monolithic(c, mode="state_vector", seed=SEED)
    res = {b: dq.simulate_monolithic(c, mode="mps", seed=SEED, max_bond_dimension=b) for b in (2, 3, 4, 8)}

    def l1(a, b):
        return sum(abs(a.get(k, 0.0) - b.get(k, 0.0)) for k in set(a) | set(b))

    for qubits in (None, [1, 3], [3, 1], [5, 2, 0]):
        errs = [l1(sv.probabilities(qubits), res[b].probabilities(qubits)) for b in (2, 3, 4, 8)]
        assert errs[1] < errs[0] and errs[2] < errs[1] and errs[3] < 1e-10
    ins = [C.instruction("h", qubit=q) for q in range(4)]
    for layer in range(2):
        for ctl, tgt in [(0, 3), (1, 2), (0, 2), (3, 1)]:
            ins.append(C.instruction("cx", control=ctl, target=tgt))
        for q in range(4):
            ins.append(C.instruction("ry", qubit=q, theta=0.3 * (layer + 1) * (q + 1)))
    ins += [C.instruction("measure", qubit=q, cbit=q) for q in range(4)]
    c = C.circuit(4, ins, n_cbits=4)
    shots = 300
    svc = dq.simulate_monolithic_shots(c, mode="state_vector", shots=shots, seed=SEED)
    errs = []
    for b in (1, 2, 4):
        mc = dq.simulate_monolithic_shots(c, mode="mps", shots=shots, seed=SEED, max_bond_dimension=b)
        assert sum(mc.values()) == shots
        errs.append(sum(abs(svc.get(k, 0) - mc.get(k, 0)) / shots for k in set(svc) | set(mc)))
    assert errs[1] < errs[0] and errs[2] < 1e-10
    # and the counts equal the oracle's MPS shots exactly at full bond
    assert dq.simulate_monolithic_shots(c, mode="mps", shots=shots, seed=SEED, max_bond_dimension=4) == \
        orc.simulate_mps_shots_json(c.model_dump_json(), shots, SEED, max_bond=4)


def test_random_circuits_against_oracles():
    checked = 0
    for seed in range(10):
        rng = np.random.default_rng(seed)
        n = int(rng.integers(3, 10))
        c = W.mps_generic(n, int(rng.integers(10, 60)), seed)
        text = c.model_dump_json()
        exact = orc.OracleSimulator(seed=SEED).simulate_json(text).statevector
        got = dq.MpsSimulator(seed=SEED).simulate(c).statevector
        assert np.max(np.abs(got - exact)) <= 1e-10
        for mb in (1, 2, 4):
            want, _, om = orc.simulate_mps_json(text, SEED, max_bond=mb)
            if om.degenerate_cuts:
                continue
            got = dq.MpsSimulator(seed=SEED, max_bond_dimension=mb).simulate(c).statevector
            assert np.max(np.abs(got - want)) <= 1e-10, (seed, mb)
            checked += 1
    assert checked >= 15


def test_measurements_conditionals_resets_match_oracle():
    rng = np.random.default_rng(9)
    for seed in range(6):
        n = 5
        base = W.mps_generic(n, 25, 50 + seed)
        ins = list(base.instructions)
        out = []
        for i, g in enumerate(ins):
            out.append(g)
            if i % 9 == 8:
                q = int(rng.integers(n))
                out.append(C.instruction("measure", qubit=q, cbit=q))
                out.append(C.instruction("conditional", condition=C.Condition(q, 1, 1),
                                         op=C.instruction("ry", qubit=(q + 1) % n, theta=0.7)))
                out.append(C.instruction("reset", qubit=(q + 2) % n))
        c = C.circuit(n, out, n_cbits=n)
        text = c.model_dump_json()
        for mb in (None, 2):
            want, cb, om = orc.simulate_mps_json(text, seed, max_bond=mb)
            if om.degenerate_cuts:
                continue
            r = dq.MpsSimulator(seed=seed, max_bond_dimension=mb).simulate(c)
            assert r.classical_bits == cb
            assert np.max(np.abs(r.statevector - want)) <= 1e-9
        assert dq.MpsSimulator(seed=seed).simulate_shots(c, shots=40) == \
            orc.simulate_mps_shots_json(text, 40, seed)


def test_shots_profile_file(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    c = C.circuit(2, [C.instruction("h", qubit=0), C.instruction("cx", control=0, target=1),
                      C.instruction("measure", qubit=0, cbit=0), C.instruction("measure", qubit=1, cbit=1)],
                  n_cbits=2)
    sim = dq.MpsSimulator(seed=SEED)
    counts = sim.simulate_shots(c, shots=50)
    assert sum(counts.values()) == 50 and set(counts) <= {"00", "11"}
    assert not (tmp_path / "dqsim_profiles").exists()
    counts = sim.simulate_shots(c, shots=64, profile=True)
    assert sum(counts.values()) == 64
    files = list((tmp_path / "dqsim_profiles").glob("mps_shots_profile_*.json"))
    assert len(files) == 1
    data = json.loads(files[0].read_text())
    assert data["num_shots"] == 64 and len(data["shot_times"]) == 64
    for k in ("preprocessing_time", "shots_total_time", "total_time", "svd_time"):
        assert isinstance(data[k], float) and data[k] >= 0.0
    assert isinstance(data["svd_calls"], int) and data["svd_calls"] > 0


def test_single_site_and_measure_everything():
    c = C.circuit(1, [C.instruction("h", qubit=0), C.instruction("measure", qubit=0, cbit=0)], n_cbits=1)
    text = c.model_dump_json()
    for seed in range(6):
        want, cb, _ = orc.simulate_mps_json(text, seed)
        r = dq.MpsSimulator(seed=seed).simulate(c)
        assert r.classical_bits == cb and np.max(np.abs(r.statevector - want)) <= 1e-15
    assert sum(dq.MpsSimulator(seed=1).simulate_shots(c, shots=200).values()) == 200
