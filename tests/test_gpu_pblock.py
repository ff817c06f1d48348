"""GPU tests of the P-block simulator (src/distributed/pblock): blocks as
device states, merged by dq_state_tensor, epochs drawing from per-member
ChaCha8 streams; shots through the batched-shots engine.

Oracle: ``simulate_pblock`` / ``simulate_pblock_shots`` / ``merge_blocks``
(pinned in test_pblock_host.py), the oracle statevector simulator for runs
without measurement, and the statevector shots of the monolithic circuit
(the reference's own invariant, tests/test_pblock_conditional.py).
Tolerance: 1e-12 on amplitudes; records and counts exact.
"""
from __future__ import annotations

import json

import numpy as np
import pytest

from oracle import oracle as orc

import pblock_cases as P  # noqa: E402

pytestmark = pytest.mark.gpu

capi = pytest.importorskip("dqsim_bosonic_amd.capi")
dq = pytest.importorskip("dqsim_bosonic_amd")
from dqsim_bosonic_amd import circuit as C  # noqa: E402

SEED = 42


@pytest.mark.parametrize("dtype,tol", [(capi.C128, 0.0), (capi.C64, 1e-6)])
def test_tensor_kernel_against_merge_blocks(dtype, tol):
    rng = np.random.default_rng(3)
    for aq, bq in (([0], [1]), ([1], [0]), ([0, 2, 5], [1, 3, 4]), ([4, 9], [0, 1, 2, 3, 5, 6, 7, 8]),
                   ([], [0, 1, 2]), ([0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11], [12, 13, 14, 15, 16, 17, 18, 19, 20])):
        a = rng.normal(size=1 << len(aq)) + 1j * rng.normal(size=1 << len(aq))
        b = rng.normal(size=1 << len(bq)) + 1j * rng.normal(size=1 << len(bq))
        want, qs = orc.merge_blocks(a, aq, b, bq)
        sa, sb, sd = capi.State(len(aq), dtype), capi.State(len(bq), dtype), capi.State(len(qs), dtype)
        sa.set(a)
        sb.set(b)
        sd.tensor(sa, [qs.index(q) for q in aq], sb, [qs.index(q) for q in bq])
        got = sd.get()
        if tol == 0.0:
            assert np.array_equal(got, want), (aq, bq)
        else:
            assert np.max(np.abs(got - want)) <= tol * max(1.0, np.max(np.abs(want)))
    with pytest.raises(capi.DqError):
        sd = capi.State(2)
        sd.tensor(capi.State(1), [0], capi.State(1), [0])  # position named twice


def test_reference_conditional_shots_equal_statevector_shots(tmp_path, monkeypatch):
    # tests/test_pblock_conditional.py of the reference
    assert dq.PBlockSimulator(seed=SEED).simulate_shots(P.symbolic_distributed(), shots=1000) == \
        dq.StatevectorSimulator(seed=SEED).simulate_shots(P.original_circuit(), shots=1000)
    assert dq.PBlockSimulator(seed=SEED).simulate_shots(P.symbolic_parametric_distributed(), shots=1000) == \
        dq.StatevectorSimulator(seed=SEED).simulate_shots(P.original_parametric_circuit(), shots=1000)
    monkeypatch.chdir(tmp_path)
    counts = dq.PBlockSimulator(seed=SEED).simulate_shots(P.symbolic_distributed(), shots=64, profile=True)
    assert sum(counts.values()) == 64
    files = list((tmp_path / "dqsim_profiles").glob("pblock_shots_profile_*.json"))
    assert len(files) == 1
    data = json.loads(files[0].read_text())
    assert data["num_shots"] == 64 and isinstance(data["merge_calls"], int) and data["merge_calls"] > 0
    for k in ("preprocessing_time", "fusion_time", "shots_total_time", "total_time"):
        assert isinstance(data[k], float)
    assert len(data["shot_times"]) == 64


@pytest.mark.parametrize("seed", range(8))
def test_random_distributed_simulate_matches_oracle(seed):
    circ, dist = P.random_distributed(seed, nodes=3, per_node=3, count=60)
    st, qs, cbits = orc.simulate_pblock(dist, seed)
    r = dq.PBlockSimulator(seed=seed).simulate(dist)
    assert r.physical_qubits == qs and r.num_qubits == len(qs)
    assert r.classical_bits == cbits
    assert np.max(np.abs(r.statevector - st)) <= 1e-12
    want = orc.marginal_probs(st, len(qs), [len(qs) - 1 - j for j in range(len(qs))])
    pr = r.probabilities()
    assert max(abs(pr.get(j, 0.0) - float(p)) for j, p in enumerate(want)) <= 1e-12
    pr = r.probabilities([4, 0])
    want2 = orc.marginal_probs(st, len(qs), [qs.index(4), qs.index(0)])
    assert max(abs(pr.get(j, 0.0) - float(p)) for j, p in enumerate(want2)) <= 1e-12


@pytest.mark.parametrize("seed", range(4))
def test_random_distributed_without_measurement_equals_statevector(seed):
    circ, dist = P.random_distributed(100 + seed, nodes=4, per_node=3, count=80, measure=False)
    r = dq.simulate_distributed(dist, mode="p-block", seed=1)
    want = dq.StatevectorSimulator(seed=1).simulate(circ).statevector
    assert r.physical_qubits == list(range(12))
    assert np.max(np.abs(r.statevector - want)) <= 1e-12


@pytest.mark.parametrize("seed", range(4))
def test_random_distributed_shots_match_oracle_and_statevector(seed, tmp_path, monkeypatch):
    circ, dist = P.random_distributed(200 + seed, nodes=2, per_node=3, count=40)
    want, merges = orc.simulate_pblock_shots(dist, 150, seed)
    monkeypatch.chdir(tmp_path)
    got = dq.simulate_distributed_shots(dist, shots=150, seed=seed, profile=True)
    assert got == want
    assert got == dq.StatevectorSimulator(seed=seed).simulate_shots(circ, shots=150)
    data = json.loads(next((tmp_path / "dqsim_profiles").glob("pblock_shots_profile_*.json")).read_text())
    assert data["merge_calls"] == 150 * merges


def test_epoch_draws_and_unsupported_gates():
    q, c = {"q": C.Register("q", 2, 0)}, {"c": C.Register("c", 2, 0)}
    ins = [C.instruction("h", qubit=0), C.instruction("h", qubit=1), C.instruction("measure", qubit=0, cbit=0),
           C.instruction("measure", qubit=1, cbit=1)]
    d = C.distribute(C.Circuit(q, c, ins), {0: [0], 1: [1]})
    for seed in (1, 2, 3, 42, 2**63 + 11):
        _, _, cbits = orc.simulate_pblock(d, seed)
        assert dq.PBlockSimulator(seed=seed).simulate(d).classical_bits == cbits
    for name, exc in (("teleport", NotImplementedError), ("remote_swap", NotImplementedError)):
        bad = C.distribute(C.Circuit(q, c, [C.GateInstruction(name=name, qubits=[0, 1], params=[])]),
                           {0: [0], 1: [1]})
        with pytest.raises(exc):
            dq.PBlockSimulator(seed=1).simulate(bad)
        with pytest.raises(RuntimeError):
            dq.PBlockSimulator(seed=1).simulate_shots(bad, shots=4)
    bad = C.distribute(C.Circuit(q, c, [C.GateInstruction(name="remote_rzz", qubits=[0, 1], params=[])]),
                       {0: [0], 1: [1]})
    with pytest.raises(ValueError, match="remote_rzz requires one theta parameter"):
        dq.PBlockSimulator(seed=1).simulate(bad)


def test_empty_node_and_untouched_blocks():
    """A node without qubits is a 0-qubit block; blocks no instruction merges
    are tensored in node order at the end."""
    q, c = {"q": C.Register("q", 5, 0)}, {"c": C.Register("c", 5, 0)}
    ins = [C.instruction("h", qubit=0), C.instruction("ry", qubit=3, theta=0.7),
           C.instruction("cx", control=0, target=1), C.instruction("measure", qubit=1, cbit=1)]
    qpn = {0: [0, 1], 1: [], 2: [2, 3], 3: [4]}
    d = C.distribute(C.Circuit(q, c, ins), {k: v for k, v in qpn.items() if v})
    d.qubits_per_node = qpn
    d.circuits[1] = C.Circuit(q, c, [])
    for seed in (1, 2, 3):
        st, qs, cb = orc.simulate_pblock(d, seed)
        r = dq.PBlockSimulator(seed=seed).simulate(d)
        assert r.physical_qubits == qs == [0, 1, 2, 3, 4]
        assert r.classical_bits == cb
        assert np.max(np.abs(r.statevector - st)) <= 1e-14
