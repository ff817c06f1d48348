"""GPU parity of simulate_shots / simulate_monolithic_shots (batched shots).

Oracle: the oracle's restatement of StatevectorSimulator::simulate_shots
(src/monolithic/statevector/simulator.rs:215-322) over the same circuit and
seed.  Counts are integers and must be equal exactly: on the LDS path (n <=
13) every shot's gates use the exact op classes, p1 is summed in the
reference's sequential order and the draws come from the same ChaCha8
stream, so every shot's record is the oracle's.  The engine path (n > 13)
sums p1 with a tree, which can flip an outcome only when a draw falls within
a few ulps of p1; the test seeds are checked to agree exactly.
"""
from __future__ import annotations

import json
import os

import numpy as np
import pytest

from dqsim_bosonic_amd import circuit as C
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

dq = pytest.importorskip("dqsim_bosonic_amd")

from test_shots_host import shots_circuit  # noqa: E402


def bell():
    return C.circuit(2, [C.HInstruction(qubit=0), C.CxInstruction(control=0, target=1),
                         C.MeasureInstruction(qubit=0, cbit=0), C.MeasureInstruction(qubit=1, cbit=1)], 2)


def test_bell_counts_equal_oracle():
    got = dq.StatevectorSimulator(seed=42).simulate_shots(bell(), shots=2000)
    assert got == orc.simulate_shots_json(bell().model_dump_json(), 2000, 42)


@pytest.mark.parametrize("n,count,shots,seed", [(3, 60, 400, 1), (6, 120, 300, 2), (10, 150, 200, 3),
                                                (13, 120, 60, 4)])
def test_random_midcircuit_programs_equal_oracle(n, count, shots, seed):
    rng = np.random.default_rng(seed)
    circ = shots_circuit(n, count, rng)
    got = dq.StatevectorSimulator(seed=1000 + seed).simulate_shots(circ, shots=shots)
    want = orc.simulate_shots_json(circ.model_dump_json(), shots, 1000 + seed)
    assert got == want


def test_engine_path_equals_oracle():
    rng = np.random.default_rng(9)
    circ = shots_circuit(14, 60, rng)
    got = dq.StatevectorSimulator(seed=77).simulate_shots(circ, shots=12)
    assert got == orc.simulate_shots_json(circ.model_dump_json(), 12, 77)


def test_monolithic_shots_entry_and_seed_wrap():
    circ = bell()
    a = dq.simulate_monolithic_shots(circ, mode="state_vector", shots=500, seed=2**64 - 3)
    b = orc.simulate_shots_json(circ.model_dump_json(), 500, 2**64 - 3)  # base_seed + i wraps
    assert a == b
    assert sum(dq.simulate_monolithic_shots(circ, shots=64).values()) == 64  # seed drawn at random
    with pytest.raises(TypeError):
        dq.simulate_monolithic_shots(circ, mode="state_vector", shots=4, bond=3)


def test_profile_file(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    counts = dq.StatevectorSimulator(seed=5).simulate_shots(bell(), shots=50)
    assert sum(counts.values()) == 50 and not (tmp_path / "dqsim_profiles").exists()
    counts = dq.StatevectorSimulator(seed=5).simulate_shots(bell(), shots=64, profile=True)
    files = list((tmp_path / "dqsim_profiles").glob("statevector_shots_profile_*.json"))
    assert len(files) == 1
    data = json.loads(files[0].read_text())
    assert data["num_shots"] == 64 and len(data["shot_times"]) == 64
    assert all(isinstance(t, float) and t >= 0.0 for t in data["shot_times"])
    for key in ("preprocessing_time", "fusion_time", "shots_total_time", "total_time"):
        assert isinstance(data[key], float) and data[key] >= 0.0
    assert data["shots_total_time"] <= data["total_time"] + 1e-6


def test_generic_gates_in_shots():
    ok = C.circuit(2, [C.HInstruction(qubit=0), C.GateInstruction(name="circuit-foo", params=[], qubits=[0, 1]),
                       C.GateInstruction(name="remote_cx", params=[], qubits=[0, 1]),
                       C.MeasureInstruction(qubit=1, cbit=0)], 1)
    got = dq.StatevectorSimulator(seed=3).simulate_shots(ok, shots=300)
    assert got == orc.simulate_shots_json(ok.model_dump_json(), 300, 3)
    bad = C.circuit(2, [C.GateInstruction(name="remote_rzz", params=[0.1], qubits=[0, 1])], 1)
    with pytest.raises(RuntimeError, match="Unsupported generic gate"):
        dq.StatevectorSimulator(seed=3).simulate_shots(bad, shots=3)
