"""GPU parity of StatevectorSimulator.simulate() and SimulationResult against
the CPU oracle, plus the reference's own known answers.

Known answers restated from the reference's files:
  * README.md:71-75  Bell probabilities() == {0: 0.5000000000000001, 3: 0.5000000000000001}
  * tests/test_mps_simulation.py:96-126  x; measure; conditional x -> classical_bits {0: 1},
    probabilities {3: 1.0}; x; reset -> {0: 1.0}
  * tests/test_distribution_e2e.py:86-216  small circuits (x_cx -> |11>, bell, ghz4, ...)
"""
from __future__ import annotations

import json
import math

import numpy as np
import pytest

import dqsim_bosonic_amd as dq
from dqsim_bosonic_amd import circuit as C
from dqsim_bosonic_amd import workloads as W
from oracle.oracle import OracleSimulator

pytestmark = pytest.mark.gpu


def both(circ, seed=42, fusion=1):
    """fusion=1: program-order tiles, bit-identical arithmetic to the reference."""
    return (dq.StatevectorSimulator(seed=seed, fusion=fusion).simulate(circ),
            OracleSimulator(seed=seed).simulate(circ))


def test_bell_known_answer_readme():
    c = C.circuit(2, [C.HInstruction(qubit=0), C.CxInstruction(control=0, target=1)])
    r = dq.StatevectorSimulator(seed=42).simulate(c)
    assert r.probabilities() == {0: 0.5000000000000001, 3: 0.5000000000000001}


def test_measure_conditional_known_answer():
    c = C.circuit(2, [C.XInstruction(qubit=0), C.MeasureInstruction(qubit=0, cbit=0),
                      C.ConditionalInstruction(condition=C.Condition(creg_base=0, creg_size=1, creg_value=1),
                                               op=C.XInstruction(qubit=1))], n_cbits=1)
    r = dq.simulate_monolithic(c, mode="state_vector", seed=42)
    assert r.classical_bits == {0: 1}
    assert r.probabilities() == {3: 1.0}


def test_reset_known_answer():
    c = C.circuit(1, [C.XInstruction(qubit=0), C.ResetInstruction(qubit=0)])
    assert dq.simulate_monolithic(c, seed=42).probabilities() == {0: 1.0}


E2E = {
    "x_cx": ([C.XInstruction(qubit=0), C.CxInstruction(control=0, target=1)], 2, {3: 1.0}),
    "bell": ([C.HInstruction(qubit=0), C.CxInstruction(control=0, target=1)], 2, {0: 0.5, 3: 0.5}),
    "cz_pair": ([C.HInstruction(qubit=0), C.HInstruction(qubit=1), C.CzInstruction(control=0, target=1)], 2,
                {0: .25, 1: .25, 2: .25, 3: .25}),
    "ry_cx": ([C.RyInstruction(qubit=0, theta=math.pi / 3), C.CxInstruction(control=0, target=1)], 2,
              {0: math.cos(math.pi / 6) ** 2, 3: math.sin(math.pi / 6) ** 2}),
    "ghz4": ([C.HInstruction(qubit=0)] + [C.CxInstruction(control=i, target=i + 1) for i in range(3)], 4,
             {0: .5, 15: .5}),
    "chain_cx4": ([C.HInstruction(qubit=0), C.CxInstruction(control=0, target=2),
                   C.CxInstruction(control=1, target=3)], 4, {0: .5, 5: .5}),
    "local_4h": ([C.HInstruction(qubit=i) for i in range(4)], 4, {j: 1 / 16 for j in range(16)}),
}


@pytest.mark.parametrize("name", list(E2E))
def test_reference_e2e_circuits(name):
    ins, n, want = E2E[name]
    got = dq.StatevectorSimulator(seed=42).simulate(C.circuit(n, ins)).probabilities()
    for k in set(got) | set(want):
        assert abs(got.get(k, 0.0) - want.get(k, 0.0)) < 1e-6


@pytest.mark.parametrize("n,seed", [(3, 1), (8, 2), (13, 3), (17, 4)])
def test_random_mixed_circuits_bit_exact(n, seed):
    circ = W.random_mixed(n, 150, seed)
    r, o = both(circ, seed)
    assert np.array_equal(r.statevector, o.statevector)


@pytest.mark.parametrize("fusion", [0, 2])
def test_random_mixed_other_fusion_levels(fusion):
    circ = W.random_mixed(15, 200, 77)
    r, o = both(circ, 77, fusion=fusion)
    err = np.max(np.abs(r.statevector - o.statevector))
    assert (err == 0.0) if fusion == 0 else (err <= 1e-14), err


def test_brickwork_12q_bit_exact_probs_counts():
    circ = W.brickwork_u3(12, 10, seed=42)
    r, o = both(circ)
    assert np.array_equal(r.statevector, o.statevector)
    assert r.probabilities() == o.probabilities()
    assert r.counts(shots=10000, seed=7) == o.counts(shots=10000, seed=7)
    assert r.counts(shots=500, qubits=[3, 0], seed=9) == o.counts(shots=500, qubits=[3, 0], seed=9)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_measurement_record_matches(seed):
    circ = W.random_blocks(14, 20, seed=seed, remote_per_kind=3, measure_every=5)
    r, o = both(circ, seed)
    assert r.classical_bits == o.classical_bits
    err = np.max(np.abs(r.statevector - o.statevector))
    assert err <= 1e-12, err


@pytest.mark.parametrize("n", [9, 10, 11, 12])
@pytest.mark.parametrize("fusion", [1, 2])
def test_measurement_record_every_small_tile_width(n, fusion):
    """Regression: widths whose sweeps use 9-qubit tiles (n = 9..11) measure
    with the reduction fused into the last sweep, like the wider ones."""
    circ = W.random_blocks(n, 16, seed=n, remote_per_kind=2, measure_every=4)
    r, o = both(circ, n, fusion)
    assert r.classical_bits == o.classical_bits
    assert np.max(np.abs(r.statevector - o.statevector)) <= 1e-12


def test_fidelity_and_length_mismatch():
    circ = W.brickwork_u3(10, 6, seed=3)
    r, o = both(circ)
    f = r.fidelity(o.statevector)
    assert abs(f - 1.0) < 1e-12
    assert r.fidelity(np.zeros(8, dtype=np.complex128)) == 0.0


class _JsonCircuit:
    def __init__(self, d):
        self._text = json.dumps(d)

    def model_dump_json(self):
        return self._text


def _profile_circuits():
    from tests.test_oracle import profile_case

    yield "known", _JsonCircuit(profile_case()[0])
    yield "blocks", W.random_blocks(10, 10, seed=5, remote_per_kind=2, measure_every=5)
    mixed = json.loads(W.random_mixed(7, 150, seed=9).model_dump_json())
    ins = mixed["instructions"]
    for pos, q in ((40, 2), (90, 5), (120, 0)):  # measure, reset, conditional on the measured bit
        ins.insert(pos, {"kind": "measure", "qubit": q, "cbit": 0})
        ins.insert(pos + 1, {"kind": "reset", "qubit": (q + 1) % 7})
        ins.insert(pos + 2, {"kind": "conditional", "condition": {"creg_base": 0, "creg_size": 1, "creg_value": 1},
                             "op": {"kind": "cx", "control": q, "target": (q + 3) % 7}})
    mixed["cregs"] = {"c": {"name": "c", "size": 1, "base": 0}}
    yield "mixed", _JsonCircuit(mixed)


@pytest.mark.parametrize("fusion", [1, 2])
def test_profile_counts_equal_reference_rule(fusion):
    """apply_one_qubit / apply_n_qubit / measure_qubit calls equal the counts
    of the oracle's dispatcher, which follows the reference's profiling
    wrappers (simulator.rs:890-946): a reset's X is a one-qubit call, every
    2-5 qubit kind and generic gate an n-qubit call, remote_barrier /
    remote_cu1 / id / u0 / barrier none."""
    for name, circ in _profile_circuits():
        sim = OracleSimulator(seed=5)
        ref = sim.simulate(circ)
        r = dq.StatevectorSimulator(seed=5, profile=True, fusion=fusion).simulate(circ)
        p = r.profile
        got = {"apply_one_qubit": p.apply_one_qubit_calls, "apply_n_qubit": p.apply_n_qubit_calls,
               "measure_qubit": p.measure_qubit_calls}
        assert got == sim.calls, name
        assert r.classical_bits == ref.classical_bits, name
        assert p.total_time > 0
        assert p.total_time >= p.apply_one_qubit_time + p.apply_n_qubit_time + p.measure_qubit_time - 1e-9
    assert dq.StatevectorSimulator(seed=5).simulate(circ).profile is None


def test_unsupported_generic_gate_raises():
    c = C.circuit(2, [C.GateInstruction(name="Foo_Gate", params=[], qubits=[0, 1])])
    with pytest.raises(NotImplementedError, match="Unsupported generic gate: \"foo_gate\""):
        dq.StatevectorSimulator(seed=1).simulate(c)


def test_bad_qubit_is_value_error():
    c = C.circuit(2, [C.HInstruction(qubit=5)])
    with pytest.raises(ValueError):
        dq.StatevectorSimulator(seed=1).simulate(c)


@pytest.mark.parametrize("qubits", [[3, 9, 0], [2, 2, 5], [10], [7, 1, 7], [6, 63, 1, 4]])
def test_probabilities_out_of_range_and_repeated_qubits(qubits):
    """probabilities() / counts() take the reference's marginal_probs
    semantics (src/engine.rs:368-390) for qubits at or above n (their set
    bins are empty) and for repeated qubits; checked against the oracle's
    literal restatement of that loop."""
    from oracle import oracle as orc

    n = 8
    circ = W.random_mixed(n, 60, seed=21)
    r, o = both(circ)
    want = orc.marginal_probs(o.statevector, n, qubits, literal=True)
    got = r.probabilities(qubits)
    assert set(got) == {j for j, p in enumerate(want) if p > 0.0}
    for j, p in got.items():
        assert abs(p - want[j]) <= 1e-14 * max(1.0, want[j])
    assert np.allclose(r.probabilities_array(qubits), want, rtol=1e-14, atol=1e-15)
    c = r.counts(2000, qubits=qubits, seed=4)
    assert c == orc.sample_counts(want, len(qubits), 2000, orc.Rng(4))


def test_probabilities_negative_qubit_is_overflow_error():
    r = dq.StatevectorSimulator(seed=1).simulate(C.circuit(2, [C.HInstruction(qubit=0)]))
    with pytest.raises(OverflowError):
        r.probabilities([-1])
    with pytest.raises(ValueError):
        r.counts(10, qubits=[64])
