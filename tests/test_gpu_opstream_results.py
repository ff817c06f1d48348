"""GPU tests of the opcode-stream input and of the lazy results.

* simulate_opstream runs the same dispatcher as simulate(): at fusion level 1
  the statevector and the measurement record are bit-identical to
  simulate() of the JSON circuit (itself pinned to the oracle);
  simulate_shots_opstream gives simulate_shots' counts.
* probabilities_array equals probabilities() as an array; top_k equals the
  k largest entries of the oracle's |psi|^2 (ties by smaller index),
  including a uniform state where every probability ties.
"""
from __future__ import annotations

import numpy as np
import pytest

from dqsim_bosonic_amd import circuit as C
from dqsim_bosonic_amd import opstream as O
from dqsim_bosonic_amd import workloads as W
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

dq = pytest.importorskip("dqsim_bosonic_amd")


def test_opstream_equals_json_path():
    circ = W.random_blocks(12, 20, seed=5)
    recs, nq, ncb, names = O.encode(circ)
    sim = dq.StatevectorSimulator(seed=9, fusion=1)
    a = sim.simulate(circ)
    b = sim.simulate_opstream(recs, nq, ncb, names)
    assert np.array_equal(a.statevector, b.statevector)
    assert a.classical_bits == b.classical_bits
    mixed = W.random_mixed(9, 200, seed=2)
    r2, n2, c2, nm2 = O.encode(mixed)
    assert np.array_equal(sim.simulate(mixed).statevector, sim.simulate_opstream(r2, n2, c2, nm2).statevector)


def test_shots_opstream_equals_shots():
    circ = W.shots_blocks(8, 12, seed=4)
    recs, nq, ncb, names = O.encode(circ)
    sim = dq.StatevectorSimulator(seed=21)
    assert sim.simulate_shots(circ, shots=300) == sim.simulate_shots_opstream(recs, nq, ncb, names, shots=300)


def test_probabilities_array_and_top_k():
    circ = W.random_mixed(14, 200, seed=8, kinds=["h", "rx", "ry", "cx", "cz", "rzz", "u3"])
    res = dq.StatevectorSimulator(seed=1).simulate(circ)
    arr = res.probabilities_array()
    d = res.probabilities()
    dense = np.zeros(1 << 14)
    for k, v in d.items():
        dense[k] = v
    assert np.array_equal(arr, dense)
    assert np.array_equal(res.probabilities_array([3, 0, 7]),
                          np.array([res.probabilities([3, 0, 7]).get(i, 0.0) for i in range(8)]))
    want = orc.OracleSimulator(seed=1).simulate(circ).statevector
    p = np.abs(want) ** 2
    for k in (1, 7, 100, 4096):
        got = res.top_k(k)
        order = sorted(range(p.size), key=lambda i: (-p[i], i))[:k]
        assert [g[0] for g in got] == order
        assert max(abs(g[1] - p[i]) for g, i in zip(got, order)) <= 1e-15


def test_top_k_ties_take_smallest_indices():
    n = 12
    res = dq.StatevectorSimulator(seed=1).simulate(C.circuit(n, [C.HInstruction(qubit=q) for q in range(n)]))
    got = res.top_k(10)
    assert [g[0] for g in got] == list(range(10))
    assert all(abs(g[1] - 2.0 ** -n) < 1e-18 for g in got)
    with pytest.raises(Exception):
        res.top_k(0)
