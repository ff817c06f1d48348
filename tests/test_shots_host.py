"""CPU tests of the shots path's host side and of the oracle's restatement.

simulate_shots (src/monolithic/statevector/simulator.rs:215-322) first fuses
runs of single-qubit gates per qubit (src/types.rs:254-293) with its own
single-qubit matrices (gate_matrix_1q, types.rs:164-230), drops products
within 1e-10 of the identity, then runs every shot.  Checks:
  * the product's fusion (_core._fuse_circuit) equals the oracle's
    restatement bit for bit on random circuits;
  * the oracle's single-qubit matrices agree with the gate matrices of
    gates.rs (oracle.gate_matrix) to 1e-15, and the fused list applied to a
    state equals the unfused circuit (the mathematics of the fusion);
  * the identity-drop rule and format_cbits' bit order on known cases;
  * deterministic circuits give their known counts in the oracle.
"""
from __future__ import annotations

import json
import math

import numpy as np
import pytest

from dqsim_bosonic_amd import circuit as C
from oracle import oracle as orc

core = pytest.importorskip("dqsim_bosonic_amd._core")

ONE_Q = ["id", "u0", "x", "y", "z", "h", "s", "sdg", "t", "tdg", "sx", "sxdg", "u3", "u", "u2", "u1", "p",
         "rx", "ry", "rz"]
PARAMS = {"u3": ("theta", "phi", "lam"), "u": ("theta", "phi", "lam"), "u2": ("phi", "lam"), "u1": ("lam",),
          "p": ("lam",), "rx": ("theta",), "ry": ("theta",), "rz": ("phi",)}


def one_q(kind, q, rng):
    kw = {f: float(rng.uniform(-7, 7)) for f in PARAMS.get(kind, ())}
    return C.instruction(kind, qubit=q, **kw)


def shots_circuit(n, count, rng, n_cbits=None):
    n_cbits = n_cbits or n
    ins = []
    for _ in range(count):
        r = rng.random()
        if r < 0.5:
            ins.append(one_q(ONE_Q[int(rng.integers(len(ONE_Q)))], int(rng.integers(n)), rng))
        elif r < 0.7 and n >= 2:
            a, b = (int(x) for x in rng.choice(n, size=2, replace=False))
            g = ["cx", "cz", "swap", "cp"][int(rng.integers(4))]
            if g == "swap":
                ins.append(C.SwapInstruction(a=a, b=b))
            elif g == "cp":
                ins.append(C.CpInstruction(control=a, target=b, lam=float(rng.uniform(-3, 3))))
            else:
                ins.append(C.instruction(g, control=a, target=b))
        elif r < 0.8:
            ins.append(C.MeasureInstruction(qubit=int(rng.integers(n)), cbit=int(rng.integers(n_cbits))))
        elif r < 0.87:
            ins.append(C.ResetInstruction(qubit=int(rng.integers(n))))
        elif r < 0.95:
            base = int(rng.integers(n_cbits))
            size = int(rng.integers(1, 3))
            inner = one_q(["x", "h", "rz", "sx"][int(rng.integers(4))], int(rng.integers(n)), rng)
            ins.append(C.ConditionalInstruction(condition=C.Condition(creg_base=base, creg_size=size,
                                                                      creg_value=int(rng.integers(1 << size))),
                                                op=inner))
        else:
            ins.append(C.BarrierInstruction())
    return C.circuit(n, ins, n_cbits)


def oracle_fused(text):
    insts = json.loads(text)["instructions"]
    out = []
    for item in orc.fuse_circuit(insts):
        if item[0] == "fused":
            out.append(("fused", item[1], tuple(item[2])))
        else:
            out.append(("orig", next(i for i, x in enumerate(insts) if x is item[1])))
    return out


@pytest.mark.parametrize("seed", range(12))
def test_host_fusion_matches_oracle_bit_for_bit(seed):
    rng = np.random.default_rng(seed)
    text = shots_circuit(6, 120, rng).model_dump_json()
    got = [tuple(x) if x[0] == "orig" else ("fused", x[1], tuple(tuple(e) for e in x[2]))
           for x in core._fuse_circuit(text)]
    want = oracle_fused(text)
    assert got == want


@pytest.mark.parametrize("kind", ONE_Q)
def test_fused_matrices_agree_with_gate_matrices(kind):
    rng = np.random.default_rng(1)
    inst = one_q(kind, 0, rng).model_dump()
    m = np.array([complex(*e) for e in orc.fused_gate_matrix_1q(inst)]).reshape(2, 2)
    params = [float(inst[f]) for f in PARAMS.get(kind, ())]
    g = orc.gate_matrix("id" if kind == "u0" else kind, params)
    assert np.max(np.abs(m - g)) <= 1e-15


def test_fused_list_equals_unfused_circuit():
    rng = np.random.default_rng(4)
    n = 5
    circ = shots_circuit(n, 200, rng)
    insts = [i for i in json.loads(circ.model_dump_json())["instructions"]
             if i["kind"] not in ("measure", "reset", "conditional")]
    psi = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    psi /= np.linalg.norm(psi)
    sim = orc.OracleSimulator(seed=0)
    want = psi.copy()
    for inst in insts:
        sim._run(want, n, inst, {}, None, [])
    got = psi.copy()
    for item in orc.fuse_circuit(insts):
        if item[0] == "fused":
            orc.apply_one_qubit(got, n, np.array([complex(*e) for e in item[2]]).reshape(2, 2), item[1])
        else:
            sim._run(got, n, item[1], {}, None, [])
    assert np.max(np.abs(got - want)) <= 1e-12


def test_identity_products_are_dropped():
    insts = [{"kind": "x", "qubit": 0}, {"kind": "x", "qubit": 0},
             {"kind": "h", "qubit": 1}, {"kind": "h", "qubit": 1},
             {"kind": "rz", "qubit": 2, "phi": 0.3}, {"kind": "rz", "qubit": 2, "phi": -0.3},
             {"kind": "s", "qubit": 3}, {"kind": "sdg", "qubit": 3},
             {"kind": "t", "qubit": 4}, {"kind": "measure", "qubit": 4, "cbit": 0}]
    fused = orc.fuse_circuit(insts)
    assert [x[0] for x in fused] == ["fused", "orig"] and fused[0][1] == 4
    assert orc.is_identity_2x2(((1.0, 0.0), (5e-11, 0.0), (0.0, 0.0), (1.0, 5e-11)))
    assert not orc.is_identity_2x2(((1.0, 0.0), (2e-10, 0.0), (0.0, 0.0), (1.0, 0.0)))


def test_format_cbits_bit_order():
    assert orc.format_cbits({0: 1}, 3) == "001"
    assert orc.format_cbits({2: 1, 0: 1}, 3) == "101"
    assert orc.format_cbits({5: 1}, 3) == "000"
    assert orc.format_cbits({}, 0) == ""


def test_oracle_deterministic_shots():
    # x on q0, measure into cbit 2 of 3; conditional x on q1 when cbit 2 is set; measure q1 into cbit 0
    circ = C.circuit(2, [C.XInstruction(qubit=0), C.MeasureInstruction(qubit=0, cbit=2),
                         C.ConditionalInstruction(condition=C.Condition(creg_base=2, creg_size=1, creg_value=1),
                                                  op=C.XInstruction(qubit=1)),
                         C.MeasureInstruction(qubit=1, cbit=0)], 3)
    assert orc.simulate_shots_json(circ.model_dump_json(), 50, 7) == {"101": 50}
    bell = C.circuit(2, [C.HInstruction(qubit=0), C.CxInstruction(control=0, target=1),
                         C.MeasureInstruction(qubit=0, cbit=0), C.MeasureInstruction(qubit=1, cbit=1)], 2)
    counts = orc.simulate_shots_json(bell.model_dump_json(), 2000, 42)
    assert set(counts) <= {"00", "11"} and sum(counts.values()) == 2000
    assert abs(counts["00"] - 1000) < 4 * math.sqrt(500)
