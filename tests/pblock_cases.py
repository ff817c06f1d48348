"""Distributed circuits shared by the P-block tests: the reference's
tests/test_pblock_conditional.py cases and random distributed circuits."""
from __future__ import annotations

import math

import numpy as np

from dqsim_bosonic_amd import circuit as C


def _regs():
    return {"q": C.Register("q", 2, 0)}, {"c": C.Register("c", 2, 0)}


def original_circuit():
    q, c = _regs()
    return C.Circuit(q, c, [C.instruction("h", qubit=0), C.instruction("measure", qubit=0, cbit=0),
                            C.instruction("conditional", condition=C.Condition(0, 1, 1),
                                          op=C.instruction("cx", control=0, target=1)),
                            C.instruction("measure", qubit=1, cbit=1)])


def symbolic_distributed():
    q, c = _regs()
    h = C.instruction("h", qubit=0)
    mc = C.instruction("measure", qubit=0, cbit=0)
    cond = C.instruction("conditional", condition=C.Condition(0, 1, 1),
                         op=C.GateInstruction(name="remote_cx", qubits=[0, 1], params=[]))
    mt = C.instruction("measure", qubit=1, cbit=1)
    d = C.DistributedCircuit({0: [0], 1: [1]}, {0: C.Circuit(q, c, [h, mc, cond]), 1: C.Circuit(q, c, [cond, mt])})
    d._instruction_index = {id(h): 0, id(mc): 1, id(cond): 2, id(mt): 3}
    return d


def original_parametric_circuit():
    q, c = _regs()
    return C.Circuit(q, c, [C.instruction("h", qubit=0), C.instruction("measure", qubit=0, cbit=0),
                            C.instruction("h", qubit=1),
                            C.instruction("conditional", condition=C.Condition(0, 1, 1),
                                          op=C.instruction("rzz", a=0, b=1, theta=math.pi)),
                            C.instruction("h", qubit=1), C.instruction("measure", qubit=1, cbit=1)])


def symbolic_parametric_distributed():
    q, c = _regs()
    hc = C.instruction("h", qubit=0)
    mc = C.instruction("measure", qubit=0, cbit=0)
    hb = C.instruction("h", qubit=1)
    cond = C.instruction("conditional", condition=C.Condition(0, 1, 1),
                         op=C.GateInstruction(name="remote_rzz", qubits=[0, 1], params=[math.pi]))
    ha = C.instruction("h", qubit=1)
    mt = C.instruction("measure", qubit=1, cbit=1)
    d = C.DistributedCircuit({0: [0], 1: [1]}, {0: C.Circuit(q, c, [hc, mc, cond]), 1: C.Circuit(q, c, [hb, cond, ha, mt])})
    d._instruction_index = {id(hc): 0, id(mc): 1, id(hb): 2, id(cond): 3, id(ha): 4, id(mt): 5}
    return d


ONE_Q = ["h", "x", "s", "t", "sx", "u3", "rx", "ry", "rz", "p", "u2"]
TWO_Q = ["cx", "cz", "cp", "rzz", "rxx", "swap", "cry", "cu3"]
REMOTE = ["remote_cx", "nonlocal_cz", "remote_cz", "remote_rzz", "remote_link_phi_plus", "remote_link_psi_minus",
          "remote_link_psi_plus", "remote_epr", "remote_barrier", "remote_cu1", "circuit-sub"]
PARAMS = {"u3": ("theta", "phi", "lam"), "rx": ("theta",), "ry": ("theta",), "rz": ("phi",), "p": ("lam",),
          "u2": ("phi", "lam"), "cp": ("lam",), "rzz": ("theta",), "rxx": ("theta",), "cry": ("theta",),
          "cu3": ("theta", "phi", "lam")}


def _gate(kind, qs, rng):
    kw = {f: float(rng.uniform(-3, 3)) for f in PARAMS.get(kind, ())}
    if kind in ("swap", "rzz", "rxx"):
        return C.instruction(kind, a=qs[0], b=qs[1], **kw)
    if len(qs) == 2:
        return C.instruction(kind, control=qs[0], target=qs[1], **kw)
    return C.instruction(kind, qubit=qs[0], **kw)


def random_distributed(seed: int, nodes: int = 3, per_node: int = 3, count: int = 40, measure: bool = True):
    """Random monolithic circuit over nodes x per_node qubits (local gates,
    cross-node standard and generic remote gates, and with `measure`
    mid-circuit measure / conditional / reset), split by node; returns
    (the monolithic equivalent for the statevector simulator, distributed
    circuit)."""
    rng = np.random.default_rng(seed)
    n = nodes * per_node
    qpn = {k: list(range(k * per_node, (k + 1) * per_node)) for k in range(nodes)}
    ins = []
    mono = {}
    ncb = n
    for _ in range(count):
        r = rng.random()
        if r < 0.45:
            ins.append(_gate(ONE_Q[int(rng.integers(len(ONE_Q)))], [int(rng.integers(n))], rng))
        elif r < 0.7:
            ins.append(_gate(TWO_Q[int(rng.integers(len(TWO_Q)))], [int(x) for x in rng.choice(n, 2, replace=False)],
                             rng))
        elif r < 0.85:
            name = REMOTE[int(rng.integers(len(REMOTE)))]
            a, b = (int(x) for x in rng.choice(n, 2, replace=False))
            theta = float(rng.uniform(-3, 3))
            g = C.GateInstruction(name=name, qubits=[a, b], params=[theta] if name == "remote_rzz" else [])
            ins.append(g)
            # the statevector dispatcher has no remote_rzz / circuit-* in simulate():
            # its equivalent is rzz / nothing
            if name == "remote_rzz":
                mono[len(ins) - 1] = C.instruction("rzz", a=a, b=b, theta=theta)
            elif name.startswith("circuit-"):
                mono[len(ins) - 1] = None
        elif measure:
            q = int(rng.integers(n))
            kind = rng.integers(3)
            if kind == 0:
                ins.append(C.instruction("measure", qubit=q, cbit=q))
            elif kind == 1:
                t = int(rng.integers(n))
                ins.append(C.instruction("conditional", condition=C.Condition(q, 1, int(rng.integers(2))),
                                         op=_gate("ry", [t], rng)))
            else:
                ins.append(C.instruction("reset", qubit=q))
    if measure:
        ins += [C.instruction("measure", qubit=q, cbit=q) for q in range(n)]
    circ = C.circuit(n, ins, n_cbits=ncb if measure else 0)
    equiv = [mono.get(i, g) for i, g in enumerate(ins)]
    equiv = C.circuit(n, [g for g in equiv if g is not None], n_cbits=ncb if measure else 0)
    return equiv, C.distribute(circ, qpn)
