"""Synthetic circuit families used by the benchmark and the parity tests.

Each generator returns a ``circuit.Circuit`` in the reference's wire schema
(src/types.rs:60-135), drawn from numpy's PCG64 ``default_rng(seed)``.

* ``brickwork_u3``: depth-d brickwork, u3 on every qubit per layer, then cx on
  (2i, 2i+1) on even layers and (2i+1, 2i+2) on odd layers (20 qubits x 20
  layers is the reference's CPU-runnable plain workload).
* ``random_blocks``: the at-scale plain workload: per layer one random gate
  from {h, s, t, sx, rx, ry, rz} on every qubit, {cz, rzz, cp} on a random
  perfect matching, and 8 remote_link_psi_plus + 8 nonlocal_cz (generic
  ``gate`` kind) between the two halves of the register; after every 10th
  layer, measure two random qubits, reset them, and apply one x conditioned
  on the first measured bit.
* ``shots_blocks``: ``random_blocks`` plus a final measurement of every qubit,
  the shots workload.
* ``hea`` / ``hea_batch``: hardware-efficient ansatz circuits for batching:
  per layer ry then rz on every qubit, then a cx ring cx(i, i+1), i = 0..n-2,
  closed by cx(n-1, 0); angles U[0, 2pi).  ``hea_batch`` draws 4096 widths
  n ~ U{10..24} and depths d ~ U{8..64} by default.
"""
from __future__ import annotations

import math

import numpy as np

from . import circuit as C

TWO_PI = 2.0 * math.pi


def brickwork_u3(n: int = 20, depth: int = 20, seed: int = 42) -> C.Circuit:
    rng = np.random.default_rng(seed)
    ins = []
    for layer in range(depth):
        ang = rng.uniform(0.0, TWO_PI, size=(n, 3))
        for q in range(n):
            ins.append(C.U3Instruction(qubit=q, theta=float(ang[q, 0]), phi=float(ang[q, 1]),
                                       lam=float(ang[q, 2])))
        start = 0 if layer % 2 == 0 else 1
        for a in range(start, n - 1, 2):
            ins.append(C.CxInstruction(control=a, target=a + 1))
    return C.circuit(n, ins)


_ONE_Q = ("h", "s", "t", "sx", "rx", "ry", "rz")
_PARAM = {"rx": "theta", "ry": "theta", "rz": "phi"}
_TWO_Q = ("cz", "rzz", "cp")


def random_blocks(n: int = 30, depth: int = 40, seed: int = 42, remote_per_kind: int = 8,
                  measure_every: int = 10) -> C.Circuit:
    rng = np.random.default_rng(seed)
    half = n // 2
    block_a = list(range(half))
    block_b = list(range(half, n))
    ins = []
    n_meas = 0
    for layer in range(depth):
        for q in range(n):
            g = _ONE_Q[int(rng.integers(len(_ONE_Q)))]
            if g in _PARAM:
                ins.append(C.instruction(g, qubit=q, **{_PARAM[g]: float(rng.uniform(0.0, TWO_PI))}))
            else:
                ins.append(C.instruction(g, qubit=q))
        perm = rng.permutation(n)
        for i in range(0, n - 1, 2):
            a, b = int(perm[i]), int(perm[i + 1])
            g = _TWO_Q[int(rng.integers(len(_TWO_Q)))]
            if g == "cz":
                ins.append(C.CzInstruction(control=a, target=b))
            elif g == "rzz":
                ins.append(C.RzzInstruction(a=a, b=b, theta=float(rng.uniform(0.0, TWO_PI))))
            else:
                ins.append(C.CpInstruction(control=a, target=b, lam=float(rng.uniform(0.0, TWO_PI))))
        for name in ("remote_link_psi_plus", "nonlocal_cz"):
            for _ in range(remote_per_kind):
                a = int(rng.choice(block_a))
                b = int(rng.choice(block_b))
                ins.append(C.GateInstruction(name=name, params=[], qubits=[a, b]))
        if measure_every and (layer + 1) % measure_every == 0:
            q1, q2, q3 = (int(x) for x in rng.choice(n, size=3, replace=False))
            c1, c2 = n_meas, n_meas + 1
            n_meas += 2
            ins.append(C.MeasureInstruction(qubit=q1, cbit=c1))
            ins.append(C.MeasureInstruction(qubit=q2, cbit=c2))
            ins.append(C.ResetInstruction(qubit=q1))
            ins.append(C.ResetInstruction(qubit=q2))
            ins.append(C.ConditionalInstruction(
                condition=C.Condition(creg_base=c1, creg_size=1, creg_value=1),
                op=C.XInstruction(qubit=q3)))
    return C.circuit(n, ins, n_cbits=max(n_meas, 1))


def gate_count(circ: C.Circuit) -> int:
    """Unitary gates in a circuit (measure/reset/barrier excluded; a
    conditional counts as its gate)."""
    cnt = 0
    for i in circ.instructions:
        if i.kind in ("measure", "reset", "barrier", "classical", "id", "u0"):
            continue
        cnt += 1
    return cnt


def random_mixed(n: int, count: int, seed: int, kinds=None) -> C.Circuit:
    """Random circuit over every gate kind of the reference's dispatcher
    (simulator.rs:394-600), used by the parity tests."""
    rng = np.random.default_rng(seed)
    kinds = kinds or [k for k in C._KIND_FIELDS if k not in ("measure", "reset", "conditional", "barrier",
                                                            "classical", "gate")]
    generic = ["remote_link_phi_plus", "remote_link_psi_minus", "remote_link_psi_plus", "nonlocal_cz",
               "remote_cz", "remote_cx", "remote_epr", "remote_barrier", "remote_cu1"]
    ins = []
    qubit_fields = {"qubit", "control", "target", "a", "b", "control1", "control2", "control3", "control4",
                    "target1", "target2"}
    for _ in range(count):
        if rng.random() < 0.1:
            name = generic[int(rng.integers(len(generic)))]
            qs = [int(x) for x in rng.choice(n, size=2, replace=False)]
            ins.append(C.GateInstruction(name=name, params=[], qubits=qs))
            continue
        k = kinds[int(rng.integers(len(kinds)))]
        fields = C._KIND_FIELDS[k]
        qf = [f for f in fields if f in qubit_fields]
        if len(qf) > n:
            continue
        qs = [int(x) for x in rng.choice(n, size=len(qf), replace=False)]
        kw = dict(zip(qf, qs))
        for f in fields:
            if f not in qubit_fields:
                kw[f] = float(rng.uniform(-TWO_PI, TWO_PI))
        ins.append(C.instruction(k, **kw))
    return C.circuit(n, ins)


def hea(n: int, depth: int, angles) -> C.Circuit:
    """HEA circuit; ``angles`` has shape (depth, n, 2): (ry, rz) per qubit."""
    ins = []
    for layer in range(depth):
        for q in range(n):
            ins.append(C.RyInstruction(qubit=q, theta=float(angles[layer][q][0])))
        for q in range(n):
            ins.append(C.RzInstruction(qubit=q, phi=float(angles[layer][q][1])))
        for q in range(n - 1):
            ins.append(C.CxInstruction(control=q, target=q + 1))
        if n > 1:
            ins.append(C.CxInstruction(control=n - 1, target=0))
    return C.circuit(n, ins)


def hea_batch_arrays(count: int = 4096, seed: int = 42, n_range=(10, 24), d_range=(8, 64)):
    """(ns, ds, angles) of ``count`` HEA circuits, n ~ U{n_range}, d ~ U{d_range}
    (inclusive); angles concatenated per instance in (layer, qubit, (ry, rz))
    order, as ``_core.simulate_hea_batch`` takes them."""
    rng = np.random.default_rng(seed)
    ns = rng.integers(n_range[0], n_range[1] + 1, size=count)
    ds = rng.integers(d_range[0], d_range[1] + 1, size=count)
    angles = [rng.uniform(0.0, TWO_PI, size=(int(d), int(n), 2)) for n, d in zip(ns, ds)]
    return ns, ds, angles


def hea_batch(count: int = 4096, seed: int = 42, n_range=(10, 24), d_range=(8, 64)):
    """The same circuits as ``hea_batch_arrays`` as circuit objects."""
    ns, ds, angles = hea_batch_arrays(count, seed, n_range, d_range)
    return [hea(int(n), int(d), a) for n, d, a in zip(ns, ds, angles)]


def shots_blocks(n: int = 12, depth: int = 20, seed: int = 42) -> C.Circuit:
    """``random_blocks`` (mid-circuit measure / reset / conditional x every 10
    layers) followed by a measurement of every qubit into its own cbit; the
    classical register holds those n bits and the mid-circuit ones."""
    base = random_blocks(n, depth, seed=seed)
    mid = sum(1 for i in base.instructions if i.kind == "measure")
    ins = list(base.instructions)
    for q in range(n):
        ins.append(C.MeasureInstruction(qubit=q, cbit=mid + q))
    return C.circuit(n, ins, mid + n)


def mps_generic(n: int, count: int, seed: int, pair_kinds=("cry", "crx", "cu3", "rxx", "cx")) -> C.Circuit:
    """Circuit for the matrix-product-state simulator: u3 on every qubit,
    then `count` gates, 40% one-qubit {u3, rx, ry} on a random qubit, else a
    two-qubit gate from ``pair_kinds`` on a random (not necessarily
    adjacent) pair; angles U[0.3, 2.8] so no gate is close to the identity
    or to a Clifford."""
    rng = np.random.default_rng(seed)
    ang = lambda: float(rng.uniform(0.3, 2.8))  # noqa: E731
    ins = [C.instruction("u3", qubit=q, theta=ang(), phi=ang(), lam=ang()) for q in range(n)]
    for _ in range(count):
        if rng.random() < 0.4:
            k = ["u3", "rx", "ry"][int(rng.integers(3))]
            q = int(rng.integers(n))
            ins.append(C.instruction("u3", qubit=q, theta=ang(), phi=ang(), lam=ang()) if k == "u3"
                       else C.instruction(k, qubit=q, theta=ang()))
            continue
        k = pair_kinds[int(rng.integers(len(pair_kinds)))]
        a, b = (int(x) for x in rng.choice(n, size=2, replace=False))
        if k == "rxx":
            ins.append(C.instruction(k, a=a, b=b, theta=ang()))
        elif k == "cu3":
            ins.append(C.instruction(k, control=a, target=b, theta=ang(), phi=ang(), lam=ang()))
        elif k in ("crx", "cry"):
            ins.append(C.instruction(k, control=a, target=b, theta=ang()))
        else:
            ins.append(C.instruction(k, control=a, target=b))
    return C.circuit(n, ins)


def random_mps(n: int, chi: int, seed: int = 42) -> list:
    """Random matrix-product state of n sites: bond q = min(2^q, 2^(n-q), chi),
    complex normal entries, scaled so that sum |psi|^2 = 1 (the norm is
    computed by the O(n chi^3) transfer-matrix contraction)."""
    rng = np.random.default_rng(seed)
    bond = [min(1 << min(q, n - q), chi) for q in range(n + 1)]
    ts = [(rng.normal(size=(bond[q], 2, bond[q + 1])) + 1j * rng.normal(size=(bond[q], 2, bond[q + 1])))
          / np.sqrt(2.0 * bond[q]) for q in range(n)]
    e = np.ones((1, 1), dtype=np.complex128)
    for a in ts:
        e = a[:, 0, :].conj().T @ e @ a[:, 0, :] + a[:, 1, :].conj().T @ e @ a[:, 1, :]
    s = float(e[0, 0].real) ** (0.5 / n)
    return [np.ascontiguousarray(a / s) for a in ts]


def distributed_layers(nodes: int = 4, per_node: int = 7, depth: int = 20, seed: int = 42):
    """Distributed circuit for the P-block simulator: per layer u3 on every
    qubit, a cx chain inside every node, one remote_cx and one nonlocal_cz
    between random nodes; after every 5th layer a measurement of a random
    qubit and an ry conditioned on it.  Returns (monolithic circuit,
    DistributedCircuit) with the nodes owning consecutive qubits."""
    rng = np.random.default_rng(seed)
    n = nodes * per_node
    ins = []
    for layer in range(depth):
        for q in range(n):
            ins.append(C.instruction("u3", qubit=q, theta=float(rng.uniform(0, TWO_PI)),
                                     phi=float(rng.uniform(0, TWO_PI)), lam=float(rng.uniform(0, TWO_PI))))
        for k in range(nodes):
            for q in range(k * per_node, (k + 1) * per_node - 1):
                ins.append(C.instruction("cx", control=q, target=q + 1))
        for name in ("remote_cx", "nonlocal_cz"):
            a, b = (int(x) for x in rng.choice(nodes, size=2, replace=False))
            qa = a * per_node + int(rng.integers(per_node))
            qb = b * per_node + int(rng.integers(per_node))
            ins.append(C.GateInstruction(name=name, qubits=[qa, qb], params=[]))
        if layer % 5 == 4:
            q = int(rng.integers(n))
            ins.append(C.instruction("measure", qubit=q, cbit=layer // 5))
            ins.append(C.instruction("conditional", condition=C.Condition(layer // 5, 1, 1),
                                     op=C.instruction("ry", qubit=int(rng.integers(n)), theta=0.9)))
    circ = C.circuit(n, ins, n_cbits=max(1, depth // 5))
    qpn = {k: list(range(k * per_node, (k + 1) * per_node)) for k in range(nodes)}
    return circ, C.distribute(circ, qpn)
