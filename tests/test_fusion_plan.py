"""CPU tests of the planner's gate fusion (no device needed).

Fusion level 3 groups gates into blocks on at most 5 qubits and multiplies
each block into one 32x32 matrix in FP64; pair fusion (levels 2 and 3)
merges neighbouring gates on at most 2 qubits.  Gates moved past others
must commute with them.  The check is the mathematics: applying the rewritten list (fused
blocks plus the gates left as they are) to a random state must give the same
state as the original gate sequence, to FP64 rounding, computed with the
oracle's gate application.  A wrong embedding order, a reordering across a
shared qubit or a dropped gate fails by O(1).
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as orc

core = pytest.importorskip("dqsim_bosonic_amd._core")

GATES_1Q = ["h", "s", "t", "sx", "u3", "rx", "ry", "rz", "x"]
GATES_2Q = ["cx", "cz", "cp", "rzz", "rxx", "swap", "cry"]
NPARAM = {"u3": 3, "rx": 1, "ry": 1, "rz": 1, "cp": 1, "rzz": 1, "rxx": 1, "cry": 1}


def random_ops(n, count, rng):
    ops = []
    for _ in range(count):
        r = rng.random()
        if r < 0.45:
            g = GATES_1Q[int(rng.integers(len(GATES_1Q)))]
            ops.append(([int(rng.integers(n))], orc.gate_matrix(g, rng.uniform(-4, 4, NPARAM.get(g, 0))), []))
        elif r < 0.8:
            g = GATES_2Q[int(rng.integers(len(GATES_2Q)))]
            qs = [int(x) for x in rng.choice(n, size=2, replace=False)]
            ops.append((qs, orc.gate_matrix(g, rng.uniform(-4, 4, NPARAM.get(g, 0))), []))
        elif r < 0.9:
            qs = [int(x) for x in rng.choice(n, size=3, replace=False)]
            ops.append((qs, orc.gate_matrix("ccx", []), []))
        else:
            qs = [int(x) for x in rng.choice(n, size=3, replace=False)]
            ops.append((qs, orc.gate_matrix("u3", rng.uniform(-4, 4, 3)), [int(rng.integers(2)), int(rng.integers(2))]))
    return ops


def apply_op(psi, n, qs, u, cv):
    if not cv:
        if len(qs) == 1:
            orc.apply_one_qubit(psi, n, u, qs[0])
        else:
            orc.apply_n_qubit(psi, n, u, qs)
        return psi
    idx = np.arange(1 << n)
    mask = np.ones(1 << n, dtype=bool)
    for q, v in zip(qs[:-1], cv):
        mask &= ((idx >> q) & 1) == v
    sub = psi.copy()
    orc.apply_one_qubit(sub, n, u, qs[-1])
    return np.where(mask, sub, psi)


@pytest.mark.parametrize("seed", range(8))
@pytest.mark.parametrize("n,tile", [(8, list(range(8))), (9, [0, 1, 2, 3, 5, 7, 8]), (10, [9, 8, 7, 6, 5, 4])])
def test_fused_list_equals_program(seed, n, tile):
    rng = np.random.default_rng(seed)
    ops = random_ops(n, 60, rng)
    plan = core._fuse_blocks(tile, [(qs, u.ravel(), cv) for qs, u, cv in ops])
    psi = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    psi /= np.linalg.norm(psi)
    want = psi.copy()
    for qs, u, cv in ops:
        want = apply_op(want, n, qs, u, cv)
    got = psi.copy()
    seen, fused = [], 0
    for item in plan:
        if len(item) == 1:
            qs, u, cv = ops[item[0]]
            got = apply_op(got, n, qs, u, cv)
            seen.append(item[0])
        else:
            qs, m, cnt, _ = item
            m = m[0]
            assert len(set(qs)) == 5 and all(q in tile for q in qs)
            assert np.allclose(m.conj().T @ m, np.eye(32), atol=1e-12)
            orc.apply_n_qubit(got, n, m, qs)
            fused += cnt
    assert len(seen) + fused == len(ops)
    assert len(set(seen)) == len(seen)
    assert np.max(np.abs(got - want)) <= 1e-12


def test_blocks_form_on_layered_circuit():
    """A brickwork of 1-qubit layers and neighbour cz on 12 tile qubits
    fuses most gates into blocks of at least 4 gates."""
    n = 12
    rng = np.random.default_rng(1)
    ops = []
    for layer in range(6):
        for q in range(n):
            ops.append(([q], orc.gate_matrix("u3", rng.uniform(-4, 4, 3)).ravel(), []))
        for q in range(layer % 2, n - 1, 2):
            ops.append(([q, q + 1], orc.gate_matrix("cz", []).ravel(), []))
    plan = core._fuse_blocks(list(range(n)), ops)
    fused = sum(item[2] for item in plan if len(item) == 4)
    blocks = sum(1 for item in plan if len(item) == 4)
    assert blocks > 0 and fused >= 0.8 * len(ops), (blocks, fused, len(ops))


def pass_ops(n, tile, count, rng):
    """Ops as a tile pass holds them: non-diagonal action only on tile
    qubits; qubits outside the tile appear only in diagonal gates and as
    controls (with control values) of controlled 2x2 gates."""
    outer = [q for q in range(n) if q not in tile]
    ops = []
    for _ in range(count):
        r = rng.random()
        if r < 0.3:
            g = GATES_1Q[int(rng.integers(len(GATES_1Q)))]
            ops.append(([int(rng.choice(tile))], orc.gate_matrix(g, rng.uniform(-4, 4, NPARAM.get(g, 0))), []))
        elif r < 0.5:
            g = GATES_2Q[int(rng.integers(len(GATES_2Q)))]
            qs = [int(x) for x in rng.choice(tile, size=2, replace=False)]
            ops.append((qs, orc.gate_matrix(g, rng.uniform(-4, 4, NPARAM.get(g, 0))), []))
        elif r < 0.75:  # diagonal gate on any qubits
            g = ["cz", "cp", "rzz", "t", "rz", "s"][int(rng.integers(6))]
            u = orc.gate_matrix(g, rng.uniform(-4, 4, NPARAM.get(g, 0)))
            k = int(np.log2(u.shape[0]))
            ops.append(([int(x) for x in rng.choice(n, size=k, replace=False)], u, []))
        elif r < 0.9 and outer:  # controlled 2x2, target in the tile, controls anywhere
            t = int(rng.choice(tile))
            cs = [int(x) for x in rng.choice([q for q in range(n) if q != t], size=2, replace=False)]
            ops.append((cs + [t], orc.gate_matrix("u3", rng.uniform(-4, 4, 3)),
                        [int(rng.integers(2)), int(rng.integers(2))]))
        else:
            qs = [int(x) for x in rng.choice(tile, size=3, replace=False)]
            ops.append((qs, orc.gate_matrix("ccx", []), []))
    return ops


def apply_selected(psi, n, qs, mats, sq):
    """One matrix per value of the selector qubits sq (bit j = qubit sq[j])."""
    idx = np.arange(1 << n)
    out = psi.copy()
    for sv in range(mats.shape[0]):
        mask = np.ones(1 << n, dtype=bool)
        for j, q in enumerate(sq):
            mask &= ((idx >> q) & 1) == ((sv >> j) & 1)
        sub = psi.copy()
        if len(qs) == 1:
            orc.apply_one_qubit(sub, n, mats[sv], qs[0])
        else:
            orc.apply_n_qubit(sub, n, mats[sv], qs)
        out = np.where(mask, sub, out)
    return out


@pytest.mark.parametrize("seed", range(10))
@pytest.mark.parametrize("n,tile", [(8, list(range(8))), (9, [0, 1, 2, 3, 5, 7, 8]), (10, [9, 8, 7, 6, 2, 1])])
def test_pair_fused_list_equals_program(seed, n, tile):
    rng = np.random.default_rng(100 + seed)
    ops = pass_ops(n, tile, 80, rng)
    plan = core._fuse_blocks(tile, [(qs, u.ravel(), cv) for qs, u, cv in ops], max_qubits=2)
    psi = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    psi /= np.linalg.norm(psi)
    want = psi.copy()
    for qs, u, cv in ops:
        want = apply_op(want, n, qs, u, cv)
    got = psi.copy()
    seen = []
    for item in plan:
        if len(item) == 1:
            qs, u, cv = ops[item[0]]
            got = apply_op(got, n, qs, u, cv)
            seen.append(item[0])
        else:
            qs, mats, _, sq = item
            assert all(q not in tile for q in sq) and len(sq) <= 3
            if len(qs) + len(sq) > 2 or sq:
                assert 1 <= len(qs) <= 2
            for m in mats:
                assert np.allclose(m.conj().T @ m, np.eye(m.shape[0]), atol=1e-12)
            got = apply_selected(got, n, qs, mats, sq)
    assert len(set(seen)) == len(seen)
    assert len(plan) < len(ops)
    assert np.max(np.abs(got - want)) <= 1e-12


def test_pair_fusion_on_layered_circuit():
    """1q layers + cz pairs: each pair's gates become one op."""
    n = 10
    rng = np.random.default_rng(2)
    ops = []
    for layer in range(4):
        for q in range(n):
            ops.append(([q], orc.gate_matrix("u3", rng.uniform(-4, 4, 3)).ravel(), []))
        for q in range(layer % 2, n - 1, 2):
            ops.append(([q, q + 1], orc.gate_matrix("cz", []).ravel(), []))
    plan = core._fuse_blocks(list(range(n)), ops, max_qubits=2)
    assert len(plan) <= len(ops) // 2, (len(plan), len(ops))


@pytest.mark.parametrize("seed", range(8))
def test_pairs_then_blocks_equal_program(seed):
    """Fusion level 3's pipeline: pair fusion, then 5-qubit blocks, on ops
    that touch outer qubits diagonally (selector ops must stay unblocked)."""
    n, tile = 10, [0, 1, 2, 3, 4, 5, 6, 7]
    rng = np.random.default_rng(300 + seed)
    ops = pass_ops(n, tile, 90, rng)
    plan = core._fuse_blocks(tile, [(qs, u.ravel(), cv) for qs, u, cv in ops], min_cost=0.0, max_qubits=7)
    psi = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    psi /= np.linalg.norm(psi)
    want = psi.copy()
    for qs, u, cv in ops:
        want = apply_op(want, n, qs, u, cv)
    got = psi.copy()
    for item in plan:
        if len(item) == 1:
            qs, u, cv = ops[item[0]]
            got = apply_op(got, n, qs, u, cv)
        else:
            qs, mats, _, sq = item
            got = apply_selected(got, n, qs, mats, sq)
    assert np.max(np.abs(got - want)) <= 1e-12


def _decode_phase_form(v, enc):
    """The 4x4 matrix that a V_ROWPH (4) / V_COLPH (5) encoding applies: R from
    the real parts, phase k from the imaginary parts of entries 4k and 4k + 1."""
    e = np.asarray(enc, dtype=np.float64).reshape(16, 2)
    r = e[:, 0].reshape(4, 4)
    ph = np.array([e[4 * k, 1] + 1j * e[4 * k + 1, 1] for k in range(4)])
    return ph[:, None] * r if v == 4 else r * ph[None, :]


def _flat(u):
    return np.stack([u.real, u.imag], axis=-1).ravel().tolist()


@pytest.mark.parametrize("seed", range(6))
def test_phase_forms_factor_and_reproduce_the_matrix(seed):
    """U = D R (rows with one phase each) is V_ROWPH, U = R D is V_COLPH, and
    the encoding reproduces U to rounding; a generic unitary stays V_GEN.  A
    wrong phase index, a transposed factor or a dropped conjugate fails by
    O(1)."""
    rng = np.random.default_rng(seed)
    rq, _ = np.linalg.qr(rng.normal(size=(4, 4)))  # real orthogonal
    d = np.exp(1j * rng.uniform(-np.pi, np.pi, size=4))
    for u, want in ((d[:, None] * rq, 4), (rq * d[None, :], 5)):
        v, enc = core._phase_factor4(_flat(u))
        assert v == want
        assert np.max(np.abs(_decode_phase_form(v, enc) - u)) <= 4e-16
    g, _ = np.linalg.qr(rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4)))
    v, enc = core._phase_factor4(_flat(g))
    assert v == 0 and np.array_equal(np.asarray(enc), np.asarray(_flat(g)))
    # the HEA step RZ_q RY_q CX(q, q+1) is a diagonal times a real matrix
    ry = orc.gate_matrix("ry", [rng.uniform(-4, 4)])
    rz = orc.gate_matrix("rz", [rng.uniform(-4, 4)])
    u = np.kron(rz @ ry, np.eye(2)) @ orc.gate_matrix("cx", [])
    v, enc = core._phase_factor4(_flat(u))
    assert v == 4 and np.max(np.abs(_decode_phase_form(v, enc) - u)) <= 4e-16


def test_phase_forms_several_selector_blocks_and_a_near_miss():
    """Every selector block must factor (then each is encoded); an imaginary
    part above the 4e-15 tolerance in one row keeps the op general."""
    rng = np.random.default_rng(11)
    blocks = []
    for _ in range(2):
        rq, _ = np.linalg.qr(rng.normal(size=(4, 4)))
        blocks.append(np.exp(1j * rng.uniform(-3, 3, size=4))[:, None] * rq)
    v, enc = core._phase_factor4(_flat(blocks[0]) + _flat(blocks[1]))
    assert v == 4
    for b in range(2):
        assert np.max(np.abs(_decode_phase_form(v, enc[32 * b:32 * b + 32]) - blocks[b])) <= 4e-16
    bad = blocks[1].copy()
    bad[2, 1] *= np.exp(1j * 1e-9)
    v, _ = core._phase_factor4(_flat(blocks[0]) + _flat(bad))
    assert v == 0


def test_phase_forms_only_with_pair_fusion():
    """The HEA family's fused 4x4 ops take the row-phase form at fusion 2;
    fusion 1 (the exact kernels, reference arithmetic) never uses it."""
    rng = np.random.default_rng(3)
    n = 14
    ops = []
    for _ in range(4):
        for q in range(n):
            ops.append(([q], orc.gate_matrix("ry", [rng.uniform(0, 6)]).ravel(), []))
        for q in range(n):
            ops.append(([q], orc.gate_matrix("rz", [rng.uniform(0, 6)]).ravel(), []))
        for q in range(n - 1):
            ops.append(([q, q + 1], orc.gate_matrix("cx", []).ravel(), []))
    variants = {}
    for fuse in (1, 2):
        vs = [o[1] for ps in core._plan_passes(n, ops, fuse, 12, 4, 4) for o in ps["ops"] if o[0] == 2]
        variants[fuse] = vs
    assert variants[2].count(4) >= len(variants[2]) // 2
    assert 4 not in variants[1] and 5 not in variants[1]


def test_phase_forms_reject_non_finite_entries():
    """A NaN or infinite entry keeps a 4x4 general (its encoding would not
    reproduce it)."""
    rq, _ = np.linalg.qr(np.random.default_rng(2).normal(size=(4, 4)))
    for bad in (np.nan, np.inf):
        u = rq.astype(np.complex128)
        u[1, 2] = bad
        v, _ = core._phase_factor4(_flat(u))
        assert v == 0
