"""CPU tests of the pass builder (no device needed).

A pass is one HBM sweep over tiles of b qubits; the builder may move an op
ahead of earlier ops only when they commute (disjoint qubits, or both
diagonal).  Checks: every op lands in exactly one pass; each op's
non-diagonal qubits are tile qubits of its pass (a controlled 2x2 needs
only its target); applying the ops in pass order to a random state gives the
program's state to FP64 rounding (oracle gate application), so a reordering
across a shared qubit fails by O(1); and on layered circuits the tile choice
needs no more passes than taking qubits in order of first use.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as orc
from tests.test_fusion_plan import apply_op, random_ops

core = pytest.importorskip("dqsim_bosonic_amd._core")


def is_diag(u, cv):
    if cv:
        return False
    d = np.asarray(u).reshape(int(np.sqrt(np.asarray(u).size)), -1)
    return np.count_nonzero(d - np.diag(np.diag(d))) == 0


def is_controlled(u):
    """Identity outside the bottom-right 2x2 block: a 2x2 on qubits[-1]
    under controls qubits[:-1] all 1 (the planner classifies it so)."""
    d = np.asarray(u).reshape(int(np.sqrt(np.asarray(u).size)), -1)
    if d.shape[0] < 4:
        return False
    e = d.copy()
    e[-2:, -2:] = np.eye(2)
    return np.array_equal(e, np.eye(d.shape[0]))


def need_mask(qs, u, cv):
    """Tile qubits an op needs, by the planner's classification order:
    diagonal (none), then controlled 2x2 (the target), else all."""
    if is_diag(u, cv):
        return 0
    if cv or is_controlled(u):
        return 1 << qs[-1]
    return sum(1 << q for q in qs)


def layered_ops(n, depth, rng):
    """Random 1q gate on every qubit (h/sx/ry dense, rz/t diagonal), then a
    random pairing with diagonal 2q gates, per layer."""
    ops = []
    for _ in range(depth):
        for q in range(n):
            g = ["h", "sx", "ry", "rz", "t"][int(rng.integers(5))]
            ops.append(([q], orc.gate_matrix(g, rng.uniform(0, 6, 1 if g in ("ry", "rz") else 0)), []))
        perm = rng.permutation(n)
        for i in range(0, n - 1, 2):
            ops.append(([int(perm[i]), int(perm[i + 1])], orc.gate_matrix("rzz", rng.uniform(0, 6, 1)), []))
    return ops


def first_use_passes(ops, n, b, low):
    """Pass count when a pass takes tile qubits in order of first use (the
    builder's rule at fusion level 1, with level 2's lookahead)."""
    rem = list(range(len(ops)))
    count = 0
    while rem:
        S = (1 << low) - 1
        nd = dd = 0
        rest = []
        for i in rem:
            qs, u, cv = ops[i]
            qm = sum(1 << q for q in qs)
            diag = is_diag(u, cv)
            block = nd if diag else (nd | dd)
            if qm & block:
                if diag:
                    dd |= qm
                else:
                    nd |= qm
                rest.append(i)
                continue
            need = need_mask(qs, u, cv)
            if bin(S | need).count("1") <= b:
                S |= need
            else:
                nd |= qm
                rest.append(i)
        count += 1
        rem = rest
    return count


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("n,tile_bits,low_bits", [(12, 9, 4), (13, 9, 3), (14, 12, 4)])
def test_passes_equal_program(seed, n, tile_bits, low_bits):
    rng = np.random.default_rng(100 + seed)
    ops = random_ops(n, 150, rng)
    # a share of diagonal ops on arbitrary qubits (they may move past each other)
    for _ in range(30):
        qs = [int(x) for x in rng.choice(n, size=2, replace=False)]
        ops.insert(int(rng.integers(len(ops))), (qs, orc.gate_matrix("cp", rng.uniform(-4, 4, 1)), []))
    passes = core._schedule(n, [(qs, np.asarray(u).ravel(), cv) for qs, u, cv in ops], 2, tile_bits, low_bits)
    seen = [i for _, idx in passes for i in idx]
    assert sorted(seen) == list(range(len(ops)))
    for S, idx in passes:
        assert bin(S).count("1") == tile_bits
        assert S & ((1 << low_bits) - 1) == (1 << low_bits) - 1
        for i in idx:
            qs, u, cv = ops[i]
            assert need_mask(qs, u, cv) & ~S == 0, (i, qs, hex(S))
    psi0 = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    psi0 /= np.linalg.norm(psi0)
    want = psi0.copy()
    for qs, u, cv in ops:
        want = apply_op(want, n, qs, u, cv)
    got = psi0.copy()
    for i in seen:
        got = apply_op(got, n, *ops[i])
    assert np.max(np.abs(got - want)) < 1e-12


@pytest.mark.parametrize("seed", range(3))
@pytest.mark.parametrize("n,tile_bits,low_bits", [(20, 12, 4), (24, 12, 4), (24, 12, 3)])
def test_tile_choice_beats_first_use(seed, n, tile_bits, low_bits):
    rng = np.random.default_rng(7 + seed)
    ops = layered_ops(n, 10, rng)
    passes = core._schedule(n, [(qs, np.asarray(u).ravel(), cv) for qs, u, cv in ops], 2, tile_bits, low_bits)
    assert sorted(i for _, idx in passes for i in idx) == list(range(len(ops)))
    assert len(passes) <= first_use_passes(ops, n, tile_bits, low_bits)


def brute_swizzle(i, b):
    """The kernel's LDS slot of tile index i: the low nibble XORed with each
    higher nibble of the b tile bits."""
    s = 0
    sh = 4
    while sh < b:
        s ^= (i >> sh) & 15
        sh += 4
    return i ^ s


@pytest.mark.parametrize("b", [9, 12, 13])
def test_swizzle_is_a_permutation_linear_over_xor(b):
    idx = np.arange(1 << b)
    slots = np.array([core._tile_swizzle(int(i), b) for i in idx])
    assert sorted(slots) == list(idx)
    assert all(core._tile_swizzle(int(i), b) == brute_swizzle(int(i), b) for i in idx[:: 37])
    rng = np.random.default_rng(b)
    for _ in range(200):
        x, y = (int(v) for v in rng.integers(1 << b, size=2))
        assert core._tile_swizzle(x ^ y, b) == core._tile_swizzle(x, b) ^ core._tile_swizzle(y, b)


@pytest.mark.parametrize("b,elem", [(9, 16), (12, 16), (12, 8), (13, 16)])
def test_layout_table_matches_register_bits(b, elem):
    """Entry j of a layout's table is the swizzled slot of j's register bits
    (register slot m -> tile bit rbit[m], packed ascending) times the
    amplitude size: slots within the tile, distinct over j, and XOR-linear, so
    that swz(thread bits) ^ table[j] is the slot of the full index."""
    rng = np.random.default_rng(b * elem)
    for _ in range(20):
        rbits = sorted(int(v) for v in rng.choice(b, size=4, replace=False))
        rb = sum(q << (4 * m) for m, q in enumerate(rbits))
        table = core._layout_table(rb, b, elem)
        assert len(table) == 16
        for j in range(16):
            v = sum(1 << rbits[m] for m in range(4) if (j >> m) & 1)
            assert table[j] == brute_swizzle(v, b) * elem
        assert len(set(table)) == 16 and max(table) < (elem << b)
        for j1 in range(16):
            for j2 in range(16):
                assert table[j1 ^ j2] == table[j1] ^ table[j2]


def test_tile_choice_cost_at_flush_size():
    """The engine flushes at 4096 queued ops; planning them for a 30-qubit
    state stays far below the time of one sweep (about 10 ms)."""
    import time

    rng = np.random.default_rng(5)
    ops = layered_ops(30, 90, rng)[:4096]
    args = [(qs, np.asarray(u).ravel(), cv) for qs, u, cv in ops]
    t0 = time.perf_counter()
    passes = core._schedule(30, args, 2, 12, 4)
    elapsed = time.perf_counter() - t0
    assert sorted(i for _, idx in passes for i in idx) == list(range(len(ops)))
    assert elapsed < 0.5, elapsed


@pytest.mark.parametrize("seed", range(8))
@pytest.mark.parametrize("n,tile", [(9, list(range(9))), (12, [0, 1, 2, 3, 5, 6, 8, 9, 10, 11])])
def test_layout_order_equals_program(seed, n, tile):
    """Ops reordered inside a pass so that layouts serve more ops: only
    commuting ops (disjoint tile qubits, or both diagonal; outer qubits are
    touched diagonally) change places, so the state equals the program's."""
    rng = np.random.default_rng(300 + seed)
    ops = []
    for qs, u, cv in random_ops(n, 50, rng):
        d = np.asarray(u).reshape(int(np.sqrt(np.asarray(u).size)), -1)
        diag_only = is_diag(u, cv)
        outer = [q for q in qs if q not in tile]
        # in a tile pass, ops touch outer qubits only diagonally (or as controls)
        if outer and not diag_only and not (cv and qs[-1] in tile):
            continue
        ops.append((qs, u, cv))
    for _ in range(12):
        qs = [int(x) for x in rng.choice(n, size=2, replace=False)]
        ops.insert(int(rng.integers(len(ops) + 1)), (qs, orc.gate_matrix("cp", rng.uniform(-4, 4, 1)), []))
    order = core._order_for_layouts(tile, [(qs, np.asarray(u).ravel(), cv) for qs, u, cv in ops])
    assert sorted(order) == list(range(len(ops)))
    psi0 = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    psi0 /= np.linalg.norm(psi0)
    want = psi0.copy()
    for qs, u, cv in ops:
        want = apply_op(want, n, qs, u, cv)
    got = psi0.copy()
    for i in order:
        got = apply_op(got, n, *ops[i])
    assert np.max(np.abs(got - want)) < 1e-12


def test_layout_order_groups_register_needs():
    """Two interleaved streams of gates on disjoint pairs need a relayout
    per gate in program order; reordered, each pair's gates run together."""
    n = 12
    tile = list(range(n))
    h = orc.gate_matrix("h", [])
    cx = orc.gate_matrix("cx", [])
    ops = []
    for _ in range(4):
        for a, b in ((4, 5), (6, 7), (8, 9)):
            ops.append(([a], h, []))
            ops.append(([a, b], cx, []))
    order = core._order_for_layouts(tile, [(qs, np.asarray(u).ravel(), cv) for qs, u, cv in ops])

    def relayouts(seq):
        """Layout changes a 4-bit register layout needs for this op order,
        each new layout covering the longest run of ops ahead (the planner's
        rule)."""
        seq = list(seq)
        cur, count = set(), 0
        for i, k in enumerate(seq):
            need = set(ops[k][0])
            if need <= cur:
                continue
            cur = set(need)
            for k2 in seq[i + 1:]:
                if set(ops[k2][0]) <= cur:
                    continue
                if len(cur | set(ops[k2][0])) > 4:
                    break
                cur |= set(ops[k2][0])
            count += 1
        return count

    assert relayouts(order) < relayouts(range(len(ops)))


def spreading_ops(n, count, rng):
    """Ops whose qubits spread upward from the low ones (op i draws from the
    lowest 3 + i // 8 qubits), plus diagonals and controlled 2x2s whose
    controls may sit on qubits no op has touched yet (control value 0 or 1)."""
    ops = []
    for i in range(count):
        hi = min(n, 3 + i // 8)
        r = rng.random()
        if r < 0.15:
            qs = [int(x) for x in rng.choice(n, size=2, replace=False)]
            ops.append((qs, orc.gate_matrix("cp", rng.uniform(-4, 4, 1)), []))
        elif r < 0.3:
            c = int(rng.integers(n))
            t = int(rng.choice([q for q in range(hi) if q != c]))
            ops.append(([c, t], orc.gate_matrix("u3", rng.uniform(-4, 4, 3)), [int(rng.integers(2))]))
        else:
            ops.extend(random_ops(hi, 1, rng))
    return ops


def apply_any(psi, n, qs, u, cv):
    """apply_op, or a collapse ([q], [], [outcome]) as the engine applies it:
    the other half set to exact zeros, the kept half divided by its norm."""
    if len(u) == 0:
        idx = np.arange(1 << n)
        keep = ((idx >> qs[0]) & 1) == cv[0]
        nrm = np.sqrt(np.sum(np.abs(psi[keep]) ** 2))
        psi = psi.copy()
        psi[~keep] = 0.0
        if nrm > 0:
            psi[keep] /= nrm
        return psi
    return apply_op(psi, n, qs, u, cv)


def with_collapses(n, ops, rng, count):
    """ops with `count` collapses inserted at random places, each on a qubit
    that an earlier op acted on, to outcome 0 or 1 (one with probability
    above 1e-6), a collapse to 1 followed by an x or a y on the qubit (as a
    reset's)."""
    at = set(int(x) for x in rng.choice(np.arange(10, len(ops)), size=count, replace=False))
    out = []
    psi = np.zeros(1 << n, dtype=complex)
    psi[0] = 1.0
    idx = np.arange(1 << n)
    for i, op in enumerate(ops):
        if i in at:
            q = int(rng.choice(sorted({q for o in out if len(o[1]) for q in o[0]})))
            p0 = float(np.sum(np.abs(psi[((idx >> q) & 1) == 0]) ** 2))
            p1 = float(np.sum(np.abs(psi[((idx >> q) & 1) == 1]) ** 2))
            c = ([q], [], [0 if p0 > 1e-6 and (p1 <= 1e-6 or rng.random() < 0.5) else 1])
            out.append(c)
            psi = apply_any(psi, n, *c)
            if c[2][0] == 1:  # a reset's x, or a y: the qubit is 0 everywhere after it
                x = ([q], orc.gate_matrix("x" if rng.random() < 0.7 else "y", []), [])
                out.append(x)
                psi = apply_any(psi, n, *x)
        out.append(op)
        psi = apply_any(psi, n, *op)
    return out


def check_zero_plan(n, ops, tile_bits, low_bits):
    """Applied to |0...0>, the qubits a flush takes as 0 before a pass are 0
    in every amplitude there (exact zeros in the oracle's statevector), the
    outer index puts them after the others (ascending within each group) and
    n_outer counts the others.  Returns (plan, outer qubits skipped over all
    passes, qubits taken as 0 again after an op acted on them)."""
    plan = core._zero_passes(n, [(qs, np.asarray(u).ravel(), cv) for qs, u, cv in ops], 2, tile_bits, low_bits)
    idx = np.arange(1 << n)
    psi = np.zeros(1 << n, dtype=complex)
    psi[0] = 1.0
    skipped = regained = 0
    touched = 0
    for S, pass_ops, zero_q, n_outer, oq in plan:
        regained += bin(zero_q & touched).count("1")
        outer = [q for q in range(n) if not (S >> q) & 1]
        live = [q for q in outer if not (zero_q >> q) & 1]
        assert oq == live + [q for q in outer if (zero_q >> q) & 1]
        assert n_outer == len(live)
        assert np.all(psi[(idx & zero_q) != 0] == 0)
        skipped += len(outer) - n_outer
        for i in sorted(pass_ops):
            psi = apply_any(psi, n, *ops[i])
        touched |= S
    want = np.zeros(1 << n, dtype=complex)
    want[0] = 1.0
    for op in ops:
        want = apply_any(want, n, *op)
    assert np.max(np.abs(psi - want)) < 1e-12
    return plan, skipped, regained


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("n,tile_bits,low_bits", [(14, 9, 4), (16, 10, 3)])
def test_zero_qubits_before_each_pass(seed, n, tile_bits, low_bits):
    """The check above on programs whose qubits spread upward; without
    collapses the set only shrinks, and the first pass visits one tile."""
    rng = np.random.default_rng(300 + seed)
    ops = spreading_ops(n, 120, rng)
    plan, skipped, _ = check_zero_plan(n, ops, tile_bits, low_bits)
    assert plan[0][3] == 0
    assert skipped > 0
    prev = (1 << n) - 1
    for _, _, zero_q, _, _ in plan:
        assert zero_q & ~prev == 0
        prev = zero_q


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("n,tile_bits,low_bits", [(14, 9, 4), (16, 10, 3)])
def test_zero_qubits_with_collapses(seed, n, tile_bits, low_bits):
    """With collapses between the ops: a qubit collapsed to 0, or collapsed
    to 1 and then flipped by an anti-diagonal 2x2, with no later op acting on
    it other than diagonally, is taken as 0 again, and the check above
    holds."""
    rng = np.random.default_rng(400 + seed)
    ops = with_collapses(n, spreading_ops(n, 120, rng), rng, 24)
    _, _, regained = check_zero_plan(n, ops, tile_bits, low_bits)
    assert regained > 0


def test_zero_qubit_invariant_catches_a_wrong_support():
    """The invariant above fails when a pass's ops are taken as leaving a
    qubit they act on at 0: a hadamard on qubit 0 makes amplitudes with
    bit 0 set nonzero."""
    n = 4
    psi = np.zeros(1 << n, dtype=complex)
    psi[0] = 1.0
    psi = apply_op(psi, n, [0], orc.gate_matrix("h", []), [])
    idx = np.arange(1 << n)
    wrong_zero_q = 0b0001
    assert not np.all(psi[(idx & wrong_zero_q) != 0] == 0)
