"""GPU tests of the complex64 precision form.

The oracle for complex64 results is complex128: the engine's complex128 path
at fusion level 1 is bit-identical to the oracle's restatement of
src/engine.rs (pinned in test_gpu_capi.py), and small cases here are also
checked against the oracle directly.  Tolerances come from FP32 rounding:
unit roundoff 6e-8, a few hundred gates, so amplitude errors of order 1e-6;
the bounds below are 2e-5 absolute per amplitude and 1e-5 fidelity deficit.

Fusion level 3 replaces runs of gates on at most 5 qubits with their 32x32
product (formed in FP64 on the host) applied as complex GEMMs on FP32 MFMA
(v_mfma_f32_32x32x2f32).  A transposed operand, a wrong output-lane mapping,
a dropped gate or a wrong embedding of a gate in the block fails these tests
by O(1), far above the tolerance.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

capi = pytest.importorskip("dqsim_bosonic_amd.capi")

GATES_1Q = ["h", "s", "t", "sx", "u3", "rx", "ry", "rz", "x", "y"]
GATES_2Q = ["cx", "cz", "cp", "rzz", "rxx", "swap", "cry", "cu3"]
GATES_3Q = ["ccx", "cswap", "rccx"]
NPARAM = {"u3": 3, "rx": 1, "ry": 1, "rz": 1, "cp": 1, "rzz": 1, "rxx": 1, "cry": 1, "cu3": 3}
AMP_TOL = 2e-5
FID_TOL = 1e-5


def local_program(n, count, rng, window=(0, None), controls=True):
    """Random gates, mostly inside a qubit window (so that runs fuse), with
    some 3-qubit gates and controlled 1-qubit gates with control values."""
    lo, hi = window[0], window[1] if window[1] is not None else n
    prog = []
    for _ in range(count):
        r = rng.random()
        if r < 0.5:
            g = GATES_1Q[int(rng.integers(len(GATES_1Q)))]
            prog.append(("1q", orc.gate_matrix(g, rng.uniform(-4, 4, NPARAM.get(g, 0))), [int(rng.integers(lo, hi))]))
        elif r < 0.85:
            g = GATES_2Q[int(rng.integers(len(GATES_2Q)))]
            qs = [int(x) for x in rng.choice(np.arange(lo, hi), size=2, replace=False)]
            prog.append(("nq", orc.gate_matrix(g, rng.uniform(-4, 4, NPARAM.get(g, 0))), qs))
        elif r < 0.93 or not controls:
            g = GATES_3Q[int(rng.integers(len(GATES_3Q)))]
            qs = [int(x) for x in rng.choice(np.arange(lo, hi), size=3, replace=False)]
            prog.append(("nq", orc.gate_matrix(g, []), qs))
        else:
            qs = [int(x) for x in rng.choice(np.arange(lo, hi), size=3, replace=False)]
            u = orc.gate_matrix("u3", rng.uniform(-4, 4, 3))
            prog.append(("ctrl", u, qs, [int(rng.integers(2)), int(rng.integers(2))]))
    return prog


def apply(st, prog):
    for item in prog:
        if item[0] == "1q":
            st.apply_1q(item[1], item[2][0])
        elif item[0] == "nq":
            st.apply_nq(item[1], item[2])
        else:
            st.apply_1q(item[1], item[2][2], controls=list(zip(item[2][:2], item[3])))


def oracle_run(n, prog, psi):
    ref = psi.copy()
    for item in prog:
        if item[0] == "1q":
            orc.apply_one_qubit(ref, n, item[1], item[2][0])
        elif item[0] == "nq":
            orc.apply_n_qubit(ref, n, item[1], item[2])
        else:
            c0, c1, t = item[2]
            v0, v1 = item[3]
            idx = np.arange(1 << n)
            mask = (((idx >> c0) & 1) == v0) & (((idx >> c1) & 1) == v1)
            sub = ref.copy()
            orc.apply_one_qubit(sub, n, item[1], t)
            ref = np.where(mask, sub, ref)
    return ref


def random_state(n, rng):
    psi = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    return psi / np.linalg.norm(psi)


def run_c64(n, prog, psi, fuse, tile_bits=0, min_cost=0.0, pair=False):
    """pair=False: blocks are formed from the gates as given (pair fusion,
    on by default, would first merge neighbouring gates)."""
    st = capi.State(n, capi.C64)
    st.set_options(fuse, tile_bits)
    st.set_fuse_min_cost(min_cost)
    st.set_pair_fusion(pair)
    st.set(psi)
    apply(st, prog)
    return st


@pytest.mark.parametrize("n,tile_bits,window", [(12, 12, (0, 12)), (14, 12, (0, 8)), (14, 13, (3, 11)),
                                                 (16, 13, (0, 16)), (16, 12, (9, 16))])
def test_fused5_matches_oracle(n, tile_bits, window):
    rng = np.random.default_rng(100 + n + tile_bits)
    psi = random_state(n, rng)
    prog = local_program(n, 160, rng, window)
    want = oracle_run(n, prog, psi)
    st = run_c64(n, prog, psi, fuse=3, tile_bits=tile_bits)
    got = st.get()  # flushes the queue; the block counters are updated then
    blocks, fused_gates = st.fused_stats()
    assert blocks > 0 and fused_gates >= 4 * blocks, (blocks, fused_gates)
    err = np.max(np.abs(got - want))
    assert err <= AMP_TOL, err
    # unfused complex64 for the same program: same error scale
    err2 = np.max(np.abs(run_c64(n, prog, psi, fuse=2, tile_bits=tile_bits).get() - want))
    assert err2 <= AMP_TOL, err2


def test_fused5_single_block_known_answer():
    """One block on 5 qubits: 4 single-qubit gates and a cz.  The block's
    product is checked against the oracle on a basis state, so every output
    amplitude is one column of the fused matrix."""
    n = 12
    prog = [("1q", orc.gate_matrix("h", []), [2]), ("1q", orc.gate_matrix("rx", [0.3]), [5]),
            ("1q", orc.gate_matrix("t", []), [7]), ("nq", orc.gate_matrix("cz", []), [2, 7]),
            ("1q", orc.gate_matrix("ry", [1.1]), [11]), ("nq", orc.gate_matrix("rxx", [0.7]), [5, 11])]
    for basis in (0, 1 << 2, (1 << 5) | (1 << 11), 0xFFF, 0x5A5):
        psi = np.zeros(1 << n, dtype=complex)
        psi[basis] = 1.0
        st = run_c64(n, prog, psi, fuse=3, tile_bits=12)
        got = st.get()
        assert st.fused_stats() == (1, len(prog))
        want = oracle_run(n, prog, psi)
        assert np.max(np.abs(got - want)) <= 1e-6


def test_fused5_with_measurement_and_outer_ops():
    """Blocks interleaved with collapses and gates on outer qubits."""
    n = 15
    rng = np.random.default_rng(7)
    psi = random_state(n, rng)
    ref = capi.State(n)
    ref.set_options(1)
    ref.set(psi)
    st = capi.State(n, capi.C64)
    st.set_options(3, 12)
    st.set_fuse_min_cost(0.0)
    st.set_pair_fusion(False)
    st.set(psi)
    for rnd in range(3):
        prog = local_program(n, 60, rng, (0, n))
        apply(ref, prog)
        apply(st, prog)
        q = int(rng.integers(n))
        p1r, p1 = ref.prob_one(q), st.prob_one(q)
        assert abs(p1r - p1) <= 1e-5
        out = 1 if p1r >= 0.5 else 0
        nr = np.sqrt(p1r if out else 1 - p1r)
        ref.collapse(q, out, nr)
        st.collapse(q, out, np.sqrt(p1 if out else 1 - p1))
    cmp = st.compare(ref)
    assert cmp["max_abs"] <= AMP_TOL
    assert 1 - cmp["fidelity"] <= FID_TOL
    assert st.fused_stats()[0] > 0


def test_fusion3_requires_complex64():
    st = capi.State(10)
    with pytest.raises(capi.DqError):
        st.set_options(3)


def test_compare_states_against_numpy():
    n = 13
    rng = np.random.default_rng(3)
    a, b = random_state(n, rng), random_state(n, rng)
    b = 0.9 * a + 0.1 * b
    b /= np.linalg.norm(b)
    for da, db in ((capi.C128, capi.C128), (capi.C64, capi.C128), (capi.C128, capi.C64)):
        sa, sb = capi.State(n, da), capi.State(n, db)
        sa.set(a)
        sb.set(b)
        ha, hb = sa.get(), sb.get()
        c = sa.compare(sb)
        assert abs(c["max_abs"] - np.max(np.abs(ha - hb))) <= 1e-12
        assert abs(c["inner"] - np.vdot(ha, hb)) <= 1e-12
        assert abs(c["tvd"] - 0.5 * np.sum(np.abs(np.abs(ha) ** 2 - np.abs(hb) ** 2))) <= 1e-12


def test_complex64_renormalisation_keeps_norm():
    """200 collapse/rotate rounds on complex64: the FP64 renormalisation
    keeps the norm at 1 to FP32 rounding instead of drifting."""
    n = 14
    rng = np.random.default_rng(11)
    st = capi.State(n, capi.C64)
    st.set(random_state(n, rng))
    for _ in range(200):
        q = int(rng.integers(n))
        st.apply_1q(orc.gate_matrix("ry", [float(rng.uniform(0.3, 1.2))]), q)
        p1 = st.prob_one(q)
        out = int(p1 >= 0.5)
        st.collapse(q, out, np.sqrt(p1 if out else 1 - p1))
    norm = np.linalg.norm(st.get())
    assert abs(norm - 1.0) <= 1e-5, norm


def test_fuse_cost_threshold():
    """The default cost threshold leaves cheap blocks unfused; a threshold of
    0 fuses them; both give the oracle's state to FP32 rounding."""
    n = 13
    rng = np.random.default_rng(21)
    psi = random_state(n, rng)
    prog = local_program(n, 120, rng, (0, 6))
    want = oracle_run(n, prog, psi)
    counts = []
    for cost in (0.0, 128.0, 1e9):
        st = run_c64(n, prog, psi, fuse=3, tile_bits=12, min_cost=cost)
        assert np.max(np.abs(st.get() - want)) <= AMP_TOL
        counts.append(st.fused_stats()[0])
    assert counts[0] >= counts[1] >= counts[2] == 0 and counts[0] > 0, counts
    with pytest.raises(capi.DqError):
        capi.State(8, capi.C64).set_fuse_min_cost(-1.0)


@pytest.mark.parametrize("fuse", [2, 3])
def test_pair_fusion_then_blocks_match_oracle(fuse):
    """Default planning (pair fusion on) at fusion levels 2 and 3."""
    n = 14
    rng = np.random.default_rng(31)
    psi = random_state(n, rng)
    prog = local_program(n, 200, rng, (0, 9))
    want = oracle_run(n, prog, psi)
    st = run_c64(n, prog, psi, fuse=fuse, tile_bits=12, min_cost=0.0, pair=True)
    assert np.max(np.abs(st.get() - want)) <= AMP_TOL


def test_blocks_with_outer_qubit_gates():
    """Gates across tile and outer qubits: pair fusion gives ops with one
    matrix per outer-qubit value; those must stay out of 5-qubit blocks."""
    n = 16
    rng = np.random.default_rng(41)
    psi = random_state(n, rng)
    prog = []
    for layer in range(12):
        for q in range(n):
            g = GATES_1Q[int(rng.integers(len(GATES_1Q)))]
            prog.append(("1q", orc.gate_matrix(g, rng.uniform(-4, 4, NPARAM.get(g, 0))), [q]))
        perm = rng.permutation(n)
        for i in range(0, n - 1, 2):
            g = ["cz", "cp", "rzz"][int(rng.integers(3))]
            prog.append(("nq", orc.gate_matrix(g, rng.uniform(-4, 4, NPARAM.get(g, 0))), [int(perm[i]), int(perm[i + 1])]))
    want = oracle_run(n, prog, psi)
    for cost in (0.0, 128.0):
        st = run_c64(n, prog, psi, fuse=3, tile_bits=12, min_cost=cost, pair=True)
        assert np.max(np.abs(st.get() - want)) <= AMP_TOL
