"""GPU parity of the C ABI (include/dqsim_gpu.h) against the CPU oracle.

Every entry point is called through ctypes on libdqsim_gpu.so and compared
with the oracle's restatement of src/engine.rs on the same seeded inputs.
Gate applications are bit-exact (np.array_equal on the complex values; the
only allowed difference is the sign of a zero), in both the fused (LDS tile)
and the unfused (one sweep per gate) modes and for several tile sizes.
Reductions whose summation order differs from the reference's are checked to
1e-14 relative.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

capi = pytest.importorskip("dqsim_bosonic_amd.capi")

GATES_1Q = ["x", "y", "z", "h", "s", "sdg", "t", "tdg", "sx", "sxdg", "u3", "u2", "u1", "rx", "ry", "rz"]
GATES_NQ = ["cx", "cz", "cy", "ch", "swap", "csx", "crx", "cry", "crz", "cu1", "cp", "cu3", "cu", "rxx",
            "rzz", "ccx", "cswap", "rccx", "rc3x", "c3x", "c3sqrtx", "c4x", "remote_link_psi_minus",
            "remote_link_psi_plus", "remote_link_phi_plus"]
NPARAM = {"u3": 3, "u2": 2, "u1": 1, "rx": 1, "ry": 1, "rz": 1, "crx": 1, "cry": 1, "crz": 1, "cu1": 1,
          "cp": 1, "cu3": 3, "cu": 4, "rxx": 1, "rzz": 1}


def random_state(n: int, rng) -> np.ndarray:
    psi = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    return psi / np.linalg.norm(psi)


def random_program(n: int, count: int, rng):
    prog = []
    for _ in range(count):
        if rng.random() < 0.5:
            g = GATES_1Q[int(rng.integers(len(GATES_1Q)))]
            prog.append((g, rng.uniform(-7, 7, NPARAM.get(g, 0)), [int(rng.integers(n))]))
        else:
            g = GATES_NQ[int(rng.integers(len(GATES_NQ)))]
            u = orc.gate_matrix(g, rng.uniform(-7, 7, NPARAM.get(g, 0)))
            k = int(np.log2(u.shape[0]))
            if k > n:
                continue
            prog.append((g, None, [int(x) for x in rng.choice(n, size=k, replace=False)], u))
    return prog


def run_both(n, prog, psi, fuse=1, tile_bits=0):
    st = capi.State(n)
    st.set_options(fuse, tile_bits)
    st.set(psi)
    ref = psi.copy()
    for item in prog:
        if item[1] is not None:
            g, params, qs = item
            u = orc.gate_matrix(g, params)
            st.apply_1q(u, qs[0])
            orc.apply_one_qubit(ref, n, u, qs[0])
        else:
            g, _, qs, u = item
            st.apply_nq(u, qs)
            orc.apply_n_qubit(ref, n, u, qs)
    out = st.get()
    st.close()
    return out, ref


@pytest.mark.parametrize("n", [1, 2, 5, 9, 13, 16])
@pytest.mark.parametrize("fuse", [0, 1, 2])
def test_random_programs(n, fuse):
    """fuse 0 (one sweep per gate) and 1 (program-order tiles) are bit-exact;
    fuse 2 reorders gates on disjoint qubits, which changes rounding only."""
    rng = np.random.default_rng(100 + n)
    psi = random_state(n, rng)
    prog = random_program(n, 60, rng)
    out, ref = run_both(n, prog, psi, fuse=fuse)
    if fuse < 2:
        assert np.array_equal(out, ref), np.max(np.abs(out - ref))
    else:
        assert np.max(np.abs(out - ref)) <= 1e-14


@pytest.mark.parametrize("tile_bits", [6, 8, 10, 12, 13])
@pytest.mark.parametrize("fuse", [1, 2])
def test_tile_sizes(tile_bits, fuse):
    n = 17
    rng = np.random.default_rng(7 + tile_bits)
    psi = random_state(n, rng)
    prog = random_program(n, 80, rng)
    out, ref = run_both(n, prog, psi, fuse=fuse, tile_bits=tile_bits)
    if fuse == 1:
        assert np.array_equal(out, ref), np.max(np.abs(out - ref))
    else:
        assert np.max(np.abs(out - ref)) <= 1e-14


@pytest.mark.parametrize("g", GATES_1Q)
def test_each_1q_gate_every_target(g):
    n = 14
    rng = np.random.default_rng(3)
    psi = random_state(n, rng)
    prog = [(g, rng.uniform(-7, 7, NPARAM.get(g, 0)), [q]) for q in range(n)]
    out, ref = run_both(n, prog, psi, fuse=1)
    assert np.array_equal(out, ref)


@pytest.mark.parametrize("g", GATES_NQ)
def test_each_nq_gate(g):
    n = 15
    rng = np.random.default_rng(4)
    psi = random_state(n, rng)
    prog = []
    for _ in range(6):
        u = orc.gate_matrix(g, rng.uniform(-7, 7, NPARAM.get(g, 0)))
        k = int(np.log2(u.shape[0]))
        prog.append((g, None, [int(x) for x in rng.choice(n, size=k, replace=False)], u))
    out, ref = run_both(n, prog, psi, fuse=1)
    assert np.array_equal(out, ref)


def test_dense_random_unitaries_k1_to_k5():
    n = 12
    rng = np.random.default_rng(9)
    psi = random_state(n, rng)
    prog = []
    for k in (1, 2, 3, 4, 5, 3, 2):
        m = rng.normal(size=(1 << k, 1 << k)) + 1j * rng.normal(size=(1 << k, 1 << k))
        q, _ = np.linalg.qr(m)
        prog.append(("dense", None, [int(x) for x in rng.choice(n, size=k, replace=False)], q))
    for fuse in (0, 1, 2):
        out, ref = run_both(n, prog, psi, fuse=fuse)
        assert np.array_equal(out, ref) if fuse < 2 else np.max(np.abs(out - ref)) <= 1e-14


def test_controlled_1q_with_control_values():
    n = 11
    rng = np.random.default_rng(12)
    psi = random_state(n, rng)
    st = capi.State(n)
    st.set_options(1, 0, 0)
    st.set(psi)
    ref = psi.copy()
    for _ in range(20):
        qs = [int(x) for x in rng.choice(n, size=4, replace=False)]
        nc = int(rng.integers(1, 4))
        ctrls = [(qs[1 + j], int(rng.integers(2))) for j in range(nc)]
        u = orc.gate_matrix("u3", rng.uniform(-7, 7, 3))
        st.apply_1q(u, qs[0], ctrls)
        orc.apply_one_qubit(ref, n, u, qs[0], ctrls)
    assert np.array_equal(st.get(), ref)


@pytest.mark.parametrize("n", [1, 6, 15])
def test_prob_one_and_collapse(n):
    rng = np.random.default_rng(20 + n)
    psi = random_state(n, rng)
    st = capi.State(n)
    for q in range(n):
        st.set(psi)
        p1 = st.prob_one(q)
        p1_ref = orc.prob_one(psi, n, q)
        assert abs(p1 - p1_ref) <= 1e-14 * max(p1_ref, 1e-300) + 1e-300
        for outcome in (0, 1):
            st.set(psi)
            norm = max(np.sqrt(p1_ref), 1e-15) if outcome else max(np.sqrt(max(1 - p1_ref, 0.0)), 1e-15)
            st.collapse(q, outcome, norm)
            ref = psi.copy()
            orc.collapse(ref, n, q, outcome, p1_ref)
            assert np.array_equal(st.get(), ref)


@pytest.mark.parametrize("n,qubits", [
    (8, [7, 6, 5, 4, 3, 2, 1, 0]),   # full distribution, identity order
    (8, [0, 3, 7, 1, 2, 4, 6, 5]),   # full distribution, permuted
    (10, [4, 1, 9, 0]),              # 64 terms per bin: sequential path
    (14, [2, 11]),                   # many terms per bin: tree path
    (12, []),                        # no qubits: the total
])
def test_marginal_probs(n, qubits):
    rng = np.random.default_rng(n + len(qubits))
    psi = random_state(n, rng)
    st = capi.State(n)
    st.set(psi)
    got = st.marginal_probs(qubits)
    ref = orc.marginal_probs(psi, n, qubits)
    if (1 << (n - len(qubits))) <= 64:
        assert np.array_equal(got, ref)  # ascending per-bin order, as the reference
    else:
        np.testing.assert_allclose(got, ref, rtol=1e-14, atol=0)


def test_inner_product():
    n = 13
    rng = np.random.default_rng(31)
    a, b = random_state(n, rng), random_state(n, rng)
    st = capi.State(n)
    st.set(a)
    z = st.inner(b)
    zr = orc.inner(a, b)
    assert abs(z - zr) <= 1e-14


def test_argument_errors_are_reported_not_faults():
    st = capi.State(4)
    u = np.eye(2, dtype=np.complex128)
    with pytest.raises(capi.DqError) as e:
        st.apply_1q(u, 4)
    assert e.value.code == capi.ERR_ARG
    with pytest.raises(capi.DqError):
        st.apply_nq(np.eye(4), [1, 1])
    with pytest.raises(capi.DqError):
        st.apply_nq(np.eye(64), [0, 1, 2, 3, 4, 5])
    with pytest.raises(capi.DqError):
        st.prob_one(9)
    with pytest.raises(capi.DqError):
        st.collapse(0, 2, 1.0)
    with pytest.raises(capi.DqError) as e:
        capi.State(41)
    assert e.value.code == capi.ERR_ARG


def test_fusion_reduces_sweeps():
    n = 20
    st = capi.State(n)
    h = orc.gate_matrix("h")
    for q in range(12):
        st.apply_1q(h, q)
    st.synchronize()
    sweeps, launches, gates = st.stats()
    assert gates == 12 and sweeps == 1


@pytest.mark.parametrize("n", [10, 16])
@pytest.mark.parametrize("fuse", [1, 2])
def test_prob_one_with_pending_elementwise_ops(n, fuse):
    """p1 with queued diagonals/collapses pending (reduced without writing the
    state) equals p1 reduced after those ops are applied: bit for bit with the
    exact sweeps (fusion 1); to rounding with the contracted ones (fusion 2)."""
    rng = np.random.default_rng(40 + n)
    psi = random_state(n, rng)
    results = []
    for sync_first in (False, True):
        st = capi.State(n)
        st.set_options(fuse, 0)
        st.set(psi)
        st.apply_nq(orc.gate_matrix("rzz", [0.7]), [1, n - 1])
        st.apply_1q(orc.gate_matrix("t"), 3)
        st.collapse(2, 1, 0.9)
        st.apply_nq(orc.gate_matrix("cp", [1.1]), [n - 2, 0])
        if sync_first:
            st.synchronize()
        results.append([st.prob_one(q) for q in (0, n - 1, 5)])
        st.close()
    if fuse == 1:
        assert results[0] == results[1]
    else:
        assert np.max(np.abs(np.array(results[0]) - np.array(results[1]))) <= 1e-15


@pytest.mark.parametrize("n", [10, 16])
@pytest.mark.parametrize("fuse", [1, 2])
def test_prob_one_leaves_out_trailing_unitaries_on_other_qubits(n, fuse):
    """After a collapse followed by unitaries that do not act on the measured
    qubit (the x of a reset of another qubit, say), p1 is reduced without
    applying those unitaries, which stay queued.  p1 and the final state equal
    the oracle's with every op applied.  A qubit that a queued unitary does
    act on, and a non-unitary matrix on another qubit, go through the flush."""
    rng = np.random.default_rng(70 + n)
    psi = random_state(n, rng)
    x, h, cx = orc.gate_matrix("x"), orc.gate_matrix("h"), orc.gate_matrix("cx")
    shrink = np.array([[1.0, 0.3], [0.2, 0.9]], dtype=np.complex128)  # not unitary
    tails = {"unitary": [(x, [2]), (h, [n - 1]), (cx, [0, 4])], "non-unitary": [(x, [2]), (shrink, [3])],
             "unitary, no collapse": [(h, [3]), (cx, [0, 4])]}
    for tail, ops in tails.items():
        for q in (5, 2, n - 1):
            st = capi.State(n)
            st.set_options(fuse, 0)
            st.set(psi)
            ref = psi.copy()
            if tail != "unitary, no collapse":
                p2 = orc.prob_one(ref, n, 2)
                st.collapse(2, 1, max(np.sqrt(p2), 1e-15))
                orc.collapse(ref, n, 2, 1, p2)
            for u, qs in ops:
                if len(qs) == 1:
                    st.apply_1q(u, qs[0])
                    orc.apply_one_qubit(ref, n, u, qs[0])
                else:
                    st.apply_nq(u, qs)
                    orc.apply_n_qubit(ref, n, u, qs)
            p = st.prob_one(q)
            p_ref = orc.prob_one(ref, n, q)
            assert abs(p - p_ref) <= 1e-14 * max(1.0, abs(p_ref)), (tail, q, p, p_ref)
            out = st.get()
            st.close()
            if fuse == 1:
                assert np.array_equal(out, ref), (tail, q)
            else:
                assert np.max(np.abs(out - ref)) <= 1e-13, (tail, q)


@pytest.mark.parametrize("reg_bits", [4, 3])
def test_phase_form_ops_match_oracle(reg_bits):
    """Fused 4x4 ops of the forms U = D R (ry, rz then cx chains) and U = R D
    (cp or rzz followed by h / ry on both qubits) run as a real 4x4 plus
    phases in the contracted kernels; the state equals the oracle's to 1e-12.
    The plan is checked to contain both forms, so the test exercises them."""
    from dqsim_bosonic_amd import _core

    n = 17
    rng = np.random.default_rng(90 + reg_bits)
    psi = random_state(n, rng)
    prog = []
    for layer in range(5):
        for q in range(n):
            prog.append((orc.gate_matrix("ry", [rng.uniform(0, 6)]), [q]))
        for q in range(n):
            prog.append((orc.gate_matrix("rz", [rng.uniform(0, 6)]), [q]))
        for q in range(n - 1):
            prog.append((orc.gate_matrix("cx"), [q, q + 1]))
        perm = [int(x) for x in rng.permutation(n)]
        for i in range(0, n - 1, 2):
            a, b = perm[i], perm[i + 1]
            d = orc.gate_matrix("cp", [rng.uniform(0, 6)]) if i % 4 else orc.gate_matrix("rzz", [rng.uniform(0, 6)])
            prog.append((d, [a, b]))
            prog.append((orc.gate_matrix("h"), [a]))
            prog.append((orc.gate_matrix("ry", [rng.uniform(0, 6)]), [b]))
    plan = _core._plan_passes(n, [(qs, u.ravel(), []) for u, qs in prog], 2, 12, 4, reg_bits)
    variants = {o[1] for ps in plan for o in ps["ops"] if o[0] == 2}
    assert {4, 5} <= variants, variants
    st = capi.State(n)
    st.set_options(2, 0)
    st.set_reg_bits(reg_bits)
    st.set(psi)
    ref = psi.copy()
    for u, qs in prog:
        if len(qs) == 1:
            st.apply_1q(u, qs[0])
            orc.apply_one_qubit(ref, n, u, qs[0])
        else:
            st.apply_nq(u, qs)
            orc.apply_n_qubit(ref, n, u, qs)
    out = st.get()
    st.close()
    assert np.max(np.abs(out - ref)) <= 1e-12


@pytest.mark.parametrize("fuse", [0, 1, 2])
def test_prob_one_fused_into_sweep(fuse):
    """p1 summed by the last sweep of a flush agrees with the oracle."""
    n = 17
    rng = np.random.default_rng(55)
    psi = random_state(n, rng)
    prog = random_program(n, 40, rng)
    for q in (0, 7, 16):
        st = capi.State(n)
        st.set_options(fuse, 0)
        st.set(psi)
        ref = psi.copy()
        for item in prog:
            if item[1] is not None:
                u = orc.gate_matrix(item[0], item[1])
                st.apply_1q(u, item[2][0])
                orc.apply_one_qubit(ref, n, u, item[2][0])
            else:
                st.apply_nq(item[3], item[2])
                orc.apply_n_qubit(ref, n, item[3], item[2])
        p = st.prob_one(q)
        assert abs(p - orc.prob_one(ref, n, q)) <= 1e-14
        st.close()


@pytest.mark.parametrize("fuse", [0, 1, 2])
def test_prob_one_fused_into_sweep_every_width(fuse):
    """Regression: the reduction summed by the last sweep of a flush, at
    every state width (tiles of 9 qubits run 32-thread workgroups, half a
    wave, at n = 9..11)."""
    for n in range(1, 17):
        rng = np.random.default_rng(300 + n)
        psi = random_state(n, rng)
        prog = random_program(n, 12, rng)
        for q in sorted({0, n // 2, n - 1}):
            st = capi.State(n)
            st.set_options(fuse, 0)
            st.set(psi)
            ref = psi.copy()
            for item in prog:
                if item[1] is not None:
                    u = orc.gate_matrix(item[0], item[1])
                    st.apply_1q(u, item[2][0])
                    orc.apply_one_qubit(ref, n, u, item[2][0])
                else:
                    st.apply_nq(item[3], item[2])
                    orc.apply_n_qubit(ref, n, item[3], item[2])
            p = st.prob_one(q)
            assert abs(p - orc.prob_one(ref, n, q)) <= 1e-14, (n, q)
            st.close()
