"""Sweeps that skip the tiles of qubits still 0 everywhere, against full sweeps and the oracle.

After a reset to |0...0>, a qubit on which no op has acted other than
diagonally is 0 in every amplitude.  A register sweep then visits only the
tiles where such outer qubits are 0; the other tiles hold zeros that the
sweep would leave zero (dq_set_zero_tiles switches this off).

Programs whose qubits spread upward from the low ones, with diagonals and
controlled gates whose controls sit on untouched qubits (control values 0
and 1), and measurements/collapses of touched and untouched qubits:
* the state equals the one of the same program with skipping off
  (np.array_equal at fusion level 1: the same ops on the same tiles; within
  1e-12 at level 2, where a measurement that needs no flush with skipping on
  may end a flush with it off, and a flush's ops are reordered; p1 within
  1e-13, as a sweep over fewer tiles sums its per-workgroup partials in other
  groups) and the oracle's
  (src/engine.rs restated; bit-exact at fusion level 1, within 1e-12 at
  level 2, which reorders and contracts, within 1e-5 in complex64);
* p1 of an untouched qubit is exactly 0.0, as the reference's sum of zeros;
  other p1 within 1e-13 relative (sums of up to 2^20 terms in another order);
* with skipping on, the sweeps of a state larger than one 12-qubit tile
  read and write fewer than sweeps x 2^n amplitudes; with it off, exactly
  that many;
* p1 of such a qubit is +0.0 without a launch or a flush.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

capi = pytest.importorskip("dqsim_bosonic_amd.capi")

GATES_1Q = ["x", "y", "h", "sx", "t", "u3", "rx", "ry", "rz"]
GATES_NQ = ["cx", "cz", "swap", "crx", "rxx", "rzz", "cu3", "ccx"]
NPARAM = {"u3": 3, "rx": 1, "ry": 1, "rz": 1, "crx": 1, "cu3": 3, "rxx": 1, "rzz": 1, "cp": 1}


def params(g, rng):
    return rng.uniform(-4, 4, NPARAM.get(g, 0))


def spreading_program(n, count, rng):
    """Op i acts non-diagonally only on the lowest 3 + i // 6 qubits; a
    measurement every 25 ops, of a touched or an untouched qubit."""
    prog = []
    for i in range(count):
        hi = min(n, 3 + i // 6)
        r = rng.random()
        if i % 25 == 24:
            prog.append(("m", int(rng.integers(n))))
        elif r < 0.12:
            qs = [int(x) for x in rng.choice(n, size=2, replace=False)]
            prog.append(("nq", orc.gate_matrix("cp", params("cp", rng)), qs))
        elif r < 0.25:
            t = int(rng.integers(hi))
            others = [q for q in range(n) if q != t]
            cs = [int(c) for c in rng.choice(others, size=int(rng.integers(1, 3)), replace=False)]
            prog.append(("c1", orc.gate_matrix("u3", params("u3", rng)), t, [(c, int(rng.integers(2))) for c in cs]))
        elif r < 0.6:
            g = GATES_1Q[int(rng.integers(len(GATES_1Q)))]
            prog.append(("c1", orc.gate_matrix(g, params(g, rng)), int(rng.integers(hi)), []))
        else:
            g = GATES_NQ[int(rng.integers(len(GATES_NQ)))]
            u = orc.gate_matrix(g, params(g, rng))
            k = int(np.log2(u.shape[0]))
            if k <= hi:
                prog.append(("nq", u, [int(x) for x in rng.choice(hi, size=k, replace=False)]))
    return prog


def run(n, prog, states, ref, touched_q):
    """Applies prog to every engine state and to the oracle's ref; returns
    the p1 values of each state per measurement."""
    p1s = [[] for _ in states]
    for item in prog:
        if item[0] == "m":
            q = item[1]
            p_ref = orc.prob_one(ref, n, q)
            for st, out in zip(states, p1s):
                out.append(st.prob_one(q))
            if not (touched_q >> q) & 1:
                assert p_ref == 0.0
            outcome = 1 if p_ref > 0.5 else 0
            norm = max(np.sqrt(p_ref), 1e-15) if outcome else max(np.sqrt(max(1 - p_ref, 0.0)), 1e-15)
            for st in states:
                st.collapse(q, outcome, norm)
            orc.collapse(ref, n, q, outcome, p_ref)
            for out in p1s:
                p = out[-1]
                if not (touched_q >> q) & 1:
                    assert p == 0.0, (q, p)
                # a sum of up to 2^(n-1) terms, in another order than the oracle's
                assert abs(p - p_ref) <= 1e-13 * max(p_ref, 1e-300) + 1e-300, (q, p, p_ref)
        elif item[0] == "c1":
            _, u, t, ctrls = item
            for st in states:
                st.apply_1q(u, t, ctrls)
            orc.apply_one_qubit(ref, n, u, t, ctrls)
            touched_q |= 1 << t
        else:
            _, u, qs = item
            for st in states:
                st.apply_nq(u, qs)
            orc.apply_n_qubit(ref, n, u, qs)
            if not np.array_equal(u, np.diag(np.diag(u))):
                touched_q |= sum(1 << q for q in qs)
    return p1s, touched_q


@pytest.mark.parametrize("n,fuse,reg_bits", [(12, 1, 4), (17, 1, 4), (21, 1, 4), (12, 2, 4), (17, 2, 4),
                                             (21, 2, 4), (17, 2, 3)])
def test_zero_tiles_equal_full_sweeps_and_oracle(n, fuse, reg_bits):
    rng = np.random.default_rng(1000 + 10 * n + fuse + reg_bits)
    prog = spreading_program(n, 160, rng)
    print(f"[zero tiles] n={n} fuse={fuse} reg_bits={reg_bits} ops={len(prog)}", flush=True)
    on, off = capi.State(n), capi.State(n)
    try:
        for st, z in ((on, True), (off, False)):
            st.set_options(fuse, 0)
            if fuse >= 2:
                st.set_reg_bits(reg_bits)
            st.set_zero_tiles(z)
            st.reset()
            st.set_timing(True)
        sw0 = [st.stats()[0] for st in (on, off)]
        ref = np.zeros(1 << n, dtype=complex)
        ref[0] = 1.0
        # twice, with a reset between: the second run starts from |0...0> again
        for rep in range(2):
            if rep:
                for st in (on, off):
                    st.reset()
                ref[:] = 0.0
                ref[0] = 1.0
            p1s, _ = run(n, prog, [on, off], ref, 0)
            # a sweep over fewer tiles runs fewer workgroups, so its p1 sums
            # per-workgroup partials grouped otherwise (the same terms)
            assert np.allclose(p1s[0], p1s[1], rtol=1e-13, atol=1e-300)
            got_on, got_off = on.get(), off.get()
            if fuse == 1:
                assert np.array_equal(got_on, got_off)
                assert np.array_equal(got_on, ref)
            else:
                # fusion 2 reorders and contracts within a flush, and the two
                # states may flush at different points (see below)
                assert np.max(np.abs(got_on - got_off)) <= 1e-12
                assert np.max(np.abs(got_on - ref)) <= 1e-12
        amps = [st.timed_amplitudes() for st in (on, off)]
        sweeps = [st.stats()[0] - s0 for st, s0 in zip((on, off), sw0)]
        # (the sweep counts may differ: p1 of a qubit still 0 everywhere
        # needs no flush with skipping on)
        assert amps[1] == sweeps[1] << n
        if n > 12:  # a 12-qubit state is one tile: nothing to skip
            assert amps[0] < sweeps[0] << n
        else:
            assert amps[0] == sweeps[0] << n
    finally:
        on.close()
        off.close()


def test_zero_tiles_complex64():
    n = 17
    rng = np.random.default_rng(77)
    prog = spreading_program(n, 160, rng)
    on, off = capi.State(n, capi.C64), capi.State(n, capi.C64)
    try:
        for st, z in ((on, True), (off, False)):
            st.set_options(2, 0)
            st.set_zero_tiles(z)
            st.reset()
        ref = np.zeros(1 << n, dtype=complex)
        ref[0] = 1.0
        p1s = [[] for _ in range(2)]
        for item in prog:
            if item[0] == "m":
                q = item[1]
                p_ref = orc.prob_one(ref, n, q)
                ps = [st.prob_one(q) for st in (on, off)]
                assert abs(ps[0] - ps[1]) <= 1e-5 and abs(ps[0] - p_ref) <= 1e-5
                outcome = 1 if p_ref > 0.5 else 0
                norm = max(np.sqrt(p_ref), 1e-15) if outcome else max(np.sqrt(max(1 - p_ref, 0.0)), 1e-15)
                for st in (on, off):
                    st.collapse(q, outcome, norm)
                orc.collapse(ref, n, q, outcome, p_ref)
            elif item[0] == "c1":
                _, u, t, ctrls = item
                for st in (on, off):
                    st.apply_1q(u, t, ctrls)
                orc.apply_one_qubit(ref, n, u, t, ctrls)
            else:
                _, u, qs = item
                for st in (on, off):
                    st.apply_nq(u, qs)
                orc.apply_n_qubit(ref, n, u, qs)
        got_on, got_off = on.get(), off.get()
        assert np.max(np.abs(got_on - got_off)) <= 1e-5
        assert np.max(np.abs(got_on - ref)) <= 1e-5
    finally:
        on.close()
        off.close()


def test_zero_knowledge_follows_copy_tensor_and_upload():
    """A copy keeps what is known to be 0, a tensor product combines its
    factors', an upload knows nothing: each continues to equal the oracle."""
    rng = np.random.default_rng(5)
    n = 16
    prog = spreading_program(n, 40, rng)  # touches at most the lowest 9 qubits
    a = capi.State(n)
    c = capi.State(n)
    u = capi.State(n)
    f1, f2, t = capi.State(8), capi.State(8), capi.State(n)
    try:
        a.reset()
        ref = np.zeros(1 << n, dtype=complex)
        ref[0] = 1.0
        run(n, prog, [a], ref, 0)
        c.copy_from(a)
        u.set(a.get())
        tail = [("c1", orc.gate_matrix("h", []), q, []) for q in range(n)] + \
               [("nq", orc.gate_matrix("cx", []), [q, q + 1]) for q in range(n - 1)]
        refs = [ref.copy() for _ in range(3)]
        for st, r in zip((a, c, u), refs):
            run(n, tail, [st], r, 0)
            assert np.max(np.abs(st.get() - r)) <= 1e-12
        # |f1> (x) |f2>: f1's qubits at the low positions, f2's at the high ones
        for f in (f1, f2):
            f.reset()
        f1.apply_1q(orc.gate_matrix("h", []), 0)
        f2.apply_1q(orc.gate_matrix("ry", [0.7]), 1)
        t.tensor(f1, list(range(8)), f2, list(range(8, 16)))
        r = np.zeros(1 << n, dtype=complex)
        r[0] = 1.0
        orc.apply_one_qubit(r, n, orc.gate_matrix("h", []), 0)
        orc.apply_one_qubit(r, n, orc.gate_matrix("ry", [0.7]), 9)
        run(n, tail, [t], r, 0)
        assert np.max(np.abs(t.get() - r)) <= 1e-12
    finally:
        for st in (a, c, u, f1, f2, t):
            st.close()


def test_random_blocks_engine_zero_tiles_equal():
    """The benchmark circuit family through the engine: records and the
    final state equal with skipping on and off, p1 within 1e-13 (20 qubits,
    depth 12)."""
    from dqsim_bosonic_amd import _core
    from dqsim_bosonic_amd import workloads as W
    n = 20
    text = W.random_blocks(n, 12, seed=9, remote_per_kind=4, measure_every=4).model_dump_json()
    engines = []
    for z in (True, False):
        e = _core.Engine(n, 2, 0, 0, 0)
        e.set_zero_tiles(z)
        e.load(text)
        e.set_timing(True)
        engines.append((e, e.step(21)))
    (e1, r1), (e0, r0) = engines
    assert r1 == r0
    assert np.allclose(e1.p1_record(), e0.p1_record(), rtol=1e-13, atol=1e-300)
    d = e1.compare(e0)
    assert d["max_abs"] == 0.0
    assert e1.timed_amplitudes() < e0.timed_amplitudes()


def test_p1_of_an_untouched_qubit_reads_nothing():
    """p1 of a qubit no queued or applied op has acted on other than
    diagonally is +0.0 without a launch, and the queue stays unflushed; a
    qubit a queued op acts on is reduced as usual."""
    n = 16
    st = capi.State(n)
    try:
        st.reset()
        st.apply_1q(orc.gate_matrix("h", []), 0)
        st.apply_1q(orc.gate_matrix("x", []), 15, [(0, 1)])   # cx 0 -> 15
        st.apply_nq(orc.gate_matrix("cp", [0.4]), [3, 9])     # diagonal on untouched qubits
        st.apply_1q(orc.gate_matrix("ry", [0.3]), 2, [(7, 1)])  # control on an untouched qubit
        launches = st.stats()[1]
        for q in (3, 7, 9, 14):
            p = st.prob_one(q)
            assert p == 0.0 and np.copysign(1.0, p) == 1.0
        assert st.stats()[1] == launches
        assert abs(st.prob_one(15) - 0.5) <= 1e-15
        assert st.stats()[1] > launches
        ref = np.zeros(1 << n, dtype=complex)
        ref[0] = 1.0
        orc.apply_one_qubit(ref, n, orc.gate_matrix("h", []), 0)
        orc.apply_one_qubit(ref, n, orc.gate_matrix("x", []), 15, [(0, 1)])
        orc.apply_n_qubit(ref, n, orc.gate_matrix("cp", [0.4]), [3, 9])
        orc.apply_one_qubit(ref, n, orc.gate_matrix("ry", [0.3]), 2, [(7, 1)])
        assert np.array_equal(st.get(), ref)
    finally:
        st.close()
