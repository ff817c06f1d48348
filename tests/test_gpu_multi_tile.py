"""GPU checks on states large enough that each workgroup walks several tiles.

The forward sweep launches at most 16384 workgroups of 12-qubit tiles, so
from 27 qubits on each workgroup processes more than one tile; the adjoint
reverse pass launches at most 8192 workgroups of 11-qubit tiles, so from 25
qubits.  There a tile's global offset is stepped from tile to tile (the
per-workgroup deposit ORed with a counter over the outer qubits of the high
tile-index bits) and, in the forward sweep, the next tile is prefetched
while the current one computes.  Each is compared with a path that uses
neither, itself checked against the oracle at small sizes: the unfused
one-sweep-per-gate mode for the forward sweep, the per-parameter sweep for
the adjoint gradient.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as orc
from tests.test_gpu_adjoint import wide_circuit
from tests.test_gpu_capi import NPARAM, random_program

pytestmark = pytest.mark.gpu

capi = pytest.importorskip("dqsim_bosonic_amd.capi")
A = pytest.importorskip("dqsim_bosonic_amd.adjoint")


def apply(st, prog):
    for item in prog:
        if item[1] is not None:
            g, params, qs = item
            st.apply_1q(orc.gate_matrix(g, params), qs[0])
        else:
            _, _, qs, u = item
            st.apply_nq(u, qs)


@pytest.mark.parametrize("dtype,tol", [(0, 1e-12), (1, 2e-5)])
def test_forward_sweep_with_several_tiles_per_workgroup(dtype, tol):
    n = 27
    rng = np.random.default_rng(2027 + dtype)
    prog = random_program(n, 90, rng)
    # gates on the highest qubits too, so that passes hold ops on outer-index bits
    for q in (n - 1, n - 2, n - 5):
        prog.append(("u3", rng.uniform(-7, 7, NPARAM["u3"]), [q]))
    fused, plain = capi.State(n, dtype), capi.State(n, dtype)
    try:
        fused.set_options(2)
        plain.set_options(0)
        apply(fused, prog)
        apply(plain, prog)
        for q in (0, 13, n - 1):
            p_f, p_p = fused.prob_one(q), plain.prob_one(q)
            assert abs(p_f - p_p) <= tol * max(1.0, p_p), (q, p_f, p_p)
        cmp = fused.compare(plain)
        assert cmp["max_abs"] <= tol, cmp
        assert abs(cmp["fidelity"] - 1.0) <= 100 * tol, cmp
    finally:
        fused.close()
        plain.close()


def test_adjoint_pass_with_several_tiles_per_workgroup():
    n = 25
    pc, H, th = wide_circuit(n, 25, layers=2)
    est = A.Estimator(pc, H)
    e_t, g_t = est.gradient(th)
    est._eng.set_tiled_sweep(False)
    e_p, g_p = est.gradient(th)
    assert abs(e_t - e_p) <= 1e-12
    assert np.max(np.abs(g_t - g_p)) <= 1e-12
