"""CPU tests of the matrix-product-state path: the oracle's restatement of
src/monolithic/mps/simulator.rs, the host tensor network of the product
(gates, SWAP network, Jacobi SVD, truncation) and the option parsing of
simulate_monolithic(mode="mps").

Pins:
* to_statevector (simulator.rs:246-264): the vectorised contraction equals
  the literal per-basis loop, the transfer-matrix norm and single amplitudes
  by matrix-chain product.
* Without truncation an MPS is exact, so the oracle MPS and the product's
  host network must give the oracle statevector simulator's state (itself
  pinned by the golden vectors of test_oracle.py).
* With truncation the kept state is the best rank-k approximation at every
  cut (Eckart-Young) and does not depend on the SVD routine unless a cut
  falls inside a group of equal singular values; the oracle counts such cuts
  and the comparison skips those cases (parity there is a basis choice of
  the reference's nalgebra SVD and is not pinned).
* The reference's own MPS test expectations (tests/test_mps_simulation.py,
  tests/test_mps_converges.py) hold for the oracle.
"""
from __future__ import annotations

import json
import math

import numpy as np
import pytest

from oracle import oracle as orc

core = pytest.importorskip("dqsim_bosonic_amd._core")
from dqsim_bosonic_amd import circuit as C  # noqa: E402
from dqsim_bosonic_amd import workloads as W  # noqa: E402


def rand_mps(rng, bonds):
    return [rng.normal(size=(bonds[q], 2, bonds[q + 1])) + 1j * rng.normal(size=(bonds[q], 2, bonds[q + 1]))
            for q in range(len(bonds) - 1)]


def test_oracle_contraction_matches_literal_loop():
    rng = np.random.default_rng(0)
    for bonds in ([1, 1], [1, 2, 1], [1, 2, 3, 1], [1, 2, 4, 3, 2, 1], [1, 1, 3, 3, 1, 2, 1]):
        ts = rand_mps(rng, bonds)
        psi = orc.mps_contract(ts)
        assert np.max(np.abs(psi - orc.mps_to_statevector_loop(ts))) <= 1e-13 * np.max(np.abs(psi))
        n = len(ts)
        ref = float(np.sum(np.abs(psi) ** 2))
        assert abs(orc.mps_norm2_transfer(ts) - ref) <= 1e-12 * ref
        idx = np.arange(1 << n)
        for q in range(n):
            p1 = float(np.sum(np.abs(psi[(idx >> q) & 1 == 1]) ** 2))
            assert abs(orc.mps_norm2_transfer(ts, q) - p1) <= 1e-12 * ref
        for x in (0, (1 << n) - 1, 5 % (1 << n)):
            assert abs(orc.mps_amplitude(ts, x) - psi[x]) <= 1e-13 * np.max(np.abs(psi))


def test_oracle_mps_exact_without_truncation():
    for seed, n in ((1, 2), (2, 4), (3, 6)):
        text = W.random_mixed(n, 60, seed, kinds=["h", "x", "y", "z", "s", "t", "sx", "u3", "rx", "ry", "rz",
                                                  "cx", "cz", "cy", "ch", "swap", "csx", "crx", "cry", "crz",
                                                  "cp", "cu3", "cu", "rxx", "rzz"]).model_dump_json()
        psi, _, _ = orc.simulate_mps_json(text, 1)
        ref = orc.OracleSimulator(seed=1).simulate_json(text).statevector
        assert np.max(np.abs(psi - ref)) <= 1e-12


def test_oracle_reference_expectations():
    # tests/test_mps_simulation.py: mid-circuit measurement + conditional, reset
    c = C.circuit(2, [C.instruction("x", qubit=0), C.instruction("measure", qubit=0, cbit=0),
                      C.instruction("conditional", condition=C.Condition(0, 1, 1),
                                    op=C.instruction("x", qubit=1))], n_cbits=1)
    psi, cbits, _ = orc.simulate_mps_json(c.model_dump_json(), 42)
    assert cbits == {0: 1} and abs(abs(psi[3]) ** 2 - 1.0) < 1e-15
    c = C.circuit(1, [C.instruction("x", qubit=0), C.instruction("reset", qubit=0)])
    psi, _, _ = orc.simulate_mps_json(c.model_dump_json(), 42)
    assert abs(abs(psi[0]) ** 2 - 1.0) < 1e-15
    # tests/test_mps_converges.py: L1 error to the exact distribution falls with the bond
    text = high_entanglement().model_dump_json()
    exact = np.abs(orc.OracleSimulator(seed=42).simulate_json(text).statevector) ** 2
    errs = [float(np.sum(np.abs(np.abs(orc.simulate_mps_json(text, 42, max_bond=b)[0]) ** 2 - exact)))
            for b in (2, 3, 4, 8)]
    assert errs[1] < errs[0] and errs[2] < errs[1] and errs[3] < 1e-10


def high_entanglement():
    """_high_entanglement_circuit of the reference's tests/test_mps_converges.py."""
    ins = [C.instruction("h", qubit=q) for q in range(6)]
    for layer in range(3):
        for ctl, tgt in [(0, 5), (1, 4), (2, 3), (0, 2), (3, 5), (1, 3)]:
            ins.append(C.instruction("cx", control=ctl, target=tgt))
        for q in range(6):
            ins.append(C.instruction("ry", qubit=q, theta=0.2 * (layer + 1) * (q + 1)))
    return C.circuit(6, ins)


@pytest.mark.parametrize("shape", [(1, 1), (1, 5), (5, 1), (2, 4), (4, 2), (6, 6), (8, 3), (3, 8), (32, 32),
                                   (17, 40)])
def test_jacobi_svd(shape):
    rng = np.random.default_rng(sum(shape))
    a = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    u, s, vt = core._svd_jacobi(a)
    s = np.array(s)
    k = min(shape)
    assert u.shape == (shape[0], k) and vt.shape == (k, shape[1])
    assert np.all(np.diff(s) <= 0.0)
    assert np.max(np.abs(s - np.linalg.svd(a, compute_uv=False))) <= 1e-13 * s[0]
    assert np.max(np.abs((u * s) @ vt - a)) <= 1e-13 * s[0]
    assert np.max(np.abs(u.conj().T @ u - np.eye(k))) <= 1e-13
    assert np.max(np.abs(vt @ vt.conj().T - np.eye(k))) <= 1e-13


def test_jacobi_svd_rank_deficient():
    rng = np.random.default_rng(4)
    x = rng.normal(size=(6, 2)) + 1j * rng.normal(size=(6, 2))
    a = x @ (rng.normal(size=(2, 5)) + 1j * rng.normal(size=(2, 5)))  # rank 2
    u, s, vt = core._svd_jacobi(a)
    assert np.max(np.abs((u * np.array(s)) @ vt - a)) <= 1e-13 * s[0]
    assert s[2] <= 1e-13 * s[0]
    u, s, vt = core._svd_jacobi(np.zeros((4, 4), dtype=complex))
    assert max(s) == 0.0 and np.all(np.isfinite(u)) and np.all(np.isfinite(vt))


def test_host_network_exact_without_truncation():
    for seed, n in ((1, 3), (2, 5), (3, 7)):
        text = W.random_mixed(n, 80, seed, kinds=["h", "x", "y", "z", "s", "t", "sx", "sxdg", "u3", "u2", "u1",
                                                  "p", "rx", "ry", "rz", "cx", "cz", "cy", "ch", "swap",
                                                  "csx", "crx", "cry", "crz", "cu1", "cp", "cu3", "cu", "rxx",
                                                  "rzz", "id", "u0"]).model_dump_json()
        psi = orc.mps_contract(core._mps_network(text))
        ref = orc.OracleSimulator(seed=1).simulate_json(text).statevector
        assert np.max(np.abs(psi - ref)) <= 1e-12


def test_host_network_truncation_matches_oracle():
    checked = 0
    for seed in range(24):
        rng = np.random.default_rng(100 + seed)
        n = int(rng.integers(3, 8))
        text = W.mps_generic(n, int(rng.integers(5, 40)), seed).model_dump_json()
        for mb in (1, 2, 3, 4):
            want, _, om = orc.simulate_mps_json(text, 1, max_bond=mb)
            if om.degenerate_cuts:
                continue
            got = orc.mps_contract(core._mps_network(text, mb))
            assert np.max(np.abs(got - want)) <= 1e-10, (seed, mb)
            checked += 1
    assert checked >= 60


def test_host_network_threshold_truncation():
    # a large threshold drops singular values below it (and rescales)
    text = W.mps_generic(5, 30, 3).model_dump_json()
    want, _, om = orc.simulate_mps_json(text, 1, threshold=0.2)
    assert om.degenerate_cuts == 0
    got = orc.mps_contract(core._mps_network(text, None, 0.2))
    assert np.max(np.abs(got - want)) <= 1e-10
    assert sum(t.shape[2] for t in core._mps_network(text, None, 0.2)) < \
        sum(t.shape[2] for t in core._mps_network(text))


def test_host_network_reversed_and_far_pairs():
    # tests/test_mps_simulation.py::test_non_adjacent_and_reversed_two_qubit_gates
    c = C.circuit(4, [C.instruction("h", qubit=0), C.instruction("h", qubit=3),
                      C.instruction("cx", control=3, target=1), C.instruction("cz", control=2, target=0),
                      C.instruction("swap", a=3, b=0), C.instruction("cx", control=0, target=2)])
    text = c.model_dump_json()
    ref = orc.OracleSimulator(seed=1).simulate_json(text).statevector
    assert np.max(np.abs(orc.mps_contract(core._mps_network(text)) - ref)) <= 1e-14


def test_host_network_generic_gate_names():
    gates = [C.GateInstruction(name=nm, params=[], qubits=[0, 2]) for nm in
             ("remote_link_phi_plus", "remote_epr", "EPR", "remote_link_psi_minus", "remote_link_psi_plus",
              "nonlocal_cz", "remote_cz", "remote_cx", "remote_barrier", "remote_cu1", "circuit-block-1")]
    c = C.circuit(3, [C.instruction("h", qubit=1)] + gates)
    text = c.model_dump_json()
    want, _, _ = orc.simulate_mps_json(text, 1)
    assert np.max(np.abs(orc.mps_contract(core._mps_network(text)) - want)) <= 1e-13
    bad = C.circuit(3, [C.GateInstruction(name="remote_rzz", params=[0.1], qubits=[0, 1])])
    with pytest.raises(Exception, match="does not support generic gate"):
        core._mps_network(bad.model_dump_json())
    with pytest.raises(Exception, match="one- and two-qubit"):
        core._mps_network(C.circuit(3, [C.instruction("ccx", control1=0, control2=1, target=2)]).model_dump_json())


def test_mps_options_are_validated_before_any_device_work():
    dq = pytest.importorskip("dqsim_bosonic_amd")
    c = C.circuit(2, [C.instruction("h", qubit=0)])
    with pytest.raises(ValueError, match="max_bond_dimension must be positive"):
        dq.simulate_monolithic(c, mode="mps", max_bond_dimension=0)
    with pytest.raises(ValueError, match="truncation_threshold"):
        dq.simulate_monolithic(c, mode="mps", truncation_threshold=-1.0)
    with pytest.raises(ValueError, match="truncation_threshold"):
        dq.simulate_monolithic(c, mode="mps", truncation_threshold=math.inf)
    with pytest.raises(TypeError, match="Unsupported MPS option"):
        dq.simulate_monolithic(c, mode="mps", bond=3)
    with pytest.raises(TypeError, match="Unsupported MPS option"):
        dq.simulate_monolithic_shots(c, mode="matrix-product-state", shots=3, chi=3)


def test_oracle_mps_shots_full_bond_equal_statevector_shots():
    # tests/test_mps_converges.py::test_measurement_shot_distribution_converges_with_bond_dimension:
    # at full bond the MPS draws the same outcomes as the statevector simulator
    ins = [C.instruction("h", qubit=q) for q in range(4)]
    for layer in range(2):
        for ctl, tgt in [(0, 3), (1, 2), (0, 2), (3, 1)]:
            ins.append(C.instruction("cx", control=ctl, target=tgt))
        for q in range(4):
            ins.append(C.instruction("ry", qubit=q, theta=0.3 * (layer + 1) * (q + 1)))
    ins += [C.instruction("measure", qubit=q, cbit=q) for q in range(4)]
    text = C.circuit(4, ins, n_cbits=4).model_dump_json()
    assert orc.simulate_mps_shots_json(text, 60, 42, max_bond=4) == orc.simulate_shots_json(text, 60, 42)
    json.loads(text)


def test_c_loop_equals_python_loop_bitwise():
    # the C restatement (the contraction's CPU baseline) and the Python one
    # follow the same op order, so they agree bit for bit
    rng = np.random.default_rng(12)
    for bonds in ([1, 2, 1], [1, 2, 4, 3, 2, 1], [1, 2, 4, 8, 4, 2, 1]):
        ts = rand_mps(rng, bonds)
        n = len(ts)
        want = orc.mps_to_statevector_loop(ts)
        assert np.array_equal(orc.mps_amplitudes_c(ts, 0, 1 << n), want)
        cnt = min(5, (1 << n) - 3)
        assert np.array_equal(orc.mps_amplitudes_c(ts, 3, cnt), want[3:3 + cnt])
