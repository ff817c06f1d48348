"""The benchmark workloads at their stated sizes, against the CPU oracle.

* Plain reference workload: 20 qubits, depth-20 u3 + cx brickwork (400 u3 +
  190 cx), circuit seed 42.  The statevector at fusion level 1 is bit-identical
  to the oracle's restatement of src/engine.rs; at the default level (FMA
  chains, reordered commuting ops) within 1e-12 per amplitude.
  probabilities() over all 20 qubits and counts(shots=10000, seed=7)
  (src/engine.rs:368-433, simulator.rs:148-181) equal the oracle's.
* At-scale plain workload: 30 qubits complex128, depth 40, circuit and
  measurement seed 42, checked against tests/golden/random_blocks_30q_seed42.json
  (written by the oracle's OpenMP restatement, make_random_blocks_30q.py):
  measurement record equal, every p1 within 1e-12, 1026 sampled amplitudes
  within 1e-12, and the marginals over qubits 7..0 and 29..22 within 1e-12, at
  fusion levels 1 and 2.
* The complex64 precision family: random_blocks (remote-link gates, measure,
  reset, conditional x) at 16 qubits with fusion level 3 (5-qubit blocks on
  FP32 MFMA) against complex128 at fusion level 1: records equal, amplitudes
  within 2e-5 (FP32 unit roundoff 6e-8 over ~1000 gates).  And the precision
  workload at its stated size, 32 qubits and depth 60, complex64 against
  complex128 on the same GPU.
"""
from __future__ import annotations

import gc
import json
import os

import numpy as np
import pytest

import dqsim_bosonic_amd as dq
from dqsim_bosonic_amd import _core
from dqsim_bosonic_amd import workloads as W
from oracle.oracle import OracleSimulator

pytestmark = pytest.mark.gpu

GOLDEN_30Q = os.path.join(os.path.dirname(__file__), "golden", "random_blocks_30q_seed42.json")


@pytest.fixture(scope="module")
def brickwork_20():
    circ = W.brickwork_u3(20, 20, seed=42)
    ref = OracleSimulator(seed=42, parallel=True).simulate(circ)
    return circ, ref


def test_brickwork_20q_depth20_fusion1_bit_exact(brickwork_20):
    circ, ref = brickwork_20
    assert sum(1 for i in circ.instructions if i.kind == "u3") == 400
    assert sum(1 for i in circ.instructions if i.kind == "cx") == 190
    r = dq.StatevectorSimulator(seed=42, fusion=1).simulate(circ)
    assert np.array_equal(r.statevector, ref.statevector)
    assert r.probabilities() == ref.probabilities()
    assert r.counts(10000, seed=7) == ref.counts(10000, seed=7)


def test_brickwork_20q_depth20_default_fusion(brickwork_20):
    circ, ref = brickwork_20
    r = dq.StatevectorSimulator(seed=42).simulate(circ)
    assert np.max(np.abs(r.statevector - ref.statevector)) <= 1e-12
    p = r.probabilities_array()
    pr = ref.probabilities()
    want = np.zeros(1 << 20)
    want[list(pr.keys())] = list(pr.values())
    assert np.max(np.abs(p - want)) <= 1e-12
    c = r.counts(10000, seed=7)
    assert sum(c.values()) == 10000 and all(len(k) == 20 for k in c)
    # a draw changes bins only if it falls within ~1e-12 of a CDF step
    assert c == ref.counts(10000, seed=7)


@pytest.mark.skipif(not os.path.exists(GOLDEN_30Q), reason="golden 30-qubit record not generated")
@pytest.mark.parametrize("fusion", [2, 1])
def test_random_blocks_30q_against_golden(fusion):
    with open(GOLDEN_30Q) as f:
        g = json.load(f)
    n = g["circuit"]["n"]
    circ = W.random_blocks(n, g["circuit"]["depth"], seed=g["circuit"]["seed"])
    print(f"[30q golden] fusion={fusion} n={n}", flush=True)
    eng = _core.Engine(n, fusion, 0, 0, 0)
    eng.load(circ.model_dump_json())
    rec = eng.step(g["measure_seed"])
    assert {str(k): v for k, v in rec.items()} == g["classical_bits"]
    p1 = np.array(eng.p1_record())
    assert p1.shape == (len(g["p1"]),)
    assert np.max(np.abs(p1 - np.array(g["p1"]))) <= 1e-12
    res = eng.result()
    lo = res.probabilities_array(list(range(7, -1, -1)))
    hi = res.probabilities_array(list(range(n - 1, n - 9, -1)))
    assert np.max(np.abs(lo - np.array(g["marg_low8"]))) <= 1e-12
    assert np.max(np.abs(hi - np.array(g["marg_high8"]))) <= 1e-12
    sv = res.statevector
    amps = sv[np.array(g["amp_index"])]
    del sv
    want = np.array(g["amp_re"]) + 1j * np.array(g["amp_im"])
    err = float(np.max(np.abs(amps - want)))
    print(f"[30q golden] fusion={fusion} max|damp|={err:.3e}", flush=True)
    assert err <= 1e-12
    del res, eng
    gc.collect()


def test_precision_workload_32q_depth60():
    """The precision workload at its stated size: 32 qubits, depth 60, circuit
    and measurement seed 42, complex64 at fusion level 3 (5-qubit blocks on
    FP32 MFMA; 32 GiB) against complex128 at fusion level 2 (64 GiB) on the
    same GPU: measurement records equal, p1 within 1e-5, max |amplitude
    difference| within 2e-5, fidelity deficit within 1e-5 and total
    variation distance of the probabilities within 1e-4; the MFMA blocks ran."""
    n = 32
    text = W.random_blocks(n, 60, seed=42).model_dump_json()
    try:
        e32 = _core.Engine(n, 3, 0, 0, 1)
        e64 = _core.Engine(n, 2, 0, 0, 0)
    except MemoryError:
        pytest.skip("needs about 100 GiB of device memory")
    e32.load(text)
    f0 = e32.fused_stats()
    r32 = e32.step(42)
    blocks = e32.fused_stats()[0] - f0[0]
    e64.load(text)
    r64 = e64.step(42)
    try:
        assert r32 == r64 and len(r32) == 12
        assert np.max(np.abs(np.array(e32.p1_record()) - np.array(e64.p1_record()))) <= 1e-5
        d = e32.compare(e64)
        assert d["max_abs"] <= 2e-5
        assert 1.0 - d["fidelity"] <= 1e-5
        assert d["tvd"] <= 1e-4
        assert blocks > 0
    finally:
        del e32, e64
        gc.collect()


@pytest.mark.parametrize("min_cost", [None, 0.0])
def test_c64_fusion3_random_blocks_16q(min_cost):
    n = 16
    circ = W.random_blocks(n, 30, seed=3, remote_per_kind=4, measure_every=5)
    text = circ.model_dump_json()
    e32 = _core.Engine(n, 3, 0, 0, 1)
    if min_cost is not None:
        e32.set_fuse_min_cost(min_cost)
    e64 = _core.Engine(n, 1, 0, 0, 0)
    e32.load(text)
    e64.load(text)
    r32 = e32.step(11)
    r64 = e64.step(11)
    assert r32 == r64 and len(r32) == 12
    assert np.max(np.abs(np.array(e32.p1_record()) - np.array(e64.p1_record()))) <= 1e-5
    d = e32.compare(e64)
    assert d["max_abs"] <= 2e-5
    assert 1.0 - d["fidelity"] <= 1e-5
    if min_cost == 0.0:
        blocks, gates = e32.fused_stats()
        assert blocks > 0 and gates > blocks
    ref = OracleSimulator(seed=11).simulate(circ)
    assert {int(k): v for k, v in r64.items()} == ref.classical_bits
    # p1 sums in a different order than the oracle's, so collapses may differ in the last bits
    assert np.max(np.abs(e64.result().statevector - ref.statevector)) <= 1e-12


@pytest.mark.parametrize("n", [12, 14, 20, 27])
def test_eight_amplitude_layouts_equal_sixteen(n):
    """The complex128 sweep with 3 register bits per layout (k_tile_reg8, 8
    amplitudes per thread) against the default 4 (k_tile_reg): the same ops
    on the same tiles in other register layouts, so records are equal and
    amplitudes agree to rounding (1e-12); at 12 and 14 qubits both also
    equal the oracle to 1e-12.  27 qubits: several tiles per workgroup."""
    circ = W.random_blocks(n, 12, seed=n, remote_per_kind=4, measure_every=4)
    text = circ.model_dump_json()
    engines = []
    for r in (3, 4):
        e = _core.Engine(n, 2, 0, 0, 0)
        e.set_reg_bits(r)
        e.load(text)
        engines.append((e, e.step(17)))
    (e3, r3), (e4, r4) = engines
    assert r3 == r4
    d = e3.compare(e4)
    assert d["max_abs"] <= 1e-12
    if n <= 14:
        ref = OracleSimulator(seed=17).simulate(circ)
        assert {int(k): v for k, v in r3.items()} == ref.classical_bits
        assert np.max(np.abs(e3.result().statevector - ref.statevector)) <= 1e-12

