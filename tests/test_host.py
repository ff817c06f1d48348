"""CPU tests of the host side (C++ module `_core`) and of the C ABI's library.

No compute call reaches the GPU here: the gate library, the ChaCha8 stream,
the wire-format decoder and the mode/option parsing run on the host; the C
ABI library is only loaded and its exports checked.
"""
from __future__ import annotations

import ctypes
import json
import os
import re

import numpy as np
import pytest

from oracle import oracle as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

dq = pytest.importorskip("dqsim_bosonic_amd")
from dqsim_bosonic_amd import _core, capi  # noqa: E402
from dqsim_bosonic_amd import circuit as C  # noqa: E402
from dqsim_bosonic_amd import workloads as W  # noqa: E402

from .test_oracle import ALL_GATES  # noqa: E402


# -- gate library: host (C++) == oracle (C), bit for bit -------------------------

@pytest.mark.parametrize("g", list(ALL_GATES))
def test_host_gates_bitwise_equal_oracle(g):
    rng = np.random.default_rng(sum(map(ord, g)) + 1)
    for _ in range(5):
        par = list(rng.uniform(-20, 20, ALL_GATES[g]))
        a = _core._gate_matrix(g, par)
        b = orc.gate_matrix(g, par)
        assert a.shape == b.shape and np.array_equal(a.view(np.float64), b.view(np.float64)), g


# -- RNG -------------------------------------------------------------------------

@pytest.mark.parametrize("seed", [0, 1, 7, 42, 2**63 + 5, 2**64 - 1])
def test_host_rng_stream_equals_oracle(seed):
    got = _core._rng_stream(seed, 300)
    r = orc.Rng(seed)
    assert got == [r.next_u64() for _ in range(300)]


# -- wire format -------------------------------------------------------------------

def test_decoder_fields_and_width():
    circ = W.random_mixed(9, 300, 3)
    text = circ.model_dump_json()
    n, nc, ins = _core._decode(text)
    assert n == 9 and nc == 0
    src = json.loads(text)["instructions"]
    assert len(ins) == len(src)
    for d, s in zip(ins, src):
        assert d["kind"] == s["kind"]
        if s["kind"] == "gate":
            assert d["qubits"] == s["qubits"] and d["name"] == s["name"]
            continue
        fields = C._KIND_FIELDS[s["kind"]]
        qf = [f for f in fields if f in ("qubit", "control", "target", "a", "b", "control1", "control2",
                                         "control3", "control4", "target1", "target2")]
        assert d["qubits"] == [s[f] for f in qf]
        assert d["params"] == [s[f] for f in fields if f not in qf]


def test_num_qubits_is_max_base_plus_size():
    # src/types.rs:24-30: qregs {a: base 0 size 3, b: base 5 size 2} -> 7 qubits
    text = json.dumps({"qregs": {"a": {"name": "a", "size": 3, "base": 0}, "b": {"name": "b", "size": 2, "base": 5}},
                       "instructions": []})
    n, nc, ins = _core._decode(text)
    assert (n, nc, ins) == (7, 0, [])


def test_decoder_ignores_unknown_fields_and_reads_nested_ops():
    text = json.dumps({"qregs": {"q": {"name": "q", "size": 2, "base": 0}}, "cregs": {"c": {"name": "c", "size": 2, "base": 0}},
                       "extra": 1, "instructions": [
                           {"kind": "h", "qubit": 1, "qubits": [1], "params": [], "label": None},
                           {"kind": "conditional", "condition": {"creg_base": 1, "creg_size": 1, "creg_value": 1},
                            "op": {"kind": "rzz", "a": 0, "b": 1, "theta": 0.5}, "qubits": [0, 1]},
                           {"kind": "barrier", "qubits": [0, 1]}, {"kind": "classical", "name": "foo"}]})
    n, nc, ins = _core._decode(text)
    assert nc == 2
    assert ins[1]["condition"] == (1, 1, 1) and ins[1]["op"]["params"] == [0.5]
    assert [i["kind"] for i in ins] == ["h", "conditional", "barrier", "classical"]


@pytest.mark.parametrize("bad,needle", [
    ('{"instructions": []}', "qregs"),
    ('{"qregs": {}}', "instructions"),
    ('{"qregs": {}, "instructions": [{"kind": "warp", "qubit": 0}]}', "unknown variant"),
    ('{"qregs": {}, "instructions": [{"kind": "h"}]}', "qubit"),
    ('{"qregs": {}, "instructions": [{"kind": "h", "qubit": -1}]}', "qubit"),
    ('{"qregs": {}, "instructions": [{"kind": "h", "qubit": 1.5}]}', "qubit"),
    ('{"qregs": {}, "instructions": [{"kind": "rx", "qubit": 0, "theta": "a"}]}', "theta"),
    ('{"qregs": {}, "instructions": [{"kind": "gate", "name": "remote_cx", "qubits": [0, 1]}]}', "params"),
    ('{"qregs": {}, "instructions": [', "EOF"),
    ('{"qregs": {}, "cregs": null, "instructions": []}', "map"),
])
def test_decoder_errors_are_value_errors(bad, needle):
    with pytest.raises(ValueError, match="Circuit JSON parse error") as e:
        _core._decode(bad)
    assert needle in str(e.value)


def test_simulate_parse_error_before_device():
    # decoding happens before the state is allocated: a bad circuit raises
    # ValueError even without a GPU (simulator.rs:332-334)
    with pytest.raises(ValueError, match="Circuit JSON parse error"):
        dq.StatevectorSimulator(seed=1).simulate(C.JsonCircuit('{"qregs": 3}'))


# -- mode / option parsing (reference tests/test_codecs.py) -------------------------

@pytest.mark.parametrize("mode", ["state_vector", "state-vector", "statevector", "sv", " SV "])
def test_statevector_mode_aliases_are_accepted_before_circuit_decoding(mode):
    with pytest.raises(AttributeError, match="model_dump_json"):
        dq.simulate_monolithic(object(), mode=mode)


def test_invalid_monolithic_mode_raises():
    with pytest.raises(ValueError, match="Unsupported monolithic simulation mode \"tensor_network\""):
        dq.simulate_monolithic(object(), mode="tensor_network")


def test_statevector_mode_rejects_options():
    with pytest.raises(TypeError, match=re.escape('Options ["max_bond_dimension"] are not supported '
                                                  'for monolithic mode "state_vector"')):
        dq.simulate_monolithic(object(), mode="state_vector", max_bond_dimension=1)


def test_stabilizer_is_not_this_engine():
    # mode "mps" is provided (MpsSimulator): it asks the object for its JSON
    with pytest.raises(AttributeError):
        dq.simulate_monolithic(object(), mode="mps")
    with pytest.raises(NotImplementedError, match="not supported for stabilizer"):
        dq.simulate_monolithic(object(), mode="stabilizer")


# -- C ABI library -----------------------------------------------------------------

def header_symbols():
    text = open(os.path.join(ROOT, "include", "dqsim_gpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dq_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    syms = header_symbols()
    assert len(syms) >= 15
    L = ctypes.CDLL(capi.LIB_PATH)
    for s in syms:
        assert hasattr(L, s), s
    assert sorted(capi.EXPORTS) == syms


def test_no_device_is_an_error_code_not_a_crash():
    # In a container without a GPU the create call must fail cleanly.
    try:
        import torch  # noqa: F401

        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    h = ctypes.c_void_p()
    L = capi.lib()
    rc = L.dq_state_create(3, 0, ctypes.byref(h))
    assert rc == capi.ERR_DEVICE
    assert b"device" in L.dq_last_error()
    assert L.dq_state_create(50, 0, ctypes.byref(h)) == capi.ERR_ARG


def test_product_does_not_import_oracle():
    pkg = os.path.join(ROOT, "dqsim_bosonic_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".cpp", ".hpp", ".hip", ".h")):
                text = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in text.replace("the oracle", ""), f


def test_profile_repr_format():
    """SimulationProfile.__repr__ on fixed values, written out from the
    reference's format string (simulator.rs:72-87: {:.2} ms total, {:4}
    calls, {:8.2} ms, {:.1}%; total floored at 1e-9 s)."""
    got = _core._profile_repr(3, 0.001, 2, 0.0025, 1, 0.0005, 0.005)
    assert got == ("SimulationProfile (total: 5.00 ms)\n"
                   "  apply_one_qubit :    3 calls      1.00 ms  (20.0%)\n"
                   "  apply_n_qubit   :    2 calls      2.50 ms  (50.0%)\n"
                   "  measure_qubit   :    1 calls      0.50 ms  (10.0%)")
    # widths are minimums; a zero total divides by 1e-9
    got = _core._profile_repr(12345, 0.0, 0, 0.0, 7, 123.456789, 0.0)
    assert got == ("SimulationProfile (total: 0.00 ms)\n"
                   "  apply_one_qubit : 12345 calls      0.00 ms  (0.0%)\n"
                   "  apply_n_qubit   :    0 calls      0.00 ms  (0.0%)\n"
                   f"  measure_qubit   :    7 calls  123456.79 ms  ({100.0 * 123.456789 / 1e-9:.1f}%)")


def test_dqsim_import_path():
    """The reference's own import line works (dqsim/__init__.py:1-27) and
    names the engine's classes."""
    import dqsim

    for name in ("StatevectorSimulator", "StabilizerSimulator", "MpsSimulator", "SimulationResult",
                 "SimulationProfile", "PBlockSimulator", "PBlockResult", "simulate_monolithic",
                 "simulate_distributed", "simulate_monolithic_shots", "simulate_distributed_shots"):
        assert name in dqsim.__all__ and hasattr(dqsim, name)
    assert dqsim.StatevectorSimulator is dq.StatevectorSimulator
    assert dqsim.simulate_monolithic is dq.simulate_monolithic
    with pytest.raises(NotImplementedError):
        dqsim.StabilizerSimulator(seed=1)
