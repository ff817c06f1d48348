"""CPU tests of the opcode-stream wire format (include/dqsim_opstream.h).

The stream must decode to exactly the instruction list the JSON decoder
produces for the same circuit (kinds, qubit order, parameters bit for bit,
names, conditions), for every instruction kind of the schema
(src/types.rs:60-135).
"""
from __future__ import annotations

import numpy as np
import pytest

from dqsim_bosonic_amd import circuit as C
from dqsim_bosonic_amd import opstream as O
from dqsim_bosonic_amd import workloads as W

core = pytest.importorskip("dqsim_bosonic_amd._core")


def roundtrip(circ):
    recs, nq, ncb, names = O.encode(circ)
    a = core._decode_opstream(recs, nq, ncb, names)
    b = core._decode(circ.model_dump_json())
    return a, b


@pytest.mark.parametrize("seed", range(5))
def test_every_kind_matches_json_decoding(seed):
    circ = W.random_mixed(7, 300, seed=seed)
    a, b = roundtrip(circ)
    assert a == b


def test_measure_reset_conditional_and_names():
    circ = W.random_blocks(8, 12, seed=3)  # measure / reset / conditional, generic gate names
    a, b = roundtrip(circ)
    assert a == b
    extra = C.circuit(3, [C.ClassicalInstruction(name="note"), C.BarrierInstruction(),
                          C.ConditionalInstruction(condition=C.Condition(creg_base=1, creg_size=2, creg_value=2**40 + 3),
                                                   op=C.ConditionalInstruction(condition=C.Condition(
                                                       creg_base=0, creg_size=1, creg_value=1),
                                                       op=C.RzInstruction(qubit=2, phi=-1.25)))], 3)
    a, b = roundtrip(extra)
    assert a == b


def test_record_layout_and_errors():
    assert O.RECORD.itemsize == 64
    recs, nq, ncb, names = O.encode(C.circuit(2, [C.HInstruction(qubit=1)]))
    bad = recs.copy()
    bad["opcode"][0] = 999
    with pytest.raises(ValueError, match="unknown opcode"):
        core._decode_opstream(bad, nq, ncb, names)
    cond = np.zeros(1, dtype=O.RECORD)
    cond["opcode"][0] = O.OPCODE["conditional"]
    with pytest.raises(ValueError, match="end of the stream"):
        core._decode_opstream(cond, 2, 1, [])
    with pytest.raises(ValueError):
        core._decode_opstream(np.zeros(3, dtype=np.float64), 2, 0, [])
