"""Opcode-stream wire format (include/dqsim_opstream.h) on the Python side.

``encode(circuit)`` turns a circuit (any object with ``model_dump()`` or the
JSON dict) into a numpy record array plus the table of generic gate names;
``StatevectorSimulator.simulate_opstream(records, num_qubits, num_cbits,
names)`` runs it with no JSON on either side.  Generators can also fill the
record array directly (``RECORD`` is its numpy dtype).  No reference
counterpart: the reference takes circuits only as JSON
(src/monolithic/statevector/simulator.rs:329-334).
"""
from __future__ import annotations

import numpy as np

KINDS = ["id", "x", "y", "z", "h", "s", "sdg", "t", "tdg", "sx", "sxdg", "u0",
         "u3", "u", "u2", "u1", "p", "rx", "ry", "rz",
         "cx", "cz", "cy", "ch", "csx", "swap",
         "crx", "cry", "crz", "cu1", "cp", "cu3", "cu", "rxx", "rzz",
         "ccx", "cswap", "rccx", "rc3x", "c3x", "c3sqrtx", "c4x",
         "gate", "measure", "reset", "conditional", "barrier", "classical"]
OPCODE = {k: i for i, k in enumerate(KINDS)}

# JSON field names per kind (src/types.rs:60-135), qubits then parameters
_QUBITS = {
    **{k: ("qubit",) for k in ("id", "x", "y", "z", "h", "s", "sdg", "t", "tdg", "sx", "sxdg", "u0",
                               "u3", "u", "u2", "u1", "p", "rx", "ry", "rz")},
    **{k: ("control", "target") for k in ("cx", "cz", "cy", "ch", "csx", "crx", "cry", "crz", "cu1", "cp",
                                           "cu3", "cu")},
    **{k: ("a", "b") for k in ("swap", "rxx", "rzz")},
    "ccx": ("control1", "control2", "target"), "rccx": ("control1", "control2", "target"),
    "cswap": ("control", "target1", "target2"),
    **{k: ("control1", "control2", "control3", "target") for k in ("rc3x", "c3x", "c3sqrtx")},
    "c4x": ("control1", "control2", "control3", "control4", "target"),
}
_PARAMS = {"u3": ("theta", "phi", "lam"), "u": ("theta", "phi", "lam"), "u2": ("phi", "lam"), "u1": ("lam",),
           "p": ("lam",), "rx": ("theta",), "ry": ("theta",), "rz": ("phi",), "crx": ("theta",),
           "cry": ("theta",), "crz": ("lam",), "cu1": ("lam",), "cp": ("lam",), "cu3": ("theta", "phi", "lam"),
           "cu": ("theta", "phi", "lam", "gamma"), "rxx": ("theta",), "rzz": ("theta",)}

RECORD = np.dtype([("opcode", "<u2"), ("nq", "<u2"), ("cbit", "<u4"), ("qubits", "<u4", (5,)), ("pad", "<u4"),
                   ("params", "<f8", (4,))])
assert RECORD.itemsize == 64


def _emit(inst: dict, out: list, names: list, index: dict) -> None:
    k = inst["kind"]
    rec = np.zeros((), dtype=RECORD)
    rec["opcode"] = OPCODE[k]
    if k in _QUBITS:
        qs = [int(inst[f]) for f in _QUBITS[k]]
        rec["nq"] = len(qs)
        rec["qubits"][: len(qs)] = qs
        ps = [float(inst[f]) for f in _PARAMS.get(k, ())]
        rec["params"][: len(ps)] = ps
    elif k in ("gate", "classical"):
        name = inst["name"]
        if name not in index:
            index[name] = len(names)
            names.append(name)
        rec["cbit"] = index[name]
        if k == "gate":
            qs = [int(q) for q in inst["qubits"]]
            if len(qs) > 5:
                raise ValueError("a generic gate record holds at most 5 qubits")
            rec["nq"] = len(qs)
            rec["qubits"][: len(qs)] = qs
    elif k in ("measure", "reset"):
        rec["nq"] = 1
        rec["qubits"][0] = int(inst["qubit"])
        if k == "measure":
            rec["cbit"] = int(inst["cbit"])
    elif k == "conditional":
        c = inst["condition"]
        rec["qubits"][0] = int(c["creg_base"])
        rec["qubits"][1] = int(c["creg_size"])
        rec["params"].view("<u8")[0] = int(c["creg_value"])
        out.append(rec)
        _emit(inst["op"], out, names, index)
        return
    out.append(rec)


def encode(circuit) -> tuple[np.ndarray, int, int, list[str]]:
    """(records, num_qubits, num_cbits, names) of a circuit object or JSON dict."""
    d = circuit.model_dump() if hasattr(circuit, "model_dump") else circuit
    nq = max((r["base"] + r["size"] for r in d.get("qregs", {}).values()), default=0)
    ncb = max((r["base"] + r["size"] for r in d.get("cregs", {}).values()), default=0)
    out: list = []
    names: list[str] = []
    index: dict = {}
    for inst in d["instructions"]:
        _emit(inst, out, names, index)
    recs = np.array(out, dtype=RECORD) if out else np.zeros(0, dtype=RECORD)
    return recs, nq, ncb, names
