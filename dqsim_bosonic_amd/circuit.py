"""Minimal stand-in for bosonic_model's Circuit objects.

The reference's simulators accept any object with ``model_dump_json()`` and
decode the JSON with the serde schema of src/types.rs:7-135 (tag ``kind``,
snake_case variant names, per-variant fields).  bosonic_model itself is not
installed here, so this module builds circuits that serialise to the same
schema.  Instruction classes mirror bosonic_model.instructions names
(``HInstruction(qubit=0)``, ``CxInstruction(control=0, target=1)``, ...);
extra keyword arguments such as ``qubits``/``params`` are serialised too and,
as with bosonic_model JSON, ignored by the decoder.
"""
from __future__ import annotations

import json
from typing import Any

_KIND_FIELDS: dict[str, tuple[str, ...]] = {
    **{k: ("qubit",) for k in ("id", "x", "y", "z", "h", "s", "sdg", "t", "tdg", "sx", "sxdg", "u0")},
    "u3": ("qubit", "theta", "phi", "lam"), "u": ("qubit", "theta", "phi", "lam"),
    "u2": ("qubit", "phi", "lam"), "u1": ("qubit", "lam"), "p": ("qubit", "lam"),
    "rx": ("qubit", "theta"), "ry": ("qubit", "theta"), "rz": ("qubit", "phi"),
    **{k: ("control", "target") for k in ("cx", "cz", "cy", "ch", "csx")},
    "swap": ("a", "b"),
    "crx": ("control", "target", "theta"), "cry": ("control", "target", "theta"),
    "crz": ("control", "target", "lam"), "cu1": ("control", "target", "lam"),
    "cp": ("control", "target", "lam"), "cu3": ("control", "target", "theta", "phi", "lam"),
    "cu": ("control", "target", "theta", "phi", "lam", "gamma"),
    "rxx": ("a", "b", "theta"), "rzz": ("a", "b", "theta"),
    "ccx": ("control1", "control2", "target"), "rccx": ("control1", "control2", "target"),
    "cswap": ("control", "target1", "target2"),
    **{k: ("control1", "control2", "control3", "target") for k in ("rc3x", "c3x", "c3sqrtx")},
    "c4x": ("control1", "control2", "control3", "control4", "target"),
    "gate": ("name", "params", "qubits"),
    "measure": ("qubit", "cbit"),
    "reset": ("qubit",),
    "conditional": ("condition", "op"),
    "barrier": (),
    "classical": ("name",),
}
_QUBIT_FIELDS = ("qubit", "control", "target", "a", "b", "control1", "control2", "control3", "control4",
                 "target1", "target2")


class Register:
    def __init__(self, name: str, size: int, base: int = 0):
        self.name, self.size, self.base = name, size, base

    def model_dump(self) -> dict:
        return {"name": self.name, "size": self.size, "base": self.base}


class Condition:
    def __init__(self, creg_base: int, creg_size: int, creg_value: int):
        self.creg_base, self.creg_size, self.creg_value = creg_base, creg_size, creg_value

    def model_dump(self) -> dict:
        return {"creg_base": self.creg_base, "creg_size": self.creg_size, "creg_value": self.creg_value}


class Instruction:
    kind = ""

    def __init__(self, **fields: Any):
        missing = [f for f in _KIND_FIELDS[self.kind] if f not in fields]
        if missing:
            raise TypeError(f"{type(self).__name__} missing fields {missing}")
        self.fields = fields

    def model_dump(self) -> dict:
        out: dict[str, Any] = {"kind": self.kind}
        for k, v in self.fields.items():
            out[k] = v.model_dump() if hasattr(v, "model_dump") else v
        return out


def _camel(kind: str) -> str:
    return "".join(p.capitalize() for p in kind.split("_")) + "Instruction"


for _kind in _KIND_FIELDS:
    globals()[_camel(_kind)] = type(_camel(_kind), (Instruction,), {"kind": _kind})

# bosonic_model spells the generic instruction GateInstruction
GateInstruction = globals()["GateInstruction"]


def instruction(kind: str, **fields: Any) -> Instruction:
    """Instruction of the given ``kind`` (e.g. ``instruction("cx", control=0, target=1)``)."""
    return globals()[_camel(kind)](**fields)


class Circuit:
    def __init__(self, qregs: dict, cregs: dict | None = None, instructions: list | None = None):
        self.qregs = qregs
        self.cregs = cregs or {}
        self.instructions = list(instructions or [])

    def model_dump(self) -> dict:
        return {
            "qregs": {k: r.model_dump() for k, r in self.qregs.items()},
            "cregs": {k: r.model_dump() for k, r in self.cregs.items()},
            "instructions": [i.model_dump() for i in self.instructions],
        }

    def model_dump_json(self) -> str:
        return json.dumps(self.model_dump())

    def qubits(self) -> int:
        return max((r.base + r.size for r in self.qregs.values()), default=0)


def circuit(n: int, instructions: list, n_cbits: int = 0) -> Circuit:
    """One quantum register ``q`` of n qubits (and ``c`` of n_cbits bits)."""
    cregs = {"c": Register("c", n_cbits, 0)} if n_cbits else {}
    return Circuit({"q": Register("q", n, 0)}, cregs, instructions)


class JsonCircuit:
    """Wraps ready-made JSON text (e.g. from a generator) as a circuit object."""

    def __init__(self, text: str):
        self.text = text

    def model_dump_json(self) -> str:
        return self.text


class DistributedCircuit:
    """The attributes of bosonic_model's DistributedCircuit that the P-block
    simulator reads (src/distributed/pblock/simulator.rs:286-348):
    ``qubits_per_node`` {node: [physical qubits]}, ``circuits`` {node:
    Circuit} whose instruction lists may share instruction objects (a gate
    across nodes appears in each of its nodes), and ``_instruction_index``
    {id(instruction object): global order}."""

    def __init__(self, qubits_per_node: dict, circuits: dict, instruction_index: dict | None = None):
        self.qubits_per_node = qubits_per_node
        self.circuits = circuits
        self._instruction_index = dict(instruction_index or {})


def distribute(circ: Circuit, qubits_per_node: dict) -> DistributedCircuit:
    """Splits a monolithic circuit by node: each instruction object goes to
    every node owning one of its qubits (the same object, so it is simulated
    once), in program order; barriers and classical instructions go to the
    lowest node.  Registers are kept whole in every node circuit."""
    owner = {q: node for node, qs in qubits_per_node.items() for q in qs}
    lists: dict = {node: [] for node in qubits_per_node}
    index = {}
    first = min(qubits_per_node)
    for order, ins in enumerate(circ.instructions):
        d = ins.model_dump()
        if d["kind"] == "conditional":
            d = d["op"]
        qs = [d[k] for k in _KIND_FIELDS[d["kind"]] if k in _QUBIT_FIELDS]
        if d["kind"] == "gate":
            qs = list(d["qubits"])
        nodes = sorted({owner[q] for q in qs}) or [first]
        for node in nodes:
            lists[node].append(ins)
        index[id(ins)] = order
    circuits = {node: Circuit(circ.qregs, circ.cregs, lists[node]) for node in qubits_per_node}
    return DistributedCircuit(qubits_per_node, circuits, index)
