"""CPU oracle for the statevector hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg import this module, and only as the checker or the timed CPU baseline.
The product (``dqsim_bosonic_amd``) never imports it.

The numeric kernels live in ``dqsim_oracle.c`` (built into ``liboracle.so``
by ``oracle/Makefile``); this module restates the instruction dispatcher of
the reference around them, independently of the product's C++ decoder:

* ``OracleSimulator.simulate`` restates ``StatevectorSimulator::simulate``
  (src/monolithic/statevector/simulator.rs:325-379): |0..0> init,
  ChaCha8Rng::seed_from_u64(seed), instruction loop.
* ``_run`` restates ``run_instruction`` (simulator.rs:386-632), including the
  generic ``gate`` names (:571-600), measurement (:603-606), conditional
  (:609-618) and reset (:621-626).
* ``probabilities`` / ``counts`` / ``fidelity`` restate
  simulator.rs:148-191 over ``marginal_probs`` / ``sample_counts``
  (src/engine.rs:368-433).
* ``num_qubits`` restates src/types.rs:24-30.
* ``simulate_shots_json`` restates ``StatevectorSimulator::simulate_shots``
  (simulator.rs:215-322) with ``fuse_circuit`` / ``gate_matrix_1q`` /
  ``matmul2x2`` / ``is_identity_2x2`` (src/types.rs:149-293) and
  ``format_cbits`` (types.rs:43-51).
* ``mps_to_statevector_loop`` restates ``Mps::to_statevector``
  (src/monolithic/mps/simulator.rs:246-264) literally; ``mps_contract`` is
  the same contraction vectorised, ``mps_norm2_transfer`` / ``mps_amplitude``
  independent closed forms for large cases.  ``OracleMps`` /
  ``simulate_mps_json`` / ``simulate_mps_shots_json`` restate the MPS
  simulator (mps/simulator.rs:68-563) with numpy's SVD.

Parity status of each piece is listed in the test modules that pin it.
"""
from __future__ import annotations

import ctypes
import json
import os
import subprocess
from typing import Any

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build() -> str:
    """Compile liboracle.so with oracle/Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE, "-B", "liboracle.so"], check=True)
    return _LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        L.or_apply_one_qubit.argtypes = [dp, ctypes.c_int, dp, ctypes.c_int, ip, ip, ctypes.c_int]
        L.or_apply_n_qubit.argtypes = [dp, ctypes.c_int, dp, ip, ctypes.c_int]
        L.or_apply_one_qubit_par.argtypes = [dp, ctypes.c_int, dp, ctypes.c_int]
        L.or_apply_n_qubit_par.argtypes = [dp, ctypes.c_int, dp, ip, ctypes.c_int]
        L.or_prob_one.argtypes = [dp, ctypes.c_int, ctypes.c_int]
        L.or_prob_one.restype = ctypes.c_double
        L.or_prob_one_par.argtypes = [dp, ctypes.c_int, ctypes.c_int]
        L.or_prob_one_par.restype = ctypes.c_double
        L.or_collapse.argtypes = [dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double]
        L.or_collapse_par.argtypes = [dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double]
        L.or_marginal_probs.argtypes = [dp, ctypes.c_int, ip, ctypes.c_int, dp]
        L.or_marginal_probs_literal.argtypes = [dp, ctypes.c_int, ip, ctypes.c_int, dp]
        L.or_sample_counts.argtypes = [dp, ctypes.c_int, ctypes.c_uint64, ctypes.c_void_p,
                                       ctypes.POINTER(ctypes.c_uint32)]
        L.or_inner.argtypes = [dp, dp, ctypes.c_uint64, dp, dp]
        L.or_gate_matrix.argtypes = [ctypes.c_char_p, dp, ctypes.c_int, dp]
        L.or_gate_matrix.restype = ctypes.c_int
        L.or_rng_seed_from_u64.argtypes = [ctypes.c_void_p, ctypes.c_uint64]
        L.or_rng_next_u64.argtypes = [ctypes.c_void_p]
        L.or_rng_next_u64.restype = ctypes.c_uint64
        L.or_rng_next_u32.argtypes = [ctypes.c_void_p]
        L.or_rng_next_u32.restype = ctypes.c_uint32
        L.or_rng_gen_f64.argtypes = [ctypes.c_void_p]
        L.or_rng_gen_f64.restype = ctypes.c_double
        L.or_chacha_block.argtypes = [ctypes.POINTER(ctypes.c_uint32), ctypes.c_uint64,
                                      ctypes.c_uint64, ctypes.c_int,
                                      ctypes.POINTER(ctypes.c_uint32)]
        L.or_num_threads.restype = ctypes.c_int
        L.or_mps_amplitudes.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_uint64), dp, ctypes.c_uint64,
                                        ctypes.c_uint64, dp]
        L.or_pcg32_output.argtypes = [ctypes.c_uint64]
        L.or_pcg32_output.restype = ctypes.c_uint32
        L.or_rng_from_seed.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int]
        _lib = L
    return _lib


def _dp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


class Rng:
    """ChaCha8Rng::seed_from_u64 + gen::<f64>() (restated in dqsim_oracle.c)."""

    _SIZE = 8 * 4 + 8 + 64 * 4 + 4 + 4

    def __init__(self, seed: int | None = None, seed_bytes: bytes | None = None, rounds: int = 8):
        self._buf = ctypes.create_string_buffer(self._SIZE + 16)
        if seed_bytes is not None:
            lib().or_rng_from_seed(self._buf, seed_bytes, rounds)
        else:
            lib().or_rng_seed_from_u64(self._buf, ctypes.c_uint64(seed & (2**64 - 1)))

    def next_u64(self) -> int:
        return int(lib().or_rng_next_u64(self._buf))

    def next_u32(self) -> int:
        return int(lib().or_rng_next_u32(self._buf))

    def gen_f64(self) -> float:
        return float(lib().or_rng_gen_f64(self._buf))


# -- gates.rs ----------------------------------------------------------------

_PARAM_FIELDS = {
    "u3": ("theta", "phi", "lam"), "u": ("theta", "phi", "lam"), "u2": ("phi", "lam"),
    "u1": ("lam",), "p": ("lam",), "rx": ("theta",), "ry": ("theta",), "rz": ("phi",),
    "crx": ("theta",), "cry": ("theta",), "crz": ("lam",), "cu1": ("lam",), "cp": ("lam",),
    "cu3": ("theta", "phi", "lam"), "cu": ("theta", "phi", "lam", "gamma"),
    "rxx": ("theta",), "rzz": ("theta",),
}
_QUBIT_FIELDS = {
    **{k: ("qubit",) for k in ("id", "u0", "x", "y", "z", "h", "s", "sdg", "t", "tdg", "sx",
                               "sxdg", "u3", "u", "u2", "u1", "p", "rx", "ry", "rz")},
    **{k: ("control", "target") for k in ("cx", "cz", "cy", "ch", "csx", "crx", "cry", "crz",
                                           "cu1", "cp", "cu3", "cu")},
    **{k: ("a", "b") for k in ("swap", "rxx", "rzz")},
    "ccx": ("control1", "control2", "target"), "rccx": ("control1", "control2", "target"),
    "cswap": ("control", "target1", "target2"),
    **{k: ("control1", "control2", "control3", "target") for k in ("rc3x", "c3x", "c3sqrtx")},
    "c4x": ("control1", "control2", "control3", "control4", "target"),
}
# generic `gate` names the statevector dispatcher accepts (simulator.rs:571-600)
_GENERIC = {
    "remote_link_phi_plus": "remote_link_phi_plus", "remote_link_psi_minus": "remote_link_psi_minus",
    "remote_link_psi_plus": "remote_link_psi_plus", "nonlocal_cz": "cz", "remote_cz": "cz",
    "remote_cx": "cx", "remote_epr": "remote_link_phi_plus",
}


def gate_matrix(name: str, params=()) -> np.ndarray:
    """Gate matrix from the oracle's restatement of src/gates.rs."""
    p = np.ascontiguousarray(params, dtype=np.float64)
    out = np.zeros(2 * 32 * 32, dtype=np.float64)
    k = lib().or_gate_matrix(name.encode(), _dp(p) if p.size else None, int(p.size), _dp(out))
    if k < 0:
        raise KeyError(name)
    d = 1 << k
    return out[: 2 * d * d].view(np.complex128).reshape(d, d).copy()


# -- engine.rs ---------------------------------------------------------------

def apply_one_qubit(state: np.ndarray, n: int, u: np.ndarray, target: int, controls=()):
    u = np.ascontiguousarray(u, dtype=np.complex128)
    w = np.array([c[0] for c in controls], dtype=np.int32)
    v = np.array([int(c[1]) for c in controls], dtype=np.int32)
    lib().or_apply_one_qubit(_dp(state.view(np.float64)), n, _dp(u.view(np.float64)), target,
                             _ip(w), _ip(v), len(controls))


def apply_n_qubit(state: np.ndarray, n: int, u: np.ndarray, qubits, parallel=False):
    u = np.ascontiguousarray(u, dtype=np.complex128)
    q = np.ascontiguousarray(qubits, dtype=np.int32)
    f = lib().or_apply_n_qubit_par if parallel else lib().or_apply_n_qubit
    f(_dp(state.view(np.float64)), n, _dp(u.view(np.float64)), _ip(q), len(q))


def prob_one(state: np.ndarray, n: int, qubit: int) -> float:
    return float(lib().or_prob_one(_dp(state.view(np.float64)), n, qubit))


def collapse(state: np.ndarray, n: int, qubit: int, outcome: int, p1: float):
    lib().or_collapse(_dp(state.view(np.float64)), n, qubit, outcome, p1)


def marginal_probs(state: np.ndarray, n: int, qubits, literal=False) -> np.ndarray:
    q = np.ascontiguousarray(qubits, dtype=np.int32)
    out = np.zeros(1 << len(q), dtype=np.float64)
    f = lib().or_marginal_probs_literal if literal else lib().or_marginal_probs
    f(_dp(state.view(np.float64)), n, _ip(q), len(q), _dp(out))
    return out


def sample_counts(probs: np.ndarray, k: int, shots: int, rng: Rng) -> dict:
    idx = np.zeros(shots, dtype=np.uint32)
    lib().or_sample_counts(_dp(np.ascontiguousarray(probs)), k, shots, rng._buf,
                           idx.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
    out: dict[str, int] = {}
    for v, c in zip(*np.unique(idx, return_counts=True)):
        out[format(int(v), "b").zfill(k)] = int(c)
    return out


def inner(a: np.ndarray, b: np.ndarray) -> complex:
    re, im = ctypes.c_double(), ctypes.c_double()
    lib().or_inner(_dp(a.view(np.float64)), _dp(b.view(np.float64)), a.size,
                   ctypes.byref(re), ctypes.byref(im))
    return complex(re.value, im.value)


# -- simulator.rs ------------------------------------------------------------

def num_qubits(circ: dict) -> int:
    """types.rs:24-30: max(base + size) over qregs, 0 if none."""
    return max((r["base"] + r["size"] for r in circ.get("qregs", {}).values()), default=0)


class OracleResult:
    def __init__(self, state, n, cbits, p1s):
        self.statevector = state
        self.num_qubits = n
        self.classical_bits = cbits
        self.p1_record = p1s  # p1 seen by every measurement, in program order

    def probabilities(self, qubits=None) -> dict:
        qs = list(range(self.num_qubits - 1, -1, -1)) if qubits is None else list(qubits)
        pr = marginal_probs(self.statevector, self.num_qubits, qs)
        return {j: float(p) for j, p in enumerate(pr) if p > 0.0}

    def counts(self, shots=1000, qubits=None, seed=None) -> dict:
        qs = list(range(self.num_qubits - 1, -1, -1)) if qubits is None else list(qubits)
        pr = marginal_probs(self.statevector, self.num_qubits, qs)
        return sample_counts(pr, len(qs), shots, Rng(seed))

    def fidelity(self, other: np.ndarray) -> float:
        if other.size != self.statevector.size:
            return 0.0
        z = inner(self.statevector, np.ascontiguousarray(other, dtype=np.complex128))
        return z.real * z.real + z.imag * z.imag


class OracleSimulator:
    """Restatement of StatevectorSimulator::simulate (simulator.rs:325-379)."""

    def __init__(self, seed: int, parallel: bool = False):
        self.seed = seed
        self.parallel = parallel
        # engine calls by the reference's profiling wrappers (simulator.rs:890-946):
        # do_oq / do_nq / do_mq, one count per call that the dispatcher makes
        self.calls = {"apply_one_qubit": 0, "apply_n_qubit": 0, "measure_qubit": 0}

    def simulate_json(self, text: str, max_instructions: int | None = None) -> OracleResult:
        circ = json.loads(text)
        n = num_qubits(circ)
        state = np.zeros(1 << n, dtype=np.complex128)
        state[0] = 1.0
        cbits: dict[int, int] = {}
        rng = Rng(self.seed)
        p1s: list[float] = []
        insts = circ["instructions"]
        if max_instructions is not None:
            insts = insts[:max_instructions]
        for inst in insts:
            self._run(state, n, inst, cbits, rng, p1s)
        return OracleResult(state, n, cbits, p1s)

    def simulate(self, circuit: Any) -> OracleResult:
        return self.simulate_json(circuit.model_dump_json())

    # run_instruction, simulator.rs:386-632
    def _run(self, state, n, inst, cbits, rng, p1s):
        kind = inst["kind"]
        if kind in ("id", "u0", "barrier", "classical"):
            return
        if kind == "measure":
            cbits[inst["cbit"]] = self._measure(state, n, inst["qubit"], rng, p1s)
            return
        if kind == "reset":
            if self._measure(state, n, inst["qubit"], rng, p1s) == 1:
                self._oq(state, n, gate_matrix("x"), inst["qubit"])
            return
        if kind == "conditional":
            c = inst["condition"]
            actual = 0
            for b in range(c["creg_size"]):
                actual |= cbits.get(c["creg_base"] + b, 0) << b
            if actual == c["creg_value"]:
                self._run(state, n, inst["op"], cbits, rng, p1s)
            return
        if kind == "gate":
            name = inst["name"].lower()
            if name in ("remote_barrier", "remote_cu1"):
                return
            if name not in _GENERIC:
                raise NotImplementedError(
                    f"Unsupported generic gate: {name!r}. Decompose it before simulating.")
            self._nq(state, n, gate_matrix(_GENERIC[name]), inst["qubits"])
            return
        params = [float(inst[f]) for f in _PARAM_FIELDS.get(kind, ())]
        qubits = [int(inst[f]) for f in _QUBIT_FIELDS[kind]]
        u = gate_matrix(kind, params)
        if len(qubits) == 1:
            self._oq(state, n, u, qubits[0])
        else:
            self._nq(state, n, u, qubits)

    # run_instruction_par (simulator.rs:638-866): the shots dispatcher; same
    # arms with the sequential kernels, and `circuit-*` generic names are
    # no-ops there (:825-827).
    def _run_shot(self, state, n, inst, cbits, rng):
        if inst["kind"] == "gate" and inst["name"].lower().startswith("circuit-"):
            return
        if inst["kind"] == "conditional":
            c = inst["condition"]
            actual = 0
            for b in range(c["creg_size"]):
                actual |= cbits.get(c["creg_base"] + b, 0) << b
            if actual == c["creg_value"]:
                self._run_shot(state, n, inst["op"], cbits, rng)
            return
        self._run(state, n, inst, cbits, rng, [])

    def _oq(self, state, n, u, q):
        self.calls["apply_one_qubit"] += 1
        if self.parallel and n >= 12:
            u = np.ascontiguousarray(u, dtype=np.complex128)
            lib().or_apply_one_qubit_par(_dp(state.view(np.float64)), n, _dp(u.view(np.float64)), q)
        else:
            apply_one_qubit(state, n, u, q)

    def _nq(self, state, n, u, qs):
        self.calls["apply_n_qubit"] += 1
        apply_n_qubit(state, n, u, qs, parallel=self.parallel and n >= 12)

    def _measure(self, state, n, q, rng, p1s):
        self.calls["measure_qubit"] += 1
        d = _dp(state.view(np.float64))
        if self.parallel and n >= 12:
            p1 = float(lib().or_prob_one_par(d, n, q))
        else:
            p1 = float(lib().or_prob_one(d, n, q))
        outcome = 1 if rng.gen_f64() < p1 else 0
        if self.parallel and n >= 12:
            lib().or_collapse_par(d, n, q, outcome, p1)
        else:
            lib().or_collapse(d, n, q, outcome, p1)
        p1s.append(p1)
        return outcome


# -- shots: simulate_shots (simulator.rs:215-322), fuse_circuit (types.rs:149-293)

def _cm(a, b):
    """num-complex multiply on (re, im) tuples: separately rounded products."""
    return (a[0] * b[0] - a[1] * b[1], a[0] * b[1] + a[1] * b[0])


def _ca(a, b):
    return (a[0] + b[0], a[1] + b[1])


def fused_gate_matrix_1q(inst: dict):
    """gate_matrix_1q (types.rs:164-230): the fusable single-qubit kinds and
    the matrices the shots path multiplies; None for anything else.  Returns
    ((re, im) x 4) row-major."""
    import math
    k = inst["kind"]
    r = lambda x: (x, 0.0)  # noqa: E731
    s2 = 0.70710678118654752440
    neg = lambda a: (-a[0], -a[1])  # noqa: E731
    if k in ("id", "u0"):
        return (r(1.0), r(0.0), r(0.0), r(1.0))
    if k == "x":
        return (r(0.0), r(1.0), r(1.0), r(0.0))
    if k == "y":
        return (r(0.0), (0.0, -1.0), (0.0, 1.0), r(0.0))
    if k == "z":
        return (r(1.0), r(0.0), r(0.0), r(-1.0))
    if k == "h":
        return (r(s2), r(s2), r(s2), r(-s2))
    if k == "s":
        return (r(1.0), r(0.0), r(0.0), (0.0, 1.0))
    if k == "sdg":
        return (r(1.0), r(0.0), r(0.0), (0.0, -1.0))
    if k == "t":
        return (r(1.0), r(0.0), r(0.0), (s2, s2))
    if k == "tdg":
        return (r(1.0), r(0.0), r(0.0), (s2, -s2))
    if k == "sx":
        return ((0.5, 0.5), (0.5, -0.5), (0.5, -0.5), (0.5, 0.5))
    if k == "sxdg":
        return ((0.5, -0.5), (0.5, 0.5), (0.5, 0.5), (0.5, -0.5))
    if k in ("u3", "u", "u2"):
        if k == "u2":
            theta, phi, lam = math.pi / 2.0, float(inst["phi"]), float(inst["lam"])
        else:
            theta, phi, lam = float(inst["theta"]), float(inst["phi"]), float(inst["lam"])
        ct, st = math.cos(theta / 2.0), math.sin(theta / 2.0)
        return (r(ct), neg(_cm((math.cos(lam), math.sin(lam)), r(st))),
                _cm((math.cos(phi), math.sin(phi)), r(st)),
                _cm((math.cos(phi + lam), math.sin(phi + lam)), r(ct)))
    if k in ("u1", "p"):
        lam = float(inst["lam"])
        return (r(1.0), r(0.0), r(0.0), (math.cos(lam), math.sin(lam)))
    if k == "rx":
        th = float(inst["theta"])
        cv, sv = math.cos(th / 2.0), math.sin(th / 2.0)
        return (r(cv), (0.0, -sv), (0.0, -sv), r(cv))
    if k == "ry":
        th = float(inst["theta"])
        cv, sv = math.cos(th / 2.0), math.sin(th / 2.0)
        return (r(cv), r(-sv), r(sv), r(cv))
    if k == "rz":
        ph = float(inst["phi"])
        return ((math.cos(-ph / 2.0), math.sin(-ph / 2.0)), r(0.0), r(0.0),
                (math.cos(ph / 2.0), math.sin(ph / 2.0)))
    return None


def matmul2x2(a, b):
    """types.rs:149-160, a * b."""
    return (_ca(_cm(a[0], b[0]), _cm(a[1], b[2])), _ca(_cm(a[0], b[1]), _cm(a[1], b[3])),
            _ca(_cm(a[2], b[0]), _cm(a[3], b[2])), _ca(_cm(a[2], b[1]), _cm(a[3], b[3])))


def is_identity_2x2(m) -> bool:
    """types.rs:239-244 (num-complex norm = hypot)."""
    import math
    return (math.hypot(m[0][0] - 1.0, m[0][1]) < 1e-10 and math.hypot(*m[1]) < 1e-10
            and math.hypot(*m[2]) < 1e-10 and math.hypot(m[3][0] - 1.0, m[3][1]) < 1e-10)


def inst_qubits(inst: dict) -> list:
    """Instruction::qubits (types.rs:348-418)."""
    k = inst["kind"]
    if k == "conditional":
        return inst_qubits(inst["op"])
    if k in ("barrier", "classical"):
        return []
    if k == "gate":
        return list(inst["qubits"])
    if k in ("measure", "reset"):
        return [inst["qubit"]]
    return [int(inst[f]) for f in _QUBIT_FIELDS[k]]


def fuse_circuit(insts: list) -> list:
    """types.rs:254-293: ("fused", qubit, matrix) or ("orig", inst).  The
    qubits still pending at the end are flushed in ascending order (the
    reference iterates a HashMap there; those gates act on distinct qubits
    after the last measurement, so the order changes no outcome)."""
    pending: dict = {}
    out = []
    ident = ((1.0, 0.0), (0.0, 0.0), (0.0, 0.0), (1.0, 0.0))

    def flush(q):
        m = pending.pop(q, None)
        if m is not None and not is_identity_2x2(m):
            out.append(("fused", q, m))

    for inst in insts:
        m1 = fused_gate_matrix_1q(inst) if inst["kind"] in _QUBIT_FIELDS and \
            len(_QUBIT_FIELDS[inst["kind"]]) == 1 else None
        if m1 is not None:
            q = int(inst["qubit"])
            pending[q] = matmul2x2(m1, pending.get(q, ident))
            continue
        for q in inst_qubits(inst):
            flush(q)
        out.append(("orig", inst))
    for q in sorted(pending):
        flush(q)
    return out


def format_cbits(cbits: dict, num_cbits: int) -> str:
    """types.rs:43-51."""
    return "".join(str(cbits.get(i, 0)) for i in range(num_cbits - 1, -1, -1))


def num_cbits(circ: dict) -> int:
    """types.rs:32-39."""
    return max((r["base"] + r["size"] for r in circ.get("cregs", {}).values()), default=0)


def simulate_shots_json(text: str, shots: int, base_seed: int) -> dict:
    """StatevectorSimulator::simulate_shots (simulator.rs:215-322): each shot
    from |0..0> with ChaCha8Rng::seed_from_u64(base_seed + i), the fused list
    through run_fused_par (:870-884) with the sequential kernels."""
    circ = json.loads(text)
    n = num_qubits(circ)
    ncb = num_cbits(circ)
    fused = fuse_circuit(circ["instructions"])
    sim = OracleSimulator(seed=0, parallel=False)
    counts: dict[str, int] = {}
    for i in range(shots):
        state = np.zeros(1 << n, dtype=np.complex128)
        state[0] = 1.0
        cbits: dict[int, int] = {}
        rng = Rng((base_seed + i) & (2**64 - 1))
        for item in fused:
            if item[0] == "fused":
                u = np.array([complex(*e) for e in item[2]], dtype=np.complex128).reshape(2, 2)
                apply_one_qubit(state, n, u, item[1])
            else:
                sim._run_shot(state, n, item[1], cbits, rng)
        key = format_cbits(cbits, ncb)
        counts[key] = counts.get(key, 0) + 1
    return counts


# -- MPS: src/monolithic/mps/simulator.rs --------------------------------------

def mps_to_statevector_loop(tensors) -> np.ndarray:
    """Mps::to_statevector (mps/simulator.rs:246-264), literally: for every
    basis index, work = [1]; next[r] += work[l] * A_q[l][bit q][r] over the
    sites in order; amplitude = work[0].  Pure-Python loops: small n only."""
    n = len(tensors)
    out = np.zeros(1 << n, dtype=np.complex128)
    for basis in range(1 << n):
        work = [1.0 + 0.0j]
        for q, a in enumerate(tensors):
            bit = (basis >> q) & 1
            nxt = [0.0 + 0.0j] * a.shape[2]
            for left in range(min(len(work), a.shape[0])):
                for right in range(a.shape[2]):
                    nxt[right] += work[left] * a[left, bit, right]
            work = nxt
        out[basis] = work[0]
    return out


def mps_amplitudes_c(tensors, x_begin: int, count: int) -> np.ndarray:
    """The literal loop of Mps::to_statevector in C (dqsim_oracle.c) for a
    range of basis indices: the CPU baseline of the contraction."""
    bond = np.array([1] + [int(a.shape[2]) for a in tensors], dtype=np.uint64)
    flat = np.concatenate([np.ascontiguousarray(a, dtype=np.complex128).ravel() for a in tensors])
    out = np.zeros(count, dtype=np.complex128)
    lib().or_mps_amplitudes(len(tensors), bond.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                            _dp(flat.view(np.float64)), x_begin, count, _dp(out.view(np.float64)))
    return out


def mps_contract(tensors) -> np.ndarray:
    """The same contraction vectorised over basis indices: psi[x][bond] grows
    one site at a time, the new site's bit becoming index bit q."""
    psi = np.ones((1, 1), dtype=np.complex128)
    for a in tensors:
        psi = np.einsum("xl,lbr->bxr", psi, a).reshape(-1, a.shape[2])
    return psi[:, 0].copy()


def mps_norm2_transfer(tensors, qubit: int | None = None) -> float:
    """sum_x |psi(x)|^2 (restricted to bit `qubit` = 1 when given) by the
    transfer-matrix contraction E' = sum_b A[b]^H E A[b], O(n chi^3): a closed
    form independent of the dense contraction."""
    e = np.ones((1, 1), dtype=np.complex128)
    for q, a in enumerate(tensors):
        bits = (1,) if q == qubit else (0, 1)
        e = sum(a[:, b, :].conj().T @ e @ a[:, b, :] for b in bits)
    return float(e[0, 0].real)


def mps_amplitude(tensors, x: int) -> complex:
    """One amplitude as the matrix-chain product A_0[b_0] ... A_{n-1}[b_{n-1}]."""
    v = np.ones((1, 1), dtype=np.complex128)
    for q, a in enumerate(tensors):
        v = v @ a[:, (x >> q) & 1, :]
    return complex(v[0, 0])


class OracleMps:
    """Mps (mps/simulator.rs:68-298) with numpy's LAPACK SVD (singular values
    decreasing, as nalgebra's DMatrix::svd returns them)."""

    def __init__(self, n: int, max_bond: int | None = None, threshold: float = 1e-12):
        self.t = []
        for _ in range(n):
            a = np.zeros((1, 2, 1), dtype=np.complex128)
            a[0, 0, 0] = 1.0
            self.t.append(a)
        self.max_bond = max_bond
        self.thr = threshold
        self.svd_calls = 0
        # truncations whose cut falls inside a group of equal singular values:
        # the kept subspace is then a basis choice of the SVD routine
        self.degenerate_cuts = 0

    def apply_1q(self, q: int, u: np.ndarray):
        """apply_1q, :94-110."""
        self.t[q] = np.einsum("oi,lir->lor", u, self.t[q])

    def apply_2q(self, a: int, b: int, u: np.ndarray):
        """apply_2q, :112-135: SWAP network, gate at (lo, lo + 1), SWAP back."""
        if a == b:
            return
        lo, hi = min(a, b), max(a, b)
        sw = gate_matrix("swap")
        for pos in range(hi - 1, lo, -1):
            self.apply_adjacent(pos, sw)
        if a > b:  # reverse_2q_order, :565-579
            u = u.reshape(2, 2, 2, 2).transpose(1, 0, 3, 2).reshape(4, 4)
        self.apply_adjacent(lo, u)
        for pos in range(lo + 1, hi):
            self.apply_adjacent(pos, sw)

    def apply_adjacent(self, q: int, u: np.ndarray):
        """apply_adjacent_2q, :137-244: merge, gate, SVD, truncate, rescale."""
        A, B = self.t[q], self.t[q + 1]
        L, R = A.shape[0], B.shape[2]
        p = np.einsum("lik,kjr->lijr", A, B)
        theta = np.einsum("abij,lijr->labr", u.reshape(2, 2, 2, 2), p).reshape(L * 2, 2 * R)
        uu, s, vh = np.linalg.svd(theta, full_matrices=False)
        self.svd_calls += 1
        keep = max(int(np.sum(s > self.thr)), 1)
        if self.max_bond is not None:
            keep = min(keep, max(self.max_bond, 1))
        if keep < s.size and s[keep] > self.thr and s[keep - 1] - s[keep] <= 1e-9 * s[0]:
            self.degenerate_cuts += 1
        kept = max(float(np.sqrt(np.sum(s[:keep] ** 2))), 1e-15)
        total = float(np.sqrt(np.sum(s ** 2)))
        discarded = float(np.sum(s[keep:] ** 2))
        scale = total / kept if discarded > 0.0 else 1.0
        self.t[q] = uu[:, :keep].reshape(L, 2, keep)
        self.t[q + 1] = ((s[:keep] * scale)[:, None] * vh[:keep]).reshape(keep, 2, R)

    def measure(self, q: int, rng: "Rng") -> int:
        """measure, :266-279: p1 from the dense state, clamp, draw, project."""
        state = mps_contract(self.t)
        idx = np.arange(state.size)
        p1 = float(np.sum(np.abs(state[(idx >> q) & 1 == 1]) ** 2))
        p1 = min(max(p1, 0.0), 1.0)
        outcome = 1 if rng.gen_f64() < p1 else 0
        prob = p1 if outcome else 1.0 - p1
        scale = 1.0 / np.sqrt(prob) if prob > 0.0 else 0.0
        a = self.t[q] * scale
        a[:, 1 - outcome, :] = 0.0
        self.t[q] = a
        return outcome


_MPS_GENERIC = {
    "remote_link_phi_plus": "remote_link_phi_plus", "remote_epr": "remote_link_phi_plus",
    "epr": "remote_link_phi_plus", "remote_link_psi_minus": "remote_link_psi_minus",
    "remote_link_psi_plus": "remote_link_psi_plus", "nonlocal_cz": "cz", "remote_cz": "cz",
    "remote_cx": "cx",
}


def _mps_run(mps: OracleMps, inst: dict, cbits: dict, rng: "Rng"):
    """run_instruction of the MPS simulator, mps/simulator.rs:423-563."""
    kind = inst["kind"]
    if kind in ("id", "u0", "barrier", "classical"):
        return
    if kind == "measure":
        cbits[inst["cbit"]] = mps.measure(inst["qubit"], rng)
        return
    if kind == "reset":
        if mps.measure(inst["qubit"], rng) == 1:
            mps.apply_1q(inst["qubit"], gate_matrix("x"))
        return
    if kind == "conditional":
        c = inst["condition"]
        actual = 0
        for b in range(c["creg_size"]):
            actual |= cbits.get(c["creg_base"] + b, 0) << b
        if actual == c["creg_value"]:
            _mps_run(mps, inst["op"], cbits, rng)
        return
    if kind == "gate":
        name = inst["name"].lower()
        if name in ("remote_barrier", "remote_cu1") or name.startswith("circuit-"):
            return
        if name not in _MPS_GENERIC:
            raise NotImplementedError(f"MPS simulator does not support generic gate {name!r}")
        mps.apply_2q(inst["qubits"][0], inst["qubits"][1], gate_matrix(_MPS_GENERIC[name]))
        return
    qubits = [int(inst[f]) for f in _QUBIT_FIELDS[kind]]
    if len(qubits) > 2:
        raise NotImplementedError("MPS simulator only supports one- and two-qubit unitary gates plus "
                                  "measurement, reset, and conditionals")
    u = gate_matrix(kind, [float(inst[f]) for f in _PARAM_FIELDS.get(kind, ())])
    if len(qubits) == 1:
        mps.apply_1q(qubits[0], u)
    else:
        mps.apply_2q(qubits[0], qubits[1], u)


def simulate_mps_json(text: str, seed: int, max_bond: int | None = None, threshold: float = 1e-12):
    """MpsSimulator::simulate (mps/simulator.rs:317-345): (statevector, cbits, mps)."""
    circ = json.loads(text)
    mps = OracleMps(num_qubits(circ), max_bond, threshold)
    cbits: dict[int, int] = {}
    rng = Rng(seed)
    for inst in circ["instructions"]:
        _mps_run(mps, inst, cbits, rng)
    return mps_contract(mps.t), cbits, mps


def simulate_mps_shots_json(text: str, shots: int, base_seed: int, max_bond: int | None = None,
                            threshold: float = 1e-12) -> dict:
    """MpsSimulator::simulate_shots (mps/simulator.rs:347-421)."""
    circ = json.loads(text)
    n, ncb = num_qubits(circ), num_cbits(circ)
    counts: dict[str, int] = {}
    for i in range(shots):
        mps = OracleMps(n, max_bond, threshold)
        cbits: dict[int, int] = {}
        rng = Rng((base_seed + i) & (2**64 - 1))
        for inst in circ["instructions"]:
            _mps_run(mps, inst, cbits, rng)
        key = format_cbits(cbits, ncb)
        counts[key] = counts.get(key, 0) + 1
    return counts


# -- P-block: src/distributed/pblock --------------------------------------------

def merge_blocks_loop(a_state, a_qubits, b_state, b_qubits):
    """merge_blocks (pblock/model.rs:30-58) literally: the product of the two
    factors' amplitudes over the sorted union of their qubits."""
    qubits = sorted(set(a_qubits) | set(b_qubits))
    out = np.zeros(1 << len(qubits), dtype=np.complex128)
    for i in range(out.size):
        al = 0
        for bit, q in enumerate(a_qubits):
            if (i >> qubits.index(q)) & 1:
                al |= 1 << bit
        bl = 0
        for bit, q in enumerate(b_qubits):
            if (i >> qubits.index(q)) & 1:
                bl |= 1 << bit
        x, y = complex(a_state[al]), complex(b_state[bl])
        out[i] = complex(x.real * y.real - x.imag * y.imag, x.real * y.imag + x.imag * y.real)
    return out, qubits


def merge_blocks(a_state, a_qubits, b_state, b_qubits):
    """merge_blocks vectorised over the merged index."""
    qubits = sorted(set(a_qubits) | set(b_qubits))
    idx = np.arange(1 << len(qubits), dtype=np.int64)
    al = np.zeros_like(idx)
    bl = np.zeros_like(idx)
    for bit, q in enumerate(a_qubits):
        al |= ((idx >> qubits.index(q)) & 1) << bit
    for bit, q in enumerate(b_qubits):
        bl |= ((idx >> qubits.index(q)) & 1) << bit
    x, y = a_state[al], b_state[bl]
    out = np.empty(idx.size, dtype=np.complex128)  # num-complex product, one rounding per operation
    out.real = x.real * y.real - x.imag * y.imag
    out.imag = x.real * y.imag + x.imag * y.real
    return out, qubits


class OracleBlockPool:
    """BlockPool (pblock/model.rs:60-132)."""

    def __init__(self, qpn: dict):
        self.blocks = []
        self.q2b = {}
        self.merge_count = 0
        for node in sorted(qpn):
            qs = sorted(set(qpn[node]))
            for q in qpn[node]:
                self.q2b[q] = len(self.blocks)
            st = np.zeros(1 << len(qs), dtype=np.complex128)
            st[0] = 1.0
            self.blocks.append([st, qs])

    def block_of(self, q):
        if q not in self.q2b:
            raise ValueError(f"qubit {q} is not assigned to any block")
        return self.q2b[q]

    def merge(self, keep, drop):
        if keep == drop:
            return
        a, b = self.blocks[keep], self.blocks[drop]
        self.blocks[drop] = None
        for q in b[1]:
            self.q2b[q] = keep
        self.blocks[keep] = list(merge_blocks(a[0], a[1], b[0], b[1]))
        self.merge_count += 1

    def ensure_single_block(self, qubits):
        if not qubits:
            return 0
        target = self.block_of(qubits[0])
        for q in qubits[1:]:
            bq = self.block_of(q)
            if bq != target:
                self.merge(target, bq)
        return target

    def merge_all(self):
        acc = None
        for b in self.blocks:
            if b is None:
                continue
            acc = b if acc is None else list(merge_blocks(acc[0], acc[1], b[0], b[1]))
        if acc is None:
            return [np.ones(1, dtype=np.complex128), []]
        return acc


def pblock_read(distributed):
    """The entry stream of pblock/simulator.rs:286-348: (order, node,
    local index) per distinct instruction object (Python identity), ordered by
    ``_instruction_index`` (unknown objects after the known ones, in encounter
    order); node circuits as JSON dicts; the largest cbit count."""
    index = dict(distributed._instruction_index)
    seen, entries, fallback, circuits = set(), [], (2**63 - 1) // 2, {}
    for node in sorted(distributed.circuits):
        c = distributed.circuits[node]
        circuits[node] = json.loads(c.model_dump_json())
        for i, inst in enumerate(c.instructions):
            if id(inst) in seen:
                continue
            seen.add(id(inst))
            if id(inst) in index:
                order = index[id(inst)]
            else:
                order, fallback = fallback, fallback + 1
            entries.append((order, node, i))
    entries.sort(key=lambda e: e[0])
    ncb = max((num_cbits(c) for c in circuits.values()), default=0)
    return entries, circuits, dict(distributed.qubits_per_node), ncb


_PBLOCK_GENERIC = {
    "remote_link_phi_plus": "remote_link_phi_plus", "remote_link_psi_minus": "remote_link_psi_minus",
    "remote_link_psi_plus": "remote_link_psi_plus", "nonlocal_cz": "cz", "remote_cz": "cz",
    "remote_cx": "cx", "remote_epr": "remote_link_phi_plus",
}


def _pblock_dispatch(block, inst, cbits, rng):
    """dispatch / dispatch_par (pblock/simulator.rs:463-710, :746-985) on one block."""
    state, qubits = block
    n = len(qubits)
    kind = inst["kind"]
    if kind in ("id", "u0", "barrier", "classical"):
        return
    if kind in ("measure", "reset"):
        lq = qubits.index(inst["qubit"])
        d = _dp(state.view(np.float64))
        p1 = float(lib().or_prob_one(d, n, lq))
        outcome = 1 if rng.gen_f64() < p1 else 0
        lib().or_collapse(d, n, lq, outcome, p1)
        if kind == "measure":
            cbits[inst["cbit"]] = outcome
        elif outcome == 1:
            apply_one_qubit(state, n, gate_matrix("x"), lq)
        return
    if kind == "conditional":
        c = inst["condition"]
        actual = 0
        for b in range(c["creg_size"]):
            actual |= cbits.get(c["creg_base"] + b, 0) << b
        if actual == c["creg_value"]:
            _pblock_dispatch(block, inst["op"], cbits, rng)
        return
    if kind == "gate":
        name = inst["name"].lower()
        lq = [qubits.index(q) for q in inst["qubits"]]
        if name in ("remote_barrier", "remote_cu1") or name.startswith("circuit-"):
            return
        if name == "remote_rzz":
            u = gate_matrix("rzz", [float(inst["params"][0])])
        elif name == "teleport":
            raise NotImplementedError("Symbolic 'teleport' gate cannot be simulated natively.")
        elif name in _PBLOCK_GENERIC:
            u = gate_matrix(_PBLOCK_GENERIC[name])
        else:
            raise NotImplementedError(f"Unsupported gate in pblock simulator: {name!r}")
        apply_n_qubit(state, n, u, lq)
        return
    lq = [qubits.index(int(inst[f])) for f in _QUBIT_FIELDS[kind]]
    u = gate_matrix(kind, [float(inst[f]) for f in _PARAM_FIELDS.get(kind, ())])
    if len(lq) == 1:
        apply_one_qubit(state, n, u, lq[0])
    else:
        apply_n_qubit(state, n, u, lq)


def _pblock_run(distributed, seed: int, budget_s: float | None = None):
    """The instruction loop of PBlockSimulator::simulate (pblock/simulator.rs:
    368-443): returns (pool, cbits, entries done); with budget_s it stops at
    the first epoch boundary after that many seconds."""
    import time

    entries, circuits, qpn, _ = pblock_read(distributed)
    pool = OracleBlockPool(qpn)
    cbits: dict[int, int] = {}
    rng = Rng(seed)
    inst_of = lambda e: circuits[e[1]]["instructions"][e[2]]  # noqa: E731
    cursor = 0
    t0 = time.perf_counter()
    while cursor < len(entries):
        if budget_s is not None and time.perf_counter() - t0 >= budget_s:
            break
        epoch, used, j = [], set(), cursor
        while j < len(entries):
            qs = inst_qubits(inst_of(entries[j]))
            if not qs:
                j += 1
                continue
            touched = {pool.q2b[q] for q in qs if q in pool.q2b}
            if len(touched) != 1:
                break
            blk = next(iter(touched))
            if blk in used:
                break
            epoch.append((j, blk))
            used.add(blk)
            j += 1
        if len(epoch) > 1:
            seeds = [rng.next_u64() for _ in epoch]
            local = [{} for _ in epoch]
            for k, (ej, blk) in enumerate(epoch):
                _pblock_dispatch(pool.blocks[blk], inst_of(entries[ej]), local[k], Rng(seeds[k]))
            for loc in local:
                cbits.update(loc)
            cursor = j
        else:
            inst = inst_of(entries[cursor])
            qs = inst_qubits(inst)
            if qs:
                blk = pool.ensure_single_block(qs)
                _pblock_dispatch(pool.blocks[blk], inst, cbits, rng)
            cursor += 1
    return pool, cbits, cursor


def simulate_pblock(distributed, seed: int):
    """PBlockSimulator::simulate (pblock/simulator.rs:285-460), epochs included:
    returns (statevector, physical qubits, cbits)."""
    pool, cbits, _ = _pblock_run(distributed, seed)
    st, qs = pool.merge_all()
    return st, qs, cbits


def simulate_pblock_shots(distributed, shots: int, base_seed: int):
    """PBlockSimulator::simulate_shots (pblock/simulator.rs:128-283) with
    run_shot_entries (:712-744) over fuse_pblock_entries (src/types.rs:306-345)
    -- the fusion of the globally ordered stream: returns (counts, merges per shot)."""
    entries, circuits, qpn, ncb = pblock_read(distributed)
    stream = [circuits[node]["instructions"][i] for _, node, i in entries]
    fused = fuse_circuit(stream)
    counts: dict[str, int] = {}
    merges = 0
    for i in range(shots):
        pool = OracleBlockPool(qpn)
        cbits: dict[int, int] = {}
        rng = Rng((base_seed + i) & (2**64 - 1))
        for item in fused:
            if item[0] == "fused":
                blk = pool.ensure_single_block([item[1]])
                st, qs = pool.blocks[blk]
                u = np.array([complex(*e) for e in item[2]], dtype=np.complex128).reshape(2, 2)
                apply_one_qubit(st, len(qs), u, qs.index(item[1]))
                continue
            inst = item[1]
            qs = inst_qubits(inst)
            if not qs:
                continue
            blk = pool.ensure_single_block(qs)
            _pblock_dispatch(pool.blocks[blk], inst, cbits, rng)
        key = format_cbits(cbits, ncb)
        counts[key] = counts.get(key, 0) + 1
        merges = pool.merge_count
    return counts, merges
