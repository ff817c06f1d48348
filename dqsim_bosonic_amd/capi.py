"""ctypes binding of the engine's C ABI (include/dqsim_gpu.h).

The same entry points a Rust/PyO3 host binds through ``extern "C"``; this
module exposes them to Python for direct use (tests, tools) without the C++
host.  ``State`` owns a device statevector.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdqsim_gpu.so")

C128, C64 = 0, 1
ERR_ARG, ERR_OOM, ERR_DEVICE = -1, -2, -3

# every function declared in include/dqsim_gpu.h
EXPORTS = (
    "dq_state_create", "dq_state_destroy", "dq_state_num_qubits", "dq_state_set", "dq_state_reset",
    "dq_apply_1q", "dq_apply_nq", "dq_prob_one", "dq_collapse", "dq_marginal_probs",
    "dq_marginal_probs_noalloc", "dq_inner",
    "dq_copy_to_host", "dq_synchronize", "dq_state_copy", "dq_apply_pauli_sum", "dq_pauli_inner",
    "dq_compare_states", "dq_state_tensor", "dq_adjoint_sweep",
    "dq_set_options", "dq_set_timing", "dq_get_timing", "dq_get_timed_amplitudes", "dq_set_zero_tiles",
    "dq_get_stats",
    "dq_get_fused_stats", "dq_set_fuse_min_cost", "dq_set_pair_fusion", "dq_set_reg_bits", "dq_top_k",
    "dq_batch_create", "dq_batch_destroy", "dq_batch_add_instance", "dq_batch_apply",
    "dq_batch_set_observable", "dq_batch_set_workers", "dq_batch_run", "dq_batch_stats",
    "dq_batch_set_autotune", "dq_batch_path_stats", "dq_batch_tuned",
    "dq_shots_create", "dq_shots_destroy", "dq_shots_add_gate", "dq_shots_add_measure", "dq_shots_add_reset",
    "dq_shots_add_conditional", "dq_shots_words", "dq_shots_run",
    "dq_mps_create", "dq_mps_destroy", "dq_mps_load", "dq_mps_to_state", "dq_mps_prob_one", "dq_mps_set_cut",
    "dq_mps_get_cut",
    "dq_last_error",
)

_lib = None
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built")
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        L.dq_state_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(vp)]
        L.dq_state_destroy.argtypes = [vp]
        L.dq_state_destroy.restype = None
        L.dq_state_num_qubits.argtypes = [vp]
        L.dq_state_set.argtypes = [vp, _dp]
        L.dq_state_reset.argtypes = [vp]
        L.dq_apply_1q.argtypes = [vp, _dp, ctypes.c_int, _ip, ctypes.POINTER(ctypes.c_ubyte), ctypes.c_int]
        L.dq_apply_nq.argtypes = [vp, _dp, _ip, ctypes.c_int]
        L.dq_prob_one.argtypes = [vp, ctypes.c_int, _dp]
        L.dq_collapse.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_double]
        L.dq_marginal_probs.argtypes = [vp, _ip, ctypes.c_int, _dp]
        L.dq_marginal_probs_noalloc.argtypes = [vp, _ip, ctypes.c_int, _dp]
        L.dq_inner.argtypes = [vp, _dp, _dp, _dp]
        L.dq_copy_to_host.argtypes = [vp, _dp]
        L.dq_synchronize.argtypes = [vp]
        L.dq_set_options.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        u64p = ctypes.POINTER(ctypes.c_uint64)
        L.dq_get_stats.argtypes = [vp, u64p, u64p, u64p]
        L.dq_get_fused_stats.argtypes = [vp, u64p, u64p]
        L.dq_set_fuse_min_cost.argtypes = [vp, ctypes.c_double]
        L.dq_set_pair_fusion.argtypes = [vp, ctypes.c_int]
        L.dq_set_reg_bits.argtypes = [vp, ctypes.c_int]
        L.dq_top_k.argtypes = [vp, ctypes.c_uint64, u64p, _dp]
        L.dq_batch_create.argtypes = [ctypes.c_int, ctypes.POINTER(vp)]
        L.dq_batch_destroy.argtypes = [vp]
        L.dq_batch_destroy.restype = None
        L.dq_batch_add_instance.argtypes = [vp, ctypes.c_int, _ip]
        L.dq_batch_apply.argtypes = [vp, ctypes.c_int, _dp, _ip, ctypes.c_int]
        L.dq_batch_set_observable.argtypes = [vp, ctypes.c_int, ctypes.c_int, u64p, _dp]
        L.dq_batch_set_workers.argtypes = [vp, ctypes.c_int]
        L.dq_batch_run.argtypes = [vp, _dp]
        L.dq_batch_stats.argtypes = [vp, u64p, u64p, _dp]
        L.dq_batch_set_autotune.argtypes = [vp, ctypes.c_int]
        L.dq_batch_path_stats.argtypes = [vp, u64p, _dp, _dp, _dp]
        ip = ctypes.POINTER(ctypes.c_int)
        L.dq_batch_tuned.argtypes = [ctypes.c_int, ctypes.c_int, ip, ip, ip, ip]
        L.dq_shots_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(vp)]
        L.dq_shots_destroy.argtypes = [vp]
        L.dq_shots_destroy.restype = None
        L.dq_shots_add_gate.argtypes = [vp, _dp, _ip, ctypes.c_int]
        L.dq_shots_add_measure.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.dq_shots_add_reset.argtypes = [vp, ctypes.c_int]
        L.dq_shots_add_conditional.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_int]
        L.dq_shots_words.argtypes = [vp]
        L.dq_shots_run.argtypes = [vp, ctypes.c_uint64, ctypes.c_uint64, u64p, _dp]
        L.dq_state_copy.argtypes = [vp, vp]
        L.dq_compare_states.argtypes = [vp, vp, _dp, _dp, _dp, _dp]
        L.dq_state_tensor.argtypes = [vp, vp, _ip, vp, _ip]
        L.dq_adjoint_sweep.argtypes = [vp, vp, ctypes.c_int, _ip, _ip, _dp, _ip, u64p, u64p, _dp, _dp]
        L.dq_apply_pauli_sum.argtypes = [vp, vp, ctypes.c_int, u64p, u64p, _dp]
        L.dq_pauli_inner.argtypes = [vp, vp, ctypes.c_int, u64p, u64p, _dp, _dp, _dp]
        L.dq_set_timing.argtypes = [vp, ctypes.c_int]
        L.dq_get_timing.argtypes = [vp, _dp, u64p]
        L.dq_get_timed_amplitudes.argtypes = [vp, u64p]
        L.dq_set_zero_tiles.argtypes = [vp, ctypes.c_int]
        L.dq_mps_create.argtypes = [ctypes.c_int, ctypes.POINTER(vp)]
        L.dq_mps_destroy.argtypes = [vp]
        L.dq_mps_destroy.restype = None
        L.dq_mps_load.argtypes = [vp, u64p, _dp]
        L.dq_mps_to_state.argtypes = [vp, vp]
        L.dq_mps_prob_one.argtypes = [vp, ctypes.c_int, _dp]
        L.dq_mps_set_cut.argtypes = [vp, ctypes.c_int]
        L.dq_mps_get_cut.argtypes = [vp]
        L.dq_last_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


class DqError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def _check(rc: int) -> None:
    if rc != 0:
        raise DqError(rc, lib().dq_last_error().decode())


def _d(a: np.ndarray):
    return a.ctypes.data_as(_dp)


class State:
    def __init__(self, n: int, dtype: int = C128):
        self.n = n
        self.dtype = dtype
        h = ctypes.c_void_p()
        _check(lib().dq_state_create(n, dtype, ctypes.byref(h)))
        self.h = h

    def close(self) -> None:
        if self.h:
            lib().dq_state_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set(self, psi: np.ndarray) -> None:
        psi = np.ascontiguousarray(psi, dtype=np.complex128)
        assert psi.size == 1 << self.n
        _check(lib().dq_state_set(self.h, _d(psi.view(np.float64))))

    def reset(self) -> None:
        _check(lib().dq_state_reset(self.h))

    def apply_1q(self, u: np.ndarray, target: int, controls=()) -> None:
        u = np.ascontiguousarray(u, dtype=np.complex128)
        w = (ctypes.c_int * max(1, len(controls)))(*[c[0] for c in controls])
        v = (ctypes.c_ubyte * max(1, len(controls)))(*[int(c[1]) for c in controls])
        _check(lib().dq_apply_1q(self.h, _d(u.view(np.float64)), target, w, v, len(controls)))

    def apply_nq(self, u: np.ndarray, qubits) -> None:
        u = np.ascontiguousarray(u, dtype=np.complex128)
        q = (ctypes.c_int * len(qubits))(*qubits)
        _check(lib().dq_apply_nq(self.h, _d(u.view(np.float64)), q, len(qubits)))

    def prob_one(self, qubit: int) -> float:
        p = ctypes.c_double()
        _check(lib().dq_prob_one(self.h, qubit, ctypes.byref(p)))
        return p.value

    def collapse(self, qubit: int, outcome: int, norm: float) -> None:
        _check(lib().dq_collapse(self.h, qubit, outcome, norm))

    def marginal_probs(self, qubits) -> np.ndarray:
        q = (ctypes.c_int * max(1, len(qubits)))(*qubits)
        out = np.zeros(1 << len(qubits), dtype=np.float64)
        _check(lib().dq_marginal_probs(self.h, q, len(qubits), _d(out)))
        return out

    def marginal_probs_noalloc(self, qubits) -> np.ndarray:
        """dq_marginal_probs_noalloc: same values, no device allocation."""
        q = (ctypes.c_int * max(1, len(qubits)))(*qubits)
        out = np.zeros(1 << len(qubits), dtype=np.float64)
        _check(lib().dq_marginal_probs_noalloc(self.h, q, len(qubits), _d(out)))
        return out

    def inner(self, other: np.ndarray) -> complex:
        other = np.ascontiguousarray(other, dtype=np.complex128)
        re, im = ctypes.c_double(), ctypes.c_double()
        _check(lib().dq_inner(self.h, _d(other.view(np.float64)), ctypes.byref(re), ctypes.byref(im)))
        return complex(re.value, im.value)

    def get(self) -> np.ndarray:
        out = np.empty(1 << self.n, dtype=np.complex128)
        _check(lib().dq_copy_to_host(self.h, _d(out.view(np.float64))))
        return out

    def copy_from(self, other: "State") -> None:
        _check(lib().dq_state_copy(self.h, other.h))

    @staticmethod
    def _terms(terms):
        """terms: iterable of (coeff, xmask, zmask)."""
        terms = list(terms)
        x = (ctypes.c_uint64 * max(1, len(terms)))(*[int(t[1]) for t in terms])
        z = (ctypes.c_uint64 * max(1, len(terms)))(*[int(t[2]) for t in terms])
        c = np.array([[complex(t[0]).real, complex(t[0]).imag] for t in terms] or [[0.0, 0.0]], dtype=np.float64)
        return len(terms), x, z, c

    def apply_pauli_sum(self, src: "State", terms) -> None:
        k, x, z, c = self._terms(terms)
        _check(lib().dq_apply_pauli_sum(self.h, src.h, k, x, z, _d(c)))

    def pauli_inner(self, ket: "State", terms) -> complex:
        k, x, z, c = self._terms(terms)
        re, im = ctypes.c_double(), ctypes.c_double()
        _check(lib().dq_pauli_inner(self.h, ket.h, k, x, z, _d(c), ctypes.byref(re), ctypes.byref(im)))
        return complex(re.value, im.value)

    def compare(self, other: "State") -> dict:
        m, r, i, t = (ctypes.c_double() for _ in range(4))
        _check(lib().dq_compare_states(self.h, other.h, ctypes.byref(m), ctypes.byref(r), ctypes.byref(i),
                                       ctypes.byref(t)))
        inner = complex(r.value, i.value)
        return {"max_abs": m.value, "inner": inner, "fidelity": abs(inner) ** 2, "tvd": t.value}

    def synchronize(self) -> None:
        _check(lib().dq_synchronize(self.h))

    def set_options(self, fuse: int = 1, tile_bits: int = 0, low_bits: int = 0) -> None:
        _check(lib().dq_set_options(self.h, fuse, tile_bits, low_bits))

    def stats(self) -> tuple[int, int, int]:
        a, b, c = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        _check(lib().dq_get_stats(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, c.value

    def top_k(self, k: int):
        """(indices, probabilities) of the k most probable basis states."""
        idx = np.zeros(k, dtype=np.uint64)
        p = np.zeros(k, dtype=np.float64)
        _check(lib().dq_top_k(self.h, k, idx.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), _d(p)))
        return idx, p

    def set_pair_fusion(self, on: bool) -> None:
        _check(lib().dq_set_pair_fusion(self.h, int(bool(on))))

    def set_reg_bits(self, r: int) -> None:
        _check(lib().dq_set_reg_bits(self.h, int(r)))

    def set_fuse_min_cost(self, min_cost: float) -> None:
        _check(lib().dq_set_fuse_min_cost(self.h, float(min_cost)))

    def fused_stats(self) -> tuple[int, int]:
        """(5-qubit MFMA blocks applied, gates they replaced) under fusion level 3."""
        a, b = ctypes.c_uint64(), ctypes.c_uint64()
        _check(lib().dq_get_fused_stats(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def set_zero_tiles(self, on: bool) -> None:
        """Whether register sweeps skip the tiles where a qubit still 0
        everywhere (untouched since the last reset) is 1."""
        _check(lib().dq_set_zero_tiles(self.h, int(bool(on))))

    def set_timing(self, on: bool) -> None:
        _check(lib().dq_set_timing(self.h, int(bool(on))))

    def timed_amplitudes(self) -> int:
        """Amplitudes the sweep kernels read and wrote since set_timing."""
        a = ctypes.c_uint64()
        _check(lib().dq_get_timed_amplitudes(self.h, ctypes.byref(a)))
        return a.value

    def tensor(self, a: "State", a_pos, b: "State", b_pos) -> None:
        """self = a (x) b, local qubit k of a at position a_pos[k], of b at b_pos[k]."""
        ap = (ctypes.c_int * max(1, len(a_pos)))(*a_pos)
        bp = (ctypes.c_int * max(1, len(b_pos)))(*b_pos)
        _check(lib().dq_state_tensor(self.h, a.h, ap, b.h, bp))


class Batch:
    """Owning wrapper of a dq_batch handle (many independent circuits)."""

    def __init__(self, dtype: int = C128):
        h = ctypes.c_void_p()
        _check(lib().dq_batch_create(dtype, ctypes.byref(h)))
        self.h = h
        self.count = 0

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib is not None:
            _lib.dq_batch_destroy(self.h)
            self.h = None

    def add(self, n: int) -> int:
        i = ctypes.c_int()
        _check(lib().dq_batch_add_instance(self.h, n, ctypes.byref(i)))
        self.count += 1
        return i.value

    def apply(self, index: int, u: np.ndarray, qubits) -> None:
        u = np.ascontiguousarray(u, dtype=np.complex128)
        q = (ctypes.c_int * len(qubits))(*qubits)
        _check(lib().dq_batch_apply(self.h, index, _d(u.view(np.float64)), q, len(qubits)))

    def set_observable(self, index: int, terms) -> None:
        """terms: (coeff, zmask) pairs."""
        zm = (ctypes.c_uint64 * max(1, len(terms)))(*[int(t[1]) for t in terms])
        c = np.ascontiguousarray([float(t[0]) for t in terms] or [0.0], dtype=np.float64)
        _check(lib().dq_batch_set_observable(self.h, index, len(terms), zm, _d(c)))

    def set_workers(self, w: int) -> None:
        _check(lib().dq_batch_set_workers(self.h, w))

    def run(self) -> np.ndarray:
        out = np.zeros(max(1, self.count), dtype=np.float64)
        _check(lib().dq_batch_run(self.h, _d(out)))
        return out[: self.count]

    def stats(self) -> tuple[int, int, float]:
        a, b, t = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_double()
        _check(lib().dq_batch_stats(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(t)))
        return a.value, b.value, t.value

    def set_autotune(self, on: bool) -> None:
        _check(lib().dq_batch_set_autotune(self.h, int(bool(on))))

    def path_stats(self) -> dict:
        """Instances per path (lds, persistent, engine), graph GPU ms, engine
        wall seconds and algorithmic bytes per path of the last run."""
        cnt = (ctypes.c_uint64 * 3)()
        by = (ctypes.c_double * 3)()
        g, e = ctypes.c_double(), ctypes.c_double()
        _check(lib().dq_batch_path_stats(self.h, cnt, ctypes.byref(g), ctypes.byref(e), by))
        return {"lds": cnt[0], "persistent": cnt[1], "engine": cnt[2], "graph_ms": g.value, "engine_s": e.value,
                "bytes": [by[0], by[1], by[2]]}


def batch_tuned(n: int, dtype: int = C128):
    """The batch autotuner's cached choice for (n, dtype), or None."""
    v = [ctypes.c_int() for _ in range(4)]
    rc = lib().dq_batch_tuned(n, dtype, *[ctypes.byref(x) for x in v])
    if rc != 0:
        return None
    return {"path": "persistent" if v[0].value == 1 else "engine", "fusion": v[1].value, "tile_bits": v[2].value,
            "reg_bits": v[3].value}


class Shots:
    """Owning wrapper of a dq_shots program (batched shots)."""

    def __init__(self, n: int, num_cbits: int):
        h = ctypes.c_void_p()
        _check(lib().dq_shots_create(n, num_cbits, ctypes.byref(h)))
        self.h = h
        self.num_cbits = num_cbits

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib is not None:
            _lib.dq_shots_destroy(self.h)
            self.h = None

    def gate(self, u: np.ndarray, qubits) -> None:
        u = np.ascontiguousarray(u, dtype=np.complex128)
        q = (ctypes.c_int * len(qubits))(*qubits)
        _check(lib().dq_shots_add_gate(self.h, _d(u.view(np.float64)), q, len(qubits)))

    def measure(self, qubit: int, cbit: int) -> None:
        _check(lib().dq_shots_add_measure(self.h, qubit, cbit))

    def reset(self, qubit: int) -> None:
        _check(lib().dq_shots_add_reset(self.h, qubit))

    def conditional(self, base: int, size: int, value: int, n_ops: int) -> None:
        _check(lib().dq_shots_add_conditional(self.h, base, size, value, n_ops))

    def run(self, base_seed: int, shots: int, timed: bool = False):
        words = int(lib().dq_shots_words(self.h))
        bits = np.zeros(max(1, shots * words), dtype=np.uint64)
        secs = np.zeros(max(1, shots), dtype=np.float64) if timed else None
        _check(lib().dq_shots_run(self.h, base_seed, shots, bits.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                  _d(secs) if timed else None))
        return bits[: shots * words].reshape(shots, words), secs


class Mps:
    """Owning wrapper of a dq_mps contraction context (matrix-product states)."""

    def __init__(self, n: int):
        h = ctypes.c_void_p()
        _check(lib().dq_mps_create(n, ctypes.byref(h)))
        self.h = h
        self.n = n

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib is not None:
            _lib.dq_mps_destroy(self.h)
            self.h = None

    def load(self, tensors) -> None:
        """tensors: n arrays of shape (left, 2, right), complex128."""
        assert len(tensors) == self.n
        bond = np.array([1] + [int(t.shape[2]) for t in tensors], dtype=np.uint64)
        for q, t in enumerate(tensors):
            if t.ndim != 3 or t.shape[1] != 2 or t.shape[0] != int(bond[q]):
                raise ValueError(f"site {q}: shape {t.shape} does not continue bond {int(bond[q])}")
        flat = np.concatenate([np.ascontiguousarray(t, dtype=np.complex128).ravel() for t in tensors])
        self._keep = flat
        _check(lib().dq_mps_load(self.h, bond.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                 _d(flat.view(np.float64))))

    def to_state(self, state: "State") -> None:
        _check(lib().dq_mps_to_state(self.h, state.h))

    def prob_one(self, qubit: int) -> float:
        p = ctypes.c_double()
        _check(lib().dq_mps_prob_one(self.h, qubit, ctypes.byref(p)))
        return p.value

    def set_cut(self, cut: int) -> None:
        _check(lib().dq_mps_set_cut(self.h, cut))

    def cut(self) -> int:
        return int(lib().dq_mps_get_cut(self.h))
