"""Differentiation: adjoint gradients of <psi(theta)|H|psi(theta)> on the GPU.

No reference counterpart (a new form of the statevector path).  The state
evolution uses the same engine as ``StatevectorSimulator.simulate``; the
reverse sweep applies gate daggers in place to |psi> and |lambda> = H|psi>
(uncompute instead of stored per-layer states) and reads every gradient as
Im <lambda| G |psi> for a gate U = exp(-i theta G / 2)
(Jones & Gacon, "Efficient calculation of gradients in classical simulations
of variational quantum algorithms", arXiv:2009.02823).

* ``PauliSum`` -- observables (sum of coefficient * Pauli string).
* ``ParametricCircuit`` -- gates with trainable parameter indices.
* ``Estimator`` -- device-resident evaluator: ``expectation(params)``,
  ``gradient(params)`` (adjoint), ``parameter_shift(params)``.
* ``AdjointExpectation`` -- ``torch.autograd.Function`` over an Estimator.
"""
from __future__ import annotations

import math
from typing import Iterable, Sequence

import numpy as np

from . import _core

_DTYPES = {"c128": 0, "complex128": 0, "c64": 1, "complex64": 1}
_PAULI_GATES = {"rx", "ry", "rz", "rxx", "rzz"}          # two-term shift rule
_CONTROLLED_ROT = {"crx", "cry", "crz", "cp", "cu1", "p", "u1"}


class PauliSum:
    """sum_k c_k P_k with P_k = (xmask, zmask): X on x bits, Z on z bits, Y on both."""

    def __init__(self, terms: Iterable[tuple[complex, int, int]] = ()):
        self.terms = [(complex(c), int(x), int(z)) for c, x, z in terms]

    @staticmethod
    def pauli(ops: dict[int, str], coeff: complex = 1.0) -> "PauliSum":
        x = z = 0
        for q, p in ops.items():
            p = p.upper()
            if p in ("X", "Y"):
                x |= 1 << q
            if p in ("Z", "Y"):
                z |= 1 << q
        return PauliSum([(coeff, x, z)])

    @staticmethod
    def zz_chain(n: int, coeff: float = 1.0) -> "PauliSum":
        return PauliSum([(coeff, 0, (1 << i) | (1 << (i + 1))) for i in range(n - 1)])

    @staticmethod
    def x_sum(n: int, coeff: float = 1.0) -> "PauliSum":
        return PauliSum([(coeff, 1 << i, 0) for i in range(n)])

    def __add__(self, other: "PauliSum") -> "PauliSum":
        return PauliSum(self.terms + other.terms)

    def __mul__(self, a: complex) -> "PauliSum":
        return PauliSum([(c * a, x, z) for c, x, z in self.terms])

    __rmul__ = __mul__

    def arrays(self):
        x = [t[1] for t in self.terms]
        z = [t[2] for t in self.terms]
        c = [t[0] for t in self.terms]
        return x, z, c


class ParametricCircuit:
    """Gates on n qubits; ``param=k`` makes the gate's angle the k-th parameter."""

    def __init__(self, n: int):
        self.n = n
        self.ops: list[tuple[str, list[int], int, float]] = []

    def add(self, name: str, qubits: Sequence[int], param: int | None = None, value: float = 0.0):
        self.ops.append((name, [int(q) for q in qubits], -1 if param is None else int(param), float(value)))
        return self

    def __getattr__(self, name: str):
        if name.startswith("_"):
            raise AttributeError(name)

        def gate(*qubits, param=None, value=0.0):
            return self.add(name, qubits, param, value)

        return gate

    @property
    def num_params(self) -> int:
        return max((p for _, _, p, _ in self.ops), default=-1) + 1


def hea_benchmark(n: int = 26, layers: int = 40, seed: int = 42) -> tuple[ParametricCircuit, PauliSum]:
    """The differentiation workload: per layer ry and rz on every qubit
    (trainable: 2 n parameters per layer), rzz on a brickwork and crx on every
    4th pair (fixed angles from the seed); H = sum Z_i Z_{i+1} + 0.7 sum X_i."""
    rng = np.random.default_rng(seed)
    pc = ParametricCircuit(n)
    k = 0
    for layer in range(layers):
        for q in range(n):
            pc.ry(q, param=k)
            k += 1
        for q in range(n):
            pc.rz(q, param=k)
            k += 1
        for a in range(layer % 2, n - 1, 2):
            pc.rzz(a, a + 1, value=float(rng.uniform(0, 2 * math.pi)))
        for a in range(0, n - 1, 4):
            pc.crx(a, a + 1, value=float(rng.uniform(0, 2 * math.pi)))
    return pc, PauliSum.zz_chain(n) + PauliSum.x_sum(n, 0.7)


class Estimator:
    """Device-resident <H> evaluator for one circuit and observable."""

    def __init__(self, circuit: ParametricCircuit, observable: PauliSum, dtype: str = "c128",
                 fusion: int = 2):
        self.circuit = circuit
        self.observable = observable
        self._eng = _core.Adjoint(circuit.n, _DTYPES[dtype], fusion)
        self._eng.set_circuit(circuit.ops)
        self._eng.set_observable(*observable.arrays())
        self.num_params = self._eng.num_params()

    def expectation(self, params) -> float:
        return float(self._eng.expectation(np.asarray(params, dtype=np.float64).tolist()))

    def gradient(self, params) -> tuple[float, np.ndarray]:
        """Adjoint gradient: one forward pass, then the in-place reverse sweep."""
        e, g = self._eng.gradient(np.asarray(params, dtype=np.float64).tolist())
        return float(e), np.asarray(g, dtype=np.float64)

    def parameter_shift(self, params, indices=None) -> np.ndarray:
        """Gradient from shifted forward evaluations (reference method).
        Pauli-generated gates use the two-term rule (shifts +-pi/2); controlled
        rotations the four-term rule (Anselmetti et al., NJP 23, 113010)."""
        params = np.asarray(params, dtype=np.float64)
        idx = range(self.num_params) if indices is None else indices
        kinds: dict[int, set] = {}
        for name, _, p, _ in self.circuit.ops:
            if p >= 0:
                kinds.setdefault(p, set()).add(name)
        out = np.zeros(len(list(idx)) if indices is not None else self.num_params)
        for j, k in enumerate(idx):
            names = kinds.get(k, set())
            if names and names <= _PAULI_GATES:
                out[j] = 0.5 * (self._shift(params, k, math.pi / 2) - self._shift(params, k, -math.pi / 2))
            else:
                a, b = math.pi / 2, 3 * math.pi / 2
                c1 = (math.sqrt(2) + 1) / (4 * math.sqrt(2))
                c2 = (math.sqrt(2) - 1) / (4 * math.sqrt(2))
                out[j] = (c1 * (self._shift(params, k, a) - self._shift(params, k, -a))
                          - c2 * (self._shift(params, k, b) - self._shift(params, k, -b)))
        return out

    def finite_difference(self, params, indices=None, h: float = 1e-6) -> np.ndarray:
        params = np.asarray(params, dtype=np.float64)
        idx = list(range(self.num_params) if indices is None else indices)
        return np.array([(self._shift(params, k, h) - self._shift(params, k, -h)) / (2 * h) for k in idx])

    def _shift(self, params: np.ndarray, k: int, d: float) -> float:
        p = params.copy()
        p[k] += d
        return self.expectation(p)


try:  # the autograd wrapper needs torch; the rest of the module does not
    import torch

    class AdjointExpectation(torch.autograd.Function):
        """E = <psi(theta)|H|psi(theta)> with the adjoint gradient as backward.

        ``AdjointExpectation.apply(theta, estimator)``; theta is a 1-D float64
        tensor.  The forward pass keeps the gradient from the same sweep."""

        @staticmethod
        def forward(ctx, theta, estimator):
            p = theta.detach().cpu().numpy().astype(np.float64)
            if theta.requires_grad:
                e, g = estimator.gradient(p)
                ctx.save_for_backward(torch.as_tensor(g, dtype=theta.dtype, device=theta.device))
            else:
                e = estimator.expectation(p)
                ctx.save_for_backward(torch.zeros_like(theta))
            return torch.tensor(e, dtype=theta.dtype, device=theta.device)

        @staticmethod
        def backward(ctx, grad_out):
            (g,) = ctx.saved_tensors
            return grad_out * g, None

except ImportError:  # pragma: no cover
    AdjointExpectation = None
