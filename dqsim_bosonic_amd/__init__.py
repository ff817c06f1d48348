"""MI355X-native statevector engine for dqsim's statevector hot path.

Drop-in for the reference package's statevector surface
(reference: dqsim/__init__.py:1-27, src/lib.rs:15-29):

    from dqsim_bosonic_amd import StatevectorSimulator, SimulationResult, SimulationProfile
    from dqsim_bosonic_amd import simulate_monolithic, MpsSimulator

The classes live in the native module ``_core`` (C++ host) and drive the HIP
engine ``libdqsim_gpu.so`` through the C ABI of include/dqsim_gpu.h.  There is
no CPU fallback: importing fails loudly if the native libraries are missing.
"""
from __future__ import annotations

import os as _os

_HERE = _os.path.dirname(_os.path.abspath(__file__))

try:
    from . import _core  # noqa: F401  (loads libdqsim_gpu.so via rpath)
except ImportError as _e:  # pragma: no cover - exercised only on broken installs
    raise ImportError(
        "dqsim_bosonic_amd native extension is not built; run "
        "`python -c 'import __graft_entry__ as g; g.build()'` (or "
        "`python dqsim_bosonic_amd/_build.py`) first: " + str(_e)
    ) from _e

from ._core import (  # noqa: E402
    MpsSimulator,
    PBlockResult,
    PBlockSimulator,
    SimulationProfile,
    SimulationResult,
    StatevectorSimulator,
    simulate_batch,
    simulate_distributed,
    simulate_distributed_shots,
    simulate_monolithic,
    simulate_monolithic_shots,
)

ENGINE_LIBRARY = _os.path.join(_HERE, "libdqsim_gpu.so")

__all__ = [
    "MpsSimulator",
    "PBlockResult",
    "PBlockSimulator",
    "simulate_distributed",
    "simulate_distributed_shots",
    "simulate_batch",
    "simulate_monolithic_shots",
    "StatevectorSimulator",
    "SimulationResult",
    "SimulationProfile",
    "simulate_monolithic",
]
