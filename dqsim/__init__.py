"""Drop-in import path for callers of the reference package.

The reference's callers write ``from dqsim import StatevectorSimulator``
(reference: dqsim/__init__.py:1-27, README.md:79-83).  This package gives
them the same names, backed by the MI355X engine in ``dqsim_bosonic_amd``.

``StabilizerSimulator`` (the reference's Clifford tableau simulator,
src/monolithic/stabilizer/) is a different algorithm and is not part of
this engine: the name exists so that the reference's import line works, and
constructing it raises NotImplementedError, as ``simulate_monolithic(...,
mode="stabilizer")`` does.
"""
from dqsim_bosonic_amd import (
    MpsSimulator,
    PBlockResult,
    PBlockSimulator,
    SimulationProfile,
    SimulationResult,
    StatevectorSimulator,
    simulate_distributed,
    simulate_distributed_shots,
    simulate_monolithic,
    simulate_monolithic_shots,
)


class StabilizerSimulator:
    """Placeholder for the reference's stabilizer simulator (not provided)."""

    def __init__(self, *args, **kwargs):
        raise NotImplementedError(
            "StabilizerSimulator is not provided by the MI355X statevector engine; "
            "use StatevectorSimulator or MpsSimulator")


__all__ = [
    "StatevectorSimulator",
    "StabilizerSimulator",
    "MpsSimulator",
    "SimulationResult",
    "SimulationProfile",
    "PBlockSimulator",
    "PBlockResult",
    "simulate_monolithic",
    "simulate_distributed",
    "simulate_monolithic_shots",
    "simulate_distributed_shots",
]
