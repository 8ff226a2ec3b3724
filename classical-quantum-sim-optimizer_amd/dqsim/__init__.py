"""MI355X dense statevector engine with the ``dqsim`` Python API.

Circuit inputs expose ``model_dump_json()`` using the register/instruction
schema of the reference package (src/types.rs). The names re-exported here are
those of the reference's ``dqsim/__init__.py``.
"""

from ._core import (
    MpsSimulator,
    MpsResult,
    PBlockResult,
    PBlockSimulator,
    SimulationProfile,
    SimulationResult,
    StabilizerSimulator,
    StabilizerResult,
    StatevectorSimulator,
    simulate_distributed,
    simulate_distributed_shots,
    simulate_monolithic,
    simulate_monolithic_shots,
)

__all__ = [
    "MpsSimulator",
    "MpsResult",
    "PBlockResult",
    "PBlockSimulator",
    "SimulationProfile",
    "SimulationResult",
    "StabilizerSimulator",
    "StabilizerResult",
    "StatevectorSimulator",
    "simulate_distributed",
    "simulate_distributed_shots",
    "simulate_monolithic",
    "simulate_monolithic_shots",
]

from .qasm import ImportedCircuit, QASMImportError, from_qiskit, load_qasm

__all__ += ["ImportedCircuit", "QASMImportError", "from_qiskit", "load_qasm"]
