"""OpenQASM / Qiskit import helpers.

The reference lowers Qiskit circuits to the JSON transport
(python/dqsim/qasm.py:56-221). That importer needs Qiskit and is not part of
the statevector engine, so only its names are provided: the transport JSON is
accepted directly by every simulator method.
"""


class QASMImportError(ValueError):
    """Raised for circuits the transport cannot represent."""


class ImportedCircuit:
    """Transport wrapper: ``model_dump_json()`` returns the stored document."""

    def __init__(self, data, metadata=None):
        self.data = data
        self.metadata = dict(metadata or {})

    def model_dump_json(self):
        import json

        return json.dumps(self.data)


def from_qiskit(circuit, *, max_instructions=250_000):
    raise NotImplementedError("Qiskit import is not available in this engine; pass the JSON transport directly")


def load_qasm(source, *, max_instructions=250_000):
    raise NotImplementedError("OpenQASM import is not available in this engine; pass the JSON transport directly")
