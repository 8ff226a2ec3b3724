"""P-block backend, host side: option checks, the distributed transport
(ownership, global order, shared copies) and the errors raised before any
block is allocated. Behaviours follow the reference's tests/test_pblock.py
(cases restated, see each test) and src/distributed/pblock/{input.rs,
simulator.rs}; nothing here touches a GPU."""

import copy
import math
from types import SimpleNamespace

import pytest

import dqsim
from dqsim import PBlockSimulator
from dqsim.workloads import Circuit


def inst(kind, **f):
    return {"kind": kind, **f}


def conditional(op, value, base=0, size=1):
    return inst("conditional", condition={"creg_base": base, "creg_size": size, "creg_value": value}, op=op)


def distributed(n, ops, groups=None, cbits=0):
    """Per-node circuits owning the operations whose first qubit they hold,
    ordered by the legacy `_instruction_index` of the operation objects."""
    groups = groups or {q: [q] for q in range(n)}
    owner = {q: node for node, qs in groups.items() for q in qs}
    buckets = {node: [] for node in groups}
    for op in ops:
        inner = op
        while inner["kind"] == "conditional":
            inner = inner["op"]
        q = next((inner[k] for k in ("qubit", "control", "control1", "a") if k in inner), None)
        if q is None:
            q = (inner.get("qubits") or [next(iter(owner), 0)])[0]
        buckets[owner.get(q, next(iter(groups), 0))].append(op)
    circuits = {node: Circuit(n, bucket, cbits) for node, bucket in buckets.items()}
    for c in circuits.values():
        c.instructions = c.data["instructions"]
    return SimpleNamespace(circuits=circuits, qubits_per_node=groups,
                           _instruction_index={id(op): i for i, op in enumerate(ops)})


def test_constructor_options_are_checked():
    # test_pblock.py:286-289 (simulator.rs:58-95)
    for opts in ({"max_memory_mb": 0}, {"max_parallel_shots": 0}, {"max_block_qubits": 0},
                 {"max_split_qubits": 0}, {"separation_tolerance": float("nan")},
                 {"separation_tolerance": 1.0}, {"max_block_qubits": 64}):
        with pytest.raises(ValueError):
            PBlockSimulator(**opts)
    with pytest.raises(TypeError):
        dqsim.simulate_distributed(Circuit(1), bogus=1)
    with pytest.raises(ValueError):
        dqsim.simulate_distributed(Circuit(1), mode="stabilizer")


def test_invalid_operations_fail_before_execution():
    # test_pblock.py:60-91: unknown or malformed operations, plain or distributed
    bad = [inst("gate", name="circuit-123", qubits=[0], params=[]),
           inst("gate", name="teleport", qubits=[], params=[]), inst("classical", name="unknown"),
           conditional(inst("gate", name="unknown", qubits=[], params=[]), 1),
           inst("cx", control=0, target=0), inst("x", qubit=4), inst("measure", qubit=0, cbit=7),
           inst("u3", qubit=0, theta=.5, phi=1e308, lam=1e308), conditional(inst("x", qubit=99), 1)]
    for op in bad:
        for c in (Circuit(2, [op], 1), distributed(2, [op], cbits=1)):
            with pytest.raises(ValueError):
                PBlockSimulator().simulate(c)
            with pytest.raises(ValueError):
                PBlockSimulator().simulate_shots(c, 0)


def test_ownership_and_legacy_order_errors():
    # test_pblock.py:77-79, 93-106 (input.rs:54-150)
    d = distributed(2, [inst("h", qubit=0)])
    d.qubits_per_node = {0: [0], 1: [0]}
    with pytest.raises(ValueError, match="unique ownership"):
        PBlockSimulator().simulate(d)
    ops = [inst("x", qubit=0), inst("cx", control=0, target=1)]
    d = distributed(2, ops)
    del d._instruction_index[id(ops[0])]
    with pytest.raises(ValueError, match="Missing instruction order"):
        PBlockSimulator().simulate(d)
    x = inst("x", qubit=0)
    with pytest.raises(ValueError, match="Repeated operation ID"):
        PBlockSimulator().simulate(distributed(1, [x, x]))
    d = distributed(2, ops)
    d._instruction_index = {id(op): 0 for op in ops}
    with pytest.raises(ValueError, match="order"):
        PBlockSimulator().simulate_shots(d, 0)
    d = distributed(2, [inst("h", qubit=0), inst("h", qubit=1)])
    d._instruction_index = {k: 0 for k in d._instruction_index}
    with pytest.raises(ValueError, match="Ambiguous global order"):
        PBlockSimulator().simulate(d)
    d = distributed(2, ops)
    d.circuits[0].instructions = []
    with pytest.raises(ValueError, match="different lengths"):
        PBlockSimulator().simulate(d)
    d = distributed(2, ops)
    d.qubits_per_node = {0: [0, 1]}
    with pytest.raises(ValueError, match="same node IDs"):
        PBlockSimulator().simulate(d)


def test_explicit_operation_id_errors():
    # test_pblock.py:108-126
    x = inst("x", qubit=0)
    d = SimpleNamespace(circuits={0: Circuit(1, [x, x])}, qubits_per_node={0: [0]},
                        operation_ids={0: ["first", "second"]}, operation_order={"first": 0, "second": -1})
    with pytest.raises(ValueError, match="contradicts"):
        PBlockSimulator().simulate(d)
    cx = inst("cx", control=0, target=1)
    d = SimpleNamespace(circuits={0: Circuit(2, [cx]), 1: Circuit(2, [inst("cz", control=0, target=1)])},
                        qubits_per_node={0: [0], 1: [1]}, operation_ids={0: ["gate"], 1: ["gate"]},
                        operation_order={"gate": 0})
    with pytest.raises(ValueError, match="must agree"):
        PBlockSimulator().simulate(d)
    d = SimpleNamespace(circuits={0: Circuit(1, [x])}, qubits_per_node={0: [0]},
                        operation_ids={0: ["a"]}, operation_order={})
    with pytest.raises(ValueError, match="Missing order"):
        PBlockSimulator().simulate(d)
    d = SimpleNamespace(circuits={0: Circuit(1, [x])}, qubits_per_node={0: [0]},
                        operation_ids={0: ["a", "b"]}, operation_order={"a": 0})
    with pytest.raises(ValueError, match="length must match"):
        PBlockSimulator().simulate(d)


def test_budget_errors_before_allocation():
    # test_pblock.py:290-296: wire metadata and the classical workspace are
    # charged before any amplitude exists.
    with pytest.raises(MemoryError):
        PBlockSimulator(max_memory_mb=1).simulate(Circuit(100000))
    with pytest.raises(MemoryError):
        PBlockSimulator(max_memory_mb=1).simulate_shots(Circuit(1, cbits=100000), 0)
    big = SimpleNamespace(circuits={0: Circuit(5000)}, qubits_per_node={0: list(range(5000))},
                          _instruction_index={})
    with pytest.raises(MemoryError):
        PBlockSimulator(max_memory_mb=1).simulate(big)


def test_zero_shots_do_not_execute():
    assert PBlockSimulator().simulate_shots(Circuit(3, [inst("h", qubit=0)], 3), 0) == {}
    assert dqsim.simulate_distributed_shots(Circuit(0), shots=0) == {}


def test_transport_merges_nodes_in_global_order():
    # input.rs:150-224: one register per physical qubit, node instructions
    # interleaved by their global positions, shared copies kept once.
    from dqsim import _pblock
    import json
    x, cx = inst("x", qubit=0), inst("cx", control=0, target=1)
    c0, c1 = Circuit(2, [x, cx]), Circuit(2, [copy.deepcopy(cx)])
    d = SimpleNamespace(circuits={0: c0, 1: c1}, qubits_per_node={0: [0], 1: [1]},
                        operation_ids={0: ["x", "cx"], 1: ["cx"]}, operation_order={"x": 0, "cx": 1})
    merged = json.loads(_pblock._transport(d, 1 << 30))
    assert merged["instructions"] == [x, cx]
    assert merged["qregs"] == {"0": {"name": "0", "base": 0, "size": 1}, "1": {"name": "1", "base": 1, "size": 1}}
    ops = [inst("h", qubit=1), inst("rz", qubit=0, phi=math.pi / 3), inst("cx", control=1, target=0)]
    merged = json.loads(_pblock._transport(distributed(2, ops), 1 << 30))
    assert merged["instructions"] == ops
