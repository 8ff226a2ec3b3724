"""GPU parity: the HIP engine against the CPU FP64 oracle on identical inputs.

complex128 runs must match the oracle bit for bit (same arithmetic order, same
pairwise-tree sums, same Philox stream); complex64 fused runs must stay within
1e-5 max amplitude error.
"""

import numpy as np
import pytest

import dqsim
from dqsim import workloads
from dqsim.workloads import Circuit
from oracle import oracle

pytestmark = pytest.mark.gpu


def inst(kind, **fields):
    return {"kind": kind, **fields}


def standard_gate_cases():
    """(JSON kind, operand names, parameters) for every standard gate kind
    (the gate list of the reference's tests/test_statevector.py:350-401)."""
    theta, phi, lam, gamma = 0.731, -0.412, 1.137, 0.289
    single, controlled, pair = ("qubit",), ("control", "target"), ("a", "b")
    triple = ("control1", "control2", "target")
    quad = ("control1", "control2", "control3", "target")
    return [
        ("id", single, {}), ("u0", single, {}), ("x", single, {}), ("y", single, {}), ("z", single, {}),
        ("h", single, {}), ("s", single, {}), ("sdg", single, {}), ("t", single, {}), ("tdg", single, {}),
        ("sx", single, {}), ("sxdg", single, {}),
        ("u3", single, {"theta": theta, "phi": phi, "lam": lam}), ("u2", single, {"phi": phi, "lam": lam}),
        ("u1", single, {"lam": lam}), ("u", single, {"theta": theta, "phi": phi, "lam": lam}),
        ("p", single, {"lam": lam}), ("rx", single, {"theta": theta}), ("ry", single, {"theta": theta}),
        ("rz", single, {"phi": phi}),
        ("cx", controlled, {}), ("cz", controlled, {}), ("cy", controlled, {}), ("ch", controlled, {}),
        ("swap", pair, {}), ("csx", controlled, {}), ("crx", controlled, {"theta": theta}),
        ("cry", controlled, {"theta": theta}), ("crz", controlled, {"lam": lam}), ("cu1", controlled, {"lam": lam}),
        ("cp", controlled, {"lam": lam}), ("cu3", controlled, {"theta": theta, "phi": phi, "lam": lam}),
        ("cu", controlled, {"theta": theta, "phi": phi, "lam": lam, "gamma": gamma}),
        ("rxx", pair, {"theta": theta}), ("rzz", pair, {"theta": theta}),
        ("ccx", triple, {}), ("cswap", ("control", "target1", "target2"), {}), ("rccx", triple, {}),
        ("rc3x", quad, {}), ("c3x", quad, {}), ("c3sqrtx", quad, {}),
        ("c4x", ("control1", "control2", "control3", "control4", "target"), {}),
    ]


def entangled_prefix(n):
    out = []
    for q in range(n):
        out.append(inst("u", qubit=q, theta=0.21 + 0.37 * q, phi=-0.13 + 0.31 * q, lam=0.7 - 0.09 * q))
    for q in range(n - 1):
        out.append(inst("cx", control=q, target=q + 1))
    return out


def gpu_state(instructions, n, **opts):
    sim = dqsim.StatevectorSimulator(seed=60, **opts)
    return sim.simulate(Circuit(n, instructions)).statevector


def oracle_state(instructions, n, seed=60):
    return oracle.simulate(Circuit(n, instructions).data, seed)[0]


@pytest.mark.parametrize("case", standard_gate_cases(), ids=lambda c: c[0])
def test_every_gate_bit_exact_on_permuted_qubits(case):
    kind, names, params = case
    n = len(names) + 1
    for qubits in (list(range(len(names))), list(range(n - 1, 0, -1))):
        insts = entangled_prefix(n) + [inst(kind, **dict(zip(names, qubits)), **params)]
        got, want = gpu_state(insts, n), oracle_state(insts, n)
        assert np.array_equal(got, want), (kind, qubits, np.abs(got - want).max())


def test_distant_qubits_13q_bit_exact():
    n = 13
    insts = [inst("u", qubit=q, theta=0.21 + 0.03 * q, phi=-0.17, lam=0.48) for q in range(n)]
    for kind, names, params in standard_gate_cases():
        qubits = [12, 0, 7, 3, 10][: len(names)]
        insts.append(inst(kind, **dict(zip(names, qubits)), **params))
    assert np.array_equal(gpu_state(insts, n), oracle_state(insts, n))


def test_random_circuits_bit_exact():
    rng = np.random.default_rng(1907)
    cases = standard_gate_cases()
    for run in range(6):
        n = 6
        insts = []
        for _ in range(60):
            kind, names, params = cases[int(rng.integers(len(cases)))]
            qubits = rng.choice(n, size=len(names), replace=False).tolist()
            insts.append(inst(kind, **dict(zip(names, qubits)), **params))
        assert np.array_equal(gpu_state(insts, n), oracle_state(insts, n)), run


def test_brickwork_20q_state_and_counts_match_oracle():
    circ = workloads.brickwork(n=20, layers=20, seed=42)
    unitary = [i for i in circ.data["instructions"] if i["kind"] != "measure"]
    sim = dqsim.StatevectorSimulator(seed=42)
    got = sim.simulate(Circuit(20, unitary)).statevector
    want = oracle_state(unitary, 20, seed=42)
    assert np.array_equal(got, want)
    counts = sim.simulate_shots(circ, shots=1000)
    ref, _ = oracle.run_shots(circ.data, 1000, 42)
    assert counts == ref


DYNAMIC = [
    Circuit(2, [inst("h", qubit=0), inst("measure", qubit=0, cbit=0),
                {"kind": "conditional", "condition": {"creg_base": 0, "creg_size": 1, "creg_value": 1},
                 "op": inst("x", qubit=1)},
                inst("x", qubit=0), inst("measure", qubit=0, cbit=1), inst("measure", qubit=1, cbit=2)], 3),
    Circuit(3, [inst("h", qubit=0), inst("cx", control=0, target=1), inst("reset", qubit=0),
                inst("ry", qubit=2, theta=0.7), inst("measure", qubit=2, cbit=0), inst("h", qubit=0),
                inst("measure", qubit=0, cbit=1), inst("measure", qubit=1, cbit=2)], 3),
    Circuit(12, [inst("h", qubit=q) for q in range(12)] + [inst("measure", qubit=5, cbit=0)]
            + [{"kind": "conditional", "condition": {"creg_base": 0, "creg_size": 1, "creg_value": 1},
                "op": inst("rx", qubit=11, theta=0.3)}] + [inst("measure", qubit=q, cbit=q) for q in range(12)], 12),
]


@pytest.mark.parametrize("idx", range(len(DYNAMIC)))
@pytest.mark.parametrize("terminal", [True, False])
def test_shot_counts_match_oracle(idx, terminal):
    circ = DYNAMIC[idx]
    for mps in (None, 1, 7):
        sim = dqsim.StatevectorSimulator(seed=38, sample_terminal=terminal, max_parallel_shots=mps)
        got = sim.simulate_shots(circ, shots=257)
        ref, _ = oracle.run_shots(circ.data, 257, 38, sample_terminal=terminal, max_parallel_shots=mps)
        assert got == ref


def test_simulate_with_measurements_matches_oracle():
    circ = DYNAMIC[1]
    for seed in (1, 2, 3, 4):
        res = dqsim.StatevectorSimulator(seed=seed).simulate(circ)
        st, bits = oracle.simulate(circ.data, seed)
        assert np.array_equal(res.statevector, st)
        assert res.classical_bits == bits


def test_probabilities_counts_and_fidelity_match_oracle():
    n = 9
    insts = entangled_prefix(n) + [inst("rxx", a=1, b=7, theta=0.4)]
    res = dqsim.StatevectorSimulator(seed=3).simulate(Circuit(n, insts))
    st = oracle_state(insts, n, 3)
    for qubits in ([3, 0, 8], [0], list(range(n))[::-1], [], [8, 1, 2, 3, 4, 5, 6, 7, 0]):
        want = oracle.marginal(st, qubits)
        got = res.probabilities(qubits)
        assert {j: p for j, p in enumerate(want) if p > 0} == got
        idx = oracle.sample_marginal(st, qubits, 500, 77)
        assert res.counts(shots=500, qubits=qubits, seed=77) == oracle.counts_from_indices(idx, len(qubits))
    other = oracle_state(entangled_prefix(n), n)
    assert res.fidelity(other) == oracle.fidelity(st, other)


@pytest.mark.parametrize("max_k", [2, 3, 4, 5])
def test_complex64_fused_within_1e5(max_k):
    circ = workloads.random_su2(n=12, depth=12, seed=5)
    want = oracle_state(circ.data["instructions"], 12, 1)
    sim = dqsim.StatevectorSimulator(seed=1, precision="complex64", fusion="greedy", max_fused_qubits=max_k)
    got = sim.simulate(circ).statevector
    assert np.abs(got - want).max() < 1e-5
    sim128 = dqsim.StatevectorSimulator(seed=1, fusion="greedy", max_fused_qubits=max_k)
    got128 = sim128.simulate(circ).statevector
    assert np.abs(got128 - want).max() < 1e-12
