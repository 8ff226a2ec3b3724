"""Pins of the CPU P-block restatement (oracle/pblock.py) before it checks
the GPU backend: its states equal the dense FP64 oracle's, and it reproduces
the structural facts the reference's tests/test_pblock.py asserts (block
sizes, merge and skip counts, splits, phases after measurement)."""

import numpy as np
import pytest

from oracle import oracle
from oracle import pblock as opb
from test_gpu_parity import standard_gate_cases
from test_pblock_host import conditional, inst
from dqsim.workloads import Circuit


def measures(n):
    return [inst("measure", qubit=q, cbit=q) for q in range(n)]


def bell_pairs(n):
    return [op for q in range(0, n, 2) for op in (inst("h", qubit=q), inst("cx", control=q, target=q + 1))]


def dense(c):
    return oracle.simulate(c.data, 1)[0]


def test_states_equal_the_dense_oracle():
    # Every standard gate after an entangling prefix (test_pblock.py:371-384)
    # and random 7-qubit circuits (test_pblock.py:399-418), with and without
    # separation.
    n = 6
    for kind, names, params in standard_gate_cases():
        ops = [inst("u", qubit=q, theta=0.21 + 0.27 * q, phi=-0.17, lam=0.61) for q in range(n)]
        ops += [inst("cx", control=q, target=q + 1) for q in range(0, n, 2)]
        ops.append(inst(kind, **dict(zip(names, [5, 0, 3, 2, 4][:len(names)])), **params))
        c = Circuit(n, ops)
        np.testing.assert_allclose(opb.simulate(c.data, 1).statevector, dense(c), atol=3e-12, err_msg=kind)
    rng = np.random.default_rng(923)
    cases = standard_gate_cases()
    for trial in range(6):
        ops = []
        for _ in range(50):
            kind, names, params = cases[int(rng.integers(len(cases)))]
            qs = [int(q) for q in rng.choice(7, len(names), replace=False)]
            ops.append(inst(kind, **dict(zip(names, qs)), **params))
        c = Circuit(7, ops)
        r = opb.simulate(c.data, 1, split_separable=bool(trial % 2))
        np.testing.assert_allclose(r.statevector, dense(c), atol=5e-12)
        want = oracle.marginal(dense(c), [6, 0, 3])
        got = r.probabilities([6, 0, 3])
        for k in range(8):
            assert abs(got.get(k, 0.0) - want[k]) < 1e-12


def test_reference_structure_facts():
    # test_pblock.py:148-158, 173-189, 197-209, 355-366
    r = opb.simulate(Circuit(8, bell_pairs(8)).data, 1)
    assert r.diagnostics["merge_calls"] == 4 and r.diagnostics["peak_block_qubits"] == 2
    assert sorted(map(len, r.block_qubits)) == [2] * 4
    ops = [inst("cx", control=q, target=q + 1) for q in range(39)]
    ops += [conditional(inst("cx", control=0, target=39), 1), inst("x", qubit=0), inst("swap", a=0, b=39)]
    r = opb.simulate(Circuit(40, ops, 1).data, 1, max_block_qubits=1)
    assert r.diagnostics["merge_calls"] == 0 and r.diagnostics["skipped_controlled_gates"] == 39
    assert r.pool.blocks[r.pool.positions[39][0]].state[1] == 1
    ops = [inst("h", qubit=0), inst("h", qubit=1), inst("cx", control=0, target=1), inst("rzz", a=0, b=1, theta=0.0)]
    r = opb.simulate(Circuit(2, ops).data, 1, max_block_qubits=1)
    assert r.diagnostics["merge_calls"] == 0
    np.testing.assert_allclose(r.statevector, [0.5] * 4, atol=1e-13)
    ops = bell_pairs(2) + [inst("cx", control=0, target=1)]
    assert len(opb.simulate(Circuit(2, ops).data, 1).block_qubits) == 1
    assert len(opb.simulate(Circuit(2, ops).data, 1, split_separable=True).block_qubits) == 2
    ent = Circuit(2, [inst("ry", qubit=0, theta=0.02), inst("cx", control=0, target=1)]).data
    r = opb.simulate(ent, 1, split_separable=True, separation_tolerance=0.001)
    assert len(r.block_qubits) == 2 and r.diagnostics["separation_residual"] > 0
    assert len(opb.simulate(ent, 1, split_separable=True).block_qubits) == 1
    with pytest.raises(MemoryError):
        opb.simulate(Circuit(20, [inst("h", qubit=0)] + [inst("cx", control=q, target=q + 1)
                                                       for q in range(19)]).data, 1, max_block_qubits=4)


def test_measurement_releases_blocks_and_keeps_phase():
    # test_pblock.py:160-171
    ops = [inst("h", qubit=0), inst("s", qubit=0), inst("cx", control=0, target=1), inst("measure", qubit=0, cbit=0)]
    seen = set()
    for seed in range(12):
        r = opb.simulate(Circuit(2, ops, 1).data, seed)
        assert sorted(map(len, r.block_qubits)) == [1, 1]
        bit = r.classical_bits[0]
        seen.add(bit)
        expected = np.zeros(4, dtype=complex)
        expected[3 if bit else 0] = 1j if bit else 1
        np.testing.assert_allclose(r.statevector, expected, atol=1e-13)
    assert seen == {0, 1}


def test_shot_strategies_and_outcome_sets():
    # test_pblock.py:311-316, 51-58, 355-366
    ops = bell_pairs(2) + [inst("measure", qubit=1, cbit=0), inst("measure", qubit=1, cbit=1),
                           inst("measure", qubit=0, cbit=0)]
    for terminal in (False, True):
        counts, strategy = opb.simulate_shots(Circuit(2, ops, 3).data, 400, 5, sample_terminal=terminal)
        assert set(counts) == {"000", "011"} and abs(counts["011"] / 400 - 0.5) < 0.08
        assert strategy == ("terminal_sampling" if terminal else "prefix_trajectories")
    ops = bell_pairs(4) + [inst("measure", qubit=0, cbit=0), conditional(inst("x", qubit=1), 1),
                           inst("reset", qubit=2)] + measures(4)
    a, _ = opb.simulate_shots(Circuit(4, ops, 4).data, 300, 91, max_parallel_shots=1)
    b, _ = opb.simulate_shots(Circuit(4, ops, 4).data, 300, 91, max_parallel_shots=4)
    assert a == b and set(a) == {"0000", "0001", "1000", "1001"}
    counts, strategy = opb.simulate_shots(Circuit(8, bell_pairs(8) + measures(8), 8).data, 100, 3)
    assert strategy == "terminal_sampling" and all(k[0:2] in ("00", "11") for k in counts)
