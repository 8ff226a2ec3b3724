"""P-block backend on the GPU: block structure, statistics and results
against the FP64 oracle and the statevector path. Cases restate the
reference's tests/test_pblock.py (line numbers per test); where that file
compares with Qiskit (absent here) the oracle's dense state is used. Counts
use Philox streams, not the reference's ChaCha8, so exact counts are
compared only where the outcome is deterministic or between two runs of
this engine."""

import json
import math
import os
from pathlib import Path
from types import SimpleNamespace

import numpy as np
import pytest

import dqsim
from dqsim import PBlockSimulator, StatevectorSimulator, simulate_distributed, simulate_distributed_shots
from dqsim.workloads import Circuit
from oracle import oracle
from test_gpu_parity import standard_gate_cases
from test_pblock_host import conditional, distributed, inst

pytestmark = pytest.mark.gpu


def measures(n):
    return [inst("measure", qubit=q, cbit=q) for q in range(n)]


def bell_pairs(n):
    return [op for q in range(0, n, 2) for op in (inst("h", qubit=q), inst("cx", control=q, target=q + 1))]


def dense(circuit):
    return oracle.simulate(circuit.data, 1)[0]


def test_cross_node_feedback_both_execution_paths():
    # test_pblock.py:40-49
    print("[pblock] cross-node feedback", flush=True)
    ops = [inst("x", qubit=0), inst("measure", qubit=0, cbit=0), conditional(inst("x", qubit=1), 1),
           inst("measure", qubit=1, cbit=1)]
    for prior in (False, True):
        current = ops[:2] + ([inst("id", qubit=0)] if prior else []) + ops[2:]
        d = distributed(2, current, cbits=2)
        sim = PBlockSimulator(seed=3)
        assert sim.simulate(d).classical_bits == {0: 1, 1: 1}
        assert sim.simulate_shots(d, 64) == {"11": 64}


def test_dynamic_feedback_correlations_and_worker_reuse():
    # test_pblock.py:51-58: counts do not depend on max_parallel_shots
    print("[pblock] feedback and workers", flush=True)
    ops = bell_pairs(4) + [inst("measure", qubit=0, cbit=0), conditional(inst("x", qubit=1), 1),
                           inst("reset", qubit=2)] + measures(4)
    c = Circuit(4, ops, 4)
    a = PBlockSimulator(seed=91, max_parallel_shots=1).simulate_shots(c, 1000)
    b = PBlockSimulator(seed=91, max_parallel_shots=4).simulate_shots(c, 1000)
    assert a == b
    assert set(a) == {"0000", "0001", "1000", "1001"}
    for v in a.values():
        assert abs(v / 1000 - 0.25) < 0.06


def test_remote_phase_gate_matches_statevector():
    # test_pblock.py:60-70
    ops = [inst("h", qubit=0), inst("h", qubit=1), inst("gate", name="remote_cu1", qubits=[0, 1], params=[math.pi]),
           inst("h", qubit=0), inst("h", qubit=1)]
    c = Circuit(2, ops)
    r = PBlockSimulator().simulate(c)
    np.testing.assert_allclose(r.statevector, StatevectorSimulator().simulate(c).statevector, atol=1e-13)
    for p in r.probabilities().values():
        assert abs(p - 0.25) < 1e-12


def test_transport_results():
    # test_pblock.py:93-121
    ops = [inst("x", qubit=0), inst("cx", control=0, target=1)]
    np.testing.assert_array_equal(PBlockSimulator().simulate(distributed(2, ops)).statevector, [0, 0, 0, 1])
    x, cx = inst("x", qubit=0), inst("cx", control=0, target=1)
    d = SimpleNamespace(circuits={0: Circuit(2, [x, cx]), 1: Circuit(2, [dict(cx)])}, qubits_per_node={0: [0], 1: [1]},
                        operation_ids={0: ["x", "cx"], 1: ["cx"]}, operation_order={"x": 0, "cx": 1})
    np.testing.assert_array_equal(PBlockSimulator().simulate(d).statevector, [0, 0, 0, 1])
    d = SimpleNamespace(circuits={0: Circuit(1, [x, x])}, qubits_per_node={0: [0]},
                        operation_ids={0: ["first", "second"]}, operation_order={"first": 0, "second": 1})
    np.testing.assert_array_equal(PBlockSimulator().simulate(d).statevector, [1, 0])


def test_partition_independence_and_large_product_state():
    # test_pblock.py:128-146
    print("[pblock] 100-qubit product state", flush=True)
    n = 100
    ops = [inst("h", qubit=q) for q in range(n)]
    sim = PBlockSimulator(seed=1, max_memory_mb=1, max_block_qubits=2)
    r = sim.simulate(distributed(n, ops, {0: list(range(n))}))
    assert len(r.block_qubits) == n
    assert r.diagnostics["live_amplitudes"] == 2 * n
    assert abs(r.expectation_value("X" * n) - 1.0) < 1e-12
    assert abs(r.amplitude("0" * n) - 2 ** (-n / 2)) < 1e-25
    counts = r.counts(20, seed=3)
    assert sum(counts.values()) == 20 and all(len(k) == n for k in counts)
    with pytest.raises(MemoryError):
        _ = r.statevector
    with pytest.raises(MemoryError):
        r.probabilities()
    p = r.probabilities([99, 0])
    assert len(p) == 4 and all(abs(v - 0.25) < 1e-12 for v in p.values())
    small = ops[:8] + measures(8)
    a = sim.simulate_shots(distributed(8, small, {0: list(range(8))}, 8), 100)
    b = sim.simulate_shots(distributed(8, small, cbits=8), 100)
    assert a == b


def test_large_independent_bell_pairs():
    # test_pblock.py:148-158
    n = 100
    sim = PBlockSimulator(seed=2, max_memory_mb=1, max_block_qubits=2)
    r = sim.simulate(Circuit(n, bell_pairs(n)))
    assert sorted(map(len, r.block_qubits)) == [2] * 50
    assert r.diagnostics["peak_block_qubits"] == 2
    assert abs(r.expectation_value("X" * n) - 1.0) < 1e-12
    assert abs(r.expectation_value("Y" * n) - (-1.0 if (n // 2) % 2 else 1.0)) < 1e-12
    counts = sim.simulate_shots(Circuit(n, bell_pairs(n) + measures(n), n), 100)
    assert sum(counts.values()) == 100
    for key in counts:
        assert all(key[i] == key[i + 1] for i in range(0, n, 2))


def test_measurement_and_reset_release_blocks_and_keep_phase():
    # test_pblock.py:160-171
    ops = [inst("h", qubit=0), inst("s", qubit=0), inst("cx", control=0, target=1), inst("measure", qubit=0, cbit=0)]
    for seed in range(6):
        r = PBlockSimulator(seed=seed).simulate(Circuit(2, ops, 1))
        assert sorted(map(len, r.block_qubits)) == [1, 1]
        bit = r.classical_bits[0]
        expected = np.zeros(4, dtype=complex)
        expected[3 if bit else 0] = 1j if bit else 1
        np.testing.assert_allclose(r.statevector, expected, atol=1e-13)
    r = PBlockSimulator(seed=2).simulate(Circuit(2, ops[:-1] + [inst("reset", qubit=0)]))
    assert sorted(map(len, r.block_qubits)) == [1, 1]
    assert abs(r.probabilities([0]).get(0, 0) - 1.0) < 1e-12


def test_known_controls_conditions_and_swaps_do_not_merge():
    # test_pblock.py:173-181
    print("[pblock] known controls", flush=True)
    ops = [inst("cx", control=q, target=q + 1) for q in range(99)]
    ops += [conditional(inst("cx", control=0, target=99), 1),
            inst("gate", name="remote_barrier", qubits=list(range(100)), params=[]),
            inst("x", qubit=0), inst("swap", a=0, b=99)]
    r = PBlockSimulator(max_memory_mb=1, max_block_qubits=1).simulate(Circuit(100, ops, 1))
    assert r.diagnostics["merge_calls"] == 0
    assert r.amplitude("1" + "0" * 99) == 1
    assert r.diagnostics["skipped_controlled_gates"] == 99


def test_controlled_eigenstates_and_identity_diagonals_stay_separate():
    # test_pblock.py:183-189
    ops = [inst("h", qubit=0), inst("h", qubit=1), inst("cx", control=0, target=1), inst("rzz", a=0, b=1, theta=0.0)]
    r = PBlockSimulator(max_block_qubits=1).simulate(Circuit(2, ops))
    assert r.diagnostics["merge_calls"] == 0
    np.testing.assert_allclose(r.statevector, [0.5, 0.5, 0.5, 0.5], atol=1e-13)


def test_swaps_within_and_between_entangled_blocks():
    # test_pblock.py:191-195
    ops = bell_pairs(4) + [inst("s", qubit=0), inst("swap", a=0, b=3), inst("swap", a=1, b=3), inst("h", qubit=3),
                           inst("cx", control=3, target=2)]
    c = Circuit(4, ops)
    np.testing.assert_allclose(PBlockSimulator().simulate(c).statevector, StatevectorSimulator().simulate(c).statevector,
                               atol=1e-13)
    np.testing.assert_allclose(PBlockSimulator().simulate(c).statevector, dense(c), atol=1e-13)


def test_optional_separation_and_norm():
    # test_pblock.py:197-209
    ops = bell_pairs(2) + [inst("cx", control=0, target=1)]
    plain = PBlockSimulator().simulate(Circuit(2, ops))
    exact = PBlockSimulator(split_separable=True).simulate(Circuit(2, ops))
    assert len(plain.block_qubits) == 1
    assert len(exact.block_qubits) == 2
    np.testing.assert_allclose(exact.statevector, plain.statevector, atol=1e-13)
    entangled = Circuit(2, [inst("ry", qubit=0, theta=0.02), inst("cx", control=0, target=1)])
    r = PBlockSimulator(split_separable=True, separation_tolerance=0.001).simulate(entangled)
    assert len(r.block_qubits) == 2
    assert r.diagnostics["separation_residual"] > 0
    assert abs(r.diagnostics["norm_squared"] - 1.0) < 1e-7
    assert len(PBlockSimulator(split_separable=True).simulate(entangled).block_qubits) == 1


def test_memory_limits_fail_before_large_allocations():
    # test_pblock.py:284-296
    ops = [inst("h", qubit=0)] + [inst("cx", control=q, target=q + 1) for q in range(19)]
    for sim in (PBlockSimulator(max_block_qubits=4), PBlockSimulator(max_memory_mb=1)):
        with pytest.raises(MemoryError):
            sim.simulate(Circuit(20, ops))
    r = PBlockSimulator(max_memory_mb=1).simulate(Circuit(20))
    with pytest.raises(MemoryError):
        _ = r.statevector


def test_query_order_and_sparse_physical_labels():
    # test_pblock.py:298-309
    c = Circuit(0, [inst("x", qubit=100), inst("h", qubit=7)])
    c.data["qregs"] = {"a": {"name": "a", "base": 7, "size": 1}, "b": {"name": "b", "base": 100, "size": 1}}
    r = PBlockSimulator().simulate(c)
    assert r.physical_qubits == [7, 100]
    np.testing.assert_allclose(r.statevector, [0, 0, 2 ** -0.5, 2 ** -0.5], atol=1e-13)
    assert abs(r.probabilities([7, 100])[1] - 0.5) < 1e-12
    assert set(r.counts(100, qubits=[7, 100], seed=1)) == {"01", "11"}
    assert abs(r.amplitude("10") - 2 ** -0.5) < 1e-12
    assert abs(r.expectation_value("ZX") + 1.0) < 1e-12
    assert abs(r.probabilities([])[0] - 1.0) < 1e-12
    for qs in ([3], [7, 7]):
        with pytest.raises(ValueError):
            r.probabilities(qs)
        with pytest.raises(ValueError):
            r.counts(qubits=qs)
    with pytest.raises(ValueError):
        r.amplitude("02")
    with pytest.raises(ValueError):
        r.expectation_value("AB")


def test_terminal_repeated_measurements_and_overwrites():
    # test_pblock.py:311-316
    ops = bell_pairs(2) + [inst("measure", qubit=1, cbit=0), inst("measure", qubit=1, cbit=1),
                           inst("measure", qubit=0, cbit=0)]
    for terminal in (False, True):
        counts = PBlockSimulator(seed=5, sample_terminal=terminal).simulate_shots(Circuit(2, ops, 3), 1000)
        assert set(counts) == {"000", "011"}
        assert abs(counts["011"] / 1000 - 0.5) < 0.06


def test_tiny_gates_are_preserved():
    # test_pblock.py:318-320
    r = PBlockSimulator().simulate(Circuit(1, [inst("rx", qubit=0, theta=1e-11)]))
    assert abs(r.statevector[1].imag + 5e-12) < 1e-25


def test_empty_result_and_wrappers(tmp_path):
    # test_pblock.py:322-332
    old = os.getcwd()
    os.chdir(tmp_path)
    try:
        r = simulate_distributed(Circuit(0), profile=True, max_memory_mb=1)
        np.testing.assert_array_equal(r.statevector, [1])
        assert r.amplitude("") == 1
        assert r.probabilities() == {0: 1}
        assert r.counts(7) == {"": 7}
        assert r.expectation_value("") == 1
        assert "total_time" in r.profile
        assert simulate_distributed_shots(Circuit(0), shots=7) == {"": 7}
        assert PBlockSimulator().simulate_shots(Circuit(0), 0) == {}
        assert PBlockSimulator().simulate_shots(Circuit(1, cbits=3), 7) == {"000": 7}
    finally:
        os.chdir(old)


def _profile(tmp_path, fn):
    old = os.getcwd()
    os.chdir(tmp_path)
    try:
        out = fn()
        report = json.loads(next(Path("dqsim_profiles").glob("*.json")).read_text())
    finally:
        os.chdir(old)
    return out, report


def test_sampling_and_prefix_fall_back_when_memory_is_tight(tmp_path):
    # test_pblock.py:334-353: two 16-qubit GHZ blocks fit, their sampling
    # tables and a retained prefix do not.
    print("[pblock] memory-tight shots", flush=True)
    ops = []
    for start in (0, 16):
        ops.append(inst("h", qubit=start))
        ops.extend(inst("cx", control=q, target=q + 1) for q in range(start, start + 15))
    ops.extend([inst("measure", qubit=0, cbit=0), inst("measure", qubit=16, cbit=1)])
    counts, report = _profile(tmp_path, lambda: PBlockSimulator(seed=3, max_memory_mb=3, max_parallel_shots=4)
                              .simulate_shots(Circuit(32, ops, 2), 4, profile=True))
    assert sum(counts.values()) == 4
    assert report["execution_strategy"] == "trajectories"
    assert report["deterministic_prefix_ops"] == 0
    assert report["parallel_shots"] == 1
    assert report["peak_working_bytes"] <= 3 * 1024 ** 2
    # Unprofiled, the shots run in one lockstep batch under the same budget.
    assert PBlockSimulator(seed=3, max_memory_mb=3, max_parallel_shots=4).simulate_shots(Circuit(32, ops, 2), 4) \
        == counts
    # A trajectory that outgrows max_block_qubits raises what a lone shot raises.
    # (ry states: a CX onto |+> would be skipped as trivial and never merge)
    ops = [inst("ry", qubit=q, theta=0.3 + 0.1 * q) for q in range(6)] + [inst("measure", qubit=5, cbit=0)]
    ops += [inst("cx", control=q, target=q + 1) for q in range(4)]
    with pytest.raises(MemoryError, match="max_block_qubits"):
        PBlockSimulator(seed=3, max_block_qubits=4).simulate_shots(Circuit(6, ops, 1), 50)


def test_profiles_confirm_shared_evolution(tmp_path):
    # test_pblock.py:355-366
    _, p = _profile(tmp_path, lambda: PBlockSimulator().simulate_shots(Circuit(8, bell_pairs(8) + measures(8), 8), 100,
                                                                        profile=True))
    assert p["execution_strategy"] == "terminal_sampling"
    assert p["merge_calls"] == 4
    assert p["peak_block_qubits"] == 2
    assert p["peak_working_bytes"] < 1024 ** 2


def test_every_standard_gate_on_entangled_inputs_matches_oracle():
    # test_pblock.py:371-384 (Qiskit replaced by the FP64 oracle)
    print("[pblock] standard gates vs oracle", flush=True)
    n = 6
    for kind, names, params in standard_gate_cases():
        ops = [inst("u", qubit=q, theta=0.21 + 0.27 * q, phi=-0.17, lam=0.61) for q in range(n)]
        ops += [inst("cx", control=q, target=q + 1) for q in range(0, n, 2)]
        qs = [5, 0, 3, 2, 4][:len(names)]
        ops.append(inst(kind, **dict(zip(names, qs)), **params))
        r = PBlockSimulator().simulate(distributed(n, ops))
        np.testing.assert_allclose(r.statevector, dense(Circuit(n, ops)), atol=3e-12, err_msg=kind)


def test_parallel_merges_and_large_block_kernels():
    # test_pblock.py:386-397
    n = 13
    ops = [inst("h", qubit=q) for q in range(n)]
    ops += [inst("cz", control=q, target=(q + 5) % n) for q in range(n)]
    ops += [inst("swap", a=12, b=0), inst("rxx", a=11, b=2, theta=0.47),
            inst("c4x", control1=1, control2=12, control3=6, control4=3, target=8)]
    c = Circuit(n, ops)
    r = PBlockSimulator(max_memory_mb=8).simulate(c)
    assert r.diagnostics["peak_block_qubits"] == n
    np.testing.assert_allclose(r.statevector, StatevectorSimulator().simulate(c).statevector, atol=3e-12)
    np.testing.assert_allclose(r.statevector, dense(c), atol=3e-12)


def test_random_circuits_and_compact_queries():
    # test_pblock.py:399-418 (Qiskit replaced by the FP64 oracle)
    print("[pblock] random circuits", flush=True)
    rng = np.random.default_rng(923)
    cases = standard_gate_cases()
    mats = {"I": np.eye(2), "X": np.array([[0, 1], [1, 0]]), "Y": np.array([[0, -1j], [1j, 0]]), "Z": np.diag([1, -1])}
    for trial in range(12):
        n = 7
        ops = []
        for _ in range(70):
            kind, names, params = cases[int(rng.integers(len(cases)))]
            qs = [int(q) for q in rng.choice(n, len(names), replace=False)]
            ops.append(inst(kind, **dict(zip(names, qs)), **params))
        c = Circuit(n, ops)
        r = PBlockSimulator(split_separable=bool(trial % 2)).simulate(c)
        v = dense(c)
        np.testing.assert_allclose(r.statevector, v, atol=5e-12)
        expected = StatevectorSimulator().simulate(c).probabilities([6, 0, 3])
        actual = r.probabilities([6, 0, 3])
        for k in set(actual) | set(expected):
            assert abs(actual.get(k, 0) - expected.get(k, 0)) < 1e-12
        pauli = "XYZIYZX"
        m = np.array([[1]])
        for ch in pauli:
            m = np.kron(m, mats[ch])
        assert abs(r.expectation_value(pauli) - np.vdot(v, m @ v).real) < 1e-12


def test_split_blocks_stay_exact_for_product_states():
    # Separation of a block whose state factors up to rounding (CZ applied
    # twice): the split factors reproduce the dense state.
    n = 6
    ops = [inst("u", qubit=q, theta=0.3 + 0.2 * q, phi=0.1 * q, lam=-0.2) for q in range(n)]
    ops += [inst("cz", control=0, target=1), inst("cz", control=0, target=1)]  # CZ twice: identity
    c = Circuit(n, ops)
    r = PBlockSimulator(split_separable=True, separation_tolerance=1e-20).simulate(c)
    np.testing.assert_allclose(r.statevector, dense(c), atol=1e-12)
    assert sorted(map(len, r.block_qubits)) == [1] * n
    assert r.diagnostics["split_calls"] >= 1 and r.diagnostics["separation_residual"] < 1e-20
