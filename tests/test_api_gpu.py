"""Statevector API behaviour on the GPU engine, restating the reference's
tests/test_statevector.py:53-347 (StatevectorTests) against ``dqsim``.

Counts are seeded Philox draws, so the exact values differ from the
reference's ChaCha8 streams; distributions are checked with the reference's
+-0.08 band and deterministic outcomes exactly.
"""

import json
import math
import os
from pathlib import Path

import numpy as np
import pytest

import dqsim
from dqsim import workloads
from dqsim.workloads import Circuit
from oracle import oracle

pytestmark = pytest.mark.gpu


def inst(kind, **fields):
    return {"kind": kind, **fields}


def conditional(op, value, base=0, size=1):
    return inst("conditional", condition={"creg_base": base, "creg_size": size, "creg_value": value}, op=op)


def simulate(circuit, **options):
    return dqsim.StatevectorSimulator(seed=1729, **options).simulate(circuit)


def assert_distribution(counts, expected, shots):
    assert sum(counts.values()) == shots
    assert set(counts) == set(expected)
    for outcome, p in expected.items():
        assert abs(counts[outcome] / shots - p) <= 0.08


@pytest.mark.parametrize("precision", ["complex128", "complex64"])
def test_little_endian_statevector_and_ordered_marginals(precision):
    result = simulate(Circuit(4, [inst("x", qubit=0), inst("x", qubit=2)]), precision=precision)
    expected = np.zeros(16, dtype=complex)
    expected[5] = 1
    np.testing.assert_allclose(result.statevector, expected, atol=1e-14)
    assert result.probabilities() == {5: 1.0}
    assert result.probabilities([0, 1]) == {2: 1.0}
    assert result.probabilities([1, 0]) == {1: 1.0}
    assert result.probabilities([]) == {0: 1.0}
    assert result.counts(shots=17, qubits=[0, 1], seed=2) == {"10": 17}
    assert result.counts(shots=17, qubits=[], seed=2) == {"": 17}
    assert result.amplitude("0101") == 1 and result.amplitude(5) == 1 and result.num_qubits == 4


def test_full_and_marginal_probabilities_match_amplitudes():
    result = simulate(Circuit(5, [inst("ry", qubit=0, theta=0.7), inst("h", qubit=4),
                                  inst("cx", control=4, target=2), inst("rx", qubit=1, theta=-0.3),
                                  inst("rxx", a=1, b=3, theta=0.8)]))
    probabilities = np.abs(result.statevector) ** 2
    actual = result.probabilities()
    np.testing.assert_allclose([actual.get(i, 0) for i in range(32)], probabilities, atol=1e-14)
    assert abs(sum(actual.values()) - 1.0) < 1e-13
    for qubits in ([3, 0, 4], [0], [4, 3, 2, 1, 0], []):
        expected = np.zeros(1 << len(qubits))
        for index, p in enumerate(probabilities):
            key = 0
            for q in qubits:
                key = (key << 1) | ((index >> q) & 1)
            expected[key] += p
        got = result.probabilities(qubits)
        np.testing.assert_allclose([got.get(i, 0) for i in range(len(expected))], expected, atol=1e-14)


def test_tiny_rotations_are_not_discarded_by_fusion():
    theta = 1e-11
    for fusion in ("none", "greedy"):
        result = simulate(Circuit(1, [inst("rx", qubit=0, theta=theta)]), fusion=fusion)
        np.testing.assert_allclose(result.statevector, [math.cos(theta / 2), -1j * math.sin(theta / 2)],
                                   rtol=1e-13, atol=1e-15)


def test_relative_controlled_x_gates_flip_the_target():
    for kind, controls in (("rccx", 2), ("rc3x", 3)):
        gate = inst(kind, target=controls)
        gate.update({f"control{i + 1}": i for i in range(controls)})
        result = simulate(Circuit(controls + 1, [*(inst("x", qubit=i) for i in range(controls)), gate]))
        assert abs(result.probabilities().get((1 << (controls + 1)) - 1, 0) - 1.0) < 1e-12


@pytest.mark.parametrize("sample_terminal", [True, False])
def test_terminal_sampling_preserves_joint_correlations_and_mapping(sample_terminal):
    circuit = Circuit(3, [inst("h", qubit=0), inst("cx", control=0, target=2), inst("measure", qubit=2, cbit=1),
                          inst("barrier"), inst("measure", qubit=0, cbit=3)], cbits=5)
    sim = dqsim.StatevectorSimulator(seed=10, sample_terminal=sample_terminal)
    counts = sim.simulate_shots(circuit, shots=4096)
    assert_distribution(counts, {"00000": 0.5, "01010": 0.5}, 4096)
    assert counts == sim.simulate_shots(circuit, shots=4096)


@pytest.mark.parametrize("sample_terminal", [True, False])
def test_repeated_measurements_and_overwritten_classical_bits(sample_terminal):
    circuit = Circuit(3, [inst("h", qubit=0), inst("x", qubit=1), inst("measure", qubit=0, cbit=0),
                          inst("measure", qubit=0, cbit=2), inst("measure", qubit=1, cbit=0),
                          inst("measure", qubit=0, cbit=1)], cbits=3)
    counts = dqsim.StatevectorSimulator(seed=21, sample_terminal=sample_terminal).simulate_shots(circuit, 4096)
    assert_distribution(counts, {"001": 0.5, "111": 0.5}, 4096)


@pytest.mark.parametrize("sample_terminal", [True, False])
def test_mid_circuit_measurement_and_feedback(sample_terminal):
    circuit = Circuit(2, [inst("h", qubit=0), inst("measure", qubit=0, cbit=0), conditional(inst("x", qubit=1), 1),
                          inst("x", qubit=0), inst("measure", qubit=0, cbit=1), inst("measure", qubit=1, cbit=2)],
                      cbits=3)
    sim = dqsim.StatevectorSimulator(seed=15, sample_terminal=sample_terminal)
    counts = sim.simulate_shots(circuit, shots=4096)
    assert_distribution(counts, {"010": 0.5, "101": 0.5}, 4096)
    assert counts == sim.simulate_shots(circuit, shots=4096)


def test_reset_and_nested_classical_conditionals():
    circuit = Circuit(2, [inst("x", qubit=0), inst("measure", qubit=0, cbit=2), inst("reset", qubit=0),
                          inst("measure", qubit=0, cbit=0),
                          conditional(conditional(inst("x", qubit=1), 0), 2, base=1, size=2),
                          inst("measure", qubit=1, cbit=1)], cbits=3)
    result = simulate(circuit)
    assert result.classical_bits == {0: 0, 1: 1, 2: 1}
    np.testing.assert_allclose(result.statevector, [0, 0, 1, 0], atol=1e-14)
    assert dqsim.StatevectorSimulator(seed=22).simulate_shots(circuit, shots=64) == {"110": 64}


def test_reset_of_entangled_qubit_retains_other_qubits_mixture():
    circuit = Circuit(2, [inst("h", qubit=0), inst("cx", control=0, target=1), inst("reset", qubit=0),
                          inst("measure", qubit=0, cbit=0), inst("measure", qubit=1, cbit=1)], cbits=2)
    counts = dqsim.StatevectorSimulator(seed=24).simulate_shots(circuit, shots=4096)
    assert_distribution(counts, {"00": 0.5, "10": 0.5}, 4096)


@pytest.mark.parametrize("sample_terminal", [True, False])
def test_zero_shots_and_unwritten_classical_bits(sample_terminal):
    sim = dqsim.StatevectorSimulator(seed=30, sample_terminal=sample_terminal)
    for cbits, outcome in ((0, ""), (3, "000")):
        circuit = Circuit(1, [inst("h", qubit=0)], cbits=cbits)
        assert sim.simulate_shots(circuit, shots=0) == {}
        assert sim.simulate_shots(circuit, shots=13) == {outcome: 13}


def test_result_sampling_is_seeded():
    result = simulate(Circuit(2, [inst("h", qubit=0), inst("cx", control=0, target=1)]))
    counts = result.counts(shots=4096, seed=77)
    assert counts == result.counts(shots=4096, seed=77)
    assert_distribution(counts, {"00": 0.5, "11": 0.5}, 4096)
    assert result.counts(shots=0) == {}
    assert sum(result.counts(shots=100).values()) == 100


def test_invalid_probability_queries_fail():
    result = simulate(Circuit(2))
    for qubits in ([2], [0, 0], [1, 0, 1]):
        for method in (result.probabilities, result.counts):
            with pytest.raises(ValueError):
                method(qubits=qubits)


def test_worker_limit_preserves_seeded_trajectory_results():
    circuit = Circuit(2, [inst("h", qubit=0), inst("measure", qubit=0, cbit=0), conditional(inst("x", qubit=1), 1),
                          inst("measure", qubit=1, cbit=1)], cbits=2)
    counts = [dqsim.StatevectorSimulator(seed=38, max_memory_mb=1, max_parallel_shots=w,
                                         sample_terminal=False).simulate_shots(circuit, shots=257)
              for w in (1, 2, 4)]
    assert counts[0] == counts[1] == counts[2]


def test_wrapper_forwards_statevector_options():
    circuit = Circuit(1, [inst("x", qubit=0), inst("measure", qubit=0, cbit=0)], cbits=1)
    options = {"seed": 11, "max_memory_mb": 1, "max_parallel_shots": 1, "sample_terminal": False}
    result = dqsim.simulate_monolithic(circuit, **options)
    np.testing.assert_allclose(result.statevector, [0, 1], atol=1e-14)
    assert dqsim.simulate_monolithic_shots(circuit, shots=19, **options) == {"1": 19}


def test_profiles_confirm_terminal_sampling_and_dynamic_worker_budget(tmp_path, monkeypatch):
    terminal = Circuit(2, [inst("h", qubit=0), inst("h", qubit=1), inst("measure", qubit=0, cbit=0),
                           inst("measure", qubit=1, cbit=1)], cbits=2)
    dynamic = Circuit(15, [inst("h", qubit=0), inst("measure", qubit=0, cbit=0), conditional(inst("x", qubit=14), 1),
                           inst("measure", qubit=14, cbit=1)], cbits=2)
    monkeypatch.chdir(tmp_path)
    sim = dqsim.StatevectorSimulator(seed=12, max_memory_mb=1)
    counts = sim.simulate_shots(terminal, shots=4096, profile=True)
    assert_distribution(counts, {"00": .25, "01": .25, "10": .25, "11": .25}, 4096)
    data = json.loads(next(Path("dqsim_profiles").glob("*.json")).read_text())
    assert data["execution_strategy"] == "terminal_sampling"
    assert data["shot_times"] == []
    assert data["working_bytes"] <= 1024 ** 2
    sim.simulate_shots(dynamic, shots=16, profile=True)
    data = json.loads(sorted(Path("dqsim_profiles").glob("*.json"))[-1].read_text())
    assert data["execution_strategy"] == "trajectories"
    assert data["parallel_shots"] == 1
    assert len(data["shot_times"]) == 16
    assert data["working_bytes"] <= 1024 ** 2


def test_simulation_profile_counts_kernel_categories():
    circ = Circuit(3, [inst("h", qubit=0), inst("cx", control=0, target=1), inst("rzz", a=1, b=2, theta=0.3),
                       inst("measure", qubit=2, cbit=0), conditional(inst("x", qubit=0), 5, 0, 3)], cbits=3)
    res = dqsim.StatevectorSimulator(seed=4, profile=True).simulate(circ)
    p = res.profile
    assert p.apply_one_qubit_calls == 1 and p.apply_n_qubit_calls == 2 and p.measure_qubit_calls == 1
    assert p.total_time >= p.apply_one_qubit_time + p.apply_n_qubit_time >= 0
    assert "SimulationProfile" in repr(p)
    assert simulate(circ).profile is None


def test_budget_working_bytes_and_fallback_match_reference_rule(tmp_path, monkeypatch):
    # shots.rs:118-185: a terminal CDF that does not fit the budget falls
    # back to trajectories with budget-limited workers. 17 qubits complex64:
    # state 1 MiB, CDF 1 MiB, budget 2 MiB -> trajectories, one worker.
    monkeypatch.chdir(tmp_path)
    n = 17
    circ = Circuit(n, [inst("h", qubit=q) for q in range(n)] + [inst("measure", qubit=q, cbit=q) for q in range(n)],
                   cbits=n)
    counts = dqsim.StatevectorSimulator(seed=5, max_memory_mb=2, precision="complex64").simulate_shots(
        circ, shots=10, profile=True)
    data = json.loads(next(Path("dqsim_profiles").glob("*.json")).read_text())
    ref, info = oracle.run_shots(circ.data, 10, 5, budget_mb=2, prec=8)
    assert data["execution_strategy"] == "trajectories" and data["parallel_shots"] == 1
    for key in ("execution_strategy", "parallel_shots", "working_bytes", "deterministic_prefix_ops"):
        assert data[key] == info[key], key
    assert sum(counts.values()) == 10 and sum(ref.values()) == 10


def test_batched_kernel_launch_limits_are_readable():
    # No launch: the device LDS limits and the single-launch kernel's
    # attributes at its largest LDS request (12 qubits: 32 KiB of state plus
    # 2 KiB of chunk sums, within the default 64 KiB), and that the kernel has
    # no private stack (every helper inlined).
    from dqsim import batch
    info = batch.device_info(256)
    print(info, flush=True)
    assert info["device_count"] >= 1 and info["max_lds_per_block"] >= 65536
    assert info["lds_qubits"] == 12 and info["set_attribute_rc"] == 0 and info["kernel_max_threads"] >= 256
    assert info["lds_per_cu"] >= info["max_lds_per_block"] and info["kernel_private_bytes"] == 0
    assert info["occupancy_rc"] == 0 and info["blocks_per_cu"] >= 4


def test_fp32_matrix_core_microbenchmark_runs():
    # The register-only MFMA loop measures the FP32 matrix-core ceiling the
    # bench reports beside the 157.3 TFLOP/s datasheet figure.
    from dqsim import _core
    tflops, ms = _core.mfma_fp32_peak(1)
    print("mfma fp32", tflops, ms, flush=True)
    assert ms > 0 and 20.0 < tflops < 200.0


_TORCH_THEN_DQSIM = """
import sys
sys.path[:0] = {paths!r}
import torch
x = torch.ones(1 << 20, device="cuda") * 2
torch.cuda.synchronize()
import dqsim
from dqsim.workloads import Circuit
c = Circuit(3, [{{"kind": "h", "qubit": 0}}, {{"kind": "cx", "control": 0, "target": 2}}])
sv = dqsim.StatevectorSimulator(seed=1).simulate(c).statevector
assert abs(abs(sv[0]) ** 2 - 0.5) < 1e-12 and abs(abs(sv[5]) ** 2 - 0.5) < 1e-12
assert float(x.sum()) == 2 * (1 << 20)
print("ok")
"""

_DQSIM_THEN_TORCH = """
import sys
sys.path[:0] = {paths!r}
import dqsim
from dqsim.workloads import Circuit
c = Circuit(3, [{{"kind": "h", "qubit": 0}}, {{"kind": "cx", "control": 0, "target": 2}}])
sv = dqsim.StatevectorSimulator(seed=1).simulate(c).statevector
import torch
x = torch.ones(1 << 20, device="cuda") * 2
torch.cuda.synchronize()
assert float(x.sum()) == 2 * (1 << 20)
sv2 = dqsim.StatevectorSimulator(seed=1).simulate(c).statevector
assert (sv == sv2).all()
print("ok")
"""


@pytest.mark.parametrize("script", [_TORCH_THEN_DQSIM, _DQSIM_THEN_TORCH], ids=["torch-first", "dqsim-first"])
def test_one_hip_runtime_shared_with_torch(script):
    # PyTorch and the engine in one process, in either import order.
    import subprocess
    import sys

    pkg = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                       "classical-quantum-sim-optimizer_amd")
    out = subprocess.run([sys.executable, "-c", script.format(paths=[pkg])], capture_output=True, text=True,
                         timeout=300)
    assert out.returncode == 0 and out.stdout.strip().endswith("ok"), out.stderr[-2000:]
