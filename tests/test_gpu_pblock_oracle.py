"""GPU P-block backend against the CPU P-block restatement (oracle/pblock.py)
on seeded random dynamic circuits: the same Philox draws give the same
classical bits, block structure, integer statistics and shot counts; states
agree to 1e-12 (the separation test's sums are pairwise on the GPU and
sequential in the oracle, so split factors may differ in the last bits)."""

import numpy as np
import pytest

from dqsim import PBlockSimulator
from dqsim._pblock import last_shots_stats
from dqsim.workloads import Circuit
from oracle import pblock as opb
from test_gpu_parity import standard_gate_cases
from test_pblock_host import conditional, inst

pytestmark = pytest.mark.gpu

INT_STATS = ("merge_calls", "merge_output_amplitudes", "split_calls", "separation_checks",
             "skipped_controlled_gates", "peak_block_qubits", "peak_live_amplitudes", "peak_working_bytes")


def dynamic_circuit(n, steps, seed):
    """Standard gates on random qubits with mid-circuit measurements,
    resets and gates conditioned on measured bits."""
    rng = np.random.default_rng(seed)
    cases = standard_gate_cases()
    ops, measured = [], []
    for _ in range(steps):
        u = rng.random()
        if u < 0.08:
            q = int(rng.integers(n))
            ops.append(inst("measure", qubit=q, cbit=q))
            measured.append(q)
        elif u < 0.12:
            ops.append(inst("reset", qubit=int(rng.integers(n))))
        elif u < 0.2 and measured:
            c = measured[int(rng.integers(len(measured)))]
            body = inst("x", qubit=int(rng.integers(n))) if rng.random() < 0.5 else \
                inst("cx", control=int(rng.integers(n // 2)), target=n // 2 + int(rng.integers(n - n // 2)))
            ops.append(conditional(body, int(rng.integers(2)), base=c))
        else:
            kind, names, params = cases[int(rng.integers(len(cases)))]
            qs = [int(q) for q in rng.choice(n, len(names), replace=False)]
            ops.append(inst(kind, **dict(zip(names, qs)), **params))
    return Circuit(n, ops, n)


@pytest.mark.parametrize("split", [False, True])
def test_simulate_matches_oracle(split):
    for trial in range(10):
        c = dynamic_circuit(8, 40, 500 + trial)
        seed = 77 + trial
        print(f"[pblock-oracle] simulate trial={trial} split={split}", flush=True)
        g = PBlockSimulator(seed=seed, split_separable=split).simulate(c)
        o = opb.simulate(c.data, seed, split_separable=split)
        assert g.classical_bits == o.classical_bits
        assert g.block_qubits == o.block_qubits
        gd, od = g.diagnostics, o.diagnostics
        for k in INT_STATS:
            assert gd[k] == od[k], (k, gd[k], od[k])
        np.testing.assert_allclose(g.statevector, o.statevector, atol=1e-12)
        gp, op_ = g.probabilities([7, 2, 4]), o.probabilities([7, 2, 4])
        for k in set(gp) | set(op_):
            assert abs(gp.get(k, 0.0) - op_.get(k, 0.0)) < 1e-12


@pytest.mark.parametrize("terminal", [True, False])
def test_shot_counts_match_oracle(terminal):
    # Blocks stay within 1024 amplitudes, where the GPU's grouped CDF is the
    # sequential sum, so counts are equal draw for draw.
    for trial in range(4):
        ops = dynamic_circuit(8, 30, 900 + trial).data["instructions"]
        ops = [op for op in ops if op["kind"] not in ("measure", "reset", "conditional")]
        if not terminal:
            ops = ops[:15] + [inst("measure", qubit=3, cbit=3), conditional(inst("h", qubit=5), 1, base=3)] + ops[15:]
        c = Circuit(8, ops + [inst("measure", qubit=q, cbit=q) for q in range(8)], 8)
        print(f"[pblock-oracle] shots trial={trial} terminal={terminal}", flush=True)
        g = PBlockSimulator(seed=11 + trial).simulate_shots(c, 300)
        o, strategy = opb.simulate_shots(c.data, 300, 11 + trial)
        assert strategy == ("terminal_sampling" if terminal else "prefix_trajectories")
        assert g == o


def test_result_counts_match_oracle():
    c = dynamic_circuit(8, 30, 1234)
    g = PBlockSimulator(seed=5).simulate(c)
    o = opb.simulate(c.data, 5)
    assert g.counts(500, qubits=[0, 5, 6], seed=19) == o.counts(500, qubits=[0, 5, 6], seed=19)
    assert g.counts(200, seed=3) == o.counts(200, seed=3)


@pytest.mark.parametrize("trial", range(4))
def test_lockstep_trajectories_match_oracle_and_lone_runs(trial, tmp_path, monkeypatch):
    # Dynamic circuits whose shots part ways in structure (controls on
    # measured wires, conditional gates that merge blocks, resets): the
    # lockstep batches give the oracle's counts draw for draw, and so do the
    # lone per-shot runs (a profiled call runs every shot on its own pool).
    # Shots whose decisions differ leave the batch and run alone; most stay.
    c = dynamic_circuit(6, 40, 3100 + trial)
    c = Circuit(6, c.data["instructions"] + [inst("measure", qubit=q, cbit=q) for q in range(6)], 6)
    seed = 41 + trial
    print(f"[pblock-oracle] lockstep trial={trial}", flush=True)
    g = PBlockSimulator(seed=seed, sample_terminal=False).simulate_shots(c, 120)
    st = last_shots_stats()
    print(f"[pblock-oracle] lockstep trial={trial} {st}", flush=True)
    assert st["lockstep_shots"] + st["lone_shots"] == 120 and st["lockstep_shots"] >= 60
    o, _ = opb.simulate_shots(c.data, 120, seed, sample_terminal=False)
    assert g == o
    small = PBlockSimulator(seed=seed, sample_terminal=False, max_memory_mb=16384, max_parallel_shots=1)
    assert small.simulate_shots(c, 120) == g
    monkeypatch.chdir(tmp_path)
    assert PBlockSimulator(seed=seed, sample_terminal=False).simulate_shots(c, 120, profile=True) == g


def entangled_trajectory_circuit(n, seed):
    """A GHZ chain and a U3 layer over n wires (one block of n qubits), then
    every wire measured: with sample_terminal=False every shot runs the same
    structure, so all of them can stay in lockstep."""
    rng = np.random.default_rng(seed)
    ops = [inst("h", qubit=0)] + [inst("cx", control=q, target=q + 1) for q in range(n - 1)]
    ops += [inst("u3", qubit=q, theta=float(t), phi=float(p), lam=float(l))
            for q, (t, p, l) in enumerate(rng.uniform(0, 3, size=(n, 3)))]
    return Circuit(n, ops + [inst("measure", qubit=q, cbit=q) for q in range(n)], n)


def test_uniform_trajectories_stay_in_lockstep():
    c = entangled_trajectory_circuit(8, 3)
    g = PBlockSimulator(seed=7, sample_terminal=False).simulate_shots(c, 200)
    st = last_shots_stats()
    assert st == {**st, "lockstep_shots": 200, "lone_shots": 0, "batch_size": 200}, st
    o, _ = opb.simulate_shots(c.data, 200, 7, sample_terminal=False)
    assert g == o


def test_lockstep_batches_stay_within_the_budget():
    # 14 wires in one block: a slot needs 2^14 x 16 B = 256 KiB (plus the
    # reference's 256 B per wire), so a 1 MB budget holds at most 3 slots at
    # once; the batch shrinks to fit and the counts are unchanged.
    n = 14
    c = entangled_trajectory_circuit(n, 5)
    roomy = PBlockSimulator(seed=9, sample_terminal=False, max_memory_mb=4096).simulate_shots(c, 48)
    big = last_shots_stats()
    tight = PBlockSimulator(seed=9, sample_terminal=False, max_memory_mb=1).simulate_shots(c, 48)
    small = last_shots_stats()
    print(f"[pblock-budget] roomy {big} tight {small}", flush=True)
    assert tight == roomy
    assert big["batch_size"] == 48 and big["lockstep_shots"] == 48
    assert small["peak_batch_bytes"] <= 1 << 20 and 1 <= small["batch_size"] <= 3
    assert small["lockstep_shots"] + small["lone_shots"] == 48
    o, _ = opb.simulate_shots(c.data, 48, 9, sample_terminal=False, memory_mb=4096)
    assert roomy == o
