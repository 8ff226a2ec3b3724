"""LDS tile sweeps (complex64 + fusion): fused blocks grouped into sweeps over
2^L-amplitude tiles held in LDS, against the FP64 oracle (max amplitude error
<= 1e-5) and against the one-launch-per-block fused plan."""

import numpy as np
import pytest

import dqsim
from dqsim import workloads
from dqsim.workloads import Circuit
from oracle import oracle

pytestmark = pytest.mark.gpu


def mixed_circuit(n, depth, seed):
    """U3 layers plus seeded multi-qubit gates of every lowering kind."""
    return workloads.mixed(n, depth, seed)


CASES = [
    ("su2", 16, 10, 2), ("su2", 16, 10, 3), ("su2", 16, 10, 4), ("su2", 16, 10, 5),
    ("su2", 17, 11, 4), ("su2", 18, 12, 4), ("su2", 20, 12, 5), ("su2", 20, 12, 3),
    ("qft", 16, 12, 4), ("mixed", 17, 12, 3), ("mixed", 18, 11, 5), ("mixed", 16, 10, 4),
    ("su2", 24, 12, 3), ("mixed", 22, 12, 3),
    # 13-qubit tiles (out-of-place sweeps only: 512 threads, the whole 64 KiB of LDS)
    ("su2", 17, 13, 3), ("su2", 20, 13, 4), ("su2", 22, 13, 2), ("qft", 17, 13, 4), ("mixed", 18, 13, 3),
    ("mixed", 19, 13, 5),
]


def build(kind, n):
    if kind == "su2":
        return workloads.random_su2(n=n, depth=12, seed=n)
    if kind == "qft":
        return workloads.qft(n, seed=3, measure=False)[0]
    if kind == "brickwork":
        return workloads.brickwork(n=n, layers=12, seed=n, measure=False)
    return mixed_circuit(n, 6, seed=n)


@pytest.mark.parametrize("kind,n,L,k", CASES, ids=lambda v: str(v))
def test_tile_sweeps_match_oracle(kind, n, L, k):
    print(f"[tiles] {kind} n={n} L={L} k={k}", flush=True)
    circ = build(kind, n)
    plan = dqsim._core.plan_json(circ, precision="complex64", fusion="greedy", max_fused_qubits=k, tile_qubits=L)
    assert any(x["kind"] == "tile" for x in plan["launches"])
    want = oracle.simulate(circ.data, 1)[0]
    sim = dqsim.StatevectorSimulator(seed=1, precision="complex64", fusion="greedy", max_fused_qubits=k,
                                     tile_qubits=L)
    got = sim.simulate(circ).statevector
    assert np.abs(got - want).max() < 1e-5
    flat = dqsim.StatevectorSimulator(seed=1, precision="complex64", fusion="greedy", max_fused_qubits=k,
                                      tile_qubits=0)
    assert np.abs(flat.simulate(circ).statevector - got).max() < 1e-5


def test_tile_sweeps_with_trajectory_batches():
    # Mid-circuit measurements split the circuit into fused runs; every slot
    # of a trajectory batch is swept (grid.y = slot).
    n = 16
    ins = workloads.random_su2(n=n, depth=6, seed=2).data["instructions"]
    ins = ins + [{"kind": "measure", "qubit": 3, "cbit": 0}] + \
        workloads.random_su2(n=n, depth=4, seed=3).data["instructions"] + \
        [{"kind": "measure", "qubit": q, "cbit": q} for q in range(n)]
    circ = Circuit(n, ins, cbits=n)
    res = {}
    for tile in (0, 12):
        sim = dqsim.StatevectorSimulator(seed=9, precision="complex64", fusion="greedy", max_fused_qubits=4,
                                         tile_qubits=tile, sample_terminal=False, max_parallel_shots=64)
        res[tile] = sim.simulate_shots(circ, shots=64)
        assert sum(res[tile].values()) == 64
    # Same Philox draws on states that agree to ~1e-6: outcomes may differ only
    # where a draw falls within that distance of a CDF boundary.
    common = sum(min(res[0].get(b, 0), res[12].get(b, 0)) for b in set(res[0]) | set(res[12]))
    assert common >= 60


@pytest.mark.parametrize("L,k", [(12, 3), (12, 4), (11, 3), (10, 2)])
def test_relabelled_sweeps_match_oracle_and_fixed_low(L, k):
    # Sweeps that store their tile with bit positions exchanged (lane swaps
    # and permuted store positions) and the restore passes that end each run:
    # same state as sweeps that pin qubits 0..3, within the FP32 tolerance.
    n = 20
    circ = workloads.random_su2(n=n, depth=24, seed=40 + L + k)
    print(f"[tiles] relabel su2 n={n} L={L} k={k}", flush=True)
    prev = dqsim._core.set_sweep_relabel(True)
    try:
        plan = dqsim._core.plan_json(circ, precision="complex64", fusion="greedy", max_fused_qubits=k,
                                     tile_qubits=L)
    finally:
        dqsim._core.set_sweep_relabel(prev)
    tiles = [x for x in plan["launches"] if x["kind"] == "tile"]
    assert any(len(x["swaps"]) == 2 for x in tiles) and any(x["gates"] == 0 for x in tiles)
    want = oracle.simulate(circ.data, 1)[0]
    states = {}
    for on in (True, False):
        prev = dqsim._core.set_sweep_relabel(on)
        try:
            sim = dqsim.StatevectorSimulator(seed=1, precision="complex64", fusion="greedy", max_fused_qubits=k,
                                             tile_qubits=L)
            states[on] = sim.simulate(circ).statevector
        finally:
            dqsim._core.set_sweep_relabel(prev)
        assert np.abs(states[on] - want).max() < 1e-5
    assert np.abs(states[True] - states[False]).max() < 1e-5


@pytest.mark.parametrize("mode", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("kind,n,L,k", [("su2", 20, 12, 3), ("su2", 18, 11, 4), ("mixed", 17, 12, 3),
                                        ("qft", 16, 10, 2)], ids=lambda v: str(v))
def test_sweep_modes_match_oracle(mode, kind, n, L, k):
    # In-place sweeps (0) and out-of-place sweeps (1-4: persistent or
    # one-shot workgroups of 256 or 512 threads; 512 only for 11-12 qubit
    # tiles, 256 otherwise) all within the complex64 tolerance of the oracle,
    # with and without room for the second state in the budget.
    print(f"[sweep-mode] {mode} {kind} n={n} L={L} k={k}", flush=True)
    circ = build(kind, n)
    want = oracle.simulate(circ.data, 1)[0]
    old = dqsim._core.set_sweep_mode(mode)
    try:
        state_mb = (1 << n) * 8 >> 20
        for budget in (max(1, state_mb) + 1, 4 * state_mb + 64):
            sim = dqsim.StatevectorSimulator(seed=1, precision="complex64", fusion="greedy", max_fused_qubits=k,
                                             tile_qubits=L, max_memory_mb=budget)
            got = sim.simulate(circ).statevector
            assert np.abs(got - want).max() < 1e-5, budget
    finally:
        dqsim._core.set_sweep_mode(old)


@pytest.mark.parametrize("pinned", [False, True], ids=["free", "pinned"])
@pytest.mark.parametrize("kind,n,L,k", [("su2", 20, 12, 3), ("su2", 19, 11, 4), ("qft", 18, 12, 2),
                                        ("mixed", 18, 12, 3), ("brickwork", 20, 12, 3), ("brickwork", 21, 12, 4)],
                         ids=lambda v: str(v))
def test_tile_rules_match_oracle(pinned, kind, n, L, k):
    # Out-of-place sweeps under both tile rules (qubits 0..3 in every tile,
    # or only in a run's first and last tile); with room for one state only,
    # the same call sweeps in place.
    print(f"[tile-rule] pinned={pinned} {kind} n={n} L={L} k={k}", flush=True)
    circ = build(kind, n)
    want = oracle.simulate(circ.data, 1)[0]
    prev = dqsim._core.set_sweep_pinned(pinned)
    try:
        plan = dqsim._core.plan_json(circ, precision="complex64", fusion="greedy", max_fused_qubits=k,
                                     tile_qubits=L, out_of_place=True)
        xs = [x for x in plan["launches"] if x.get("out_of_place")]
        assert xs and all(len(x["local_qubits"]) == L for x in xs)
        state_mb = (1 << n) * 8 >> 20
        for budget in (max(1, state_mb) + 1, 4 * state_mb + 64):
            sim = dqsim.StatevectorSimulator(seed=1, precision="complex64", fusion="greedy", max_fused_qubits=k,
                                             tile_qubits=L, max_memory_mb=budget)
            got = sim.simulate(circ).statevector
            assert np.abs(got - want).max() < 1e-5, budget
    finally:
        dqsim._core.set_sweep_pinned(prev)


@pytest.mark.parametrize("mode,arith", [(2, 1), (2, 0), (1, 0), (1, 1), (3, 1)], ids=lambda v: str(v))
@pytest.mark.parametrize("kind,n,L,k", [("su2", 20, 12, 3), ("su2", 18, 11, 4), ("mixed", 17, 12, 3),
                                        ("qft", 16, 10, 2), ("brickwork", 20, 12, 5), ("mixed", 18, 10, 3)],
                         ids=lambda v: str(v))
def test_wave_runs_bit_identical(mode, arith, kind, n, L, k):
    # Blocks grouped into wave-owned runs (barriers only after a run) give
    # the same bits as a barrier after every block: each group goes through
    # the same operations, only the wave that computes it changes. Sweep
    # modes 1 and 2 (256-thread out-of-place workgroups) use the runs; mode
    # 3 (512 threads) keeps a barrier per block either way. Both states are
    # also within the complex64 tolerance of the oracle.
    print(f"[wave-runs] mode={mode} arith={arith} {kind} n={n} L={L} k={k}", flush=True)
    circ = build(kind, n)
    want = oracle.simulate(circ.data, 1)[0]
    old_mode, old_arith = dqsim._core.set_sweep_mode(mode), dqsim._core.set_block_arith(arith)
    states = {}
    try:
        for runs in (1, 0):
            prev = dqsim._core.set_wave_runs(runs)
            try:
                sim = dqsim.StatevectorSimulator(seed=1, precision="complex64", fusion="greedy",
                                                 max_fused_qubits=k, tile_qubits=L)
                states[runs] = sim.simulate(circ).statevector
            finally:
                dqsim._core.set_wave_runs(prev)
    finally:
        dqsim._core.set_sweep_mode(old_mode)
        dqsim._core.set_block_arith(old_arith)
    assert np.array_equal(states[1], states[0])
    assert np.abs(states[1] - want).max() < 1e-5
