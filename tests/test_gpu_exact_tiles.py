"""Exact complex128 tile sweeps (the default for complex128 without fusion
at 15+ qubits): consecutive compiled ops applied in LDS with the per-op
kernels' arithmetic, in program order, out of place between two state
buffers when the budget holds both (the default budget does at these sizes),
else in place. Results must equal the oracle bit for bit, like one launch
per op, in both placements."""

import numpy as np
import pytest

import dqsim
from dqsim import workloads
from dqsim.workloads import Circuit
from oracle import oracle
from test_gpu_parity import entangled_prefix, inst, standard_gate_cases
from test_gpu_tiles import mixed_circuit

pytestmark = pytest.mark.gpu


def exact_state(insts, n, tile=12, seed=60):
    """The state from out-of-place sweeps, checked equal (bit for bit up to
    the sign of zeros) to the state from in-place sweeps."""
    sim = dqsim.StatevectorSimulator(seed=seed, tile_qubits=tile)
    got = sim.simulate(Circuit(n, insts)).statevector
    prev = dqsim._core.set_sweep_mode(0)
    try:
        in_place = sim.simulate(Circuit(n, insts)).statevector
    finally:
        dqsim._core.set_sweep_mode(prev)
    assert np.array_equal(got, in_place)
    return got


def has_exact_tiles(insts, n, cbits=0):
    plan = dqsim._core.plan_json(Circuit(n, insts, cbits=cbits))
    xplan = dqsim._core.plan_json(Circuit(n, insts, cbits=cbits), out_of_place=True)
    return (any(x["kind"] == "tile" for x in plan["launches"])
            and any(x.get("out_of_place") and x["exact"] for x in xplan["launches"]))


def test_every_gate_kind_bit_exact_in_sweeps():
    # All standard gates at several placements, many per sweep, after an
    # entangling prefix, at 16 qubits (12 ... 15 appear as tile qubits).
    n = 16
    rng = np.random.default_rng(11)
    insts = entangled_prefix(n)
    for kind, names, params in standard_gate_cases():
        for _ in range(3):
            qubits = rng.choice(n, size=len(names), replace=False).tolist()
            insts.append(inst(kind, **dict(zip(names, qubits)), **params))
    assert has_exact_tiles(insts, n)
    want = oracle.simulate(Circuit(n, insts).data, 60)[0]
    assert np.array_equal(exact_state(insts, n), want)
    assert np.array_equal(exact_state(insts, n, tile=10), want)
    assert np.array_equal(exact_state(insts, n, tile=0), want)


@pytest.mark.parametrize("n,seed", [(16, 1), (17, 2), (18, 3)])
def test_mixed_circuits_bit_exact(n, seed):
    insts = mixed_circuit(n, 5, seed).data["instructions"]
    assert has_exact_tiles(insts, n)
    want = oracle.simulate(Circuit(n, insts).data, 60)[0]
    assert np.array_equal(exact_state(insts, n), want)


def test_headline_family_at_24_qubits_bit_exact():
    # The benchmark's circuit family (Haar SU(2) layers + random CZ/CX
    # matchings) at 24 qubits: 2^13 tiles per sweep, every sweep kind of the
    # 32-qubit run (register chains, dense pairs, diagonals, swaps of pairs).
    circ = workloads.random_su2(n=24, depth=10, seed=77)
    insts = circ.data["instructions"]
    print("[exact] random_su2 24q depth 10", flush=True)
    assert has_exact_tiles(insts, 24)
    want = oracle.simulate(circ.data, 60)[0]
    assert np.array_equal(exact_state(insts, 24), want)


def test_qft_and_brickwork_bit_exact():
    for circ in (workloads.qft(17, seed=4, measure=False)[0], workloads.brickwork(n=18, layers=8, measure=False)):
        n = circ.num_qubits
        insts = circ.data["instructions"]
        assert has_exact_tiles(insts, n)
        want = oracle.simulate(circ.data, 60)[0]
        assert np.array_equal(exact_state(insts, n), want)


@pytest.mark.parametrize("terminal", [True, False])
def test_shots_with_mid_circuit_measurement_match_oracle(terminal):
    # Unitary runs between measurements are swept; trajectory batches sweep
    # every slot. Exact states give the oracle's counts exactly.
    n = 16
    ins = workloads.random_su2(n=n, depth=3, seed=5).data["instructions"]
    ins = ins + [inst("measure", qubit=2, cbit=0)] + workloads.random_su2(n=n, depth=2, seed=6).data["instructions"]
    ins = ins + [inst("measure", qubit=q, cbit=q) for q in range(n)]
    circ = Circuit(n, ins, cbits=n)
    assert has_exact_tiles(ins, n, cbits=n)
    sim = dqsim.StatevectorSimulator(seed=21, sample_terminal=terminal, max_parallel_shots=16,
                                     max_memory_mb=4096)
    got = sim.simulate_shots(circ, shots=40)
    ref, _ = oracle.run_shots(circ.data, 40, 21, budget_mb=4096, sample_terminal=terminal, max_parallel_shots=16)
    assert got == ref


def test_relabelled_exact_sweeps_bit_identical():
    # Exact sweeps that relabel bit positions (lane swaps, permuted store
    # positions, restore passes) change where amplitudes are stored, never
    # their arithmetic: the state equals the oracle's and the fixed-low
    # sweeps' bit for bit, for 11- and 10-qubit tiles.
    n = 20
    circ = workloads.random_su2(n=n, depth=30, seed=52)
    insts = circ.data["instructions"]
    print("[exact] relabel random_su2 20q depth 30", flush=True)
    prev = dqsim._core.set_sweep_relabel(True)
    try:
        plan = dqsim._core.plan_json(circ)
    finally:
        dqsim._core.set_sweep_relabel(prev)
    tiles = [x for x in plan["launches"] if x["kind"] == "tile"]
    assert any(len(x["swaps"]) == 2 for x in tiles) and any(x["gates"] == 0 for x in tiles)
    want = oracle.simulate(circ.data, 60)[0]
    for tile in (11, 10):
        for on in (True, False):
            prev = dqsim._core.set_sweep_relabel(on)
            try:
                got = exact_state(insts, n, tile=tile)
            finally:
                dqsim._core.set_sweep_relabel(prev)
            assert np.array_equal(got, want), (tile, on)
