"""Batched small circuits.

The stream-pool runner (one host thread, HIP stream and device arena per
worker) runs every circuit through simulate_shots' terminal-sampling path with
the circuit's own seed, so its counts must equal those of the per-circuit call
exactly, for any worker count; final states are pinned by the FP64 oracle.
Those tests set lds_qubits=0 so that every circuit takes the pool.

The single-launch kernel (one workgroup per circuit, state in LDS, one
hipGraph launch) runs complex64 circuits of up to 12 qubits that measure every
qubit: their states are pinned by the FP64 oracle and their outcomes, draw for
draw, by the oracle's restatement of its chunked sampler on those states.

The earlier persistent kernel is removed and its test skipped: it stalled (no
progress for 180 s) on its first GPU run and no cause has been shown (a CPU
emulation of its logic on that test's program stayed in bounds and
reproduced the oracle)."""

import numpy as np
import pytest

from dqsim import _core, batch, workloads
from dqsim.workloads import Circuit
from oracle import oracle

pytestmark = pytest.mark.gpu


def circuits(count, nmin, nmax, dmin=2, dmax=8, base=77):
    return [workloads.batch_circuit(i, base_seed=base, nmin=nmin, nmax=nmax, dmin=dmin, dmax=dmax)
            for i in range(count)]


def per_circuit_counts(c, shots, seed, k, tile, precision="complex64"):
    sim = _core.StatevectorSimulator(seed=seed, precision=precision, fusion="greedy", max_fused_qubits=k,
                                     tile_qubits=tile)
    return sim.simulate_shots(c, shots=shots)


@pytest.mark.parametrize("workers,k", [(1, 3), (4, 2), (8, 3), (3, 4)])
def test_batch_counts_equal_per_circuit_calls(workers, k):
    # 3..12 qubits (one launch per block) and 16..18 qubits (LDS tile sweeps).
    cs = circuits(10, 3, 12) + circuits(4, 16, 18, base=5)
    print("batch", workers, k, [c.num_qubits for c in cs], flush=True)
    res = batch.run_batch(cs, shots=300, base_seed=11, workers=workers, max_fused_qubits=k, autotune=False,
                          lds_qubits=0)
    assert res.lds_circuits == 0
    assert res.workers == workers and res.max_fused_qubits == k and res.tile_qubits == 12
    for i, c in enumerate(cs):
        got = res.counts(i)
        assert sum(got.values()) == 300
        assert got == per_circuit_counts(c, 300, 11 + i, k, 12), i


def test_batch_states_match_oracle_and_indices_match_counts():
    cs = circuits(8, 3, 12) + circuits(2, 16, 17, base=5)
    res = batch.run_batch(cs, shots=200, base_seed=3, workers=4, max_fused_qubits=3, autotune=False,
                          keep_states=True, lds_qubits=0)
    for i, c in enumerate(cs):
        n = c.num_qubits
        unitary = [x for x in c.data["instructions"] if x["kind"] != "measure"]
        want, _ = oracle.simulate(Circuit(n, unitary).data, 1)
        got = res.statevector(i)
        assert np.abs(got - want).max() < 1e-5, (i, n)
        # Every qubit is measured into the cbit of the same number: the outcome
        # index is the register value, so the counts are its histogram.
        idx = res.indices(i)
        hist = {}
        for x in idx:
            key = format(int(x), f"0{n}b")
            hist[key] = hist.get(key, 0) + 1
        assert hist == res.counts(i)


def test_batch_complex128_counts_equal_per_circuit_calls():
    cs = circuits(6, 4, 12, base=21) + circuits(2, 16, 16, base=8)
    res = batch.run_batch(cs, shots=128, base_seed=40, workers=2, max_fused_qubits=2, autotune=False,
                          precision="complex128", keep_states=True)
    for i, c in enumerate(cs):
        assert res.counts(i) == per_circuit_counts(c, 128, 40 + i, 2, 12, precision="complex128"), i
        unitary = [x for x in c.data["instructions"] if x["kind"] != "measure"]
        want, _ = oracle.simulate(Circuit(c.num_qubits, unitary).data, 1)
        assert np.abs(res.statevector(i) - want).max() < 1e-12, i


def test_batch_zero_shots_and_unmeasured_circuit():
    measured = circuits(2, 3, 6, base=2)
    bare = Circuit(3, [{"kind": "h", "qubit": 0}, {"kind": "cx", "control": 0, "target": 2}], 0)
    res = batch.run_batch(measured + [bare], shots=[0, 50, 40], base_seed=1, workers=2, max_fused_qubits=3,
                          autotune=False, lds_qubits=0)
    assert res.counts(0) == {}
    assert sum(res.counts(1).values()) == 50
    assert res.counts(2) == {"": 40}


@pytest.mark.parametrize("workers", [0, -1])
def test_batch_autotune_and_latency(workers):
    cs = circuits(40, 4, 14, dmin=4, dmax=12, base=9)
    res = batch.run_batch(cs, shots=64, base_seed=3, autotune=True, lds_qubits=0, workers=workers)
    print("tuned", workers, res.workers, res.max_fused_qubits, res.wall_ms, flush=True)
    assert 2 <= res.max_fused_qubits <= 4 and res.tile_qubits in (0, 10, 11, 12)
    if workers == 0:
        # 4 workers; 3 widths at the default tile, then 3 more tile shapes.
        assert res.workers == 4 and len(res.tuning) == 6
    else:
        # 9 (workers, width) pairs at the default tile, then 3 more tile shapes.
        assert res.workers in (2, 4, 8) and len(res.tuning) == 12
    assert all(ms > 0 for ms in res.tuning.values())
    assert f"workers={res.workers},k<={res.max_fused_qubits},tile={res.tile_qubits}" in res.tuning
    assert {int(k.rsplit("=", 1)[1]) for k in res.tuning} == {0, 10, 11, 12}
    assert res.wall_ms > 0 and len(res.latency_us) == 40 and (res.latency_us > 0).all()
    for i, c in enumerate(cs):
        assert res.counts(i) == per_circuit_counts(c, 64, 3 + i, res.max_fused_qubits, res.tile_qubits), i


@pytest.mark.parametrize("k", [2, 3, 4])
def test_single_launch_states_and_samples_match_oracle(k):
    # 3..12 qubits in the single launch, 16..17 on the pool, both in one call.
    cs = circuits(14, 3, 12, base=31) + circuits(2, 16, 17, base=5)
    print("[batch-lds]", k, [c.num_qubits for c in cs], flush=True)
    res = batch.run_batch(cs, shots=300, base_seed=11, workers=2, max_fused_qubits=k, autotune=False,
                          keep_states=True)
    assert res.lds_circuits == 14 and res.lds_ms > 0
    for i, c in enumerate(cs):
        n = c.num_qubits
        unitary = [x for x in c.data["instructions"] if x["kind"] != "measure"]
        want, _ = oracle.simulate(Circuit(n, unitary).data, 1)
        got = res.statevector(i)
        assert np.abs(got - want).max() < 1e-5, (i, n)
        if n <= 12:
            idx = oracle.batch_sample(got, 300, 11 + i, 256)
            assert np.array_equal(res.indices(i).astype(np.uint64), idx), i
            assert res.latency_us[i] > 0
        else:
            assert res.counts(i) == per_circuit_counts(c, 300, 11 + i, k, 12), i


def test_single_launch_edge_cases():
    # Zero shots, one qubit, a circuit that measures only some qubits (pool),
    # and a different shot count per circuit.
    one = Circuit(1, [{"kind": "h", "qubit": 0}, {"kind": "measure", "qubit": 0, "cbit": 0}], 1)
    part = Circuit(4, [{"kind": "h", "qubit": 0}, {"kind": "cx", "control": 0, "target": 3},
                       {"kind": "measure", "qubit": 3, "cbit": 0}], 1)
    cs = [one, part] + circuits(3, 5, 9, base=44)
    shots = [257, 40, 0, 1, 1000]
    res = batch.run_batch(cs, shots=shots, base_seed=2, workers=1, max_fused_qubits=3, autotune=False,
                          keep_states=True)
    assert res.lds_circuits == 4
    assert sum(res.counts(0).values()) == 257 and set(res.counts(0)) <= {"0", "1"}
    assert res.counts(1) == per_circuit_counts(part, 40, 3, 3, 12)
    assert res.counts(2) == {}
    for i in (0, 3, 4):
        idx = oracle.batch_sample(res.statevector(i), shots[i], 2 + i, 256)
        assert np.array_equal(res.indices(i).astype(np.uint64), idx), i


@pytest.mark.skip(reason="the persistent batched kernel stalled (no progress for 180 s) on its first GPU run in "
                         "test_batch_states_and_samples_match_oracle[256-2]; no cause has been shown, so it raises "
                         "and this case stays off the GPU")
@pytest.mark.parametrize("block,k", [(256, 2), (512, 3), (1024, 4), (256, 4)])
def test_batch_states_and_samples_match_oracle(block, k):
    # The persistent kernel with `block` threads per workgroup and fusion
    # width k; its chunked sampler is restated by oracle.batch_sample.
    cs = circuits(12, 3, 12) + circuits(3, 15, 16, base=5)
    res = batch.run_batch(cs, shots=300, base_seed=11, max_fused_qubits=k, autotune=False, keep_states=True,
                          persistent=True)
    for i, c in enumerate(cs):
        n = c.num_qubits
        unitary = [x for x in c.data["instructions"] if x["kind"] != "measure"]
        want, _ = oracle.simulate(Circuit(n, unitary).data, 1)
        got = res.statevector(i)
        assert np.abs(got - want).max() < 1e-5, (i, n)
        idx = oracle.batch_sample(got, 300, 11 + i, block)
        assert np.array_equal(res.indices(i).astype(np.uint64), idx), i
