"""complex64 sampling against the oracle, draw for draw.

A complex64 state is upcast exactly by ``SimulationResult.statevector``; the
engine squares the upcast FP64 values, sums them in the canonical pairwise
tree and draws from the same Philox streams as the oracle, so the oracle
sampling that upcast state must reproduce every shot:
  - terminal sampling (shots.rs:86-174) over the prefix state the engine
    evolved (simulate() of the unitary prefix runs the same launch plan);
  - trajectories (shots.rs:178-259) whose remaining instructions are
    measurements: each collapse scales in FP64 and rounds the kept amplitudes
    to FP32, which ``oracle.measure(..., f32=True)`` restates;
  - SimulationResult.counts(qubits, seed) (simulator.rs:164-176,
    engine.rs:457-507), including the blocked path above 2^24 outcomes;
  - the batched runner's sampled indices.
"""

import numpy as np
import pytest

import dqsim
from dqsim import batch, workloads
from dqsim.workloads import Circuit
from oracle import oracle

pytestmark = pytest.mark.gpu

C64 = dict(precision="complex64", fusion="greedy")


def unitary_part(circ):
    return [i for i in circ.data["instructions"] if i["kind"] != "measure"]


def measured(circ):
    return [(i["qubit"], i["cbit"]) for i in circ.data["instructions"] if i["kind"] == "measure"]


def terminal_counts_from_state(state, circ, shots, seed):
    """Terminal sampling of `circ`'s measurements over `state` by the oracle
    (measured qubits highest first; cbit -> qubit, last write wins)."""
    n, nc = oracle.circuit_dims(circ.data)
    cmap = {}
    for q, c in measured(circ):
        cmap[c] = q
    qubits = sorted(set(cmap.values()), reverse=True)
    idx = oracle.sample_marginal(state, qubits, shots, seed)
    out = {}
    for x in idx.tolist():
        bits = ["0"] * nc
        for c, q in cmap.items():
            bits[nc - 1 - c] = str((x >> (len(qubits) - 1 - qubits.index(q))) & 1)
        key = "".join(bits)
        out[key] = out.get(key, 0) + 1
    return out


@pytest.mark.parametrize("opts", [dict(max_fused_qubits=3), dict(max_fused_qubits=4, tile_qubits=0),
                                  dict(fusion="none")], ids=["tiles_k3", "blocks_k4", "unfused"])
def test_brickwork_20q_terminal_counts(opts):
    circ = workloads.brickwork(n=20, layers=20, seed=42)
    kw = {**C64, **opts}
    print(f"[c64] brickwork terminal {kw}", flush=True)
    sim = dqsim.StatevectorSimulator(seed=42, **kw)
    state = sim.simulate(Circuit(20, unitary_part(circ))).statevector
    assert np.array_equal(state, state.astype(np.complex64).astype(np.complex128))
    got = sim.simulate_shots(circ, shots=1000)
    assert got == terminal_counts_from_state(state, circ, 1000, 42)


@pytest.mark.parametrize("mps", [None, 7])
def test_brickwork_20q_trajectories(mps):
    circ = workloads.brickwork(n=20, layers=20, seed=42)
    print(f"[c64] brickwork trajectories mps={mps}", flush=True)
    sim = dqsim.StatevectorSimulator(seed=42, sample_terminal=False, max_parallel_shots=mps, **C64)
    prefix = dqsim.StatevectorSimulator(seed=42, **C64).simulate(Circuit(20, unitary_part(circ))).statevector
    got = sim.simulate_shots(circ, shots=48)
    want = oracle.terminal_trajectories(prefix, measured(circ), 20, 48, 42, f32=True)
    assert got == want


@pytest.mark.parametrize("mps", [None, 1, 5])
def test_repeated_measurements_trajectories_14q(mps):
    # Tile-swept prefix, then measurements that revisit qubits and overwrite
    # cbits; each collapse changes the next measurement's probabilities.
    n = 14
    unitary = workloads.random_su2(n=n, depth=6, seed=8).data["instructions"]
    meas = [(3, 0), (13, 1), (3, 2), (0, 1), (7, 3), (12, 4), (13, 0)]
    circ = Circuit(n, unitary + [{"kind": "measure", "qubit": q, "cbit": c} for q, c in meas], 5)
    print(f"[c64] 14q trajectories mps={mps}", flush=True)
    prefix = dqsim.StatevectorSimulator(seed=3, max_fused_qubits=3, **C64).simulate(Circuit(n, unitary)).statevector
    for terminal in (False, True):
        sim = dqsim.StatevectorSimulator(seed=3, max_fused_qubits=3, sample_terminal=terminal,
                                         max_parallel_shots=mps, **C64)
        got = sim.simulate_shots(circ, shots=200)
        if terminal:
            want = terminal_counts_from_state(prefix, circ, 200, 3)
        else:
            want = oracle.terminal_trajectories(prefix, meas, 5, 200, 3, f32=True)
        assert got == want, terminal


def test_blocked_terminal_sampling_25q():
    # 2^25 outcomes: per-2048 block totals, binary search, in-block scan.
    n = 25
    unitary = workloads.random_su2(n=n, depth=3, seed=31).data["instructions"]
    circ = Circuit(n, unitary + workloads.measure_all(n), n)
    print("[c64] blocked terminal 25q", flush=True)
    sim = dqsim.StatevectorSimulator(seed=17, max_memory_mb=4096, max_fused_qubits=3, **C64)
    state = sim.simulate(Circuit(n, unitary)).statevector
    got = sim.simulate_shots(circ, shots=3000)
    assert sum(got.values()) == 3000
    assert got == terminal_counts_from_state(state, circ, 3000, 17)


@pytest.mark.parametrize("n", [16, 25])
def test_result_counts(n):
    circ = workloads.random_su2(n=n, depth=4, seed=40 + n)
    print(f"[c64] result counts n={n}", flush=True)
    res = dqsim.StatevectorSimulator(seed=5, max_memory_mb=4096, max_fused_qubits=3, **C64).simulate(circ)
    st = res.statevector
    rng = np.random.default_rng(n)
    perm = rng.permutation(n).tolist()
    for qubits in (list(range(n))[::-1], perm, perm[:7], [n - 1], []):
        idx = oracle.sample_marginal(st, qubits, 700, 77)
        assert res.counts(shots=700, qubits=qubits, seed=77) == oracle.counts_from_indices(idx, len(qubits)), qubits


def test_batch_indices():
    cs = [workloads.batch_circuit(i, base_seed=77, nmin=3, nmax=12, dmin=2, dmax=8) for i in range(8)]
    cs += [workloads.batch_circuit(i, base_seed=5, nmin=16, nmax=17, dmin=2, dmax=8) for i in range(2)]
    print("[c64] batch indices", [c.num_qubits for c in cs], flush=True)
    res = batch.run_batch(cs, shots=300, base_seed=9, workers=3, max_fused_qubits=3, autotune=False,
                          keep_states=True, lds_qubits=0)
    for i, c in enumerate(cs):
        n = c.num_qubits
        st = res.statevector(i)
        want = oracle.sample_marginal(st, list(range(n))[::-1], 300, 9 + i)
        assert np.array_equal(res.indices(i).astype(np.uint64), want), i
