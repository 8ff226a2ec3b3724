"""Sampling above 2^24 outcomes (BASELINE config-2 sizes use this path): only
per-2048-outcome block totals are materialised, a shot finds its block by
binary search and scans it sequentially. Counts must equal the oracle's
restatement of the same procedure exactly (complex128 states are
bit-identical)."""

import numpy as np
import pytest

import dqsim
from dqsim import workloads
from dqsim.workloads import Circuit
from oracle import oracle

pytestmark = pytest.mark.gpu

N = 25  # 2^25 outcomes > 2^24


def test_blocked_terminal_sampling_matches_oracle():
    circ = workloads.random_su2(n=N, depth=3, seed=31)
    ins = circ.data["instructions"] + workloads.measure_all(N)
    c = Circuit(N, ins, cbits=N)
    sim = dqsim.StatevectorSimulator(seed=17, max_memory_mb=4096)
    got = sim.simulate_shots(c, shots=3000)
    want, _ = oracle.run_shots(c.data, 3000, 17, budget_mb=4096)
    assert sum(got.values()) == 3000
    assert got == want


def test_blocked_result_counts_match_oracle():
    circ = workloads.random_su2(n=N, depth=3, seed=32)
    res = dqsim.StatevectorSimulator(seed=5, max_memory_mb=4096).simulate(circ)
    st = oracle.simulate(circ.data, 5)[0]
    assert np.array_equal(res.statevector, st)
    for qubits in (list(range(N))[::-1], [24, 3, 17, 0, 9, 11, 12, 13, 14, 15, 16, 1, 2, 4, 5, 6, 7, 8, 10, 18, 19,
                                          20, 21, 22, 23]):
        idx = oracle.sample_marginal(st, qubits, 700, 77)
        assert res.counts(shots=700, qubits=qubits, seed=77) == oracle.counts_from_indices(idx, len(qubits))
