"""Benchmark workloads at their full sizes.

QFT after a U3 layer (the precision workload): 30 qubits in complex64 and 33
qubits in complex128 (a 128 GiB state), checked against the exact closed form
amp[k] = N^-1/2 prod_q (a_q + b_q w^(k 2^q mod N)) at 65,600 sampled indices
(max error <= 1e-5 for complex64, <= 1e-12 for complex128), plus a 10-qubit
marginal of the complex64 state against the complex128 state at 30 qubits.

The fused-block workload (Haar SU(2) layers + CZ/CX on random matchings,
depth 40) with blocks of <= 3 qubits in 12-qubit LDS tile sweeps, against the
FP64 oracle at 26 qubits (max amplitude error <= 1e-5)."""

import numpy as np
import pytest

import dqsim
from dqsim import workloads
from oracle import oracle

pytestmark = pytest.mark.gpu


def sample_indices(n, rng, count=1 << 16):
    return np.unique(np.concatenate([np.arange(64), rng.integers(0, 1 << n, size=count)])).astype(np.uint64)


def test_qft_30q_complex64_and_marginal():
    n = 30
    circ, angles = workloads.qft(n, seed=7, measure=False)
    idx = sample_indices(n, np.random.default_rng(7))
    exact = workloads.qft_closed_form(angles, idx)
    mem = (1 << n) * 24 // (1 << 20) + 4096
    print("[full] qft 30q complex64", flush=True)
    r64 = dqsim.StatevectorSimulator(seed=1, precision="complex64", fusion="greedy", max_fused_qubits=4,
                                     max_memory_mb=mem).simulate(circ)
    err64 = np.abs(r64.amplitudes(idx) - exact).max()
    print(f"[full] qft 30q complex64 max err {err64:.3e}", flush=True)
    assert err64 <= 1e-5
    r128 = dqsim.StatevectorSimulator(seed=1, precision="complex128", max_memory_mb=mem).simulate(circ)
    err128 = np.abs(r128.amplitudes(idx) - exact).max()
    assert err128 <= 1e-12
    # A 10-qubit marginal (1,024 outcomes) of the complex64 state against the
    # complex128 one: total-variation distance of the exact marginals.
    qubits = [29, 23, 17, 13, 11, 7, 5, 3, 1, 0]
    p64, p128 = r64.probabilities(qubits), r128.probabilities(qubits)
    tvd = 0.5 * sum(abs(p64.get(k, 0.0) - p128.get(k, 0.0)) for k in set(p64) | set(p128))
    print(f"[full] 10-qubit marginal TVD {tvd:.3e}", flush=True)
    assert tvd < 1e-5
    max_abs, l2 = r64.compare(r128)
    assert max_abs <= 1e-5 and l2 <= 1e-4


def test_qft_33q_complex128():
    n = 33
    circ, angles = workloads.qft(n, seed=7, measure=False)
    idx = sample_indices(n, np.random.default_rng(33))
    exact = workloads.qft_closed_form(angles, idx)
    print("[full] qft 33q complex128 (128 GiB state)", flush=True)
    res = dqsim.StatevectorSimulator(seed=1, precision="complex128",
                                     max_memory_mb=(1 << n) * 16 // (1 << 20) + 4096).simulate(circ)
    err = np.abs(res.amplitudes(idx) - exact).max()
    print(f"[full] qft 33q complex128 max err {err:.3e}", flush=True)
    assert err <= 1e-12


def test_fused_su2_depth40_26q_against_oracle():
    n = 26
    circ = workloads.random_su2(n=n, depth=40, seed=1234)
    print("[full] random su2 26q depth 40 (oracle ~10 s)", flush=True)
    want, _ = oracle.simulate(circ.data, 1)
    res = dqsim.StatevectorSimulator(seed=1, precision="complex64", fusion="greedy", max_fused_qubits=3,
                                     tile_qubits=12, max_memory_mb=4096).simulate(circ)
    got = res.statevector
    err = np.abs(got - want).max()
    l2 = np.linalg.norm(got - want)
    print(f"[full] 26q depth 40 complex64 max err {err:.3e} L2 {l2:.3e}", flush=True)
    assert err <= 1e-5 and l2 <= 1e-4
