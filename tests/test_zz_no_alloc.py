"""The allocation-free evolution (dq_sv_evolve_noalloc) against dq_sv_simulate.

The caller owns the state buffer (here a torch device tensor); the library
keeps the plan constants in its static per-GPU scratch array. With only the
caller's buffer, complex64 tile sweeps run in place, so the state must equal
the statevector of dq_sv_simulate on the same options with in-place sweeps
(sweep mode 0) bit for bit, and the default out-of-place plan's within the
complex64 tolerance (1e-5 max amplitude difference), on
consecutive calls (the second smaller than the first, so stale scratch
contents from the first call would show), for complex64 tile sweeps and the
exact complex128 sweeps. Oversized plans, non-unitary circuits and wrong
buffer lengths are rejected before any device work.
"""

import ctypes

import numpy as np
import pytest

import dqsim
from dqsim import _core, workloads
from dqsim.workloads import Circuit

pytestmark = pytest.mark.gpu

PRECS = {"complex64": (np.complex64, 4), "complex128": (np.complex128, 8)}


def simulate_in_place(sim, circ):
    """dq_sv_simulate with complex64 sweeps in place (the plan the
    allocation-free evolution runs), then the default mode again."""
    prev = _core.set_sweep_mode(0)
    try:
        return sim.simulate(circ).statevector
    finally:
        _core.set_sweep_mode(prev)


def evolve_noalloc(sim, circ, n, precision, stream=None):
    import torch

    dtype, _ = PRECS[precision]
    real = torch.float32 if dtype == np.complex64 else torch.float64
    state = torch.full((2 << n,), float("nan"), dtype=real, device="cuda")
    torch.cuda.synchronize()
    js = _core._circuit_json(circ)
    rc = _core._lib.dq_sv_evolve_noalloc(sim._h, js, len(js), ctypes.c_void_p(state.data_ptr()), 1 << n, stream)
    _core._check(rc)
    return state.cpu().numpy().view(dtype).astype(np.complex128)


CASES = [
    # (precision, options, first circuit, second circuit)
    ("complex64", dict(fusion="greedy", max_fused_qubits=3, tile_qubits=12),
     workloads.random_su2(n=22, depth=12, seed=4), workloads.random_su2(n=14, depth=6, seed=5)),
    ("complex64", dict(fusion="greedy", max_fused_qubits=4, tile_qubits=0),
     workloads.mixed(18, 4, seed=6), workloads.random_su2(n=9, depth=5, seed=7)),
    ("complex128", dict(),
     workloads.random_su2(n=20, depth=10, seed=8), workloads.mixed(12, 4, seed=9)),
    ("complex128", dict(fusion="greedy", max_fused_qubits=5),
     workloads.mixed(16, 5, seed=10), workloads.qft(11, seed=3, measure=False)[0]),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}-{c[1]}")
def test_noalloc_matches_simulate_on_consecutive_calls(case):
    precision, opts, first, second = case
    sim = dqsim.StatevectorSimulator(seed=1, precision=precision, max_memory_mb=8192, **opts)
    for circ in (first, second):
        n = circ.num_qubits
        print(f"[noalloc] {precision} {opts} n={n}", flush=True)
        got = evolve_noalloc(sim, circ, n, precision)
        want = simulate_in_place(sim, circ)
        assert np.array_equal(got, want), (n, np.abs(got - want).max())
        default = sim.simulate(circ).statevector
        assert np.abs(got - default).max() <= (1e-5 if precision == "complex64" else 0.0), n


def test_noalloc_on_a_caller_stream():
    import torch

    circ = workloads.random_su2(n=16, depth=8, seed=11)
    sim = dqsim.StatevectorSimulator(seed=1, precision="complex64", fusion="greedy", max_fused_qubits=3)
    s = torch.cuda.Stream()
    got = evolve_noalloc(sim, circ, 16, "complex64", stream=ctypes.c_void_p(s.cuda_stream))
    assert np.array_equal(got, simulate_in_place(sim, circ))


def test_noalloc_rejects_before_device_work():
    import torch

    sim = dqsim.StatevectorSimulator(seed=1)
    buf = torch.zeros(2 << 4, dtype=torch.float64, device="cuda")
    ptr = ctypes.c_void_p(buf.data_ptr())
    measured = Circuit(4, [{"kind": "h", "qubit": 0}, {"kind": "measure", "qubit": 0, "cbit": 0}], 1)
    for circ, length, text in ((measured, 16, "unitary circuits only"),
                               (Circuit(4, [{"kind": "h", "qubit": 1}]), 8, "amplitudes")):
        js = _core._circuit_json(circ)
        with pytest.raises(ValueError, match=text):
            _core._check(_core._lib.dq_sv_evolve_noalloc(sim._h, js, len(js), ptr, length, None))
    # Plan constants beyond the scratch array: 70,000 dense two-qubit gates
    # (256 bytes of matrix each) need more than 16 MiB.
    ins = [{"kind": "rxx", "a": i % 3, "b": (i % 3) + 1, "theta": 0.001 * (i % 97)} for i in range(70000)]
    js = _core._circuit_json(Circuit(4, ins))
    with pytest.raises(ValueError, match="allocation-free evolution holds at most"):
        _core._check(_core._lib.dq_sv_evolve_noalloc(sim._h, js, len(js), ptr, 16, None))
    assert torch.count_nonzero(buf).item() == 0  # nothing was written
