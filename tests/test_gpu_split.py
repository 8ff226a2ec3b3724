"""Split-precision 3-qubit blocks in complex64 tile sweeps (block arithmetic
mode 1): each FP32 tile amplitude and embedding entry enters the
16x16x32 f16 matrix core as two FP16 parts, and all four partial products
accumulate in FP32.

The FP64 oracle pins every state (max amplitude error <= 1e-5). That bound
alone would let a lost low part pass (an FP16-only contraction errs by about
2^-11 relative, far below 1e-5 absolute at these sizes), so each case also
asks the split state's error against the oracle to stay within 4x the error of
the FP32 matrix-core path on the same plan: a dropped partial product or a
wrong low part would multiply it by hundreds."""

import numpy as np
import pytest

import dqsim
from dqsim import _core, workloads
from oracle import oracle

pytestmark = pytest.mark.gpu

CASES = [
    # (family, n, tile qubits, widest block, sweep mode)
    ("su2", 16, 12, 3, 2), ("su2", 18, 12, 3, 0), ("su2", 20, 11, 3, 2), ("su2", 20, 10, 3, 0),
    ("su2", 22, 12, 4, 2), ("qft", 18, 12, 3, 2), ("mixed", 18, 12, 5, 2), ("brickwork", 20, 12, 3, 0),
    ("su2", 24, 12, 3, 2), ("su2", 20, 12, 3, 3),
]


def build(kind, n):
    if kind == "su2":
        return workloads.random_su2(n=n, depth=12, seed=n)
    if kind == "qft":
        return workloads.qft(n, seed=3, measure=False)[0]
    if kind == "brickwork":
        return workloads.brickwork(n=n, layers=12, seed=n, measure=False)
    return workloads.mixed(n, 6, seed=n)


def evolve(circ, k, L, mode, arith):
    prev_mode, prev_arith = _core.set_sweep_mode(mode), _core.set_block_arith(arith)
    try:
        sim = dqsim.StatevectorSimulator(seed=1, precision="complex64", fusion="greedy", max_fused_qubits=k,
                                         tile_qubits=L, max_memory_mb=4096)
        plan = _core.plan_json(circ, precision="complex64", fusion="greedy", max_fused_qubits=k, tile_qubits=L,
                               out_of_place=mode != 0)
        return sim.simulate(circ).statevector, plan
    finally:
        _core.set_sweep_mode(prev_mode)
        _core.set_block_arith(prev_arith)


@pytest.mark.parametrize("kind,n,L,k,mode", CASES, ids=lambda v: str(v))
def test_split_blocks_match_oracle_within_fp32_error(kind, n, L, k, mode):
    print(f"[split] {kind} n={n} L={L} k={k} mode={mode}", flush=True)
    circ = build(kind, n)
    want = oracle.simulate(circ.data, 1)[0]
    split, plan = evolve(circ, k, L, mode, 1)
    fp32, _ = evolve(circ, k, L, mode, 0)
    assert any(x["kind"] == "tile" for x in plan["launches"])
    err_split = np.abs(split - want).max()
    err_fp32 = np.abs(fp32 - want).max()
    print(f"[split] max err split {err_split:.3e}, fp32 {err_fp32:.3e}", flush=True)
    assert err_split < 1e-5
    assert err_split <= 4 * err_fp32 + 1e-12


def test_split_mode_switch_and_default():
    prev = _core.set_block_arith(1)
    try:
        assert _core.set_block_arith(prev) == 1
    finally:
        _core.set_block_arith(prev)
    with pytest.raises(ValueError):
        _core.set_block_arith(2)


# 4-qubit blocks in split precision (32x32x16 f16) next to 3-qubit ones, in
# place and out of place, at every tile size.
CASES4 = [("su2", 18, 12, 4, 2), ("su2", 20, 12, 4, 0), ("mixed", 18, 11, 4, 2), ("su2", 17, 10, 4, 2),
          ("qft", 18, 12, 4, 0), ("brickwork", 22, 12, 4, 2)]


@pytest.mark.parametrize("kind,n,L,k,mode", CASES4, ids=lambda v: str(v))
def test_split_4q_blocks_match_oracle_within_fp32_error(kind, n, L, k, mode):
    print(f"[split4] {kind} n={n} L={L} k={k} mode={mode}", flush=True)
    circ = build(kind, n)
    want = oracle.simulate(circ.data, 1)[0]
    split, plan = evolve(circ, k, L, mode, 1)
    fp32, _ = evolve(circ, k, L, mode, 0)
    blocks = [b for x in plan["launches"] if x["kind"] == "tile" for b in x["blocks"]]
    assert any(len(b["qubits"]) == 4 and b["split"] for b in blocks)
    err_split = np.abs(split - want).max()
    err_fp32 = np.abs(fp32 - want).max()
    print(f"[split4] max err split {err_split:.3e}, fp32 {err_fp32:.3e}", flush=True)
    assert err_split < 1e-5
    assert err_split <= 4 * err_fp32 + 1e-12


@pytest.mark.parametrize("kind,n,k", [("su2", 22, 3), ("qft", 20, 4), ("brickwork", 21, 3), ("mixed", 19, 4)])
def test_zero_start_is_bit_identical(kind, n, k):
    # Plans opening with out-of-place sweeps start from a synthesised |0...0>
    # and store the all-zero tiles of their first sweeps without reading
    # them: the same state bit for bit as initialising and reading.
    circ = build(kind, n)
    sim = dqsim.StatevectorSimulator(seed=1, precision="complex64", fusion="greedy", max_fused_qubits=k,
                                     tile_qubits=12, max_memory_mb=4096)
    prev = _core.set_zero_start(True)
    try:
        fast = sim.simulate(circ).statevector
        _core.set_zero_start(False)
        plain = sim.simulate(circ).statevector
    finally:
        _core.set_zero_start(prev)
    assert np.array_equal(fast, plain)
    want = oracle.simulate(circ.data, 1)[0]
    assert np.abs(fast - want).max() < 1e-5
