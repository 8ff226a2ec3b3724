"""Batched execution of many small circuits.

No reference counterpart: the reference would run such a workload as a loop of
``StatevectorSimulator.simulate_shots`` calls. Every circuit must end in
terminal measurements; each gets its own shot count and seed. A pool of host
threads, each with its own HIP stream and stream-ordered device arena (no
allocation or device-wide synchronisation per circuit), takes the circuits
largest first; each circuit follows ``simulate_shots``' terminal-sampling
path with its own seed, so its counts equal that call's.

Small complex64 circuits (up to ``lds_qubits`` qubits, default 12, every qubit
measured) instead run in one single-launch kernel captured in a hipGraph, one
workgroup per circuit with its whole state in LDS, concurrently with the pool:
their final states match the FP64 oracle within complex64 rounding and their
outcomes are the chunked sampler's draws on those states (restated by
``oracle.batch_sample``), not the per-circuit call's kernels. The earlier
persistent kernel is removed (``persistent=True`` raises): it stalled on its
first GPU run and no cause has been shown.
"""

from __future__ import annotations

import ctypes
from typing import Dict, List, Optional, Sequence

import numpy as np

from ._core import _BatchOpts, _check, _circuit_json, _CountTable, _lib

_PREC = {"complex64": 8, "complex128": 16}


class BatchResult:
    """Counts, sampled outcomes, timings and the tuned settings of a batch."""

    def __init__(self, handle, ncirc: int, shots: List[int], nqubits: List[int]):
        self._h = handle
        self.num_circuits = ncirc
        self._shots = shots
        self._nq = nqubits
        total = ctypes.c_double()
        workers, fused, tile = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        streams = ctypes.c_uint32()
        lat = np.zeros(max(1, ncirc), dtype=np.float64)
        _check(_lib.dq_batch_info(handle, ctypes.byref(total), ctypes.byref(workers), ctypes.byref(fused),
                                  ctypes.byref(tile), ctypes.byref(streams), lat.ctypes.data, ncirc))
        self.wall_ms = total.value
        self.workers = workers.value
        self.max_fused_qubits = fused.value
        self.tile_qubits = tile.value
        self.streams = streams.value
        self.latency_us = lat[:ncirc]
        cnt, lms = ctypes.c_uint64(), ctypes.c_double()
        _check(_lib.dq_batch_lds_info(handle, ctypes.byref(cnt), ctypes.byref(lms)))
        self.lds_circuits = cnt.value  # circuits run by the single launch
        self.lds_ms = lms.value        # its device time
        self.tuning = {}
        name, ms = ctypes.create_string_buffer(64), ctypes.c_double()
        for i in range(int(_lib.dq_batch_tuning_count(handle))):
            _check(_lib.dq_batch_tuning_entry(handle, i, name, 64, ctypes.byref(ms)))
            self.tuning[name.value.decode()] = ms.value

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            _lib.dq_batch_free(h)
            self._h = None

    def counts(self, i: int) -> Dict[str, int]:
        tab = _CountTable()
        _check(_lib.dq_batch_counts(self._h, i, tab.cb, tab.user))
        return tab.result()

    def indices(self, i: int) -> np.ndarray:
        """Sampled outcomes over the measured qubits (highest measured qubit =
        most significant bit)."""
        out = np.empty(max(1, self._shots[i]), dtype=np.uint32)
        _check(_lib.dq_batch_indices(self._h, i, out.ctypes.data, self._shots[i]))
        return out[: self._shots[i]]

    def statevector(self, i: int) -> np.ndarray:
        out = np.empty(1 << self._nq[i], dtype=np.complex128)
        _check(_lib.dq_batch_state_c128(self._h, i, out.ctypes.data, out.size))
        return out


def run_batch(circuits: Sequence, shots=1024, seeds: Optional[Sequence[int]] = None, base_seed: int = 0,
              workers: int = 0, max_fused_qubits: int = 0, autotune: bool = True, tile_qubits: Optional[int] = None,
              precision: str = "complex64", keep_states: bool = False, persistent: bool = False,
              lds_qubits: int = 12) -> BatchResult:
    """Run circuits as one batch. ``shots`` is an int or a list; ``seeds``
    defaults to base_seed + index. ``max_fused_qubits`` of 0 and
    ``tile_qubits`` of None (LDS tile sweeps of 10, 11 or 12 qubits, or none)
    are chosen by timing candidates on a sample when ``autotune`` is set; an
    untuned tile is 12. ``workers`` of 0 runs 4 host threads (each with its
    own HIP stream), -1 times 2, 4 and 8 on the sample as well (a small
    sample favours fewer workers than the whole call does). Complex64 circuits of up to ``lds_qubits`` qubits
    (0..12; 0 = none) that measure every qubit run in the single launch."""
    if precision not in _PREC:
        raise ValueError("precision must be 'complex64' or 'complex128'")
    texts = [_circuit_json(c) for c in circuits]
    n = len(texts)
    shots_l = [int(shots)] * n if isinstance(shots, (int, np.integer)) else [int(s) for s in shots]
    seeds_l = [base_seed + i for i in range(n)] if seeds is None else [int(s) for s in seeds]
    bufs = [ctypes.create_string_buffer(t, len(t)) for t in texts]
    ptrs = (ctypes.c_void_p * max(1, n))(*[ctypes.addressof(b) for b in bufs])
    lens = (ctypes.c_size_t * max(1, n))(*[len(t) for t in texts])
    sh = (ctypes.c_uint64 * max(1, n))(*shots_l)
    sd = (ctypes.c_uint64 * max(1, n))(*seeds_l)
    o = _BatchOpts(workers, max_fused_qubits, 1 if autotune else 0, -1 if tile_qubits is None else int(tile_qubits),
                   _PREC[precision],
                   1 if keep_states else 0, 1 if persistent else 0, int(lds_qubits))
    h = ctypes.c_void_p()
    _check(_lib.dq_batch_run(ptrs, lens, n, sh, sd, ctypes.byref(o), ctypes.byref(h)))
    nq = [int(_lib.dq_batch_num_qubits(h, i)) for i in range(n)]
    return BatchResult(h, n, shots_l, nq)


def program_text(circuits: Sequence, shots=1024, seeds: Optional[Sequence[int]] = None, base_seed: int = 0,
                 max_fused_qubits: int = 3, lds_qubits: int = 14) -> str:
    """Host-only dump of the program ``run_batch`` would upload (no GPU)."""
    texts = [_circuit_json(c) for c in circuits]
    n = len(texts)
    shots_l = [int(shots)] * n if isinstance(shots, (int, np.integer)) else [int(s) for s in shots]
    seeds_l = [base_seed + i for i in range(n)] if seeds is None else [int(s) for s in seeds]
    bufs = [ctypes.create_string_buffer(t, len(t)) for t in texts]
    ptrs = (ctypes.c_void_p * max(1, n))(*[ctypes.addressof(b) for b in bufs])
    lens = (ctypes.c_size_t * max(1, n))(*[len(t) for t in texts])
    sh = (ctypes.c_uint64 * max(1, n))(*shots_l)
    sd = (ctypes.c_uint64 * max(1, n))(*seeds_l)
    out = ctypes.c_void_p()
    _check(_lib.dq_batch_program_text(ptrs, lens, n, sh, sd, max_fused_qubits, lds_qubits, ctypes.byref(out)))
    try:
        return ctypes.string_at(out.value).decode()
    finally:
        _lib.dq_string_free(out)


def device_info(block: int = 256) -> Dict[str, int]:
    """Launch limits of the single-launch batched kernel (256 threads; no
    launch): device LDS per block and per CU, the largest LDS-resident
    circuit, whether its LDS request stays within the default limit (0), the
    kernel's own attributes (its per-lane private segment included) and the
    runtime's count of resident workgroups per CU at that LDS size."""
    info = (ctypes.c_int32 * 11)()
    _check(_lib.dq_batch_device_info(block, info))
    keys = ("max_lds_per_block", "set_attribute_rc", "kernel_max_dynamic_lds", "kernel_static_lds",
            "kernel_max_threads", "lds_qubits", "device_count", "kernel_private_bytes", "lds_per_cu",
            "occupancy_rc", "blocks_per_cu")
    return dict(zip(keys, (int(v) for v in info)))
