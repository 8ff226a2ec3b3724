"""Learned fusion planner.

A 16 -> 64 -> 64 -> 1 FP32 MLP predicts the device time of one launch (log of
milliseconds per 2^24 amplitudes) from the launch's features:

   0 k/5              1 lowest qubit/34      2 highest qubit/34    3 qubits < 4, /4
   4 ops folded/32    5 diagonal-op share    6 Clifford-op share   7 all ops diagonal
   8 n/34             9 k >= 4 (MFMA path)  10 span/34            11 share of amplitudes touched
  12 tile sweep      13 tile qubits/12      14 3-qubit-path blocks/16  15 4-qubit blocks/16

(0-11 describe a one-block launch and are zero for a tile sweep; 12-15 are
zero for a one-block launch.) With ``fusion="learned"`` and weights
installed, complex64 plans with LDS tile sweeps are chosen among candidate
(block width, tile size) pairs by the model's predicted total time; plans
without sweeps ask the model before every merge and merge blocks a and b only
when t(a) + t(b) > t(a+b). Without weights both fall back to plain greedy.

Training data are the engine's own per-launch HIP-event timings
(``StatevectorSimulator.evolve_timed``) on seeded random circuits, and the
network is trained on the GPU by the engine's forward/backward/Adam kernels
(csrc/planner.hip).
"""

from __future__ import annotations

import ctypes
import math
from typing import Iterable, List, Optional, Sequence, Tuple

import numpy as np

from . import _core, workloads
from ._core import _check, _lib

N_FEATURES = 16
HIDDEN = 64
N_PARAMS = int(_lib.dq_mlp_param_count())
FEATURE_NAMES = ["k", "lowest_qubit", "highest_qubit", "low_qubits", "ops", "diagonal_share", "clifford_share",
                 "all_diagonal", "n", "mfma", "span", "touched_share", "tile", "tile_qubits", "tile_blocks3",
                 "tile_blocks4"]


def init_weights(seed: int = 0) -> np.ndarray:
    """He-initialised packed weights (biases zero)."""
    rng = np.random.default_rng(seed)
    w1 = rng.normal(0.0, math.sqrt(2.0 / N_FEATURES), size=(N_FEATURES, HIDDEN))
    w2 = rng.normal(0.0, math.sqrt(2.0 / HIDDEN), size=(HIDDEN, HIDDEN))
    w3 = rng.normal(0.0, math.sqrt(1.0 / HIDDEN), size=HIDDEN)
    parts = [w1.ravel(), np.zeros(HIDDEN), w2.ravel(), np.zeros(HIDDEN), w3, np.zeros(1)]
    return np.ascontiguousarray(np.concatenate(parts), dtype=np.float32)


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def grad(w: np.ndarray, X: np.ndarray, y: np.ndarray) -> Tuple[np.ndarray, float]:
    """Gradient of the mean squared error and the loss, on the GPU."""
    w, X, y = _f32(w), _f32(X), _f32(y)
    g = np.empty(N_PARAMS, dtype=np.float32)
    loss = np.zeros(1, dtype=np.float32)
    _check(_lib.dq_mlp_grad(w.ctypes.data, X.ctypes.data, y.ctypes.data, len(y), g.ctypes.data, loss.ctypes.data))
    return g, float(loss[0])


def predict(w: np.ndarray, X: np.ndarray) -> np.ndarray:
    w, X = _f32(w), _f32(X)
    out = np.empty(max(1, len(X)), dtype=np.float32)
    _check(_lib.dq_mlp_predict(w.ctypes.data, X.ctypes.data, len(X), out.ctypes.data))
    return out[: len(X)]


def train(X: np.ndarray, y: np.ndarray, epochs: int = 200, batch: int = 256, lr: float = 2e-3, seed: int = 0,
          w0: Optional[np.ndarray] = None) -> Tuple[np.ndarray, np.ndarray]:
    """Train on the GPU; returns (weights, per-epoch loss of the last batch)."""
    X, y = _f32(X), _f32(y)
    w = _f32(init_weights(seed) if w0 is None else w0).copy()
    curve = np.zeros(epochs, dtype=np.float32)
    _check(_lib.dq_mlp_train(w.ctypes.data, X.ctypes.data, y.ctypes.data, len(y), epochs, batch, lr, seed,
                             curve.ctypes.data))
    return w, curve


def install(w: Optional[np.ndarray]) -> None:
    """Use these weights for fusion="learned" (None restores greedy)."""
    _core.set_planner_weights(w)


def launch_records(circuit, max_fused_qubits: int, precision: str = "complex64", fusion: str = "greedy",
                   tile_qubits: int = 0):
    sim = _core.StatevectorSimulator(seed=1, precision=precision, fusion=fusion, max_fused_qubits=max_fused_qubits,
                                     max_memory_mb=1 << 20, tile_qubits=tile_qubits)
    rows, launches, gates, total_ms, _ = sim.evolve_timed(circuit, repeats=1, record=True)
    return rows, total_ms


def dataset(circuits: Iterable, widths: Sequence[int] = (2, 3, 4, 5), seed: int = 0, log=None,
            tiles: Sequence[int] = (0, 0, 0, 10, 11, 12)):
    """(X, y): features and log(ms per 2^24 amplitudes) of every launch of the
    circuits, each evolved with a block width drawn from `widths` and a tile
    size drawn from `tiles` (0: one launch per block; sweeps use widths <= 4)."""
    rng = np.random.default_rng(seed)
    X: List[List[float]] = []
    y: List[float] = []
    for i, c in enumerate(circuits):
        if log is not None and i % 25 == 0:
            log(f"[planner] dataset: circuit {i}, {len(y)} launches so far")
        n = c.num_qubits
        k = int(widths[int(rng.integers(len(widths)))])
        tile = int(tiles[int(rng.integers(len(tiles)))])
        if tile and c.num_qubits < tile + 4:
            tile = 0
        rows, _ = launch_records(c, min(k, 4) if tile else k, tile_qubits=tile)
        for r in rows:
            if r["ms"] <= 0:
                continue
            X.append(r["feat"])
            y.append(math.log(r["ms"] * 2.0 ** (24 - n)))
    return np.asarray(X, dtype=np.float32), np.asarray(y, dtype=np.float32)


FAMILIES = ("su2", "qft", "brickwork", "mixed")


def family_circuit(family: str, n: int, depth: int, seed: int):
    """One seeded circuit of a workload family (random SU(2) + matchings, QFT
    after a U3 layer, nearest-neighbour brickwork, mixed multi-qubit gates)."""
    if family == "su2":
        return workloads.random_su2(n=n, depth=depth, seed=seed)
    if family == "qft":
        return workloads.qft(n, seed=seed, measure=False)[0]
    if family == "brickwork":
        return workloads.brickwork(n=n, layers=depth, seed=seed, measure=False)
    return workloads.mixed(n, max(1, depth // 4), seed)


def training_circuits(count: int, nmin: int = 24, nmax: int = 28, dmin: int = 20, dmax: int = 80, seed: int = 0,
                      families: Sequence[str] = FAMILIES):
    """Seeded circuits of uneven size, rotating through `families`."""
    rng = np.random.default_rng(seed)
    for i in range(count):
        n = int(rng.integers(nmin, nmax + 1))
        d = int(rng.integers(dmin, dmax + 1))
        yield family_circuit(families[i % len(families)], n, d, seed * 1_000_003 + i)
