"""Learned fusion planner.

A 16 -> 64 -> 64 -> 1 FP32 MLP predicts the device time of one launch (log of
milliseconds per 2^24 amplitudes) from the launch's features. A one-block
launch and an in-place tile sweep use

   0 k/5              1 lowest qubit/34      2 highest qubit/34    3 qubits < 4, /4
   4 ops folded/32    5 diagonal-op share    6 Clifford-op share   7 all ops diagonal
   8 n/34             9 k >= 4 (MFMA path)  10 span/34            11 share of amplitudes touched
  12 tile sweep      13 tile qubits/12      14 3-qubit-path blocks/16  15 4-qubit blocks/16

(0-11 describe a one-block launch and are zero for an in-place tile sweep
except 4 and 8; 12-15 are zero for a one-block launch). An out-of-place
(permuting) tile sweep, the complex64 default when the budget holds a second
state, uses

   0 1 (out of place)  1 3-qubit blocks on a tile bit < 4, /16
   2 2-qubit blocks on a tile bit < 6, /16      3 2-qubit blocks/16
   4 gates/32          5 gather (first sweep of a run)
   6 destination positions 0..3 stored from tile bits >= 4, /4
   7 lowest destination positions stored from tile bits, /12
   8 n/34              9 512-thread workgroups 10 persistent workgroups
  11 blob bytes/32768 12 1 (tile sweep)       13 tile qubits/12
  14 3-qubit blocks/16 15 4-qubit blocks/16

With ``fusion="learned"`` and weights installed:
  * out-of-place sweeps: the plan is chosen among candidate (tile rule, block
    width) pairs -- qubits 0..3 pinned in every tile or only shared between
    neighbouring tiles, blocks of 2..max_fused (<= 4) qubits -- by the
    model's predicted total time, the pair predicted fastest (the default,
    the process-wide tile rule with blocks of max_fused qubits, on a tie);
  * in-place sweeps: among (block width, tile size) pairs, the default
    (blocks <= 3 qubits, the largest tile) unless another pair is predicted
    at least 3 % faster;
  * plans without sweeps ask the model before every merge and merge blocks a
    and b only when t(a) + t(b) > t(a+b).
Without weights all three fall back to plain greedy. ``StatevectorSimulator``
with ``fusion="learned"`` installs the weights shipped with the package
(``planner_weights.npy``, trained by ``python -m dqsim.planner``) unless
weights were installed (or cleared) explicitly.

Training data are the engine's own per-launch HIP-event timings
(``StatevectorSimulator.evolve_timed``) on seeded random circuits, and the
network is trained on the GPU by the engine's forward/backward/Adam kernels
(csrc/planner.hip).
"""

from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import sys
import time
from typing import Dict, Iterable, List, Optional, Sequence, Tuple

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


WEIGHTS_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "planner_weights.npy")


def load_default(path: str = WEIGHTS_FILE) -> Optional[np.ndarray]:
    """The shipped weights (None if the file is absent)."""
    if not os.path.exists(path):
        return None
    w = np.load(path, allow_pickle=False)
    if w.dtype != np.float32 or w.shape != (N_PARAMS,):
        raise ValueError(f"{path}: expected {N_PARAMS} float32 weights, got {w.dtype} {w.shape}")
    return w


def install_default(path: str = WEIGHTS_FILE) -> bool:
    """Install the shipped weights; False (greedy stays) if there are none."""
    w = load_default(path)
    if w is None:
        return False
    install(w)
    return True


def launch_records(circuit, max_fused_qubits: int, precision: str = "complex64", fusion: str = "greedy",
                   tile_qubits: int = 0, out_of_place: bool = True):
    """Per-launch timings of one evolution. out_of_place=False gives the call
    a budget that holds one state only, so its tile sweeps run in place."""
    if out_of_place:
        mem = 1 << 20
    else:
        state_mb = ((1 << circuit.num_qubits) * (8 if precision == "complex64" else 16)) >> 20
        mem = state_mb + state_mb // 2 + 64
    sim = _core.StatevectorSimulator(seed=1, precision=precision, fusion=fusion, max_fused_qubits=max_fused_qubits,
                                     max_memory_mb=mem, tile_qubits=tile_qubits)
    rows, launches, gates, total_ms, _ = sim.evolve_timed(circuit, repeats=1, record=True)
    return rows, total_ms


def sweep_dataset(circuits: Iterable, widths: Sequence[int] = (2, 3, 4), seed: int = 0, log=None,
                  block_share: float = 0.1, in_place_share: float = 0.15, time_limit_s: Optional[float] = None):
    """(X, y, info): features and log(ms per 2^24 amplitudes) of every launch
    of the circuits, complex64. Each circuit is evolved once: with
    probability block_share one launch per block (width 2..5), with
    in_place_share in-place 12-qubit tile sweeps, otherwise out-of-place
    12-qubit sweeps under a tile rule (pinned or shared) and block width drawn
    at random -- the choices the learned planner makes. Sweeps that the zero
    start shortens (a plan's first sweeps from |0...0>, whose recorded bytes
    fall below one full read and write of the state) are left out: their
    time depends on the start state, not on the features. Collection stops
    after time_limit_s seconds if given."""
    rng = np.random.default_rng(seed)
    X: List[List[float]] = []
    y: List[float] = []
    info = {"circuits": 0, "block_plans": 0, "in_place_plans": 0, "out_of_place_plans": 0, "qubits": {}}
    t0 = time.perf_counter()
    last = t0
    for i, c in enumerate(circuits):
        n = c.num_qubits
        u = float(rng.random())
        k = int(widths[int(rng.integers(len(widths)))])
        pinned = bool(rng.integers(2))
        if u < block_share:
            rows, _ = launch_records(c, int(rng.integers(2, 6)), tile_qubits=0)
            info["block_plans"] += 1
        elif u < block_share + in_place_share:
            rows, _ = launch_records(c, k, tile_qubits=12, out_of_place=False)
            info["in_place_plans"] += 1
        else:
            prev = _core.set_sweep_pinned(pinned)
            try:
                rows, _ = launch_records(c, k, tile_qubits=12)
            finally:
                _core.set_sweep_pinned(prev)
            info["out_of_place_plans"] += 1
        info["circuits"] += 1
        info["qubits"][str(n)] = info["qubits"].get(str(n), 0) + 1
        full = 2.0 * 8.0 * 2.0 ** n  # one complex64 state read and written
        for r in rows:
            if r["ms"] <= 0 or (r["kind"] == 11 and r["bytes"] < full):
                continue
            X.append(r["feat"])
            y.append(math.log(r["ms"] * 2.0 ** (24 - n)))
        now = time.perf_counter()
        if log is not None and (now - last > 30 or i == 0):
            last = now
            log(f"[planner] dataset: circuit {i + 1} ({n} qubits), {len(y)} launches, {now - t0:.0f} s")
        if time_limit_s is not None and now - t0 > time_limit_s:
            break
    info["collect_s"] = time.perf_counter() - t0
    return np.asarray(X, dtype=np.float32), np.asarray(y, dtype=np.float32), info


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


# Share of training circuits per qubit range: the time of a launch per 2^24
# amplitudes is what the model learns, so most circuits are small (fast to
# record) and the range the planner is evaluated on (28-33) is still covered.
SIZE_MIX = ((24, 27, 0.62), (28, 30, 0.25), (31, 32, 0.10), (33, 33, 0.03))


def mixed_size_circuits(count: int, seed: int, dmin: int = 20, dmax: int = 80,
                        mix: Sequence[Tuple[int, int, float]] = SIZE_MIX, families: Sequence[str] = FAMILIES):
    """Seeded training circuits: qubit count drawn from `mix`, depth uniform in
    [dmin, dmax], families in rotation."""
    rng = np.random.default_rng(seed)
    p = np.array([m[2] for m in mix], dtype=np.float64)
    p /= p.sum()
    for i in range(count):
        lo, hi, _ = mix[int(rng.choice(len(mix), p=p))]
        n = int(rng.integers(lo, hi + 1))
        d = int(rng.integers(dmin, dmax + 1))
        yield family_circuit(families[i % len(families)], n, d, seed * 1_000_003 + i)


def collect(out_dir: str, circuits: int = 4000, seed: int = 11, time_limit_s: float = 600, log=print) -> Dict:
    """Record engine timings of seeded training circuits and write them to
    out_dir/planner_dataset.npz (X, y) with planner_dataset.json (what was
    recorded)."""
    os.makedirs(out_dir, exist_ok=True)
    X, y, info = sweep_dataset(mixed_size_circuits(circuits, seed), seed=seed, log=log, time_limit_s=time_limit_s)
    log(f"[planner] {info['circuits']} circuits, {len(y)} launches in {info['collect_s']:.0f} s")
    np.savez_compressed(os.path.join(out_dir, "planner_dataset.npz"), X=X, y=y)
    info = dict(info, seed=seed, samples=int(len(y)))
    with open(os.path.join(out_dir, "planner_dataset.json"), "w") as f:
        json.dump(info, f, indent=1)
    return info


def train_and_save(out_dir: str, circuits: int = 4000, seed: int = 11, epochs: int = 60, time_limit_s: float = 600,
                   log=print, datasets: Sequence[str] = ()) -> Dict:
    """Record engine timings (unless circuits is 0), add the samples of
    earlier recordings (`datasets`: planner_dataset.npz files, each with its
    planner_dataset.json beside it), train the MLP on the GPU, and write
    planner_weights.npy, planner_weights.json (how it was made) and
    planner_dataset.npz (this run's samples) into out_dir."""
    os.makedirs(out_dir, exist_ok=True)
    parts_X, parts_y, runs = [], [], []
    info = {"circuits": 0, "block_plans": 0, "in_place_plans": 0, "out_of_place_plans": 0, "qubits": {},
            "collect_s": 0.0}
    if circuits > 0:
        X, y, info = sweep_dataset(mixed_size_circuits(circuits, seed), seed=seed, log=log,
                                   time_limit_s=time_limit_s)
        log(f"[planner] {info['circuits']} circuits, {len(y)} launches in {info['collect_s']:.0f} s")
        np.savez_compressed(os.path.join(out_dir, "planner_dataset.npz"), X=X, y=y)
        parts_X.append(X)
        parts_y.append(y)
        runs.append({"seed": seed, "circuits": info["circuits"], "samples": int(len(y))})
    for path in datasets:
        with np.load(path) as d:
            parts_X.append(d["X"].astype(np.float32))
            parts_y.append(d["y"].astype(np.float32))
        side = os.path.join(os.path.dirname(path), "planner_dataset.json")
        extra = json.load(open(side)) if os.path.exists(side) else {}
        for key in ("circuits", "block_plans", "in_place_plans", "out_of_place_plans", "collect_s"):
            info[key] = info.get(key, 0) + extra.get(key, 0)
        for q, c in extra.get("qubits", {}).items():
            info["qubits"][q] = info["qubits"].get(q, 0) + c
        runs.append({"seed": extra.get("seed"), "circuits": extra.get("circuits"), "samples": int(len(parts_y[-1]))})
    if not parts_y:
        raise ValueError("no samples: record circuits or give datasets")
    X, y = np.concatenate(parts_X), np.concatenate(parts_y)
    perm = np.random.default_rng(seed).permutation(len(y))
    cut = int(0.9 * len(y))
    tr, te = perm[:cut], perm[cut:]
    t0 = time.perf_counter()
    w, curve = train(X[tr], y[tr], epochs=epochs, batch=256, lr=2e-3, seed=seed)
    train_s = time.perf_counter() - t0
    pred = predict(w, X[te])
    meta = dict(info)
    meta.update({"seed": seed, "recordings": runs, "size_mix": [list(m) for m in SIZE_MIX], "families": list(FAMILIES),
                 "depth_range": [20, 80], "train_samples": int(len(tr)), "test_samples": int(len(te)),
                 "epochs": epochs, "train_s": train_s, "final_train_loss": float(curve[-1]),
                 "test_mse": float(np.mean((pred - y[te]) ** 2)),
                 "test_mse_predict_mean": float(np.mean((y[te] - y[tr].mean()) ** 2))})
    np.save(os.path.join(out_dir, "planner_weights.npy"), w)
    with open(os.path.join(out_dir, "planner_weights.json"), "w") as f:
        json.dump(meta, f, indent=1)
    log(f"[planner] trained in {train_s:.0f} s: test mse {meta['test_mse']:.4f} "
        f"(predicting the mean: {meta['test_mse_predict_mean']:.4f})")
    return meta


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="Train the learned planner on engine-recorded launch timings.")
    ap.add_argument("--out", required=True, help="directory for planner_weights.npy/.json and the dataset")
    ap.add_argument("--circuits", type=int, default=4000)
    ap.add_argument("--seed", type=int, default=11)
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--time-limit", type=float, default=600, help="seconds of timing collection at most")
    ap.add_argument("--collect-only", action="store_true", help="record timings into --out, do not train")
    ap.add_argument("--datasets", nargs="*", default=[],
                    help="planner_dataset.npz files of earlier recordings to train on as well")
    a = ap.parse_args(argv)
    log = lambda m: print(m, flush=True)  # noqa: E731
    if a.collect_only:
        info = collect(a.out, a.circuits, a.seed, a.time_limit, log=log)
        print(json.dumps({k: v for k, v in info.items() if k != "qubits"}), flush=True)
        return 0
    meta = train_and_save(a.out, a.circuits, a.seed, a.epochs, a.time_limit, log=log, datasets=a.datasets)
    print(json.dumps({k: v for k, v in meta.items() if k != "qubits"}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
