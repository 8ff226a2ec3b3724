"""Seeded synthetic circuits in the JSON transport (src/types.rs:7-296).

Every generator documents its draw order, so the same seed always yields the
same instruction list; tests, the benchmark and the CPU oracle all consume the
identical JSON. Random numbers come from numpy's PCG64 ``default_rng(seed)``.

  brickwork ..... U3 on every qubit then CX on alternating pairs
                  (20 qubits, 20 layers, seed 42: 590 gates)
  qft ........... one U3 layer then the textbook QFT (H, controlled-phase
                  cascade, final swaps)
  random_su2 .... per layer Haar SU(2) on every qubit (as U3 angles) then
                  CZ/CX on a random perfect matching (32 qubits, depth 40:
                  1,920 gates)
  batch ......... many small U3 + random-matching-CX circuits of uneven size
"""

from __future__ import annotations

import json
import math
from typing import Dict, List, Optional, Tuple

import numpy as np

TAU = 2.0 * math.pi


class Circuit:
    """Minimal transport object exposing ``model_dump_json()``."""

    def __init__(self, qubits: int, instructions=(), cbits: int = 0):
        self.data = {
            "qregs": {"q": {"name": "q", "base": 0, "size": qubits}},
            "cregs": {"c": {"name": "c", "base": 0, "size": cbits}},
            "instructions": list(instructions),
        }

    @property
    def num_qubits(self) -> int:
        return self.data["qregs"]["q"]["size"]

    def model_dump_json(self) -> str:
        return json.dumps(self.data)


def gate_count(instructions) -> int:
    """Unitary gate instructions (measure, reset, barrier excluded)."""
    return sum(1 for i in instructions if i["kind"] not in ("measure", "reset", "barrier", "conditional"))


def measure_all(n: int) -> List[dict]:
    return [{"kind": "measure", "qubit": q, "cbit": q} for q in range(n)]


def brickwork(n: int = 20, layers: int = 20, seed: int = 42, measure: bool = True) -> Circuit:
    """Layer L: U3(theta, phi, lam) on q = 0..n-1, angles uniform in [0, 2pi)
    drawn in the order (L, q, theta, phi, lam); then CX on (0,1),(2,3),... for
    even L and (1,2),(3,4),... for odd L."""
    rng = np.random.default_rng(seed)
    insts: List[dict] = []
    for layer in range(layers):
        for q in range(n):
            th, ph, la = (float(x) for x in rng.uniform(0.0, TAU, size=3))
            insts.append({"kind": "u3", "qubit": q, "theta": th, "phi": ph, "lam": la})
        start = 0 if layer % 2 == 0 else 1
        for a in range(start, n - 1, 2):
            insts.append({"kind": "cx", "control": a, "target": a + 1})
    if measure:
        insts += measure_all(n)
    return Circuit(n, insts, n if measure else 0)


def qft_instructions(n: int) -> List[dict]:
    """QFT: for j = n-1..0: H(j), then CP(pi / 2^(j-k)) on (k, j) for
    k = j-1..0; then SWAP(i, n-1-i) for i < n/2."""
    insts: List[dict] = []
    for j in range(n - 1, -1, -1):
        insts.append({"kind": "h", "qubit": j})
        for k in range(j - 1, -1, -1):
            insts.append({"kind": "cp", "control": k, "target": j, "lam": math.pi / (1 << (j - k))})
    for i in range(n // 2):
        insts.append({"kind": "swap", "a": i, "b": n - 1 - i})
    return insts


def qft(n: int, seed: int = 7, measure: bool = True) -> Tuple[Circuit, np.ndarray]:
    """One U3 layer (theta, phi, lam uniform in [0, 2pi), drawn per qubit in
    order) followed by the QFT. Also returns the U3 angles (n x 3)."""
    rng = np.random.default_rng(seed)
    angles = rng.uniform(0.0, TAU, size=(n, 3))
    insts = [{"kind": "u3", "qubit": q, "theta": float(a[0]), "phi": float(a[1]), "lam": float(a[2])}
             for q, a in enumerate(angles)]
    insts += qft_instructions(n)
    if measure:
        insts += measure_all(n)
    return Circuit(n, insts, n if measure else 0), angles


def qft_closed_form(angles: np.ndarray, indices: np.ndarray) -> np.ndarray:
    """Exact amplitudes of QFT applied to the U3 product state.

    With (a_q, b_q) = U3|0> = (cos(theta/2), e^{i phi} sin(theta/2)) the QFT
    output is amp[k] = N^{-1/2} prod_q (a_q + b_q w^{(k 2^q) mod N}),
    w = e^{2 pi i / N}: the product form of the DFT of a product state.
    """
    n = angles.shape[0]
    N = 1 << n
    a = np.cos(angles[:, 0] / 2.0)
    b = np.exp(1j * angles[:, 1]) * np.sin(angles[:, 0] / 2.0)
    out = np.ones(len(indices), dtype=np.complex128)
    for q in range(n):
        ph = (indices.astype(np.uint64) * np.uint64(1 << q)) % np.uint64(N)
        out *= a[q] + b[q] * np.exp(2j * np.pi * ph.astype(np.float64) / N)
    return out / math.sqrt(N)


def haar_u3(rng: np.random.Generator) -> Tuple[float, float, float]:
    """Haar-random SU(2) as U3 angles: theta = arccos(1 - 2u), phi and lam
    uniform in [0, 2pi) (ZYZ Euler angles with the sin(theta) measure)."""
    u, ph, la = rng.random(3)
    return float(math.acos(1.0 - 2.0 * u)), float(ph * TAU), float(la * TAU)


def random_su2(n: int = 32, depth: int = 40, seed: int = 1234, measure: bool = False) -> Circuit:
    """Per layer: haar_u3 for q = 0..n-1 (in order), then a perfect matching
    from rng.permutation(n) paired as (p0,p1),(p2,p3),...; for each pair one
    draw picks CZ (< 0.5) or CX, a second draw swaps control and target."""
    rng = np.random.default_rng(seed)
    insts: List[dict] = []
    for _ in range(depth):
        for q in range(n):
            th, ph, la = haar_u3(rng)
            insts.append({"kind": "u3", "qubit": q, "theta": th, "phi": ph, "lam": la})
        perm = rng.permutation(n)
        for i in range(0, n - 1, 2):
            a, b = int(perm[i]), int(perm[i + 1])
            kind = "cz" if rng.random() < 0.5 else "cx"
            if rng.random() < 0.5:
                a, b = b, a
            insts.append({"kind": kind, "control": a, "target": b})
    if measure:
        insts += measure_all(n)
    return Circuit(n, insts, n if measure else 0)


def mirror(circ: Circuit) -> Circuit:
    """C followed by C^dagger (for gate sets used here): the result is |0...0>."""
    insts = [i for i in circ.data["instructions"] if i["kind"] != "measure"]
    inv: List[dict] = []
    for i in reversed(insts):
        k = i["kind"]
        if k == "u3":
            inv.append({"kind": "u3", "qubit": i["qubit"], "theta": -i["theta"], "phi": -i["lam"],
                        "lam": -i["phi"]})
        elif k in ("cx", "cz", "h", "x", "swap"):
            inv.append(dict(i))
        elif k == "cp":
            inv.append({**i, "lam": -i["lam"]})
        else:
            raise ValueError(f"mirror: unsupported kind {k}")
    n = circ.num_qubits
    return Circuit(n, insts + inv, 0)


def batch_circuit(index: int, base_seed: int = 2024, nmin: int = 10, nmax: int = 22, dmin: int = 10,
                  dmax: int = 60) -> Circuit:
    """Circuit `index` of the batched workload: seed = base_seed + index;
    n ~ U{nmin..nmax}, depth ~ U{dmin..dmax}; per layer U3 (uniform angles) on
    every qubit then CX on a random matching; all qubits measured."""
    rng = np.random.default_rng(base_seed + index)
    n = int(rng.integers(nmin, nmax + 1))
    depth = int(rng.integers(dmin, dmax + 1))
    insts: List[dict] = []
    for _ in range(depth):
        for q in range(n):
            th, ph, la = (float(x) for x in rng.uniform(0.0, TAU, size=3))
            insts.append({"kind": "u3", "qubit": q, "theta": th, "phi": ph, "lam": la})
        perm = rng.permutation(n)
        for i in range(0, n - 1, 2):
            insts.append({"kind": "cx", "control": int(perm[i]), "target": int(perm[i + 1])})
    insts += measure_all(n)
    return Circuit(n, insts, n)


def mixed(n: int, depth: int, seed: int = 0) -> Circuit:
    """U3 layers plus seeded multi-qubit gates of every lowering kind (dense,
    controlled, diagonal, swap, 3-5 qubit controls): n // 2 gates per layer,
    kind and qubits drawn uniformly."""
    rng = np.random.default_rng(seed)
    ins: List[dict] = []
    for _ in range(depth):
        for q in range(n):
            t, p, l = rng.uniform(-np.pi, np.pi, 3)
            ins.append({"kind": "u3", "qubit": q, "theta": float(t), "phi": float(p), "lam": float(l)})
        for _ in range(n // 2):
            kind = str(rng.choice(["cx", "cz", "swap", "rzz", "cp", "ccx", "cswap", "c3x", "c4x", "ch"]))
            q = [int(x) for x in rng.choice(n, 5, replace=False)]
            if kind in ("cx", "cz", "ch"):
                ins.append({"kind": kind, "control": q[0], "target": q[1]})
            elif kind == "cp":
                ins.append({"kind": kind, "control": q[0], "target": q[1], "lam": float(rng.uniform(0, 3))})
            elif kind == "swap":
                ins.append({"kind": kind, "a": q[0], "b": q[1]})
            elif kind == "rzz":
                ins.append({"kind": kind, "a": q[0], "b": q[1], "theta": float(rng.uniform(0, 3))})
            elif kind == "ccx":
                ins.append({"kind": kind, "control1": q[0], "control2": q[1], "target": q[2]})
            elif kind == "cswap":
                ins.append({"kind": kind, "control": q[0], "target1": q[1], "target2": q[2]})
            elif kind == "c3x":
                ins.append({"kind": kind, "control1": q[0], "control2": q[1], "control3": q[2], "target": q[3]})
            else:
                ins.append({"kind": kind, "control1": q[0], "control2": q[1], "control3": q[2], "control4": q[3],
                            "target": q[4]})
    return Circuit(n, ins)
