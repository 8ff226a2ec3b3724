"""Python access to the CPU FP64 oracle (libsv_oracle.so) -- TEST INFRASTRUCTURE ONLY.

Only tests, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` import this module, and only as the checker. The product path
(``dqsim`` / ``libdqsim_sv.so``) never imports, links or calls it.

The C library restates the reference algorithm (see sv_oracle.h for the
file:line list). This module converts the JSON transport (src/types.rs:60-296)
into the oracle's flat instruction records and wraps the entry points.
"""

from __future__ import annotations

import ctypes
import json
import os
import subprocess
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libsv_oracle.so")

KINDS = [
    "id", "x", "y", "z", "h", "s", "sdg", "t", "tdg", "sx", "sxdg",
    "u3", "u2", "u1", "u", "p", "rx", "ry", "rz", "u0",
    "cx", "cz", "cy", "ch", "swap", "csx",
    "crx", "cry", "crz", "cu1", "cp", "cu3", "cu", "rxx", "rzz",
    "ccx", "cswap", "rccx", "rc3x", "c3x", "c3sqrtx", "c4x",
    "gate", "measure", "reset", "barrier",
]
KIND = {k: i for i, k in enumerate(KINDS)}

# Operand fields in Instruction::qubits() order (src/types.rs:346-464).
OPERANDS = {
    **{k: ("qubit",) for k in ("id", "x", "y", "z", "h", "s", "sdg", "t", "tdg", "sx", "sxdg", "u3", "u2", "u1",
                               "u", "p", "rx", "ry", "rz", "u0", "measure", "reset")},
    **{k: ("control", "target") for k in ("cx", "cz", "cy", "ch", "csx", "crx", "cry", "crz", "cu1", "cp", "cu3",
                                          "cu")},
    **{k: ("a", "b") for k in ("swap", "rxx", "rzz")},
    "ccx": ("control1", "control2", "target"),
    "rccx": ("control1", "control2", "target"),
    "cswap": ("control", "target1", "target2"),
    "rc3x": ("control1", "control2", "control3", "target"),
    "c3x": ("control1", "control2", "control3", "target"),
    "c3sqrtx": ("control1", "control2", "control3", "target"),
    "c4x": ("control1", "control2", "control3", "control4", "target"),
    "barrier": (),
}


class Inst(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("nq", ctypes.c_int32),
        ("q", ctypes.c_uint64 * 8),
        ("theta", ctypes.c_double),
        ("phi", ctypes.c_double),
        ("lam", ctypes.c_double),
        ("gamma", ctypes.c_double),
        ("params", ctypes.c_double * 2),
        ("nparams", ctypes.c_int32),
        ("name", ctypes.c_char * 32),
        ("cbit", ctypes.c_uint64),
        ("ncond", ctypes.c_int32),
        ("cbase", ctypes.c_uint64 * 8),
        ("csize", ctypes.c_uint64 * 8),
        ("cvalue", ctypes.c_uint64 * 8),
    ]


class ShotInfo(ctypes.Structure):
    _fields_ = [
        ("strategy", ctypes.c_int32),
        ("parallel_shots", ctypes.c_uint64),
        ("working_bytes", ctypes.c_uint64),
        ("prefix_ops", ctypes.c_uint64),
    ]


STRATEGIES = {0: "empty", 1: "terminal_sampling", 2: "trajectories"}


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _load():
    if not os.path.exists(LIB):
        build()
    lib = ctypes.CDLL(LIB)
    P, U64, I32, I64, D = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int32, ctypes.c_int64, ctypes.c_double
    lib.or_uniform53.restype = D
    lib.or_uniform53.argtypes = [U64, U64, U64]
    lib.or_philox.argtypes = [P, ctypes.c_uint32, ctypes.c_uint32]
    lib.or_gate_1q.restype = ctypes.c_int
    lib.or_gate_1q.argtypes = [ctypes.c_int, D, D, D, P]
    lib.or_compile_kinds.restype = I64
    lib.or_compile_kinds.argtypes = [P, I64, P, I64]
    lib.or_simulate.restype = ctypes.c_int
    lib.or_simulate.argtypes = [P, I64, ctypes.c_int, U64, U64, P, P]
    lib.or_run_shots.restype = ctypes.c_int
    lib.or_run_shots.argtypes = [P, I64, ctypes.c_int, U64, U64, U64, U64, ctypes.c_int, U64, ctypes.c_int, U64,
                                 ctypes.c_int, P, ctypes.POINTER(ShotInfo)]
    lib.or_apply_one.argtypes = [P, ctypes.c_int, ctypes.c_uint32, P, P, P, ctypes.c_int]
    lib.or_apply_x.argtypes = [P, ctypes.c_int, ctypes.c_uint32, P, P, ctypes.c_int]
    lib.or_apply_swap.argtypes = [P, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32, P, P, ctypes.c_int]
    lib.or_apply_diagonal.argtypes = [P, ctypes.c_int, P, ctypes.c_int, P]
    lib.or_apply_dense.argtypes = [P, ctypes.c_int, P, ctypes.c_int, P]
    lib.or_measure.restype = ctypes.c_int
    lib.or_measure.argtypes = [P, ctypes.c_int, ctypes.c_uint32, D, ctypes.POINTER(D), ctypes.POINTER(D)]
    lib.or_measure_f32.restype = ctypes.c_int
    lib.or_measure_f32.argtypes = [P, ctypes.c_int, ctypes.c_uint32, D, ctypes.POINTER(D), ctypes.POINTER(D)]
    lib.or_ptree.restype = D
    lib.or_ptree.argtypes = [P, U64]
    lib.or_marginal.argtypes = [P, ctypes.c_int, P, ctypes.c_int, P]
    lib.or_sample_marginal.restype = ctypes.c_int
    lib.or_sample_marginal.argtypes = [P, ctypes.c_int, P, ctypes.c_int, U64, U64, P]
    lib.or_fidelity.restype = D
    lib.or_fidelity.argtypes = [P, P, ctypes.c_int]
    lib.or_set_threads.argtypes = [ctypes.c_int]
    lib.or_cdf.restype = D
    lib.or_cdf.argtypes = [P, U64, P]
    lib.or_batch_sample.restype = ctypes.c_int
    lib.or_batch_sample.argtypes = [P, ctypes.c_int, U64, U64, ctypes.c_int, P]
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def set_threads(t: int) -> None:
    lib().or_set_threads(int(t))


# ---------------------------------------------------------------------------
# Transport conversion
# ---------------------------------------------------------------------------
def _convert(d: dict, conds: List[Tuple[int, int, int]]) -> Inst:
    kind = d["kind"]
    if kind == "conditional":
        c = d["condition"]
        return _convert(d["op"], conds + [(c["creg_base"], c["creg_size"], c["creg_value"])])
    rec = Inst()
    rec.kind = KIND[kind]
    if kind == "gate":
        qs = list(d["qubits"])
        rec.name = d["name"].lower().encode()[:31]
        ps = list(d["params"])
        rec.nparams = len(ps)
        for i, p in enumerate(ps[:2]):
            rec.params[i] = p
    else:
        qs = [d[f] for f in OPERANDS[kind]]
    rec.nq = len(qs)
    for i, q in enumerate(qs[:8]):
        rec.q[i] = q
    rec.theta = d.get("theta", 0.0)
    rec.phi = d.get("phi", 0.0)
    rec.lam = d.get("lam", 0.0)
    rec.gamma = d.get("gamma", 0.0)
    rec.cbit = d.get("cbit", 0)
    rec.ncond = len(conds)
    for i, (b, s, v) in enumerate(conds):
        rec.cbase[i], rec.csize[i], rec.cvalue[i] = b, s, v
    return rec


def to_insts(instructions: Sequence[dict]):
    arr = (Inst * max(1, len(instructions)))()
    for i, d in enumerate(instructions):
        arr[i] = _convert(d, [])
    return arr, len(instructions)


def circuit_dims(data: dict) -> Tuple[int, int]:
    n = max((r["base"] + r["size"] for r in data["qregs"].values()), default=0)
    nc = max((r["base"] + r["size"] for r in data.get("cregs", {}).values()), default=0)
    return n, nc


def _data(circuit) -> dict:
    if isinstance(circuit, dict):
        return circuit
    return json.loads(circuit.model_dump_json())


# ---------------------------------------------------------------------------
# Entry points
# ---------------------------------------------------------------------------
def simulate(circuit, seed: int) -> Tuple[np.ndarray, Dict[int, int]]:
    data = _data(circuit)
    n, nc = circuit_dims(data)
    arr, count = to_insts(data["instructions"])
    state = np.empty(1 << n, dtype=np.complex128)
    cb = (ctypes.c_uint8 * max(1, nc))()
    rc = lib().or_simulate(arr, count, n, nc, seed, state.ctypes.data, cb)
    if rc:
        raise RuntimeError(f"oracle simulate failed ({rc})")
    return state, {i: int(cb[i]) for i in range(nc)}


def run_shots(circuit, shots: int, seed: int, budget_mb: int = 1024, max_parallel_shots: Optional[int] = None,
              sample_terminal: bool = True, prec: int = 16, max_batch: int = 1024):
    data = _data(circuit)
    n, nc = circuit_dims(data)
    arr, count = to_insts(data["instructions"])
    ncs = max(1, nc)
    bits = np.zeros((max(1, shots), ncs), dtype=np.uint8)
    info = ShotInfo()
    rc = lib().or_run_shots(arr, count, n, nc, shots, seed, budget_mb << 20, 0 if max_parallel_shots is None else 1,
                            max_parallel_shots or 0, 1 if sample_terminal else 0, max_batch, prec,
                            bits.ctypes.data, ctypes.byref(info))
    if rc == 2:
        raise MemoryError("oracle: budget exceeded")
    if rc:
        raise RuntimeError(f"oracle run_shots failed ({rc})")
    counts: Dict[str, int] = {}
    for s in range(shots):
        key = "".join(str(int(b)) for b in bits[s, :nc][::-1])
        counts[key] = counts.get(key, 0) + 1
    return counts, {"execution_strategy": STRATEGIES[info.strategy], "parallel_shots": info.parallel_shots,
                    "working_bytes": info.working_bytes, "deterministic_prefix_ops": info.prefix_ops}


def marginal(state: np.ndarray, qubits: Sequence[int]) -> np.ndarray:
    state = np.ascontiguousarray(state, dtype=np.complex128)
    n = int(state.size).bit_length() - 1
    q = (ctypes.c_uint64 * max(1, len(qubits)))(*qubits)
    out = np.empty(1 << len(qubits), dtype=np.float64)
    lib().or_marginal(state.ctypes.data, n, q, len(qubits), out.ctypes.data)
    return out


def sample_marginal(state: np.ndarray, qubits: Sequence[int], shots: int, seed: int) -> np.ndarray:
    state = np.ascontiguousarray(state, dtype=np.complex128)
    n = int(state.size).bit_length() - 1
    q = (ctypes.c_uint64 * max(1, len(qubits)))(*qubits)
    out = np.empty(max(1, shots), dtype=np.uint64)
    if lib().or_sample_marginal(state.ctypes.data, n, q, len(qubits), shots, seed, out.ctypes.data):
        raise RuntimeError("oracle: invalid normalisation")
    return out[:shots]


def counts_from_indices(idx: np.ndarray, width: int) -> Dict[str, int]:
    out: Dict[str, int] = {}
    for x in idx.tolist():
        key = format(int(x), f"0{width}b") if width else ""
        out[key] = out.get(key, 0) + 1
    return out


def fidelity(a: np.ndarray, b: np.ndarray) -> float:
    a = np.ascontiguousarray(a, dtype=np.complex128)
    b = np.ascontiguousarray(b, dtype=np.complex128)
    n = int(a.size).bit_length() - 1
    return lib().or_fidelity(a.ctypes.data, b.ctypes.data, n)


def uniform53(seed: int, stream: int, draw: int) -> float:
    return lib().or_uniform53(seed, stream, draw)


def philox(ctr: Sequence[int], key: Sequence[int]) -> List[int]:
    c = (ctypes.c_uint32 * 4)(*ctr)
    lib().or_philox(c, key[0], key[1])
    return list(c)


def gate_1q(kind: str, theta=0.0, phi=0.0, lam=0.0) -> np.ndarray:
    out = (ctypes.c_double * 8)()
    if not lib().or_gate_1q(KIND[kind], theta, phi, lam, out):
        raise ValueError(kind)
    return np.array([complex(out[2 * i], out[2 * i + 1]) for i in range(4)]).reshape(2, 2)


def compile_kinds(instructions: Sequence[dict]) -> List[str]:
    arr, count = to_insts(instructions)
    kinds = (ctypes.c_int32 * max(1, 4 * count + 8))()
    m = lib().or_compile_kinds(arr, count, kinds, 4 * count + 8)
    if m < 0:
        raise ValueError("oracle compile failed")
    names = ["one", "x", "swap", "diagonal", "dense", "measure", "reset", "conditional"]
    return [names[kinds[i]] for i in range(m)]


# Kernel-level helpers (complex128 numpy arrays, modified in place).
def _ctl(controls):
    cq = (ctypes.c_uint32 * 5)(*[c for c, _ in controls])
    cv = (ctypes.c_uint8 * 5)(*[1 if v else 0 for _, v in controls])
    return cq, cv, len(controls)


def apply_one(state, target, u, controls=()):
    u = np.asarray(u, dtype=np.complex128).reshape(4)
    ud = (ctypes.c_double * 8)(*[x for c in u for x in (c.real, c.imag)])
    n = state.size.bit_length() - 1
    lib().or_apply_one(state.ctypes.data, n, target, ud, *_ctl(controls))


def apply_x(state, target, controls=()):
    n = state.size.bit_length() - 1
    lib().or_apply_x(state.ctypes.data, n, target, *_ctl(controls))


def apply_swap(state, a, b, controls=()):
    n = state.size.bit_length() - 1
    lib().or_apply_swap(state.ctypes.data, n, a, b, *_ctl(controls))


def apply_diagonal(state, qubits, diag):
    n = state.size.bit_length() - 1
    d = np.ascontiguousarray(diag, dtype=np.complex128)
    q = (ctypes.c_uint32 * 5)(*qubits)
    lib().or_apply_diagonal(state.ctypes.data, n, q, len(qubits), d.ctypes.data)


def apply_dense(state, qubits, mat):
    n = state.size.bit_length() - 1
    m = np.ascontiguousarray(mat, dtype=np.complex128)
    q = (ctypes.c_uint32 * 5)(*qubits)
    lib().or_apply_dense(state.ctypes.data, n, q, len(qubits), m.ctypes.data)


def measure(state, qubit, r, f32=False):
    """measure_qubit with draw r (in place). f32: the state holds complex64
    values and the collapse rounds the kept amplitudes to FP32."""
    n = state.size.bit_length() - 1
    p0, p1 = ctypes.c_double(), ctypes.c_double()
    fn = lib().or_measure_f32 if f32 else lib().or_measure
    out = fn(state.ctypes.data, n, qubit, r, ctypes.byref(p0), ctypes.byref(p1))
    return out, p0.value, p1.value


def terminal_trajectories(prefix: np.ndarray, measures: Sequence[Tuple[int, int]], nc: int, shots: int, seed: int,
                          f32: bool = False) -> Dict[str, int]:
    """Trajectory shots (shots.rs:178-259) of a circuit whose unitary prefix
    has already been evolved to `prefix` and whose remaining instructions are
    the unconditional measurements `measures` = [(qubit, cbit), ...]: shot s
    copies the prefix and measures in order with draws (seed, stream s,
    draw 0, 1, ...). f32 collapses as a complex64 state does."""
    counts: Dict[str, int] = {}
    for s in range(shots):
        st = np.array(prefix, dtype=np.complex128, copy=True)
        bits = [0] * max(1, nc)
        for d, (q, c) in enumerate(measures):
            out, _, _ = measure(st, q, uniform53(seed, s, d), f32=f32)
            if out < 0:
                raise RuntimeError("oracle: invalid normalisation")
            bits[c] = out
        key = "".join(str(b) for b in bits[:nc][::-1])
        counts[key] = counts.get(key, 0) + 1
    return counts


def ptree(x: np.ndarray) -> float:
    x = np.ascontiguousarray(x, dtype=np.float64)
    return lib().or_ptree(x.ctypes.data, x.size)


def cdf(x: np.ndarray) -> np.ndarray:
    """The canonical inclusive prefix the samplers use."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty(max(1, x.size), dtype=np.float64)
    if x.size:
        lib().or_cdf(x.ctypes.data, x.size, out.ctypes.data)
    return out[: x.size]


def batch_sample(state: np.ndarray, shots: int, seed: int, block: int) -> np.ndarray:
    """Sample indices exactly as the batched-circuit kernel does."""
    state = np.ascontiguousarray(state, dtype=np.complex128)
    n = int(state.size).bit_length() - 1
    out = np.empty(max(1, shots), dtype=np.uint64)
    lib().or_batch_sample(state.ctypes.data, n, shots, seed, block, out.ctypes.data)
    return out[:shots]
