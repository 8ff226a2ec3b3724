"""P-block backend: PBlockSimulator / PBlockResult and the distributed
wrappers, over the engine's C ABI (dq_pb_*).

Follows the reference's ``src/distributed/pblock/`` (simulator.rs:24-322,
result.rs:21-293) and, for distributed inputs, ``input.rs:54-224``: the
per-node circuits are merged here into one circuit (one quantum register per
physical qubit, the classical ranges of every node, instructions in global
order) and the engine runs that circuit. All amplitude work happens on the
GPU; this module only validates and reshapes inputs and results.
"""

import ctypes
import json
import random
from typing import Dict, List, Optional

import numpy as np

from . import _core
from ._core import _as_bool, _as_u64, _check, _circuit_json, _CountTable, _lib, _u64_array

_USIZE_MAX = (1 << 64) - 1


class _PbOpts(ctypes.Structure):
    _fields_ = [
        ("max_memory_mb", ctypes.c_uint64),
        ("has_max_block_qubits", ctypes.c_int32),
        ("max_block_qubits", ctypes.c_uint64),
        ("max_parallel_shots", ctypes.c_uint64),
        ("sample_terminal", ctypes.c_int32),
        ("split_separable", ctypes.c_int32),
        ("max_split_qubits", ctypes.c_uint64),
        ("separation_tolerance", ctypes.c_double),
        ("profile", ctypes.c_int32),
        ("has_seed", ctypes.c_int32),
        ("seed", ctypes.c_uint64),
    ]


def _bind():
    P, SZ, U64, I32 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint64, ctypes.c_int32
    D = ctypes.c_double
    OP = ctypes.POINTER(_PbOpts)
    sigs = {
        "dq_pb_opts_default": (None, [OP]),
        "dq_pb_check_opts": (ctypes.c_int, [OP]),
        "dq_pb_simulate": (ctypes.c_int, [OP, ctypes.c_char_p, SZ, ctypes.POINTER(P)]),
        "dq_pb_simulate_shots": (ctypes.c_int, [OP, ctypes.c_char_p, SZ, U64, I32, _core._COUNT_CB, P]),
        "dq_pb_result_free": (None, [P]),
        "dq_pb_result_json": (ctypes.c_int, [P, I32, ctypes.POINTER(P)]),
        "dq_pb_statevector_len": (ctypes.c_int, [P, ctypes.POINTER(U64)]),
        "dq_pb_statevector": (ctypes.c_int, [P, P, U64]),
        "dq_pb_amplitude": (ctypes.c_int, [P, ctypes.c_char_p, SZ, ctypes.POINTER(D)]),
        "dq_pb_probabilities": (ctypes.c_int, [P, P, U64, I32, ctypes.POINTER(P)]),
        "dq_pb_counts": (ctypes.c_int, [P, U64, P, U64, I32, U64, _core._COUNT_CB, P]),
        "dq_pb_expectation": (ctypes.c_int, [P, ctypes.c_char_p, SZ, ctypes.POINTER(D)]),
        "dq_pb_last_shots_stats": (ctypes.c_int, [P]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(_lib, name)
        fn.restype = res
        fn.argtypes = args


_bind()


def _take_json(ptr: ctypes.c_void_p):
    try:
        return json.loads(ctypes.string_at(ptr).decode("utf-8"))
    finally:
        _lib.dq_string_free(ptr)


# ---------------------------------------------------------------------------
# Distributed transport (input.rs:54-224)
# ---------------------------------------------------------------------------
def _ranges(regs, kind: str):
    out = []
    for key, r in (regs or {}).items():
        base, size = int(r["base"]), int(r["size"])
        if base > _USIZE_MAX - size:
            raise ValueError(f'{kind} register "{key}" has an overflowing range')
        if size:
            out.append((base, base + size))
    out.sort()
    for a, b in zip(out, out[1:]):
        if b[0] < a[1]:
            raise ValueError(f"Overlapping {kind} registers")
    return out


def _parse(circuit) -> dict:
    try:
        return json.loads(_circuit_json(circuit).decode("utf-8"))
    except json.JSONDecodeError as e:
        raise ValueError(f"Circuit JSON parse error: {e}") from None


def _transport(obj, budget: int) -> bytes:
    """Circuit JSON for the engine: a plain circuit as is, a distributed
    input merged into one circuit in global operation order."""
    if not hasattr(obj, "circuits"):
        return _circuit_json(obj)
    qpn = {int(k): [int(q) for q in v] for k, v in dict(obj.qubits_per_node).items()}
    node_objects = obj.circuits
    if not isinstance(node_objects, dict):
        raise TypeError("'circuits' must be a dict")
    nodes = sorted(int(k) for k in node_objects.keys())
    if set(nodes) != set(qpn):
        raise ValueError("circuits and qubits_per_node must contain the same node IDs")
    owned = set()
    for qs in qpn.values():
        for q in qs:
            if q == _USIZE_MAX or q in owned:
                raise ValueError("Physical qubits must have unique ownership and representable indices")
            owned.add(q)
    if len(owned) * 256 > budget:
        raise MemoryError("P-block wire metadata exceeds max_memory_mb")
    physical = sorted(owned)
    explicit = hasattr(obj, "operation_ids") or hasattr(obj, "operation_order")
    ids = {int(k): list(v) for k, v in dict(obj.operation_ids).items()} if explicit else {}
    order = dict(obj.operation_order) if explicit else {}
    if explicit and set(ids) != set(nodes):
        raise ValueError("operation_ids must list every circuit node")
    legacy = {} if explicit else {int(k): int(v) for k, v in dict(obj._instruction_index).items()}
    events: Dict[str, tuple] = {}
    orders: Dict[int, str] = {}
    classical = set()
    quantum = []
    for node in nodes:
        c_obj = node_objects[node]
        circuit = _parse(c_obj)
        quantum.extend(_ranges(circuit.get("qregs"), "quantum"))
        classical.update(_ranges(circuit.get("cregs"), "classical"))
        insts = circuit.get("instructions", [])
        if explicit:
            if len(ids[node]) != len(insts):
                raise ValueError("operation_ids length must match serialized instructions")
            legacy_ids = []
        else:
            objects = list(c_obj.instructions)
            if len(objects) != len(insts):
                raise ValueError("Python instruction list and serialized instructions have different lengths")
            legacy_ids = [id(o) for o in objects]
        seen = set()
        previous = None
        for i, inst in enumerate(insts):
            if explicit:
                key = str(ids[node][i])
                if key not in order:
                    raise ValueError(f'Missing order for operation "{key}"')
                position = int(order[key])
            else:
                if legacy_ids[i] not in legacy:
                    raise ValueError("Missing instruction order; supply complete _instruction_index or "
                                     "explicit operation IDs")
                key, position = str(legacy_ids[i]), legacy[legacy_ids[i]]
            if key in seen:
                raise ValueError("Repeated operation ID within one node; each occurrence requires its own "
                                 "explicit ID")
            seen.add(key)
            if previous is not None and position <= previous:
                raise ValueError("Global operation order contradicts the node instruction order")
            previous = position
            if key in events:
                if events[key][0] != position or events[key][1] != inst:
                    raise ValueError("Copies of a shared operation must agree on order and contents")
            else:
                if position in orders:
                    raise ValueError("Ambiguous global order: different operations have the same position")
                orders[position] = key
                events[key] = (position, inst)
    for q in physical:
        if not any(a <= q < b for a, b in quantum):
            raise ValueError(f"Physical qubit {q} has no register declaration")
    merged = {
        "qregs": {str(q): {"name": str(q), "base": q, "size": 1} for q in physical},
        "cregs": {str(i): {"name": str(i), "base": a, "size": b - a} for i, (a, b) in enumerate(sorted(classical))},
        "instructions": [events[orders[p]][1] for p in sorted(orders)],
    }
    return json.dumps(merged).encode("utf-8")


# ---------------------------------------------------------------------------
# Simulator and result
# ---------------------------------------------------------------------------
class PBlockResult:
    """Product-of-blocks result; local bit i of the dense export is
    physical_qubits[i] (result.rs:21-293)."""

    def __init__(self, handle):
        self._h = handle

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            _lib.dq_pb_result_free(h)
            self._h = None

    def _json(self, what: int):
        out = ctypes.c_void_p()
        _check(_lib.dq_pb_result_json(self._h, what, ctypes.byref(out)))
        return _take_json(out)

    @property
    def num_qubits(self) -> int:
        return len(self.physical_qubits)

    @property
    def physical_qubits(self) -> List[int]:
        return self._json(0)

    @property
    def classical_bits(self) -> Dict[int, int]:
        return {int(c): int(v) for c, v in self._json(1)}

    @property
    def block_qubits(self) -> List[List[int]]:
        return self._json(2)

    @property
    def diagnostics(self) -> dict:
        return self._json(3)

    @property
    def profile(self) -> Optional[dict]:
        return self._json(4)

    @property
    def statevector(self) -> np.ndarray:
        n = ctypes.c_uint64()
        _check(_lib.dq_pb_statevector_len(self._h, ctypes.byref(n)))
        out = np.empty(n.value, dtype=np.complex128)
        _check(_lib.dq_pb_statevector(self._h, out.ctypes.data_as(ctypes.c_void_p), n.value))
        return out

    def amplitude(self, bitstring: str) -> complex:
        if not isinstance(bitstring, str):
            raise TypeError("argument 'bitstring': expected str")
        data = bitstring.encode("utf-8")
        v = (ctypes.c_double * 2)()
        _check(_lib.dq_pb_amplitude(self._h, data, len(data), v))
        return complex(v[0], v[1])

    def probabilities(self, qubits=None) -> Dict[int, float]:
        qs = [] if qubits is None else _core._qubit_list(qubits)
        arr = _u64_array(qs)
        out = ctypes.c_void_p()
        _check(_lib.dq_pb_probabilities(self._h, arr, len(qs), 1 if qubits is None else 0, ctypes.byref(out)))
        return {int(i): float(p) for i, p in _take_json(out)}

    def counts(self, shots=1000, qubits=None, seed=None) -> Dict[str, int]:
        shots = _as_u64(shots, "shots")
        qs = [] if qubits is None else _core._qubit_list(qubits)
        s = random.getrandbits(64) if seed is None else _as_u64(seed, "seed")
        arr = _u64_array(qs)
        tab = _CountTable()
        _check(_lib.dq_pb_counts(self._h, shots, arr, len(qs), 1 if qubits is None else 0, s, tab.cb, tab.user))
        return tab.result()

    def expectation_value(self, pauli: str) -> float:
        if not isinstance(pauli, str):
            raise TypeError("argument 'pauli': expected str")
        data = pauli.encode("utf-8")
        v = ctypes.c_double()
        _check(_lib.dq_pb_expectation(self._h, data, len(data), ctypes.byref(v)))
        return v.value


class PBlockSimulator:
    """Product of independent dense blocks on the GPU (simulator.rs:24-322)."""

    def __init__(self, seed=None, *, max_memory_mb=1024, max_block_qubits=None, max_parallel_shots=1,
                 sample_terminal=True, split_separable=False, max_split_qubits=12, separation_tolerance=0.0,
                 profile=False):
        o = _PbOpts()
        _lib.dq_pb_opts_default(ctypes.byref(o))
        o.max_memory_mb = _as_u64(max_memory_mb, "max_memory_mb")
        if max_block_qubits is not None:
            o.has_max_block_qubits = 1
            o.max_block_qubits = _as_u64(max_block_qubits, "max_block_qubits")
        o.max_parallel_shots = _as_u64(max_parallel_shots, "max_parallel_shots")
        o.sample_terminal = 1 if _as_bool(sample_terminal, "sample_terminal") else 0
        o.split_separable = 1 if _as_bool(split_separable, "split_separable") else 0
        o.max_split_qubits = _as_u64(max_split_qubits, "max_split_qubits")
        o.separation_tolerance = float(separation_tolerance)
        o.profile = 1 if _as_bool(profile, "profile") else 0
        if seed is not None:
            o.has_seed = 1
            o.seed = _as_u64(seed, "seed")
        _check(_lib.dq_pb_check_opts(ctypes.byref(o)))
        self._o = o

    def simulate(self, distributed) -> PBlockResult:
        data = _transport(distributed, self._o.max_memory_mb << 20)
        h = ctypes.c_void_p()
        _check(_lib.dq_pb_simulate(ctypes.byref(self._o), data, len(data), ctypes.byref(h)))
        return PBlockResult(h)

    def simulate_shots(self, distributed, shots=1000, profile=False) -> Dict[str, int]:
        shots = _as_u64(shots, "shots")
        profile = _as_bool(profile, "profile")
        data = _transport(distributed, self._o.max_memory_mb << 20)
        tab = _CountTable()
        _check(_lib.dq_pb_simulate_shots(ctypes.byref(self._o), data, len(data), shots, 1 if profile else 0, tab.cb,
                                         tab.user))
        return tab.result()


def last_shots_stats() -> Dict[str, int]:
    """How this thread's last PBlockSimulator.simulate_shots ran its
    trajectory shots (no reference counterpart): shots completed in lockstep
    batches, shots run alone, the largest batch's accounted bytes (never above
    max_memory_mb) and its slot count."""
    out = (ctypes.c_uint64 * 4)()
    _check(_lib.dq_pb_last_shots_stats(out))
    return {"lockstep_shots": int(out[0]), "lone_shots": int(out[1]), "peak_batch_bytes": int(out[2]),
            "batch_size": int(out[3])}


_PB_OPTIONS = ("max_memory_mb", "max_block_qubits", "max_parallel_shots", "sample_terminal", "split_separable",
               "max_split_qubits", "separation_tolerance")


def _pblock(mode: str, seed, profile, options) -> PBlockSimulator:
    m = str(mode).strip().lower().replace("-", "_")
    if m != "p_block":
        raise ValueError(f"Unsupported distributed simulation mode \"{m}\"; expected 'p_block'")
    for key in options:
        if key not in _PB_OPTIONS:
            raise TypeError(f"Unsupported P-block option \"{key}\"")
    return PBlockSimulator(seed, profile=profile, **options)


def simulate_distributed(distributed, mode="p_block", seed=None, profile=False, **options):
    """src/simulator.rs:67-92"""
    return _pblock(mode, seed, profile, options).simulate(distributed)


def simulate_distributed_shots(distributed, mode="p_block", shots=1000, seed=None, profile=False, **options):
    """src/simulator.rs:148-175"""
    return _pblock(mode, seed, profile, options).simulate_shots(distributed, shots, profile)
