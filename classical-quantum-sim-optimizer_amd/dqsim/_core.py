"""Drop-in replacement for the reference's native module ``dqsim._core``.

The reference builds ``_core`` with PyO3 (src/lib.rs:16-32). This module keeps
its Python surface for the statevector backend -- class and function names,
keyword names and defaults, keyword-only markers and exception types -- and
forwards every call to the C ABI of ``libdqsim_sv.so`` (include/dqsim_sv.h),
which runs the simulation on the GPU. ctypes releases the GIL around each
foreign call, as the reference does around compile, execute and queries
(src/monolithic/statevector/simulator.rs:143,169,272,278,332,336).

Reference API mirrored here:
  StatevectorSimulator ......... src/monolithic/statevector/simulator.rs:213-371
  SimulationResult ............. simulator.rs:85-202
  SimulationProfile ............ simulator.rs:43-79
  simulate_monolithic(_shots) .. src/simulator.rs:14-35,94-116
  option / mode parsing ........ src/codecs.rs:23-64
The other backends of the reference (MPS, stabilizer, P-block) are not part of
this engine; their names exist so ``import dqsim`` works, and using them
raises NotImplementedError.

Additive keywords (not in the reference): ``precision`` ("complex128" default,
or "complex64"), ``fusion`` ("none" default, "greedy", "learned") and
``max_fused_qubits`` (1..5, default 4). The defaults reproduce the reference's
FP64 results.
"""

from __future__ import annotations

import ctypes
import json
import os
from typing import Any, Dict, Iterable, List, Optional

import numpy as np

_LIB_NAME = "libdqsim_sv.so"

DQ_OK, DQ_EVALUE, DQ_EMEMORY, DQ_ETYPE, DQ_ERUNTIME, DQ_EDEVICE = range(6)
_PRECISIONS = {"complex128": 16, "complex64": 8}
_FUSIONS = {"none": 0, "greedy": 1, "learned": 2}
_U64_MAX = (1 << 64) - 1


class _Opts(ctypes.Structure):
    _fields_ = [
        ("seed", ctypes.c_uint64),
        ("has_seed", ctypes.c_int32),
        ("profile", ctypes.c_int32),
        ("max_memory_mb", ctypes.c_uint64),
        ("max_parallel_shots", ctypes.c_uint64),
        ("has_max_parallel_shots", ctypes.c_int32),
        ("sample_terminal", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("fusion", ctypes.c_int32),
        ("max_fused_qubits", ctypes.c_int32),
        ("tile_qubits", ctypes.c_int32),
    ]


class _Prof(ctypes.Structure):
    _fields_ = [
        ("apply_one_qubit_calls", ctypes.c_uint64),
        ("apply_one_qubit_time", ctypes.c_double),
        ("apply_n_qubit_calls", ctypes.c_uint64),
        ("apply_n_qubit_time", ctypes.c_double),
        ("measure_qubit_calls", ctypes.c_uint64),
        ("measure_qubit_time", ctypes.c_double),
        ("total_time", ctypes.c_double),
    ]


class OpTiming(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("k", ctypes.c_int32),
        ("low", ctypes.c_int32),
        ("gates", ctypes.c_int32),
        ("ms", ctypes.c_double),
        ("bytes", ctypes.c_double),
        ("feat", ctypes.c_float * 16),
    ]


class _BatchOpts(ctypes.Structure):
    _fields_ = [
        ("workers", ctypes.c_int32),
        ("max_fused_qubits", ctypes.c_int32),
        ("autotune", ctypes.c_int32),
        ("tile_qubits", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("dump_states", ctypes.c_int32),
        ("persistent", ctypes.c_int32),
        ("lds_qubits", ctypes.c_int32),
    ]


_COUNT_CB = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(ctypes.c_char), ctypes.c_size_t,
                             ctypes.c_uint64)


def library_path() -> str:
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), _LIB_NAME)


def _share_hip_runtime():
    """Keep one HIP runtime per process. PyTorch's ROCm libraries request the
    runtime as ``libamdhip64.so`` (their bundled copy); this library requests
    it by its soname ``libamdhip64.so.7``. Two copies in one process cannot
    both open the GPU: the second to initialise sees no device. Loading the
    runtime under the unversioned name first makes every later request bind
    to one copy: if torch is already loaded this returns torch's copy (and
    the library's soname request binds to it); otherwise the system runtime
    is loaded under that name and a later ``import torch`` binds to it."""
    for name in ("libhsa-runtime64.so", "libamdhip64.so"):  # the HSA runtime under HIP likewise
        try:
            ctypes.CDLL(name, mode=ctypes.RTLD_GLOBAL)
        except OSError:
            pass  # not on the search path under that name: the soname request decides


def _load():
    path = library_path()
    if not os.path.exists(path):
        raise ImportError(f"{path} is not built; build it with __graft_entry__.build() or `make` in the package")
    _share_hip_runtime()
    lib = ctypes.CDLL(path)
    P, U64, I32, SZ, D = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int32, ctypes.c_size_t, ctypes.c_double
    sigs = {
        "dq_sv_opts_default": (None, [ctypes.POINTER(_Opts)]),
        "dq_sv_create": (ctypes.c_int, [ctypes.POINTER(_Opts), ctypes.POINTER(P)]),
        "dq_sv_free": (None, [P]),
        "dq_sv_simulate_shots": (ctypes.c_int, [P, ctypes.c_char_p, SZ, U64, I32, _COUNT_CB, P]),
        "dq_counts_new": (P, []),
        "dq_counts_view": (ctypes.c_int, [P, ctypes.POINTER(U64), ctypes.POINTER(ctypes.c_void_p),
                                          ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_void_p)]),
        "dq_counts_free": (None, [P]),
        "dq_sv_simulate": (ctypes.c_int, [P, ctypes.c_char_p, SZ, ctypes.POINTER(P)]),
        "dq_sv_evolve_noalloc": (ctypes.c_int, [P, ctypes.c_char_p, SZ, P, U64, P]),
        "dq_result_num_qubits": (ctypes.c_uint32, [P]),
        "dq_result_num_cbits": (U64, [P]),
        "dq_result_classical_bits": (ctypes.c_int, [P, P, U64]),
        "dq_result_profile": (ctypes.c_int, [P, ctypes.POINTER(_Prof)]),
        "dq_result_statevector_c128": (ctypes.c_int, [P, P, U64]),
        "dq_result_probabilities": (ctypes.c_int, [P, P, U64, P]),
        "dq_result_counts": (ctypes.c_int, [P, U64, P, U64, I32, U64, _COUNT_CB, P]),
        "dq_result_fidelity": (ctypes.c_int, [P, P, U64, ctypes.POINTER(D)]),
        "dq_result_amplitude": (ctypes.c_int, [P, U64, P]),
        "dq_result_free": (None, [P]),
        "dq_result_gather": (ctypes.c_int, [P, P, U64, P]),
        "dq_result_compare": (ctypes.c_int, [P, P, ctypes.POINTER(D), ctypes.POINTER(D)]),
        "dq_last_error": (ctypes.c_char_p, []),
        "dq_plan_json": (ctypes.c_int, [ctypes.c_char_p, SZ, I32, I32, I32, I32, I32,
                                        ctypes.POINTER(ctypes.c_void_p)]),
        "dq_set_sweep_mode": (ctypes.c_int, [I32]),
        "dq_string_free": (None, [P]),
        "dq_device_count": (ctypes.c_int, []),
        "dq_mfma_fp32_peak": (ctypes.c_int, [I32, ctypes.POINTER(D), ctypes.POINTER(D)]),
        "dq_set_sweep_relabel": (ctypes.c_int, [I32]),
        "dq_set_sweep_pinned": (ctypes.c_int, [I32]),
        "dq_set_block_arith": (ctypes.c_int, [I32]),
        "dq_set_wave_runs": (ctypes.c_int, [I32]),
        "dq_set_zero_start": (ctypes.c_int, [I32]),
        "dq_split_f16": (ctypes.c_int, [P, U64, P, P]),
        "dq_sv_evolve_timed": (ctypes.c_int, [P, ctypes.c_char_p, SZ, I32, ctypes.POINTER(OpTiming), U64,
                                              ctypes.POINTER(U64), ctypes.POINTER(U64), ctypes.POINTER(D),
                                              ctypes.POINTER(P)]),
        "dq_set_planner_weights": (ctypes.c_int, [P, U64]),
        "dq_batch_run": (ctypes.c_int, [P, P, U64, P, P, ctypes.POINTER(_BatchOpts), ctypes.POINTER(P)]),
        "dq_batch_counts": (ctypes.c_int, [P, U64, _COUNT_CB, P]),
        "dq_batch_lds_info": (ctypes.c_int, [P, ctypes.POINTER(U64), ctypes.POINTER(D)]),
        "dq_batch_indices": (ctypes.c_int, [P, U64, P, U64]),
        "dq_batch_info": (ctypes.c_int, [P, ctypes.POINTER(D), ctypes.POINTER(I32), ctypes.POINTER(I32),
                                         ctypes.POINTER(I32), ctypes.POINTER(ctypes.c_uint32), P, U64]),
        "dq_batch_state_c128": (ctypes.c_int, [P, U64, P, U64]),
        "dq_batch_free": (None, [P]),
        "dq_batch_num_qubits": (ctypes.c_uint32, [P, U64]),
        "dq_batch_tuning_count": (U64, [P]),
        "dq_batch_tuning_entry": (ctypes.c_int, [P, U64, ctypes.c_char_p, U64, ctypes.POINTER(D)]),
        "dq_batch_device_info": (ctypes.c_int, [I32, P]),
        "dq_batch_program_text": (ctypes.c_int, [P, P, U64, P, P, I32, I32, ctypes.POINTER(ctypes.c_void_p)]),
        "dq_mlp_param_count": (U64, []),
        "dq_mlp_train": (ctypes.c_int, [P, P, P, U64, I32, I32, ctypes.c_float, U64, P]),
        "dq_mlp_grad": (ctypes.c_int, [P, P, P, I32, P, P]),
        "dq_mlp_predict": (ctypes.c_int, [P, P, U64, P]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_lib = _load()


def _raise(code: int):
    msg = (_lib.dq_last_error() or b"").decode("utf-8", "replace")
    if code == DQ_EVALUE:
        raise ValueError(msg)
    if code == DQ_EMEMORY:
        raise MemoryError(msg)
    if code == DQ_ETYPE:
        raise TypeError(msg)
    raise RuntimeError(msg)


def _check(code: int):
    if code != DQ_OK:
        _raise(code)


# ---------------------------------------------------------------------------
# Argument extraction with PyO3's conversion rules
# ---------------------------------------------------------------------------
def _as_u64(value, name: str) -> int:
    if isinstance(value, (bool, np.bool_)) or not hasattr(value, "__index__"):
        if isinstance(value, bool):
            return int(value)
        raise TypeError(f"argument '{name}': '{type(value).__name__}' object cannot be interpreted as an integer")
    v = value.__index__()
    if v < 0:
        raise OverflowError(f"argument '{name}': can't convert negative int to unsigned")
    if v > _U64_MAX:
        raise OverflowError(f"argument '{name}': int too big to convert")
    return v


def _as_bool(value, name: str) -> bool:
    if isinstance(value, (bool, np.bool_)):
        return bool(value)
    raise TypeError(f"argument '{name}': '{type(value).__name__}' object cannot be converted to 'PyBool'")


def _opt_u64(value, name: str) -> Optional[int]:
    return None if value is None else _as_u64(value, name)


def _circuit_json(circuit) -> bytes:
    text = circuit.model_dump_json()
    if not isinstance(text, str):
        raise TypeError(f"'{type(text).__name__}' object cannot be converted to 'PyString'")
    return text.encode("utf-8")


def _qubit_list(qubits, name="qubits") -> List[int]:
    if isinstance(qubits, (str, bytes)):
        raise TypeError(f"argument '{name}': Can't extract `str` to `Vec`")
    return [_as_u64(q, name) for q in qubits]


def _u64_array(values: List[int]):
    arr = (ctypes.c_uint64 * max(1, len(values)))(*values)
    return arr


class _CountTable:
    """Counts received into the library's packed table: `cb`/`user` go to an
    entry point taking a dq_count_cb, `result()` reads every outcome at once
    (no Python call per outcome)."""

    def __init__(self):
        self.user = _lib.dq_counts_new()
        if not self.user:
            raise MemoryError("count table allocation failed")
        self.cb = _COUNT_CB(ctypes.cast(_lib.dq_counts_append, ctypes.c_void_p).value)

    def result(self) -> Dict[str, int]:
        n, keys, offs, vals = ctypes.c_uint64(), ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        _check(_lib.dq_counts_view(self.user, ctypes.byref(n), ctypes.byref(keys), ctypes.byref(offs),
                                   ctypes.byref(vals)))
        m = n.value
        if m == 0:
            return {}
        o = list((ctypes.c_uint64 * (m + 1)).from_address(offs.value))
        v = list((ctypes.c_uint64 * m).from_address(vals.value))
        text = ctypes.string_at(keys.value, o[m]).decode("ascii")
        return {text[o[i]:o[i + 1]]: v[i] for i in range(m)}

    def __del__(self):
        if getattr(self, "user", None):
            _lib.dq_counts_free(self.user)
            self.user = None


# ---------------------------------------------------------------------------
# SimulationProfile (simulator.rs:43-79)
# ---------------------------------------------------------------------------
class SimulationProfile:
    """Kernel counters of one simulate() trajectory (device times, seconds)."""

    __slots__ = ("apply_one_qubit_calls", "apply_one_qubit_time", "apply_n_qubit_calls", "apply_n_qubit_time",
                 "measure_qubit_calls", "measure_qubit_time", "total_time")

    def __init__(self, p: _Prof):
        for name in self.__slots__:
            object.__setattr__(self, name, getattr(p, name))

    def __setattr__(self, name, value):
        raise AttributeError(f"attribute '{name}' of 'SimulationProfile' objects is not writable")

    def __repr__(self) -> str:
        total = max(self.total_time, 1e-9)
        return (
            f"SimulationProfile (total: {self.total_time * 1000.0:.2f} ms)\n"
            f"  apply_one_qubit : {self.apply_one_qubit_calls:4} calls  {self.apply_one_qubit_time * 1000.0:8.2f} ms"
            f"  ({100.0 * self.apply_one_qubit_time / total:.1f}%)\n"
            f"  apply_n_qubit   : {self.apply_n_qubit_calls:4} calls  {self.apply_n_qubit_time * 1000.0:8.2f} ms"
            f"  ({100.0 * self.apply_n_qubit_time / total:.1f}%)\n"
            f"  measure_qubit   : {self.measure_qubit_calls:4} calls  {self.measure_qubit_time * 1000.0:8.2f} ms"
            f"  ({100.0 * self.measure_qubit_time / total:.1f}%)"
        )


# ---------------------------------------------------------------------------
# SimulationResult (simulator.rs:85-202)
# ---------------------------------------------------------------------------
class SimulationResult:
    """One trajectory's final state, resident in GPU memory."""

    def __init__(self, handle):
        if not isinstance(handle, ctypes.c_void_p):
            raise TypeError("No constructor defined")
        self._h = handle
        self._n = int(_lib.dq_result_num_qubits(handle))
        self._nc = int(_lib.dq_result_num_cbits(handle))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            _lib.dq_result_free(h)
            self._h = None

    @property
    def num_qubits(self) -> int:
        return self._n

    @property
    def statevector(self) -> np.ndarray:
        """Independent complex128 copy of the amplitudes, shape (2^n,)."""
        out = np.empty(1 << self._n, dtype=np.complex128)
        _check(_lib.dq_result_statevector_c128(self._h, out.ctypes.data, out.size))
        return out

    @property
    def classical_bits(self) -> Dict[int, int]:
        buf = (ctypes.c_uint8 * max(1, self._nc))()
        _check(_lib.dq_result_classical_bits(self._h, buf, self._nc))
        return {i: int(buf[i]) for i in range(self._nc)}

    @property
    def profile(self) -> Optional[SimulationProfile]:
        p = _Prof()
        if _lib.dq_result_profile(self._h, ctypes.byref(p)):
            return SimulationProfile(p)
        return None

    def probabilities(self, qubits=None) -> Dict[int, float]:
        qs = list(range(self._n))[::-1] if qubits is None else _qubit_list(qubits)
        # Validated by the library before any output is sized by len(qs).
        if len(qs) > self._n or len(set(qs)) != len(qs) or any(q >= self._n for q in qs):
            raise ValueError("Query qubits must be distinct and within the statevector's range")
        out = np.zeros(1 << len(qs), dtype=np.float64)
        arr = _u64_array(qs)
        _check(_lib.dq_result_probabilities(self._h, arr, len(qs), out.ctypes.data))
        nz = np.nonzero(out > 0.0)[0]
        return {int(j): float(out[j]) for j in nz}

    def counts(self, shots=1000, qubits=None, seed=None) -> Dict[str, int]:
        shots = _as_u64(shots, "shots")
        qs = list(range(self._n))[::-1] if qubits is None else _qubit_list(qubits)
        seed = _opt_u64(seed, "seed")
        arr = _u64_array(qs)
        tab = _CountTable()
        _check(_lib.dq_result_counts(self._h, shots, arr, len(qs), 0 if seed is None else 1,
                                     0 if seed is None else seed, tab.cb, tab.user))
        return tab.result()

    def fidelity(self, other) -> float:
        if not isinstance(other, np.ndarray) or other.dtype != np.complex128 or other.ndim != 1:
            raise TypeError("argument 'other': expected a 1-dimensional complex128 numpy array")
        other = np.ascontiguousarray(other)
        out = ctypes.c_double(0.0)
        _check(_lib.dq_result_fidelity(self._h, other.ctypes.data, other.size, ctypes.byref(out)))
        return out.value

    def compare(self, other: "SimulationResult"):
        """(max |a-b|, ||a-b||_2) against another result, computed on the GPU."""
        mx, l2 = ctypes.c_double(), ctypes.c_double()
        _check(_lib.dq_result_compare(self._h, other._h, ctypes.byref(mx), ctypes.byref(l2)))
        return mx.value, l2.value

    def amplitudes(self, indices) -> np.ndarray:
        """Amplitudes at the given basis indices (complex128), gathered on the GPU."""
        idx = np.ascontiguousarray(indices, dtype=np.uint64)
        out = np.empty(idx.size, dtype=np.complex128)
        _check(_lib.dq_result_gather(self._h, idx.ctypes.data, idx.size, out.ctypes.data))
        return out

    def amplitude(self, index) -> complex:
        """One amplitude by basis index, or by bitstring with the highest qubit first."""
        if isinstance(index, str):
            if len(index) != self._n or any(c not in "01" for c in index):
                raise ValueError("Amplitude bitstrings need one '0'/'1' character per qubit")
            index = int(index, 2) if index else 0
        index = _as_u64(index, "index")
        buf = (ctypes.c_double * 2)()
        _check(_lib.dq_result_amplitude(self._h, index, buf))
        return complex(buf[0], buf[1])


# ---------------------------------------------------------------------------
# StatevectorSimulator (simulator.rs:213-371)
# ---------------------------------------------------------------------------
class StatevectorSimulator:
    """Dense statevector simulation on the GPU."""

    def __init__(self, seed=None, profile=False, *, max_memory_mb=1024, max_parallel_shots=None,
                 sample_terminal=True, precision="complex128", fusion="none", max_fused_qubits=4,
                 tile_qubits=12):
        o = _Opts()
        _lib.dq_sv_opts_default(ctypes.byref(o))
        seed = _opt_u64(seed, "seed")
        o.has_seed = 0 if seed is None else 1
        o.seed = 0 if seed is None else seed
        o.profile = 1 if _as_bool(profile, "profile") else 0
        o.max_memory_mb = _as_u64(max_memory_mb, "max_memory_mb")
        mps = _opt_u64(max_parallel_shots, "max_parallel_shots")
        o.has_max_parallel_shots = 0 if mps is None else 1
        o.max_parallel_shots = 0 if mps is None else mps
        o.sample_terminal = 1 if _as_bool(sample_terminal, "sample_terminal") else 0
        if precision not in _PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(_PRECISIONS)}")
        if fusion not in _FUSIONS:
            raise ValueError(f"fusion must be one of {sorted(_FUSIONS)}")
        o.precision = _PRECISIONS[precision]
        o.fusion = _FUSIONS[fusion]
        if fusion == "learned" and not _planner_weights_set[0]:
            from . import planner  # the shipped weights, once per process

            planner.install_default()
            _planner_weights_set[0] = True
        o.max_fused_qubits = min(_as_u64(max_fused_qubits, "max_fused_qubits"), 99)
        o.tile_qubits = min(_as_u64(tile_qubits, "tile_qubits"), 99)
        h = ctypes.c_void_p()
        _check(_lib.dq_sv_create(ctypes.byref(o), ctypes.byref(h)))
        self._h = h
        self._precision = precision
        self._fusion = fusion

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            _lib.dq_sv_free(h)
            self._h = None

    def simulate_shots(self, circuit, shots=1000, profile=False) -> Dict[str, int]:
        """Classical counts; terminal measurements share a single evolution."""
        shots = _as_u64(shots, "shots")
        profile = _as_bool(profile, "profile")
        data = _circuit_json(circuit)
        tab = _CountTable()
        _check(_lib.dq_sv_simulate_shots(self._h, data, len(data), shots, 1 if profile else 0, tab.cb, tab.user))
        return tab.result()

    def simulate(self, circuit) -> SimulationResult:
        """Execute one trajectory. Measurements collapse its returned state."""
        data = _circuit_json(circuit)
        h = ctypes.c_void_p()
        _check(_lib.dq_sv_simulate(self._h, data, len(data), ctypes.byref(h)))
        return SimulationResult(h)

    # -- additive helpers (no reference counterpart) --
    def evolve_timed(self, circuit, repeats=1, record=True, keep=False):
        """Evolve the unitary prefix on the device; returns (records, launches,
        gates, total_ms, result_or_None) with per-launch device timings."""
        data = _circuit_json(circuit)
        cap = 1 << 20
        recs = (OpTiming * cap)() if record else None
        nops, gates, total = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_double()
        keep_h = ctypes.c_void_p() if keep else None
        _check(_lib.dq_sv_evolve_timed(self._h, data, len(data), int(repeats), recs, cap if record else 0,
                                       ctypes.byref(nops), ctypes.byref(gates), ctypes.byref(total),
                                       ctypes.byref(keep_h) if keep else None))
        rows = []
        if record:
            for i in range(min(nops.value, cap)):
                r = recs[i]
                rows.append({"kind": r.kind, "k": r.k, "low": r.low, "gates": r.gates, "ms": r.ms,
                             "bytes": r.bytes, "feat": list(r.feat)})
        return rows, nops.value, gates.value, total.value, (SimulationResult(keep_h) if keep else None)


# ---------------------------------------------------------------------------
# Functional wrappers and codecs (src/simulator.rs:14-116, src/codecs.rs:23-64)
# ---------------------------------------------------------------------------
_SV_OPTIONS = ("max_memory_mb", "max_parallel_shots", "sample_terminal", "precision", "fusion", "max_fused_qubits")


def _parse_mode(mode) -> str:
    if not isinstance(mode, str):
        raise TypeError(f"argument 'mode': '{type(mode).__name__}' object cannot be converted to 'PyString'")
    m = mode.strip().lower().replace("-", "_")
    if m in ("state_vector", "statevector", "sv"):
        return "state_vector"
    raise ValueError(f"Unsupported monolithic simulation mode \"{m}\"; this engine implements 'state_vector'")


def _parse_options(options: Dict[str, Any]) -> Dict[str, Any]:
    result = {"max_memory_mb": 1024, "max_parallel_shots": None, "sample_terminal": True}
    for key, value in options.items():
        if key not in _SV_OPTIONS:
            raise TypeError(f"Unsupported statevector option \"{key}\"")
        result[key] = value
    return result


def simulate_monolithic(circuit, mode="state_vector", seed=None, profile=False, **options):
    _parse_mode(mode)
    opts = _parse_options(options)
    sim = StatevectorSimulator(seed, profile, **opts)
    return sim.simulate(circuit)


def simulate_monolithic_shots(circuit, mode="state_vector", shots=1000, seed=None, profile=False, **options):
    _parse_mode(mode)
    opts = _parse_options(options)
    sim = StatevectorSimulator(seed, False, **opts)
    return sim.simulate_shots(circuit, shots, profile)


def _not_here(name):
    def init(self, *args, **kwargs):
        raise NotImplementedError(f"{name} is not part of this dense statevector engine")

    return init


class MpsSimulator:
    __init__ = _not_here("MpsSimulator")


class MpsResult:
    __init__ = _not_here("MpsResult")


class StabilizerSimulator:
    __init__ = _not_here("StabilizerSimulator")


class StabilizerResult:
    __init__ = _not_here("StabilizerResult")


# ---------------------------------------------------------------------------
# Host-only helpers
# ---------------------------------------------------------------------------
def plan_json(circuit_or_json, precision="complex128", fusion="none", max_fused_qubits=4, tile_qubits=12,
              out_of_place=False) -> dict:
    """Launch plan of a circuit (parse, validate, compile, fuse; no GPU).
    out_of_place: the plan of a call whose budget holds a second state."""
    data = circuit_or_json if isinstance(circuit_or_json, (bytes, str)) else _circuit_json(circuit_or_json)
    if isinstance(data, str):
        data = data.encode("utf-8")
    out = ctypes.c_void_p()
    _check(_lib.dq_plan_json(data, len(data), _PRECISIONS[precision], _FUSIONS[fusion], int(max_fused_qubits),
                             int(tile_qubits), 1 if out_of_place else 0, ctypes.byref(out)))
    try:
        return json.loads(ctypes.string_at(out).decode("utf-8"))
    finally:
        _lib.dq_string_free(out)


def device_count() -> int:
    return int(_lib.dq_device_count())


def mfma_fp32_peak(reps: int = 5):
    """Measured FP32 matrix-core throughput of the current GPU:
    (TFLOP/s, ms per launch) of a register-only MFMA loop."""
    tf, ms = ctypes.c_double(), ctypes.c_double()
    _check(_lib.dq_mfma_fp32_peak(reps, ctypes.byref(tf), ctypes.byref(ms)))
    return tf.value, ms.value


def set_sweep_relabel(on: bool) -> bool:
    """Relabelled tile sweeps on or off for this process (default off: fewer
    sweeps, but slower sweeps); returns the previous setting. Results do not
    depend on it (complex128 exact sweeps stay bit-identical)."""
    return bool(_lib.dq_set_sweep_relabel(1 if on else 0))


def set_sweep_mode(mode: int) -> int:
    """Sweep mode of complex64 fused plans for this process: 0 in place; out
    of place with 1 persistent 256-thread workgroups, 2 one 256-thread
    workgroup per tile (default), 3 one 512-thread workgroup per tile, 4
    persistent 512-thread workgroups. Exact complex128 sweeps run out of
    place (one 256-thread workgroup per tile) in every mode but 0. Out-of-
    place sweeps run only when a call's budget holds a second state. Returns
    the previous mode."""
    prev = int(_lib.dq_set_sweep_mode(int(mode)))
    if prev < 0:
        raise ValueError(f"sweep mode must be 0..4, got {mode}")
    return prev


def set_sweep_pinned(on: bool) -> bool:
    """Out-of-place sweeps keep qubits 0..3 in every tile (True) or only in
    the first and last tile of a run, consecutive tiles sharing 4 qubits
    (False, the default: more tile qubits for gates, fewer sweeps). Returns
    the previous setting."""
    return bool(_lib.dq_set_sweep_pinned(1 if on else 0))


def set_block_arith(mode: int) -> int:
    """Arithmetic of 3-qubit blocks in complex64 tile sweeps for this
    process: 0 FP32 matrix cores, 1 split precision (FP32 operands as two
    FP16 parts on the 16x16x32 f16 matrix core, FP32 accumulation; the
    default). Returns the previous mode."""
    prev = int(_lib.dq_set_block_arith(int(mode)))
    if prev < 0:
        raise ValueError(f"block arithmetic must be 0 or 1, got {mode}")
    return prev


def set_wave_runs(mode: int) -> int:
    """Barriers between the blocks of 256-thread out-of-place complex64
    sweeps for this process: 1 (the default) groups consecutive blocks that
    leave two tile bits untouched into runs whose amplitudes each wave owns,
    with a workgroup barrier only after a run; 0 puts one after every block.
    The states are bit-identical. Returns the previous mode."""
    prev = int(_lib.dq_set_wave_runs(int(mode)))
    if prev < 0:
        raise ValueError(f"wave runs must be 0 or 1, got {mode}")
    return prev


def set_zero_start(on: bool) -> bool:
    """Plans that open with out-of-place sweeps start from a synthesised
    |0...0> and skip reading all-zero tiles (True, the default), or
    initialise the state and read it (False); results are identical.
    Returns the previous setting."""
    return bool(_lib.dq_set_zero_start(1 if on else 0))


_planner_weights_set = [False]  # weights installed or cleared in this process


def set_planner_weights(weights) -> None:
    """Weights of the learned planner for this process (None: greedy)."""
    _planner_weights_set[0] = True
    if weights is None:
        _check(_lib.dq_set_planner_weights(None, 0))
        return
    w = np.ascontiguousarray(weights, dtype=np.float32).ravel()
    _check(_lib.dq_set_planner_weights(w.ctypes.data, w.size))


__all__ = [
    "MpsSimulator", "MpsResult", "PBlockResult", "PBlockSimulator", "SimulationProfile", "SimulationResult",
    "StabilizerSimulator", "StabilizerResult", "StatevectorSimulator", "simulate_distributed",
    "simulate_distributed_shots", "simulate_monolithic", "simulate_monolithic_shots",
]


# P-block backend (product of dense blocks): defined in _pblock, which binds
# its own C entry points on this module's library handle.
from ._pblock import PBlockResult, PBlockSimulator, simulate_distributed, simulate_distributed_shots  # noqa: E402
