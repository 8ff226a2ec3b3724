"""CPU restatement of the reference's P-block backend (test infrastructure).

Only tests/ and bench.py's cpu_baseline leg may use this module, as the
checker; the product path never does. It restates, step by step:

  compile ........................ src/distributed/pblock/execution.rs:57-120
  block pool (merge, ensure, known states, controls, gates, measure,
  split, amplitude, sampling) .... src/distributed/pblock/model.rs:50-617
  simulate / shots driver ........ src/distributed/pblock/simulator.rs:106-322
  result queries ................. src/distributed/pblock/result.rs:118-293

Block amplitudes are numpy complex128 arrays updated by this package's FP64
C kernels (oracle.apply_*, the reference's per-op arithmetic). Sums that the
reference forms sequentially are formed sequentially here. Random draws use
the engine's Philox conventions instead of the reference's ChaCha8 streams
(whose exact counts no fixture pins): a pool draws (seed, stream, draw 0, 1,
...) for its measurements, with stream 0 for simulate() and stream s for
trajectory shot s; product sampling draws (seed, stream s, draw t) for
block t of shot s. With those draws fixed, block structure, statistics,
classical bits and counts are what the reference's rules give.
"""

import ctypes
import os
import threading
from typing import Dict, List, Optional

import numpy as np

from . import oracle as so

K_ONE, K_X, K_SWAP, K_DIAG, K_DENSE, K_MEAS, K_RESET, K_NOOP = range(8)
_X = np.array([0, 1, 1, 0], dtype=np.complex128)


class _OpDesc(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("target", ctypes.c_uint32), ("a", ctypes.c_uint32),
                ("b", ctypes.c_uint32), ("qubit", ctypes.c_uint32), ("cbit", ctypes.c_uint64),
                ("nctl", ctypes.c_int32), ("cq", ctypes.c_uint32 * 5), ("cv", ctypes.c_uint8 * 5),
                ("k", ctypes.c_int32), ("qubits", ctypes.c_uint32 * 5), ("nmat", ctypes.c_int32)]


_bound = False
_bind_lock = threading.Lock()


def _lib():
    global _bound
    lib = so.lib()
    with _bind_lock:
        if not _bound:
            lib.or_compile_one.restype = ctypes.c_int
            lib.or_compile_one.argtypes = [ctypes.c_void_p, ctypes.POINTER(_OpDesc), ctypes.c_void_p,
                                           ctypes.c_int64]
            _bound = True
    return lib


class MemErr(MemoryError):
    pass


def _fail(msg: str):
    # result.rs:11-20: budget and size messages are MemoryError.
    if "memory" in msg or "allocation" in msg or "max_block_qubits" in msg:
        raise MemErr(msg)
    raise RuntimeError(msg)


def _nsq(v) -> float:
    return v.real * v.real + v.imag * v.imag


# ---------------------------------------------------------------------------
# Compile (execution.rs:57-120)
# ---------------------------------------------------------------------------
class Op:
    __slots__ = ("kind", "q", "c", "conds", "inner", "d", "mat")

    def __init__(self, kind, q=0, c=0, conds=None, inner=None, d=None, mat=None):
        self.kind, self.q, self.c, self.conds, self.inner, self.d, self.mat = kind, q, c, conds, inner, d, mat

    def unitary(self):
        return self.kind in ("gate", "noop")

    def terminal(self):
        return self.kind in ("measure", "noop")


def _compile_rec(rec) -> Op:
    d = _OpDesc()
    mat = (ctypes.c_double * 2048)()
    r = _lib().or_compile_one(ctypes.byref(rec), ctypes.byref(d), mat, 2048)
    if r < 0:
        raise ValueError("oracle compile failed")
    if r == 0:
        return Op("noop")
    if d.kind == K_MEAS:
        return Op("measure", q=d.qubit, c=d.cbit)
    if d.kind == K_RESET:
        return Op("reset", q=d.qubit)
    m = np.array([complex(mat[2 * i], mat[2 * i + 1]) for i in range(d.nmat)], dtype=np.complex128)
    return Op("gate", d=d, mat=m)


def compile_ops(instructions) -> List[Op]:
    ident = [1 + 0j, 0j, 0j, 1 + 0j]
    pending: Dict[int, list] = {}
    out: List[Op] = []

    def flush():
        for q in sorted(pending):
            m = pending[q]
            if all(m[i] == ident[i] and (m[i].imag == ident[i].imag) for i in range(4)):
                continue
            d = _OpDesc()
            d.kind, d.target = K_ONE, q
            out.append(Op("gate", d=d, mat=np.array(m, dtype=np.complex128)))
        pending.clear()

    for inst in instructions:
        if inst["kind"] != "conditional":
            u = (ctypes.c_double * 8)()
            kind = so.KIND.get(inst["kind"])
            if kind is not None and so.lib().or_gate_1q(kind, inst.get("theta", 0.0), inst.get("phi", 0.0),
                                                        inst.get("lam", 0.0), u):
                m = [complex(u[2 * i], u[2 * i + 1]) for i in range(4)]
                q = inst["qubit"]
                old = pending.get(q, ident)
                prod = []
                for i in range(2):
                    for j in range(2):
                        acc = 0j
                        for kk in range(2):
                            acc = acc + m[i * 2 + kk] * old[kk * 2 + j]
                        prod.append(acc)
                pending[q] = prod
                continue
        flush()
        conds = []
        body = inst
        while body["kind"] == "conditional":
            c = body["condition"]
            conds.append((c["creg_base"], c["creg_size"], c["creg_value"]))
            body = body["op"]
        if not conds and body["kind"] == "barrier":
            continue
        rec = so._convert(body, [])
        op = _compile_rec(rec)
        if conds:
            out.append(Op("cond", conds=conds, inner=op))
        elif op.kind != "noop":
            out.append(op)
    flush()
    return out


# ---------------------------------------------------------------------------
# Block pool (model.rs:50-617)
# ---------------------------------------------------------------------------
def _dimension(n: int) -> int:
    if n >= 64:
        _fail("P-block dense allocation dimension exceeds addressable memory")
    return 1 << n


class Block:
    __slots__ = ("state", "qubits")

    def __init__(self, state, qubits):
        self.state, self.qubits = state, qubits


STAT_KEYS = ("merge_calls", "merge_output_amplitudes", "split_calls", "separation_checks", "separation_residual",
             "skipped_controlled_gates", "peak_block_qubits", "peak_live_amplitudes", "peak_working_bytes")


class Pool:
    def __init__(self, n: int, memory: int, max_block=None, split=False, max_split=12, tol=0.0):
        if n * 288 > memory:
            _fail("P-block initial state exceeds max_memory_mb")
        self.memory, self.max_block, self.split, self.max_split, self.tol = memory, max_block, split, max_split, tol
        self.n = n
        self.positions = [(q, 0) for q in range(n)]
        self.known: List[Optional[int]] = [0] * n
        self.reset_zero()
        self.seed = 0
        self.stream = 0
        self.draw = 0

    def reset_zero(self):
        self.blocks = [Block(np.array([1, 0], dtype=np.complex128), [q]) for q in range(self.n)]
        self.free = []
        self.positions = [(q, 0) for q in range(self.n)]
        self.known = [0] * self.n
        self.live = 2 * self.n
        self.stats = dict.fromkeys(STAT_KEYS, 0)
        self.stats["separation_residual"] = 0.0
        self.stats["peak_block_qubits"] = 1 if self.n else 0
        self.record()

    def copy(self):
        p = Pool.__new__(Pool)
        p.__dict__.update(self.__dict__)
        p.blocks = [None if b is None else Block(b.state.copy(), list(b.qubits)) for b in self.blocks]
        p.positions, p.known, p.free = list(self.positions), list(self.known), list(self.free)
        p.stats = dict(self.stats)
        return p

    def restore(self, src):
        self.blocks = [None if b is None else Block(b.state.copy(), list(b.qubits)) for b in src.blocks]
        self.positions, self.known, self.free = list(src.positions), list(src.known), list(src.free)
        self.live = src.live
        self.stats = dict.fromkeys(STAT_KEYS, 0)
        self.stats["separation_residual"] = 0.0
        self.stats["peak_block_qubits"] = max((len(b.qubits) for b in src.blocks if b is not None), default=0)
        self.record()

    def seed_rng(self, seed, stream):
        self.seed, self.stream, self.draw = seed, stream, 0

    def bytes(self):
        return self.live * 16 + self.n * 256

    def check(self, extra):
        total = self.bytes() + extra
        if total > self.memory:
            _fail("P-block working allocation exceeds max_memory_mb")
        return total

    def reserve(self, extra):
        self.stats["peak_working_bytes"] = max(self.stats["peak_working_bytes"], self.check(extra))

    def record(self):
        self.stats["peak_live_amplitudes"] = max(self.stats["peak_live_amplitudes"], self.live)
        self.stats["peak_working_bytes"] = max(self.stats["peak_working_bytes"], self.bytes())

    def merge(self, a, b):
        if a == b:
            return a
        left, right = self.blocks[a], self.blocks[b]
        n = len(left.qubits) + len(right.qubits)
        if self.max_block is not None and n > self.max_block:
            _fail("P-block merge exceeds max_block_qubits; no entanglement was discarded")
        dim = _dimension(n)
        self.reserve(dim * 16)
        state = np.outer(right.state, left.state).reshape(-1)
        qubits = left.qubits + right.qubits
        self.live = self.live - left.state.size - right.state.size + dim
        for i, q in enumerate(qubits):
            self.positions[q] = (a, i)
        self.blocks[a] = Block(state, qubits)
        self.blocks[b] = None
        self.free.append(b)
        self.stats["peak_block_qubits"] = max(self.stats["peak_block_qubits"], n)
        self.stats["merge_calls"] += 1
        self.stats["merge_output_amplitudes"] += dim
        self.record()
        return a

    def ensure(self, qs):
        bl = sorted(set(self.positions[q][0] for q in qs))
        bl.sort(key=lambda i: len(self.blocks[i].qubits))
        for i in bl[1:]:
            self.merge(bl[0], i)
        return bl[0]

    def refresh_known(self, q):
        b = self.blocks[self.positions[q][0]]
        if len(b.qubits) == 1:
            self.known[q] = 0 if b.state[1] == 0 else (1 if b.state[0] == 0 else None)
        else:
            self.known[q] = None

    def controls(self, cs):
        active = []
        for q, on in cs:
            v = self.known[q]
            if v is not None:
                if (v != 0) != on:
                    self.stats["skipped_controlled_gates"] += 1
                    return None
            else:
                active.append((q, on))
        return active

    def one(self, q, m, cs):
        active = self.controls(cs)
        if active is None:
            return
        if active:
            b = self.blocks[self.positions[q][0]]
            if len(b.qubits) == 1:
                a, z = complex(b.state[0]), complex(b.state[1])
                if m[0] * a + m[1] * z == a and m[2] * a + m[3] * z == z:
                    self.stats["skipped_controlled_gates"] += 1
                    return
        qs = [c for c, _ in active] + [q]
        idx = self.ensure(qs)
        b = self.blocks[idx]
        local = [(self.positions[c][1], on) for c, on in active]
        t = self.positions[q][1]
        if np.array_equal(m, _X):
            so.apply_x(b.state, t, local)
        else:
            so.apply_one(b.state, t, m, local)
        for w in qs:
            self.refresh_known(w)
        self.maybe_split(qs)

    def relabel_swap(self, a, b):
        (ia, pa), (ib, pb) = self.positions[a], self.positions[b]
        self.blocks[ia].qubits[pa] = b
        self.blocks[ib].qubits[pb] = a
        self.positions[a], self.positions[b] = self.positions[b], self.positions[a]
        self.known[a], self.known[b] = self.known[b], self.known[a]

    def gate(self, op: Op):
        d = op.d
        cs = [(d.cq[i], bool(d.cv[i])) for i in range(d.nctl)]
        if d.kind == K_ONE:
            self.one(d.target, op.mat, cs)
        elif d.kind == K_X:
            self.one(d.target, _X, cs)
        elif d.kind == K_SWAP:
            active = self.controls(cs)
            if active is None:
                return
            if not active:
                self.relabel_swap(d.a, d.b)
                return
            qs = [c for c, _ in active] + [d.a, d.b]
            idx = self.ensure(qs)
            local = [(self.positions[c][1], on) for c, on in active]
            so.apply_swap(self.blocks[idx].state, self.positions[d.a][1], self.positions[d.b][1], local)
            for w in qs:
                self.refresh_known(w)
            self.maybe_split(qs)
        elif d.kind == K_DIAG:
            qs = [d.qubits[i] for i in range(d.k)]
            active = [q for q in qs if self.known[q] is None]
            reduced = []
            for basis in range(1 << len(active)):
                index, cursor = 0, 0
                for q in qs:
                    if self.known[q] is not None:
                        bit = self.known[q]
                    else:
                        bit = (basis >> (len(active) - 1 - cursor)) & 1
                        cursor += 1
                    index = (index << 1) | bit
                reduced.append(op.mat[index])
            if not active or all(v == reduced[0] for v in reduced):
                b = self.blocks[self.positions[qs[0]][0]]
                b.state *= reduced[0]
            else:
                idx = self.ensure(active)
                local = [self.positions[q][1] for q in active]
                so.apply_diagonal(self.blocks[idx].state, local, np.array(reduced, dtype=np.complex128))
                self.maybe_split(active)
        elif d.kind == K_DENSE:
            qs = [d.qubits[i] for i in range(d.k)]
            idx = self.ensure(qs)
            local = [self.positions[q][1] for q in qs]
            so.apply_dense(self.blocks[idx].state, local, op.mat.reshape(1 << d.k, 1 << d.k))
            for w in qs:
                self.refresh_known(w)
            self.maybe_split(qs)

    def install_split(self, q, single, remaining):
        idx, local = self.positions[q]
        old = self.blocks[idx]
        slot = self.free.pop()
        qs = list(old.qubits)
        del qs[local]
        self.live = self.live - old.state.size + remaining.size + 2
        for i, w in enumerate(qs):
            self.positions[w] = (idx, i)
        self.blocks[idx] = Block(remaining, qs)
        self.blocks[slot] = Block(np.array(single, dtype=np.complex128), [q])
        self.positions[q] = (slot, 0)
        self.stats["split_calls"] += 1
        self.refresh_known(q)
        for w in qs:
            self.refresh_known(w)
        self.record()

    @staticmethod
    def _half(state, local, bit):
        i = np.arange(state.size // 2, dtype=np.int64)
        low = (1 << local) - 1
        return (i & low) | ((i & ~low) << 1) | (bit << local)

    def measure(self, q) -> int:
        idx, local = self.positions[q]
        b = self.blocks[idx]
        size = b.state.size
        if size > 2:
            self.reserve((size // 2 + 2) * 16)
        r = so.uniform53(self.seed, self.stream, self.draw)
        self.draw += 1
        value, _, _ = so.measure(b.state, local, r)
        if value < 0:
            raise RuntimeError("Cannot measure a state with nonfinite or zero norm")
        if size > 2:
            remaining = b.state[self._half(b.state, local, value)].copy()
            single = [0j, 0j]
            single[value] = 1 + 0j
            self.install_split(q, single, remaining)
        self.known[q] = value
        return value

    def maybe_split(self, qs):
        if not self.split:
            return
        for q in qs:
            idx, local = self.positions[q]
            b = self.blocks[idx]
            if len(b.qubits) < 2 or len(b.qubits) > self.max_split:
                continue
            size = b.state.size
            extra = (size // 2 + 2) * 16
            try:
                self.check(extra)
            except MemErr:
                continue
            self.reserve(extra)
            self.stats["separation_checks"] += 1
            i0, i1 = self._half(b.state, local, 0), self._half(b.state, local, 1)
            x0, x1 = b.state[i0], b.state[i1]
            w = [_nsq(x0[i]) + _nsq(x1[i]) for i in range(size // 2)]
            best = 0
            for i in range(1, len(w)):
                if w[i] >= w[best]:
                    best = i
            single = [complex(x0[best]), complex(x1[best])]
            norm = (_nsq(single[0]) + _nsq(single[1])) ** 0.5
            if norm == 0 or not np.isfinite(norm):
                _fail("P-block separation encountered an invalid norm")
            single = [complex(s.real / norm, s.imag / norm) for s in single]
            c0, c1 = single[0].conjugate(), single[1].conjugate()
            remaining, loss, total = [], 0.0, 0.0
            for i in range(size // 2):
                a0, a1 = complex(x0[i]), complex(x1[i])
                value = c0 * a0 + c1 * a1
                remaining.append(value)
                loss += _nsq(a0 - single[0] * value)
                total += _nsq(a0)
                loss += _nsq(a1 - single[1] * value)
                total += _nsq(a1)
            relative = loss / total
            if relative <= self.tol:
                kept = 0.0
                for v in remaining:
                    kept += _nsq(v)
                scale = (total / kept) ** 0.5
                rem = np.array([complex(v.real * scale, v.imag * scale) for v in remaining], dtype=np.complex128)
                self.install_split(q, single, rem)
                self.stats["separation_residual"] += relative

    def live_blocks(self):
        return [b for b in self.blocks if b is not None]

    def sampling_tables(self):
        self.check(self.live * 8 + self.n * 64)
        tables = []
        for b in self.live_blocks():
            cdf = np.cumsum(b.state.real * b.state.real + b.state.imag * b.state.imag)
            if not (cdf[-1] > 0) or not np.isfinite(cdf[-1]):
                _fail("P-block sampling encountered an invalid norm")
            tables.append((list(b.qubits), cdf))
        return tables


def sample_tables(tables, seed, stream, bits):
    """model.rs:606-617 with draw t of (seed, stream) for table t."""
    for t, (qs, cdf) in enumerate(tables):
        draw = so.uniform53(seed, stream, t) * cdf[-1]
        index = min(int(np.searchsorted(cdf, draw, side="right")), cdf.size - 1)
        for i, q in enumerate(qs):
            bits[q] = (index >> i) & 1


# ---------------------------------------------------------------------------
# Drivers (simulator.rs:106-322)
# ---------------------------------------------------------------------------
def _run(op: Op, pool: Pool, cbits: Dict[int, int]):
    if op.kind == "cond":
        for base, size, value in op.conds:
            actual = 0
            for bit in range(size):
                actual |= cbits.get(base + bit, 0) << bit
            if actual != value:
                return
        _run(op.inner, pool, cbits)
    elif op.kind == "measure":
        cbits[op.c] = pool.measure(op.q)
    elif op.kind == "reset":
        if pool.measure(op.q) == 1:
            pool.one(op.q, _X, [])
    elif op.kind == "gate":
        pool.gate(op)


def _dims(data):
    n = sum(r["size"] for r in data["qregs"].values())
    nc = max((r["base"] + r["size"] for r in data.get("cregs", {}).values()), default=0)
    return n, nc


def _format(cbits, nc):
    return "".join("1" if cbits.get(nc - 1 - i, 0) else "0" for i in range(nc))


class Result:
    def __init__(self, pool, cbits, n):
        self.pool, self.classical_bits, self.n = pool, cbits, n

    @property
    def block_qubits(self):
        return [list(b.qubits) for b in self.pool.live_blocks()]

    @property
    def diagnostics(self):
        d = dict(self.pool.stats)
        d["block_sizes"] = [len(b.qubits) for b in self.pool.live_blocks()]
        d["live_amplitudes"] = sum(b.state.size for b in self.pool.live_blocks())
        return d

    @property
    def statevector(self):
        size = 1 << self.n
        basis = np.arange(size, dtype=np.int64)
        acc = np.ones(size, dtype=np.complex128)
        for b in self.pool.live_blocks():
            idx = np.zeros(size, dtype=np.int64)
            for i, q in enumerate(b.qubits):
                idx |= ((basis >> q) & 1) << i
            acc = acc * b.state[idx]
        return acc

    def probabilities(self, qubits=None):
        qs = list(range(self.n))[::-1] if qubits is None else list(qubits)
        size = 1 << len(qs)
        basis = np.arange(size, dtype=np.int64)
        acc = np.ones(size)
        for b in self.pool.live_blocks():
            sel = [(i, b.qubits.index(q)) for i, q in enumerate(qs) if q in b.qubits]
            probs = so.marginal(b.state, [loc for _, loc in sel])
            index = np.zeros(size, dtype=np.int64)
            for pos, _ in sel:
                index = (index << 1) | ((basis >> (len(qs) - 1 - pos)) & 1)
            acc = acc * probs[index]
        return {int(i): float(p) for i, p in enumerate(acc) if p > 0}

    def counts(self, shots, qubits=None, seed=0):
        qs = list(range(self.n))[::-1] if qubits is None else list(qubits)
        out: Dict[str, int] = {}
        if shots == 0:
            return out
        tables = self.pool.sampling_tables()
        bits = [0] * self.n
        for s in range(shots):
            sample_tables(tables, seed, s, bits)
            key = "".join("1" if bits[q] else "0" for q in qs)
            out[key] = out.get(key, 0) + 1
        return out


def simulate(data: dict, seed: int, memory_mb=1024, max_block_qubits=None, split_separable=False,
             max_split_qubits=12, separation_tolerance=0.0) -> Result:
    """A plain circuit (one quantum register from 0) on the P-block rules."""
    n, nc = _dims(data)
    ops = compile_ops(data["instructions"])
    memory = memory_mb << 20
    if nc * 32 >= memory:
        _fail("P-block classical workspace exceeds max_memory_mb")
    pool = Pool(n, memory - nc * 32, max_block_qubits, split_separable, max_split_qubits, separation_tolerance)
    pool.seed_rng(seed, 0)
    cbits: Dict[int, int] = {}
    for op in ops:
        _run(op, pool, cbits)
    return Result(pool, cbits, n)


def simulate_shots(data: dict, shots: int, seed: int, memory_mb=1024, max_parallel_shots=1, sample_terminal=True,
                   max_block_qubits=None, split_separable=False, max_split_qubits=12, separation_tolerance=0.0):
    """Counts and the execution strategy, by the reference's shot driver."""
    n, nc = _dims(data)
    ops = compile_ops(data["instructions"])
    memory = memory_mb << 20
    classical = nc * 32
    if classical >= memory:
        _fail("P-block classical workspace exceeds max_memory_mb")
    memory -= classical
    counts: Dict[str, int] = {}
    if shots == 0:
        return counts, "empty"
    kw = dict(max_block=max_block_qubits, split=split_separable, max_split=max_split_qubits,
              tol=separation_tolerance)
    prefix_len = 0
    while prefix_len < len(ops) and ops[prefix_len].unitary():
        prefix_len += 1
    prefix = Pool(n, memory, **kw)
    prefix.seed_rng(seed, 0)
    cbits: Dict[int, int] = {}
    for op in ops[:prefix_len]:
        _run(op, prefix, cbits)
    if sample_terminal and all(op.terminal() for op in ops[prefix_len:]):
        if prefix_len == len(ops):
            return {"0" * nc: shots}, "terminal_sampling"
        try:
            tables = prefix.sampling_tables()
            bits = [0] * n
            for s in range(shots):
                sample_tables(tables, seed, s, bits)
                cb = {op.c: bits[op.q] for op in ops[prefix_len:] if op.kind == "measure"}
                key = _format(cb, nc)
                counts[key] = counts.get(key, 0) + 1
            return counts, "terminal_sampling"
        except MemErr:
            pass
    retain = prefix_len > 0 and prefix.bytes() * 2 < memory and \
        prefix.bytes() + prefix.stats["peak_working_bytes"] <= memory
    reserved = prefix.bytes() if retain else 0
    per_worker = max(max(prefix.stats["peak_working_bytes"], prefix.bytes()) + classical, 4096)
    workers = min(max_parallel_shots, shots, max(1, os.cpu_count() or 1), max((memory - reserved) // per_worker, 1))
    wmem = (memory - reserved - (workers - 1) * classical) // workers
    for w in range(workers):
        if retain:
            if prefix.bytes() > wmem:
                _fail("P-block prefix exceeds worker memory budget; reduce max_parallel_shots")
            pool = prefix.copy()
            pool.memory = wmem
        else:
            pool = Pool(n, wmem, **kw)
        for it, s in enumerate(range(w, shots, workers)):
            if it > 0:
                if retain:
                    pool.restore(prefix)
                else:
                    pool.reset_zero()
            pool.seed_rng(seed, s)
            cb: Dict[int, int] = {}
            for op in ops[prefix_len if retain else 0:]:
                _run(op, pool, cb)
            key = _format(cb, nc)
            counts[key] = counts.get(key, 0) + 1
    return counts, ("prefix_trajectories" if retain else "trajectories")
