"""Host-side checks of libdqsim_sv.so that need no GPU: the C ABI loads and
exports every declared symbol, and parsing, validation, budget checks,
option handling and compile/fusion planning behave like the reference
(src/validation.rs, src/monolithic/statevector/{validation,shots,simulator}.rs,
src/codecs.rs, tests/test_statevector.py:215-298)."""

import ctypes
import json
import os
import re

import numpy as np
import pytest

import dqsim
from dqsim import _core, workloads
from dqsim.workloads import Circuit
from oracle import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def inst(kind, **f):
    return {"kind": kind, **f}


def cond(op, value, base=0, size=1):
    return {"kind": "conditional", "condition": {"creg_base": base, "creg_size": size, "creg_value": value}, "op": op}


class Raw:
    """Circuit with an arbitrary transport document."""

    def __init__(self, data):
        self.data = data

    def model_dump_json(self):
        return json.dumps(self.data) if not isinstance(self.data, str) else self.data


def reg(name, base, size):
    return {"name": name, "base": base, "size": size}


def test_library_exports_every_declared_symbol():
    header = open(os.path.join(ROOT, "include", "dqsim_sv.h")).read()
    names = set(re.findall(r"\b(dq_[a-z0-9_]+)\s*\(", header))
    assert len(names) >= 20
    lib = ctypes.CDLL(_core.library_path())
    for name in sorted(names):
        assert hasattr(lib, name), name


def test_package_exports_reference_names():
    for name in ("StatevectorSimulator", "SimulationResult", "SimulationProfile", "simulate_monolithic",
                 "simulate_monolithic_shots", "MpsSimulator", "MpsResult", "PBlockSimulator", "PBlockResult",
                 "StabilizerSimulator", "StabilizerResult", "simulate_distributed", "simulate_distributed_shots",
                 "ImportedCircuit", "QASMImportError", "from_qiskit", "load_qasm"):
        assert hasattr(dqsim, name), name
    with pytest.raises(NotImplementedError):
        dqsim.MpsSimulator(seed=1)


INVALID = [
    Circuit(12, [inst("cx", control=0, target=12)]),
    Circuit(12, [inst("x", qubit=12)]),
    Circuit(2, [inst("cx", control=1, target=1)]),
    Circuit(3, [inst("ccx", control1=0, control2=0, target=2)]),
    Circuit(2, [inst("measure", qubit=0, cbit=2)], cbits=2),
    Circuit(1, [cond(inst("x", qubit=1), 1)], cbits=1),
    Circuit(1, [cond(inst("x", qubit=0), 0, base=1)], cbits=1),
    Circuit(2, [inst("gate", name="remote_cx", params=[], qubits=[0])]),
    Circuit(2, [inst("gate", name="remote_cu1", params=[], qubits=[0, 1])]),
]


@pytest.mark.parametrize("idx", range(len(INVALID)))
def test_invalid_qubits_and_operands_raise_value_error(idx):
    sim = dqsim.StatevectorSimulator(seed=1)
    for method in (sim.simulate, sim.simulate_shots):
        with pytest.raises(ValueError):
            method(INVALID[idx])


@pytest.mark.parametrize("op", [inst("gate", name="circuit-123", params=[], qubits=[0]),
                                inst("gate", name="unknown", params=[], qubits=[0]),
                                inst("classical", name="unsupported_classical_operation")])
def test_unsupported_operations_are_rejected_even_when_unreachable(op):
    sim = dqsim.StatevectorSimulator(seed=1)
    for instruction in (op, cond(op, 1)):
        for method in (sim.simulate, sim.simulate_shots):
            with pytest.raises(ValueError):
                method(Circuit(1, [instruction], cbits=1))


def _doc(instructions, qregs=None, cregs=None):
    return {"qregs": qregs or {"q": reg("q", 0, 3)}, "cregs": cregs or {"c": reg("c", 0, 3)},
            "instructions": instructions}


@pytest.mark.parametrize("doc", [
    _doc([inst("cx", control=0, target=3)]),
    _doc([inst("id", qubit=3)]),
    _doc([inst("measure", qubit=0, cbit=3)]),
    _doc([cond(inst("x", qubit=99), 1)]),
    _doc([inst("x", qubit=1)], qregs={"a": reg("a", 0, 1), "b": reg("b", 2, 1)}),  # hole between registers
    _doc([], qregs={"a": reg("a", 0, 2), "b": reg("b", 1, 2)}),  # overlap
    _doc([], qregs={"a": reg("a", 0, 64)}),  # n >= 64
    _doc([], qregs={"a": reg("a", 2**64 - 1, 1)}),  # overflowing range
    _doc([], cregs={"c": reg("c", 0, 3), "o": reg("o", 1, 1)}),
    _doc([inst("u3", qubit=0, theta=0.0, phi=float("nan"), lam=0.0)]),
    _doc([inst("u2", qubit=0, phi=1.7976931348623157e308, lam=1.7976931348623157e308)]),
    _doc([inst("gate", name="remote_rzz", qubits=[0, 1], params=[float("-inf")])]),
    _doc([inst("gate", name="remote_epr", qubits=[0, 1, 2], params=[])]),
    _doc([inst("gate", name="remote_cz", qubits=[0, 1], params=[0.2])]),
    _doc([inst("gate", name="remote_barrier", qubits=[1, 1], params=[])]),
    _doc([cond(inst("x", qubit=0), 0, 0, 0)]),
    _doc([cond(inst("x", qubit=0), 0, 0, 65)]),
    _doc([cond(inst("x", qubit=0), 2, 0, 1)]),
    _doc([cond(inst("x", qubit=0), 0, 2, 2)]),
    _doc([cond(inst("x", qubit=0), 0, 2**64 - 1, 2)]),
], ids=lambda d: json.dumps(d["instructions"])[:60])
def test_validation_rejections(doc):
    # src/monolithic/statevector/validation.rs:56-199 (NaN is sent as a JSON
    # number the reference cannot parse either).
    text = json.dumps(doc).replace("NaN", "1e999").replace("-Infinity", "-1e999")
    with pytest.raises(ValueError):
        _core.plan_json(text)


@pytest.mark.parametrize("doc", [
    _doc([inst("gate", name="remote_cu1", qubits=[0, 1], params=[0.4]),
          inst("gate", name="remote_rzz", qubits=[0, 1], params=[0.4]),
          inst("gate", name="REMOTE_CX", qubits=[0, 2], params=[]),
          inst("gate", name="remote_barrier", qubits=[0, 1, 2], params=[])]),
    _doc([inst("h", qubit=0), inst("cx", control=0, target=2), inst("measure", qubit=2, cbit=1),
          cond(cond(inst("reset", qubit=0), 1, 1, 1), 2, 0, 3)]),
    _doc([cond(inst("x", qubit=0), 2**64 - 1, 0, 64)], cregs={"c": reg("c", 0, 64)}),
])
def test_validation_acceptances(doc):
    _core.plan_json(json.dumps(doc))


@pytest.mark.parametrize("text", ["{", "[]", '{"qregs": {}}', '{"qregs": {}, "instructions": [{"kind": "nope"}]}',
                                  '{"qregs": {"q": {"name": "q", "base": 0, "size": 1.0}}, "instructions": []}',
                                  '{"qregs": {"q": {"name": "q", "base": 0, "size": -1}}, "instructions": []}',
                                  '{"qregs": {}, "instructions": [{"kind": "x"}]}'])
def test_malformed_json_is_a_value_error(text):
    with pytest.raises(ValueError, match="parse error"):
        _core.plan_json(text)


def test_memory_budget_rejects_before_allocation():
    sim = dqsim.StatevectorSimulator(max_memory_mb=1)
    for method in (sim.simulate, sim.simulate_shots):
        with pytest.raises(MemoryError):
            method(Circuit(20))
        with pytest.raises(MemoryError):
            method(Circuit(1, cbits=2_000_000))
    with pytest.raises(MemoryError):
        dqsim.StatevectorSimulator(max_memory_mb=16, precision="complex64").simulate(Circuit(22))


def test_zero_shots_need_no_device():
    sim = dqsim.StatevectorSimulator(seed=30)
    assert sim.simulate_shots(Circuit(1, [inst("h", qubit=0)], cbits=3), shots=0) == {}


def test_constructor_and_option_errors():
    for opts in ({"max_memory_mb": 0}, {"max_parallel_shots": 0}, {"max_fused_qubits": 6},
                 {"precision": "complex32"}, {"fusion": "magic"}):
        with pytest.raises(ValueError):
            dqsim.StatevectorSimulator(**opts)
    with pytest.raises(ValueError):
        dqsim.StatevectorSimulator(max_memory_mb=2**63)
    with pytest.raises(TypeError):
        dqsim.simulate_monolithic(Circuit(1), nonexistent_option=True)
    with pytest.raises(TypeError):
        dqsim.StatevectorSimulator(seed=1.5)
    with pytest.raises(OverflowError):
        dqsim.StatevectorSimulator(seed=-1)
    with pytest.raises(TypeError):
        dqsim.StatevectorSimulator(profile=1)
    with pytest.raises(TypeError):
        dqsim.StatevectorSimulator(1, False, 1024)  # max_memory_mb is keyword-only
    with pytest.raises(ValueError):
        dqsim.simulate_monolithic(Circuit(1), mode="tensor")
    for mode in ("state_vector", "StateVector", " sv ", "state-vector"):
        dqsim._core._parse_mode(mode)


def _product_kinds(plan):
    out = []
    for l in plan["launches"]:
        out.append("conditional" if l["conditional"] else l["kind"])
    return out


def test_compile_matches_oracle_on_random_circuits():
    rng = np.random.default_rng(7)
    kinds = ["h", "x", "rz", "cx", "cz", "cp", "swap", "ccx", "rxx", "rzz", "cswap", "rccx", "u3", "id", "measure",
             "reset", "barrier", "c3sqrtx"]
    for _ in range(40):
        n = 5
        insts = []
        for _ in range(30):
            k = kinds[int(rng.integers(len(kinds)))]
            q = rng.choice(n, size=4, replace=False).tolist()
            if k in ("h", "x", "id", "reset"):
                i = inst(k, qubit=q[0])
            elif k == "rz":
                i = inst(k, qubit=q[0], phi=0.3)
            elif k == "u3":
                i = inst(k, qubit=q[0], theta=0.3, phi=0.2, lam=0.1)
            elif k in ("cx", "cz"):
                i = inst(k, control=q[0], target=q[1])
            elif k == "cp":
                i = inst(k, control=q[0], target=q[1], lam=0.7)
            elif k == "swap":
                i = inst(k, a=q[0], b=q[1])
            elif k in ("rxx", "rzz"):
                i = inst(k, a=q[0], b=q[1], theta=0.4)
            elif k in ("ccx", "rccx"):
                i = inst(k, control1=q[0], control2=q[1], target=q[2])
            elif k == "c3sqrtx":
                i = inst(k, control1=q[0], control2=q[1], control3=q[2], target=q[3])
            elif k == "cswap":
                i = inst(k, control=q[0], target1=q[1], target2=q[2])
            elif k == "measure":
                i = inst(k, qubit=q[0], cbit=q[1])
            else:
                i = inst(k)
            if rng.random() < 0.15 and k != "barrier":
                i = cond(i, int(rng.integers(2)), int(rng.integers(n)))
            insts.append(i)
        c = Circuit(n, insts, cbits=n)
        assert _product_kinds(_core.plan_json(c)) == oracle.compile_kinds(insts)


def test_fusion_plan_conserves_gates_and_respects_width():
    circ = workloads.random_su2(n=20, depth=8, seed=3)
    plain = _core.plan_json(circ, tile_qubits=0)
    for k in (2, 3, 4, 5):
        fused = _core.plan_json(circ, precision="complex64", fusion="greedy", max_fused_qubits=k, tile_qubits=0)
        assert sum(l["gates"] for l in fused["launches"]) == plain["compiled_ops"]
        assert all(l["k"] <= k for l in fused["launches"] if l["gates"] > 1)
        assert len(fused["launches"]) < len(plain["launches"])
        assert all(l["mfma"] == (l["kind"] == "fused" and l["k"] >= 4) for l in fused["launches"])
        tiled = _core.plan_json(circ, precision="complex64", fusion="greedy", max_fused_qubits=k, tile_qubits=12)
        assert sum(l["gates"] for l in tiled["launches"]) == plain["compiled_ops"]
        assert len(tiled["launches"]) < len(fused["launches"])
        assert all(len(b["qubits"]) <= k for l in tiled["launches"] if l["kind"] == "tile" for b in l["blocks"]
                   if len(b["qubits"]) > 2)


def test_workload_gate_counts():
    assert workloads.gate_count(workloads.brickwork().data["instructions"]) == 590
    assert workloads.gate_count(workloads.qft(30, measure=False)[0].data["instructions"]) == 510
    assert workloads.gate_count(workloads.qft(33, measure=False)[0].data["instructions"]) == 610
    assert workloads.gate_count(workloads.random_su2(32, 40).data["instructions"]) == 1920


def test_learned_planner_merge_decisions_follow_the_model():
    from dqsim import planner

    circ = workloads.random_su2(n=16, depth=6, seed=2)
    greedy = _core.plan_json(circ, precision="complex64", fusion="greedy", max_fused_qubits=4, tile_qubits=0)
    w = np.zeros(planner.N_PARAMS, dtype=np.float32)
    try:
        # Constant prediction: every merge saves one launch -> greedy plan.
        w[-1] = 1.0
        planner.install(w)
        same = _core.plan_json(circ, precision="complex64", fusion="learned", max_fused_qubits=4, tile_qubits=0)
        assert [l["k"] for l in same["launches"]] == [l["k"] for l in greedy["launches"]]
        # Time growing steeply with width (feature 0 through one hidden unit):
        # only merges that do not widen a block pay off, so no launch is wider
        # than the widest compiled gate (2 qubits).
        w[:] = 0
        w[0 * 64 + 0] = 1.0       # W1[k][0]
        F = planner.N_FEATURES
        w[F * 64 + 64 + 0] = 60.0  # W2[0][0]
        w[F * 64 + 64 + 64 * 64 + 64 + 0] = 1.0  # w3[0]
        planner.install(w)
        narrow = _core.plan_json(circ, precision="complex64", fusion="learned", max_fused_qubits=4, tile_qubits=0)
        assert max(l["k"] for l in narrow["launches"]) == 2
        assert len(narrow["launches"]) > len(greedy["launches"])
        assert sum(l["gates"] for l in narrow["launches"]) == narrow["compiled_ops"]
    finally:
        planner.install(None)
    with pytest.raises(ValueError):
        planner.install(np.zeros(5, dtype=np.float32))


def _parse_batch_program(text):
    lines = text.split("\n")
    nc, nops, nmats, total_shots, slot = map(int, lines[0].split())
    circ = [list(map(int, lines[1 + i].split())) for i in range(nc)]
    order = list(map(int, lines[1 + nc].split()))
    ops = [list(map(int, lines[2 + nc + j].split())) for j in range(nops)]
    mats = np.array(lines[2 + nc + nops].split(), dtype=np.float64)
    return nc, nops, nmats, total_shots, slot, circ, order, ops, mats


@pytest.mark.parametrize("k", [2, 3, 4])
def test_batched_program_stays_inside_its_buffers(k):
    # Every index the batched kernel derives from the uploaded program lies
    # inside the buffer it addresses: op ranges, matrix ranges, qubits, output
    # ranges, the HBM slot of circuits above the LDS limit, and the launch
    # order is a permutation (largest estimated cost first).
    from dqsim import batch
    cs = ([workloads.batch_circuit(i, base_seed=77, nmin=3, nmax=12) for i in range(12)]
          + [workloads.batch_circuit(i, base_seed=5, nmin=15, nmax=16) for i in range(3)])
    text = batch.program_text(cs, shots=300, base_seed=11, max_fused_qubits=k, lds_qubits=14)
    nc, nops, nmats, total_shots, slot, circ, order, ops, mats = _parse_batch_program(text)
    assert nc == len(cs) and len(ops) == nops and mats.size == nmats and total_shots == 300 * nc
    assert sorted(order) == list(range(nc))
    cost = [(c[1] + 2) * 2.0 ** c[0] for c in circ]
    assert all(cost[order[i]] >= cost[order[i + 1]] for i in range(nc - 1))
    assert slot == max(1 << c[0] for c in circ if c[0] > 14)
    out_end = 0
    for i, (n, nop, op_off, seed, shots, out_off) in enumerate(circ):
        assert n == cs[i].num_qubits and seed == 11 + i and shots == 300 and out_off == out_end
        out_end += shots
        assert op_off + nop <= nops
        for kk, mat_off, *qs in ops[op_off:op_off + nop]:
            q, srt = qs[:5], qs[5:]
            assert 1 <= kk <= k and mat_off + 2 * 4 ** kk <= nmats
            assert len(set(q[:kk])) == kk and all(x < n for x in q[:kk]) and srt[:kk] == sorted(q[:kk])
    assert out_end == total_shots
    assert np.isfinite(mats).all()


def test_batched_runner_validates_before_touching_the_device():
    # The persistent kernel raises (disabled), and bad options or circuits
    # that are not terminal-measurement circuits fail before any GPU call.
    from dqsim import batch
    c = workloads.batch_circuit(0, base_seed=1, nmin=3, nmax=4)
    with pytest.raises(RuntimeError, match="persistent batched kernel is disabled"):
        batch.run_batch([c], shots=8, persistent=True)
    with pytest.raises(ValueError, match="workers"):
        batch.run_batch([c], shots=8, workers=17)
    with pytest.raises(ValueError, match="workers"):
        batch.run_batch([c], shots=8, workers=-2)
    with pytest.raises(ValueError, match="fusion width"):
        batch.run_batch([c], shots=8, max_fused_qubits=5)
    with pytest.raises(ValueError, match="tile_qubits"):
        batch.run_batch([c], shots=8, tile_qubits=9)
    with pytest.raises(ValueError, match="precision"):
        batch.run_batch([c], shots=8, precision="complex32")
    mid = Circuit(2, [{"kind": "h", "qubit": 0}, {"kind": "measure", "qubit": 0, "cbit": 0},
                      {"kind": "x", "qubit": 1}], 1)
    with pytest.raises(ValueError, match="terminal measurements only"):
        batch.run_batch([c, mid], shots=8)


def _mixed_tile_circuit(n, depth, seed):
    rng = np.random.default_rng(seed)
    ins = []
    for _ in range(depth):
        for q in range(n):
            ins.append({"kind": "u3", "qubit": q, "theta": float(rng.uniform(0, 3)), "phi": 0.3, "lam": -0.2})
        for _ in range(n // 2):
            q = [int(x) for x in rng.choice(n, 5, replace=False)]
            kind = ["cx", "cz", "swap", "ccx", "c4x", "rzz"][int(rng.integers(6))]
            if kind in ("cx", "cz"):
                ins.append({"kind": kind, "control": q[0], "target": q[1]})
            elif kind == "swap":
                ins.append({"kind": kind, "a": q[0], "b": q[1]})
            elif kind == "rzz":
                ins.append({"kind": kind, "a": q[0], "b": q[1], "theta": 0.4})
            elif kind == "ccx":
                ins.append({"kind": kind, "control1": q[0], "control2": q[1], "target": q[2]})
            else:
                ins.append({"kind": kind, "control1": q[0], "control2": q[1], "control3": q[2],
                            "control4": q[3], "target": q[4]})
    return Circuit(n, ins)


def _replay_layout(plan, F):
    """Replays the state's bit layout through a plan: each tile reads every
    local qubit from the position the layout holds it at, local bits 0..F-1
    are positions 0..F-1 at the load and at the store, the store is a
    permutation of the tile's own positions (in place), low position j only
    ever holds qubit j or a qubit >= F, launches other than tiles run in the
    identity layout, and the plan ends in the identity layout. Returns the
    number of tiles that relabel."""
    n = plan["num_qubits"]
    pos = list(range(n))
    relabels = 0
    for x in plan["launches"]:
        if x["kind"] != "tile":
            assert pos == list(range(n)), "launch outside a tile in a relabelled layout"
            continue
        lq, sp, op, sw = x["local_qubits"], x["load_positions"], x["store_positions"], x["swaps"]
        L = len(lq)
        assert L == x["k"] and len(set(lq)) == L and sorted(lq) == sorted(x["qubits"])
        assert [pos[q] for q in lq] == sp
        assert sp[:F] == list(range(F)) and op[:F] == list(range(F))
        assert sorted(op) == sorted(sp)
        assert len(sw) <= 2 and all(0 <= j < F for j in sw)
        content = list(lq)
        for s, j in enumerate(sw):
            content[j], content[8 + s] = content[8 + s], content[j]
        for b in range(L):
            pos[content[b]] = op[b]
        assert sorted(pos) == list(range(n))
        at = {p: q for q, p in enumerate(pos)}
        assert all(at[j] == j or at[j] >= F for j in range(F))
        relabels += bool(sw) or op != sp
        for b in x["blocks"]:
            assert set(b["qubits"]) <= set(lq)
    assert pos == list(range(n)), "plan does not end in the identity layout"
    return relabels


@pytest.fixture(params=[True, False], ids=["relabel", "fixed-low"])
def relabel(request):
    prev = _core.set_sweep_relabel(request.param)
    yield request.param
    _core.set_sweep_relabel(prev)


@pytest.mark.parametrize("L,k", [(10, 2), (10, 5), (11, 4), (12, 3), (12, 4), (12, 5)])
@pytest.mark.parametrize("kind", ["su2", "mixed", "qft"])
def test_tile_sweep_plan_is_a_valid_reordering(kind, L, k, relabel):
    # Every compiled op is applied exactly once, ops that share a qubit keep
    # their order, each sweep's blocks fit its L tile qubits (which always
    # include the qubits at positions 0..3), and the bit layout stays
    # consistent and ends as the identity.
    n = 20
    if kind == "su2":
        c = workloads.random_su2(n=n, depth=12, seed=L * 10 + k)
    elif kind == "qft":
        c = workloads.qft(n, seed=k, measure=False)[0]
    else:
        c = _mixed_tile_circuit(n, 6, seed=L + k)
    plain = _core.plan_json(c, precision="complex64", fusion="none")
    opq = {x["ops"][0]: set(x["qubits"]) for x in plain["launches"]}
    plan = _core.plan_json(c, precision="complex64", fusion="greedy", max_fused_qubits=k, tile_qubits=L)
    assert plan["compiled_ops"] == plain["compiled_ops"]
    order = [o for x in plan["launches"] for o in x["ops"]]
    assert sorted(order) == list(range(plan["compiled_ops"]))
    pos = {o: i for i, o in enumerate(order)}
    last = {}
    for o in range(plan["compiled_ops"]):  # original order
        for q in opq.get(o, ()):
            if q in last:
                assert pos[last[q]] < pos[o], (last[q], o, q)
            last[q] = o
    tiles = [x for x in plan["launches"] if x["kind"] == "tile"]
    assert tiles
    relabels = _replay_layout(plan, 4)
    assert (relabels > 0) == relabel
    for t in tiles:
        s = set(t["qubits"])
        assert len(s) == L and t["k"] == L
        if not relabel:
            assert {0, 1, 2, 3} <= s and len(t["blocks"]) >= 2
        for b in t["blocks"]:
            assert set(b["qubits"]) <= s and 1 <= len(b["qubits"]) <= 5
            assert b["mfma"]  # k <= 3 (padded to 3) on 16x16x4, k = 4 on 32x32x2
    # Block ops inside a tile: the union of the tile's blocks covers its ops.
    for t in tiles:
        for o in t["ops"]:
            assert opq[o] <= set(t["qubits"])


def test_tile_sweeps_off_and_limits():
    c = workloads.random_su2(n=18, depth=6, seed=1)
    off = _core.plan_json(c, precision="complex64", fusion="greedy", tile_qubits=0)
    assert not any(x["kind"] == "tile" for x in off["launches"])
    # complex128 and small states never use tiles.
    assert not any(x["kind"] == "tile" for x in _core.plan_json(c, fusion="greedy")["launches"])
    small = workloads.random_su2(n=15, depth=6, seed=1)
    assert not any(x["kind"] == "tile"
                   for x in _core.plan_json(small, precision="complex64", fusion="greedy")["launches"])
    for bad in (9, 14, 64):
        with pytest.raises(ValueError):
            _core.plan_json(c, precision="complex64", fusion="greedy", tile_qubits=bad)
        with pytest.raises(ValueError):
            dqsim.StatevectorSimulator(precision="complex64", tile_qubits=bad)


@pytest.mark.parametrize("kind", ["su2", "mixed", "qft"])
def test_exact_sweeps_keep_program_order(kind, relabel):
    # complex128 without fusion: sweeps are runs of consecutive compiled ops
    # (nothing reordered), each within L tile qubits that include the qubits
    # at positions 0..2, with a consistent bit layout ending as the identity.
    n = 18
    if kind == "su2":
        c = workloads.random_su2(n=n, depth=6, seed=4)
    elif kind == "qft":
        c = workloads.qft(n, seed=2, measure=False)[0]
    else:
        c = _mixed_tile_circuit(n, 5, seed=9)
    for L in (10, 11, 12):
        plan = _core.plan_json(c, tile_qubits=L)
        order = [o for x in plan["launches"] for o in x["ops"]]
        assert order == list(range(plan["compiled_ops"]))
        tiles = [x for x in plan["launches"] if x["kind"] == "tile"]
        assert tiles and len(plan["launches"]) < plan["compiled_ops"]
        assert (_replay_layout(plan, 3) > 0) == relabel
        for t in tiles:
            s = set(t["qubits"])
            assert len(s) == min(L, 11) and ({0, 1, 2} <= s or relabel)
            assert all(set(b["qubits"]) <= s for b in t["blocks"])
    assert not any(x["kind"] == "tile" for x in _core.plan_json(c, tile_qubits=0)["launches"])


def test_relabelled_sweeps_take_fewer_passes():
    # On the depth-40 random SU(2) family, relabelling the low positions
    # (restore passes included) needs fewer sweeps than pinning qubits 0..3
    # (complex64) or 0..2 (complex128) into every tile.
    c = workloads.random_su2(n=26, depth=40, seed=1234)
    counts = {}
    for on in (True, False):
        prev = _core.set_sweep_relabel(on)
        try:
            c64 = _core.plan_json(c, precision="complex64", fusion="greedy", max_fused_qubits=3, tile_qubits=12)
            c128 = _core.plan_json(c, tile_qubits=11)
        finally:
            _core.set_sweep_relabel(prev)
        counts[on] = (len(c64["launches"]), len(c128["launches"]))
        assert _replay_layout(c64, 4) >= 0 and _replay_layout(c128, 3) >= 0
    assert counts[True][0] < counts[False][0] and counts[True][1] < counts[False][1], counts


def test_exact_sweep_chains_use_compiled_signatures():
    # Compiled ops of the benchmark family come as One(a), One(b), CX/CZ(a, b)
    # triples: nearly every register chain matches a compiled signature, while
    # circuits with controlled and multi-qubit ops keep the generic chains.
    c = workloads.random_su2(n=24, depth=10, seed=5)
    tiles = [x for x in _core.plan_json(c)["launches"] if x["kind"] == "tile"]
    chains = sum(t["chains"] for t in tiles)
    fixed = sum(t["fixed_chains"] for t in tiles)
    assert chains > 0 and fixed >= 0.95 * chains
    m = _mixed_tile_circuit(18, 5, seed=9)
    tiles = [x for x in _core.plan_json(m)["launches"] if x["kind"] == "tile"]
    assert sum(t["fixed_chains"] for t in tiles) < sum(t["chains"] for t in tiles)


def test_learned_plan_selection_follows_the_model():
    # complex64 tile plans under fusion="learned" are chosen among (block
    # width, tile size) candidates by the model's predicted total time.
    from dqsim import planner

    c = workloads.random_su2(n=20, depth=10, seed=3)
    F = planner.N_FEATURES
    w = np.zeros(planner.N_PARAMS, dtype=np.float32)
    try:
        # Every launch predicts the same time: the fewest launches win (12-qubit tiles).
        planner.install(w)
        plan = _core.plan_json(c, precision="complex64", fusion="learned", max_fused_qubits=4)
        tiles = [x for x in plan["launches"] if x["kind"] == "tile"]
        assert tiles and all(len(t["qubits"]) == 12 for t in tiles)
        order = sorted(o for x in plan["launches"] for o in x["ops"])
        assert order == list(range(plan["compiled_ops"]))
        # A model that charges heavily for 4-qubit blocks in a sweep (feature 15).
        w[15 * 64 + 0] = 1.0                       # W1[15][0]
        w[F * 64 + 64 + 0] = 1.0                   # W2[0][0]
        w[F * 64 + 64 + 64 * 64 + 64 + 0] = 50.0   # w3[0]
        planner.install(w)
        plan = _core.plan_json(c, precision="complex64", fusion="learned", max_fused_qubits=4)
        assert all(len(b["qubits"]) <= 3 for x in plan["launches"] if x["kind"] == "tile" for b in x["blocks"])
        greedy4 = _core.plan_json(c, precision="complex64", fusion="greedy", max_fused_qubits=4)
        assert any(len(b["qubits"]) == 4 for x in greedy4["launches"] if x["kind"] == "tile" for b in x["blocks"])
    finally:
        planner.install(None)


def test_exact_sweeps_group_ops_into_register_chains():
    # Each exact sweep applies its ops in passes: register chains of
    # consecutive 1-qubit / X / diagonal ops on <= 2 qubits, other ops alone.
    c = workloads.random_su2(n=18, depth=6, seed=4)
    plan = _core.plan_json(c)
    tiles = [x for x in plan["launches"] if x["kind"] == "tile"]
    assert tiles and all(t["exact"] for t in tiles)
    assert sum(t["chains"] for t in tiles) < sum(len(t["ops"]) for t in tiles)
    for t in tiles:
        # Restore passes (layout only) apply no ops.
        assert (1 <= t["chains"] <= len(t["ops"])) or (t["chains"] == 0 and not t["ops"] and t["gates"] == 0)


def test_packed_count_table_round_trip():
    # The library's own dq_count_cb appends into a packed table that Python
    # reads once; keys keep emission order, any length, full uint64 counts.
    import ctypes
    from dqsim import _core
    t = _core._CountTable()
    entries = [("0101", 7), ("1", 1), ("", 0), ("1110", 2 ** 40)]
    for key, count in entries:
        b = key.encode()
        assert t.cb(t.user, ctypes.cast(ctypes.c_char_p(b), ctypes.POINTER(ctypes.c_char)), len(b), count) == 0
    assert list(t.result().items()) == entries
    assert _core._CountTable().result() == {}


def test_noalloc_evolution_validates_before_touching_the_device():
    # dq_sv_evolve_noalloc: circuit, unitarity and buffer length are checked
    # before any device call (a null buffer is never dereferenced).
    sim = dqsim.StatevectorSimulator(seed=1)
    measured = Circuit(3, [{"kind": "h", "qubit": 0}, {"kind": "measure", "qubit": 0, "cbit": 0}], 1)
    bare = Circuit(3, [{"kind": "h", "qubit": 1}])
    cond = Circuit(3, [{"kind": "conditional", "condition": {"creg_base": 0, "creg_size": 1, "creg_value": 1},
                        "op": {"kind": "x", "qubit": 0}}], 1)
    for circ, length, text in ((measured, 8, "unitary circuits only"), (cond, 8, "unitary circuits only"),
                               (bare, 4, "holds 4 amplitudes"), (bare, 8, "null")):
        js = _core._circuit_json(circ)
        with pytest.raises(ValueError, match=text):
            _core._check(_core._lib.dq_sv_evolve_noalloc(sim._h, js, len(js), None, length, None))
    with pytest.raises(ValueError):
        _core._check(_core._lib.dq_sv_evolve_noalloc(sim._h, b"{", 1, None, 8, None))


def _xsweep_move(state, launch, n, L):
    """Data movement of one out-of-place sweep launch (blocks left out): the
    kernel's read and write index maps restated with numpy."""
    t = np.arange(1 << L, dtype=np.uint64)
    tau = np.arange(1 << (n - L), dtype=np.uint64)
    rpos, spos = launch["load_positions"], launch["store_positions"]
    if launch["gather"]:
        local = sum(((t >> np.uint64(j)) & np.uint64(1)) << np.uint64(p) for j, p in enumerate(rpos))
        free = [p for p in range(n) if p not in rpos]
        base = sum(((tau >> np.uint64(k)) & np.uint64(1)) << np.uint64(p) for k, p in enumerate(free))
    else:
        assert rpos == list(range(L))
        local, base = t, tau << np.uint64(L)
    wlocal = sum(((t >> np.uint64(j)) & np.uint64(1)) << np.uint64(p) for j, p in enumerate(spos))
    wbase = sum(((tau >> np.uint64(k)) & np.uint64(1)) << np.uint64(p)
                for k, p in enumerate(launch["tile_store_positions"]))
    src = (base[:, None] | local[None, :]).ravel()
    dst = (wbase[:, None] | wlocal[None, :]).ravel()
    assert src.max() < (1 << n) and dst.max() < (1 << n)
    assert np.unique(src).size == src.size and np.unique(dst).size == dst.size
    out = np.empty_like(state)
    out[dst] = state[src]
    return out


@pytest.fixture(params=[True, False], ids=["pinned", "free"])
def sweep_pinned(request):
    prev = _core.set_sweep_pinned(request.param)
    yield request.param
    _core.set_sweep_pinned(prev)


@pytest.mark.parametrize("kind,n,L,k", [("su2", 16, 12, 3), ("su2", 18, 12, 4), ("su2", 17, 11, 3),
                                        ("mixed", 17, 12, 3), ("qft", 16, 10, 2), ("su2", 20, 12, 2),
                                        ("brickwork", 22, 12, 3), ("su2", 18, 13, 3), ("mixed", 19, 13, 4)])
def test_out_of_place_sweeps_move_every_amplitude_home(kind, n, L, k, sweep_pinned):
    # Tag every amplitude with its index and push the tags through the
    # launches' read/write maps: each launch's tile qubits must sit where the
    # current layout holds them, the blocks must act on tile qubits, and a run
    # of sweeps must end with every amplitude back at its own index. Pinned
    # tiles hold qubits 0..3 always; free tiles hold them at a run's first and
    # last sweep and share at least 4 qubits with the previous tile.
    c = {"su2": lambda: workloads.random_su2(n=n, depth=10, seed=n + k),
         "mixed": lambda: workloads.mixed(n, 5, seed=n),
         "qft": lambda: workloads.qft(n, seed=2, measure=False)[0],
         "brickwork": lambda: workloads.brickwork(n=n, layers=12, seed=n, measure=False)}[kind]()
    plan = _core.plan_json(c, precision="complex64", fusion="greedy", max_fused_qubits=k, tile_qubits=L,
                           out_of_place=True)
    launches = plan["launches"]
    runs, cur = [], []
    for x in launches:
        if x.get("out_of_place"):
            cur.append(x)
        elif cur:
            runs.append(cur)
            cur = []
    if cur:
        runs.append(cur)
    for run in runs:
        assert {0, 1, 2, 3} <= set(run[0]["local_qubits"]) and {0, 1, 2, 3} <= set(run[-1]["local_qubits"])
        for a, b in zip(run, run[1:]):
            assert len(set(a["local_qubits"]) & set(b["local_qubits"])) >= 4
    if not sweep_pinned and kind in ("su2", "brickwork") and n >= 18:
        # Free tiles leave qubits 0..3 out of some sweep.
        assert any(not {0, 1, 2, 3} <= set(x["local_qubits"]) for run in runs for x in run)
    assert any(x.get("out_of_place") for x in launches)
    # Each run of sweeps (between standalone launches) may gather its first
    # tile only, and does unless that tile is qubits 0..L-1.
    for i, x in enumerate(launches):
        if x.get("out_of_place"):
            first = i == 0 or not launches[i - 1].get("out_of_place")
            assert x["gather"] == (first and x["load_positions"] != list(range(L))), i
    tags = np.arange(1 << n, dtype=np.uint64)
    pos = list(range(n))  # qubit -> bit position
    for x in plan["launches"]:
        if not x.get("out_of_place"):
            assert pos == list(range(n))  # other launches see the identity layout
            continue
        lq = x["local_qubits"]
        assert [pos[q] for q in lq] == x["load_positions"]
        assert all(set(b["qubits"]) <= set(lq) for b in x["blocks"])
        if sweep_pinned:
            assert {0, 1, 2, 3} <= set(lq)
        tags = _xsweep_move(tags, x, n, L)
        free = [q for q in sorted(range(n), key=lambda q: pos[q]) if q not in lq]
        new = [None] * n
        for j, q in enumerate(lq):
            new[q] = x["store_positions"][j]
        for kk, q in enumerate(free):
            new[q] = x["tile_store_positions"][kk]
        pos = new
        # A tag (original index) now sits at the index its bits map to.
        probe = np.array([5, (1 << n) - 3, 12345 % (1 << n)], dtype=np.uint64)
        moved = sum(((probe >> np.uint64(q)) & np.uint64(1)) << np.uint64(pos[q]) for q in range(n))
        assert np.array_equal(tags[moved.astype(np.int64)], probe)
        # Stores move runs of at least 16 amplitudes (positions 0..3 are tile bits).
        assert set(range(4)) <= set(x["store_positions"])
    assert pos == list(range(n))
    assert np.array_equal(tags, np.arange(1 << n, dtype=np.uint64))
    # Same ops as the in-place plan, each once.
    ops = sorted(o for x in plan["launches"] for o in x["ops"])
    assert ops == list(range(plan["compiled_ops"]))


@pytest.mark.parametrize("kind,n,tile", [("su2", 16, 12), ("mixed", 17, 12), ("qft", 18, 10), ("su2", 20, 12),
                                         ("brickwork", 19, 10)])
def test_exact_out_of_place_sweeps_keep_program_order(kind, n, tile):
    # complex128 without fusion and room for a second state: runs of
    # consecutive ops (program order, nothing reordered) in out-of-place
    # sweeps of L = min(tile, 11) qubits; every amplitude ends at its own
    # index, each sweep's ops act on its tile, consecutive tiles share >= 3
    # qubits, a run's first and last tiles hold qubits 0..2, and stores move
    # runs of >= 8 amplitudes (128 bytes).
    L = min(tile, 11)
    c = {"su2": lambda: workloads.random_su2(n=n, depth=8, seed=n),
         "mixed": lambda: workloads.mixed(n, 5, seed=n),
         "qft": lambda: workloads.qft(n, seed=2, measure=False)[0],
         "brickwork": lambda: workloads.brickwork(n=n, layers=10, seed=n, measure=False)}[kind]()
    plan = _core.plan_json(c, precision="complex128", fusion="none", tile_qubits=tile, out_of_place=True)
    launches = plan["launches"]
    xs = [x for x in launches if x.get("out_of_place")]
    assert xs and all(x["exact"] and x["k"] == L for x in xs)
    order = [o for x in launches for o in x["ops"]]
    assert order == list(range(plan["compiled_ops"]))  # program order, each op once
    tags = np.arange(1 << n, dtype=np.uint64)
    pos = list(range(n))
    prev = None
    for i, x in enumerate(launches):
        if not x.get("out_of_place"):
            assert pos == list(range(n))
            prev = None
            continue
        lq = x["local_qubits"]
        assert [pos[q] for q in lq] == x["load_positions"]
        assert all(set(b["qubits"]) <= set(lq) for b in x["blocks"])
        if prev is None:
            assert {0, 1, 2} <= set(lq)
        else:
            assert len(set(prev) & set(lq)) >= 3
        last = i + 1 == len(launches) or not launches[i + 1].get("out_of_place")
        if last:
            assert {0, 1, 2} <= set(lq)
        assert set(range(3)) <= set(x["store_positions"])
        tags = _xsweep_move(tags, x, n, L)
        free = [q for q in sorted(range(n), key=lambda q: pos[q]) if q not in lq]
        new = [None] * n
        for j, q in enumerate(lq):
            new[q] = x["store_positions"][j]
        for kk, q in enumerate(free):
            new[q] = x["tile_store_positions"][kk]
        pos = new
        prev = lq
    assert pos == list(range(n))
    assert np.array_equal(tags, np.arange(1 << n, dtype=np.uint64))
    assert plan["zero_start"] == bool(launches[0].get("out_of_place"))


def test_threaded_sweep_emission_is_deterministic():
    # Plans of 24 qubits and up build their sweeps' blocks on host threads and
    # append them in sweep order: the launches and the matrix blob (hash) are
    # the same on every call, and the learned plan is the greedy plan of the
    # width it chose, blob included.
    from dqsim import planner
    c = workloads.random_su2(n=26, depth=24, seed=8)
    runs = [_core.plan_json(c, precision="complex64", fusion="greedy", max_fused_qubits=3, tile_qubits=12,
                            out_of_place=True) for _ in range(4)]
    assert sum(1 for x in runs[0]["launches"] if x.get("out_of_place")) >= 8
    assert all(r == runs[0] for r in runs[1:])
    assert runs[0]["blob_bytes"] > 0 and len(runs[0]["blob_hash"]) > 0
    planner.install(planner.load_default())
    try:
        lp = _core.plan_json(c, precision="complex64", fusion="learned", max_fused_qubits=5, tile_qubits=12,
                             out_of_place=True)
    finally:
        planner.install(None)
    (choice,) = lp["learned_choices"]
    prev = _core.set_sweep_pinned(choice["pinned"])
    try:
        gp = _core.plan_json(c, precision="complex64", fusion="greedy", max_fused_qubits=choice["k"],
                             tile_qubits=12, out_of_place=True)
    finally:
        _core.set_sweep_pinned(prev)
    assert gp["launches"] == lp["launches"] and gp["blob_hash"] == lp["blob_hash"]


def test_thirteen_qubit_tiles_are_out_of_place_only():
    # 13-qubit tiles fill the 64 KiB of LDS: out-of-place sweeps take them
    # (512 threads, one workgroup per tile), in-place sweeps of the same
    # request use 12-qubit tiles; 14 is rejected.
    c = workloads.random_su2(n=20, depth=10, seed=4)
    xp = _core.plan_json(c, precision="complex64", fusion="greedy", max_fused_qubits=3, tile_qubits=13,
                         out_of_place=True)
    xs = [x for x in xp["launches"] if x.get("out_of_place")]
    assert xs and all(x["k"] == 13 and len(x["local_qubits"]) == 13 and not x["persist"] for x in xs)
    assert all(x["features"][9] == 1.0 for x in xs)  # 512-thread workgroups
    ip = _core.plan_json(c, precision="complex64", fusion="greedy", max_fused_qubits=3, tile_qubits=13,
                         out_of_place=False)
    tiles = [x for x in ip["launches"] if x["kind"] == "tile"]
    assert tiles and all(x["k"] == 12 and not x.get("out_of_place") for x in tiles)
    with pytest.raises(ValueError):
        _core.plan_json(c, precision="complex64", fusion="greedy", max_fused_qubits=3, tile_qubits=14)
    with pytest.raises(ValueError):
        dqsim.StatevectorSimulator(precision="complex64", fusion="greedy", tile_qubits=14)


@pytest.mark.parametrize("kind,n,k", [("su2", 20, 3), ("brickwork", 22, 4), ("mixed", 18, 2)])
def test_out_of_place_sweep_features_match_their_launch(kind, n, k, sweep_pinned):
    # The learned planner's inputs for an out-of-place sweep, recomputed from
    # the launch's own layout and blocks.
    c = {"su2": lambda: workloads.random_su2(n=n, depth=12, seed=n),
         "brickwork": lambda: workloads.brickwork(n=n, layers=10, seed=n, measure=False),
         "mixed": lambda: workloads.mixed(n, 5, seed=n)}[kind]()
    plan = _core.plan_json(c, precision="complex64", fusion="greedy", max_fused_qubits=k, tile_qubits=12,
                           out_of_place=True)
    assert plan["learned_choices"] == []
    seen = 0
    for x in plan["launches"]:
        if not x.get("out_of_place"):
            continue
        seen += 1
        f = np.array(x["features"])
        pos = {q: j for j, q in enumerate(x["local_qubits"])}
        sizes = [len(b["qubits"]) for b in x["blocks"]]
        low = [min(pos[q] for q in b["qubits"]) for b in x["blocks"]]
        sp = x["store_positions"]
        run = 0
        while run in sp:
            run += 1
        want = np.zeros(16)
        want[0] = 1
        want[1] = sum(1 for s, lo in zip(sizes, low) if s == 3 and lo < 4) / 16
        want[2] = sum(1 for s, lo in zip(sizes, low) if s <= 2 and lo < 6) / 16
        want[3] = sum(1 for s in sizes if s <= 2) / 16
        want[4] = x["gates"] / 32
        want[5] = 1.0 if x["gather"] else 0.0
        want[6] = sum(1 for j, p in enumerate(sp) if p < 4 and j >= 4) / 4
        want[7] = run / 12
        want[8] = n / 34
        want[11] = f[11]  # blob bytes: checked for range below
        want[12] = 1
        want[13] = 1
        want[14] = sum(1 for s in sizes if s == 3) / 16
        want[15] = sum(1 for s in sizes if s == 4) / 16
        assert np.allclose(f, want, atol=1e-6), (f, want)
        assert 0 < f[11] <= 1
    assert seen > 0


def test_learned_choice_selects_one_of_the_greedy_plans():
    # With weights installed, the out-of-place plan of fusion="learned" is
    # the greedy plan of the (tile rule, block width) it reports choosing.
    from dqsim import planner
    rng = np.random.default_rng(4)
    prev = _core.set_sweep_pinned(False)
    try:
        for trial in range(4):
            w = planner.load_default() if trial == 0 else None
            if w is None:
                w = rng.normal(0, 0.3, planner.N_PARAMS).astype(np.float32)
            planner.install(w)
            c = workloads.random_su2(n=22, depth=16, seed=trial)
            lp = _core.plan_json(c, precision="complex64", fusion="learned", max_fused_qubits=3,
                                 out_of_place=True)
            assert len(lp["learned_choices"]) == 1
            ch = lp["learned_choices"][0]
            assert ch["k"] in (2, 3)
            planner.install(None)
            _core.set_sweep_pinned(ch["pinned"])
            gp = _core.plan_json(c, precision="complex64", fusion="greedy", max_fused_qubits=ch["k"],
                                 out_of_place=True)
            _core.set_sweep_pinned(False)
            assert [x["blocks"] for x in lp["launches"]] == [x["blocks"] for x in gp["launches"]]
            assert [x["store_positions"] for x in lp["launches"]] == [x["store_positions"] for x in gp["launches"]]
    finally:
        planner.install(None)
        _core.set_sweep_pinned(prev)


def test_shipped_planner_weights_load():
    from dqsim import planner
    w = planner.load_default()
    assert w is not None, "planner_weights.npy is missing from the package"
    assert w.dtype == np.float32 and w.shape == (planner.N_PARAMS,) and np.isfinite(w).all()
    assert planner.load_default("/nonexistent/planner_weights.npy") is None


def test_fp16_split_of_block_matrices_matches_ieee_rounding():
    # Split-precision block matrices are written as FP16 pairs hi + lo of the
    # FP64 embedding. hi must be IEEE binary16 round-to-nearest-even of x
    # (numpy's float16 conversion is that rounding) and lo the same rounding
    # of the exact residual x - hi; together they carry about 22 bits.
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.uniform(-1, 1, 20000), rng.uniform(-1e-4, 1e-4, 5000), rng.uniform(-1e-7, 1e-7, 2000),
                        np.array([0.0, -0.0, 2.0 ** -25, -2.0 ** -25, 2.0 ** -24, 3 * 2.0 ** -26, 2.0 ** -14,
                                  65504.0, 1.0, -1.0, 0.5 + 2.0 ** -12, 0.5 + 3 * 2.0 ** -12, 1 / 3])])
    x = np.ascontiguousarray(x, dtype=np.float64)
    hi = np.empty(x.size, dtype=np.uint16)
    lo = np.empty(x.size, dtype=np.uint16)
    _core._check(_core._lib.dq_split_f16(x.ctypes.data, x.size, hi.ctypes.data, lo.ctypes.data))
    assert np.array_equal(hi, x.astype(np.float16).view(np.uint16))
    hv = hi.view(np.float16).astype(np.float64)
    assert np.array_equal(lo, (x - hv).astype(np.float16).view(np.uint16))
    recon = hv + lo.view(np.float16).astype(np.float64)
    big = np.abs(x) > 2.0 ** -4
    assert np.all(np.abs(recon - x)[big] <= np.abs(x[big]) * 2.0 ** -21)
    with pytest.raises(RuntimeError):
        _core._check(_core._lib.dq_split_f16(np.array([70000.0]).ctypes.data, 1, hi.ctypes.data, lo.ctypes.data))
