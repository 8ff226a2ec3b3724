"""Measurement reports for the secondary workloads of the benchmark plan.

  precision_report ... one U3 layer + QFT: complex64 (fused) at n64 qubits and
                       complex128 (plain) at n128 qubits against the exact
                       closed form at sampled indices, the full-state L2 error
                       of complex64 against the complex128 GPU result, and the
                       total-variation distance of shots drawn with the same
                       Philox stream from both states.
  batch_report ....... thousands of small uneven circuits through the
                       stream-pool batched runner (autotuned workers and
                       fusion width): circuits/s, p50/p99 latency, gate-amp/s.
  brickwork_report ... 20-qubit brickwork (20 layers of U3 + CX, seed 42), 1000
                       shots: complex128 plain (reference semantics) and
                       complex64 fused, median of 5 runs after a warm-up.
  batch_loop_report .. the same circuits through the per-circuit API, one
                       simulate_shots call after another (the baseline the
                       batched runner is measured against).
  planner_report ..... the shipped learned planner (trained on engine-recorded
                       launch timings) vs the greedy fuser on held-out
                       28-33-qubit circuits, out-of-place tile sweeps.
"""

from __future__ import annotations

import json
import math
import os
import time
from typing import Dict, List

import numpy as np

from . import _core, batch, planner, workloads


def _tvd(a: Dict[str, int], b: Dict[str, int]) -> float:
    na, nb = sum(a.values()), sum(b.values())
    keys = set(a) | set(b)
    return 0.5 * sum(abs(a.get(k, 0) / na - b.get(k, 0) / nb) for k in keys)


def precision_report(n64: int = 30, n128: int = 33, shots: int = 4096, samples: int = 1 << 16, seed: int = 7,
                     log=print) -> dict:
    out = {}
    rng = np.random.default_rng(seed)
    circ, angles = workloads.qft(n64, seed=seed, measure=False)
    gates = workloads.gate_count(circ.data["instructions"])
    mem = (1 << n64) * 16 * 3 // (1 << 20) + 4096
    idx = np.unique(np.concatenate([np.arange(64), rng.integers(0, 1 << n64, size=samples)])).astype(np.uint64)
    exact = workloads.qft_closed_form(angles, idx)
    f = _core.StatevectorSimulator(seed=1, precision="complex64", fusion="greedy", max_fused_qubits=4,
                                   max_memory_mb=mem)
    _, fl, _, fms, fres = f.evolve_timed(circ, record=False, keep=True)
    p = _core.StatevectorSimulator(seed=1, precision="complex128", max_memory_mb=mem)
    _, pl, _, pms, pres = p.evolve_timed(circ, record=False, keep=True)
    f_amp, p_amp = fres.amplitudes(idx), pres.amplitudes(idx)
    full_max, full_l2 = fres.compare(pres)
    # With 2^30 outcomes almost every shot is distinct, so a TVD of two shot
    # histograms says nothing. Reported instead: the exact TVD of a 10-qubit
    # marginal (1,024 outcomes) of the two states, the TVD of shot histograms
    # over that marginal, and how many same-stream draws over all qubits land
    # on the same outcome.
    marg = [n64 - 1 - (3 * i) % n64 for i in range(10)]
    pm64, pm128 = fres.probabilities(marg), pres.probabilities(marg)
    marg_tvd = 0.5 * sum(abs(pm64.get(k, 0.0) - pm128.get(k, 0.0)) for k in set(pm64) | set(pm128))
    c64m = fres.counts(shots=shots, qubits=marg, seed=11)
    c128m = pres.counts(shots=shots, qubits=marg, seed=11)
    full_q = list(range(n64))[::-1]
    d64 = fres.counts(shots=shots, qubits=full_q, seed=11)
    d128 = pres.counts(shots=shots, qubits=full_q, seed=11)
    same = sum(min(d64.get(k, 0), d128.get(k, 0)) for k in d64)
    out["complex64"] = {"qubits": n64, "gates": gates, "launches": fl, "device_ms": fms,
                        "max_err_vs_closed_form": float(np.abs(f_amp - exact).max()),
                        "max_err_vs_fp64_gpu": full_max, "l2_err_vs_fp64_gpu": full_l2,
                        "marginal_qubits": marg, "marginal_tvd_exact": marg_tvd,
                        "marginal_tvd_shots": _tvd(c64m, c128m), "shots": shots,
                        "same_stream_draws_equal": same / shots,
                        "samples": int(idx.size)}
    out["complex128_same_n"] = {"qubits": n64, "launches": pl, "device_ms": pms,
                                "max_err_vs_closed_form": float(np.abs(p_amp - exact).max())}
    log(f"[precision] c64 n={n64}: {out['complex64']}")
    del fres, pres
    if n128:
        circ, angles = workloads.qft(n128, seed=seed, measure=False)
        idx = np.unique(np.concatenate([np.arange(64), rng.integers(0, 1 << n128, size=samples)])).astype(np.uint64)
        exact = workloads.qft_closed_form(angles, idx)
        mem = (1 << n128) * 16 // (1 << 20) + 4096
        p = _core.StatevectorSimulator(seed=1, precision="complex128", max_memory_mb=mem)
        _, pl, _, pms, pres = p.evolve_timed(circ, record=False, keep=True)
        amp = pres.amplitudes(idx)
        out["complex128"] = {"qubits": n128, "gates": workloads.gate_count(circ.data["instructions"]),
                             "launches": pl, "device_ms": pms,
                             "max_err_vs_closed_form": float(np.abs(amp - exact).max()),
                             "samples": int(idx.size)}
        log(f"[precision] c128 n={n128}: {out['complex128']}")
    return out


def batch_report(count: int = 4096, shots: int = 1024, base_seed: int = 2024, log=print) -> dict:
    """The batched workload (4096 circuits of 10-22 qubits, depth 10-60, 1024
    shots each, complex64): circuits of up to 12 qubits in the single-launch
    kernel (one hipGraph launch, one workgroup per circuit), the rest on the
    stream-pool runner with autotuned workers, fusion width and tile shape,
    both at once. `wall_ms` covers parsing, planning, evolution and sampling
    of every circuit (the tuning sample is timed separately). The same
    circuits also run with every circuit on the pool (`pool_only`)."""
    t0 = time.perf_counter()
    circs = [workloads.batch_circuit(i, base_seed=base_seed) for i in range(count)]
    gen_s = time.perf_counter() - t0
    work = sum(workloads.gate_count(c.data["instructions"]) * float(1 << c.num_qubits) for c in circs)
    pool = batch.run_batch(circs, shots=shots, base_seed=base_seed, autotune=True, lds_qubits=0)
    pool_only = {"wall_ms": pool.wall_ms, "circuits_per_s": count / (pool.wall_ms / 1e3),
                 "latency_us_p50": float(np.percentile(pool.latency_us, 50)),
                 "latency_us_p99": float(np.percentile(pool.latency_us, 99)), "workers": pool.workers,
                 "max_fused_qubits": pool.max_fused_qubits, "tile_qubits": pool.tile_qubits}
    del pool
    t0 = time.perf_counter()
    res = batch.run_batch(circs, shots=shots, base_seed=base_seed, autotune=True)
    wall = time.perf_counter() - t0
    for i in (0, count // 2, count - 1):
        if sum(res.counts(i).values()) != shots:
            raise RuntimeError(f"circuit {i}: wrong shot total")
    lat = np.sort(res.latency_us)
    out = {"circuits": count, "shots_each": shots, "wall_ms": res.wall_ms,
           "circuits_per_s": count / (res.wall_ms / 1e3), "gate_amp_per_s": work / (res.wall_ms / 1e3),
           "gate_amp_total": work, "latency_us_p50": float(np.percentile(lat, 50)),
           "latency_us_p99": float(np.percentile(lat, 99)), "workers": res.workers,
           "max_fused_qubits": res.max_fused_qubits, "tile_qubits": res.tile_qubits, "streams": res.streams,
           "tuning_ms_on_sample": res.tuning, "host_wall_s_incl_json_and_tuning": wall, "generate_s": gen_s,
           "single_launch_circuits": res.lds_circuits, "single_launch_device_ms": res.lds_ms,
           "pool_only": pool_only, "speedup_vs_pool_only": pool_only["wall_ms"] / res.wall_ms,
           "path": "circuits of <= 12 qubits: single-launch kernel (one hipGraph launch, one workgroup per circuit, "
                   "state in LDS); the rest: stream-pool runner (one host thread + HIP stream + device arena per "
                   "worker, simulate_shots terminal-sampling path per circuit), concurrently"}
    log(f"[batch] {out}")
    return out


PLANNER_HELD_OUT = (("su2", 28, 40), ("brickwork", 29, 40), ("qft", 30, 0), ("mixed", 31, 40), ("su2", 32, 40),
                    ("brickwork", 32, 60), ("qft", 33, 0))


def planner_report(held_out=PLANNER_HELD_OUT, repeats: int = 3, log=print) -> dict:
    """The shipped planner weights (planner_weights.npy, trained by
    ``python -m dqsim.planner`` on engine-recorded launch timings) against the
    greedy planner on held-out complex64 circuits (seeds 20_000 + n, never in
    training), out-of-place tile sweeps: greedy with blocks of <= 5 qubits
    (fused as far as the tiles allow), greedy with <= 3 (the hand-tuned
    width), and the learned choice of (tile rule, block width) at <= 5. Wall
    time per evolution (plan + upload + sweeps), median of `repeats`; max
    amplitude error of each plan against the complex128 path up to 32 qubits."""
    from . import planner

    if not planner.install_default():
        raise RuntimeError("planner_weights.npy is missing")
    meta_path = planner.WEIGHTS_FILE[:-4] + ".json"
    meta = json.load(open(meta_path)) if os.path.exists(meta_path) else {}
    configs = (("greedy_k5", "greedy", 5), ("greedy_k3", "greedy", 3), ("learned_k5", "learned", 5))
    rows = []
    try:
        for fam, n, depth in held_out:
            c = planner.family_circuit(fam, n, depth, 20_000 + n)
            mem = (1 << n) * 8 * 3 // (1 << 20) + 4096
            ref = None
            if n <= 32:
                exact = _core.StatevectorSimulator(seed=1, precision="complex128", max_memory_mb=mem * 2)
                ref = exact.evolve_timed(c, record=False, keep=True)[4]
            row = {"family": fam, "qubits": n, "depth": depth,
                   "learned_choice": _core.plan_json(c, precision="complex64", fusion="learned", max_fused_qubits=5,
                                                     out_of_place=True)["learned_choices"]}
            for name, fusion, k in configs:
                sim = _core.StatevectorSimulator(seed=1, precision="complex64", fusion=fusion, max_fused_qubits=k,
                                                 max_memory_mb=mem)
                sim.evolve_timed(c, record=False)  # warm-up
                walls = []
                for _ in range(repeats):
                    t = time.perf_counter()
                    sim.evolve_timed(c, record=False)
                    walls.append((time.perf_counter() - t) * 1e3)
                err = None
                if ref is not None:
                    res = sim.evolve_timed(c, record=False, keep=True)[4]
                    err, _ = res.compare(ref)
                    del res
                row[name] = {"ms": float(np.median(walls)), "max_err": err}
                del sim
            del ref
            row["learned_vs_greedy_k5"] = row["greedy_k5"]["ms"] / row["learned_k5"]["ms"]
            row["learned_vs_greedy_k3"] = row["greedy_k3"]["ms"] / row["learned_k5"]["ms"]
            rows.append(row)
            log(f"[planner] {row}")
    finally:
        planner.install(None)
    tot = {name: sum(r[name]["ms"] for r in rows) for name, _, _ in configs}
    out = {"weights": {k: meta.get(k) for k in ("circuits", "seed", "size_mix", "families", "depth_range",
                                                 "train_samples", "test_samples", "collect_s", "train_s",
                                                 "test_mse", "test_mse_predict_mean")},
           "eval": rows, "total_ms": tot,
           "learned_vs_greedy_k5": tot["greedy_k5"] / tot["learned_k5"],
           "learned_vs_greedy_k3": tot["greedy_k3"] / tot["learned_k5"],
           "max_err": max((v["max_err"] or 0.0) for r in rows for v in r.values()
                          if isinstance(v, dict) and "max_err" in v)}
    log(f"[planner] totals {tot}: learned vs greedy k<=5 {out['learned_vs_greedy_k5']:.3f}, "
        f"vs greedy k<=3 {out['learned_vs_greedy_k3']:.3f}")
    return out


def brickwork_report(n: int = 20, layers: int = 20, shots: int = 1000, seed: int = 42, repeats: int = 5,
                     log=print) -> dict:
    circ = workloads.brickwork(n=n, layers=layers, seed=seed)
    gates = workloads.gate_count(circ.data["instructions"])
    out = {"qubits": n, "gates": gates, "shots": shots}
    for name, kw in (("complex128_plain", {}),
                     ("complex64_fused", {"precision": "complex64", "fusion": "greedy", "max_fused_qubits": 3})):
        sim = _core.StatevectorSimulator(seed=seed, **kw)
        sim.simulate_shots(circ, shots=shots)  # warm-up
        walls = []
        for _ in range(repeats):
            t0 = time.perf_counter()
            counts = sim.simulate_shots(circ, shots=shots)
            walls.append(time.perf_counter() - t0)
        rows, launches, _, dev_ms, _ = sim.evolve_timed(circ, record=True)
        nbytes = sum(r["bytes"] for r in rows)
        med = float(np.median(walls))
        out[name] = {"wall_ms_median": med * 1e3, "gate_amp_per_s_wall": gates * float(1 << n) / med,
                     "evolve_device_ms": dev_ms, "launches": launches,
                     "evolve_hbm_gbs": nbytes / (dev_ms / 1e3) / 1e9, "distinct_outcomes": len(counts)}
        log(f"[brickwork] {name}: {out[name]}")
    return out


def batch_loop_report(count: int = 4096, shots: int = 1024, base_seed: int = 2024, log=print) -> dict:
    circs = [workloads.batch_circuit(i, base_seed=base_seed) for i in range(count)]
    work = sum(workloads.gate_count(c.data["instructions"]) * float(1 << c.num_qubits) for c in circs)
    lat = np.zeros(count)
    _core.StatevectorSimulator(seed=0, precision="complex64", fusion="greedy",
                               max_fused_qubits=3).simulate_shots(circs[0], shots=shots)  # warm-up
    t0 = time.perf_counter()
    for i, c in enumerate(circs):
        t = time.perf_counter()
        sim = _core.StatevectorSimulator(seed=base_seed + i, precision="complex64", fusion="greedy",
                                         max_fused_qubits=3)
        counts = sim.simulate_shots(c, shots=shots)
        lat[i] = time.perf_counter() - t
        if sum(counts.values()) != shots:
            raise RuntimeError(f"circuit {i}: {sum(counts.values())} shots returned")
        if i % 512 == 0:
            log(f"[batch-loop] circuit {i}/{count}")
    wall = time.perf_counter() - t0
    out = {"circuits": count, "shots_each": shots, "wall_s": wall, "circuits_per_s": count / wall,
           "gate_amp_per_s": work / wall, "latency_ms_p50": float(np.percentile(lat, 50) * 1e3),
           "latency_ms_p99": float(np.percentile(lat, 99) * 1e3), "gate_amp_total": work,
           "path": "one simulate_shots call per circuit (complex64, greedy <=3-qubit fusion, tile sweeps)"}
    log(f"[batch-loop] {out}")
    return out


# ---------------------------------------------------------------------------
# P-block backend: the reference's benchmarks/pblock_benchmark.py cases
# ---------------------------------------------------------------------------
def pblock_cases():
    """(name, simulator options, circuit or distributed input, shots): the six
    cases of the reference's P-block benchmark (max_memory_mb=64, seed 71)
    and one 24-qubit entangled block (GHZ chain then a U3 layer) sampled 1000
    times, where each gate updates 2^24 amplitudes."""
    from types import SimpleNamespace

    def pairs(n):
        return [op for q in range(0, n, 2) for op in ({"kind": "h", "qubit": q},
                                                      {"kind": "cx", "control": q, "target": q + 1})]

    def meas(n):
        return [{"kind": "measure", "qubit": q, "cbit": q} for q in range(n)]

    def distributed(n, ops, singleton):
        groups = {q: [q] for q in range(n)} if singleton else {0: list(range(n))}
        buckets = {node: [] for node in groups}
        for op in ops:
            q = op.get("qubit", op.get("control", 0))
            buckets[q if singleton else 0].append(op)
        circuits = {node: workloads.Circuit(n, b, n) for node, b in buckets.items()}
        for c in circuits.values():
            c.instructions = c.data["instructions"]
        return SimpleNamespace(circuits=circuits, qubits_per_node=groups,
                               _instruction_index={id(op): i for i, op in enumerate(ops)})

    product = [{"kind": "h", "qubit": q} for q in range(16)] + meas(16)
    dynamic = pairs(20) + [{"kind": "measure", "qubit": 0, "cbit": 0},
                           {"kind": "conditional", "condition": {"creg_base": 0, "creg_size": 1, "creg_value": 1},
                            "op": {"kind": "x", "qubit": 1}}, {"kind": "reset", "qubit": 2}] + meas(20)
    rng = np.random.default_rng(71)
    ghz = [{"kind": "h", "qubit": 0}] + [{"kind": "cx", "control": q, "target": q + 1} for q in range(23)]
    ghz += [{"kind": "u3", "qubit": q, "theta": float(rng.uniform(0, np.pi)), "phi": float(rng.uniform(0, 6.28)),
             "lam": float(rng.uniform(0, 6.28))} for q in range(24)] + meas(24)
    base = {"seed": 71, "max_memory_mb": 64}
    return [
        ("product_16q_one_node_16shots", base, distributed(16, product, False), 16),
        ("product_16q_sixteen_nodes_16shots", base, distributed(16, product, True), 16),
        ("bell_pairs_20q_terminal_1000shots", base, workloads.Circuit(20, pairs(20) + meas(20), 20), 1000),
        ("bell_pairs_20q_forced_1000shots", dict(base, sample_terminal=False),
         workloads.Circuit(20, pairs(20) + meas(20), 20), 1000),
        ("bell_pairs_100q_1000shots", base, workloads.Circuit(100, pairs(100) + meas(100), 100), 1000),
        ("dynamic_20q_100shots", base, workloads.Circuit(20, dynamic, 20), 100),
        ("ghz_u3_24q_block_1000shots", {"seed": 71, "max_memory_mb": 2048}, workloads.Circuit(24, ghz, 24), 1000),
    ]


def pblock_report(repeats: int = 5, log=print) -> dict:
    """Median wall seconds of PBlockSimulator.simulate_shots per case (one
    warm-up call first), as the reference's benchmark reports them."""
    out = {"repeats": repeats, "seconds": {}}
    for name, opts, circ, shots in pblock_cases():
        sim = _core.PBlockSimulator(**opts)
        sim.simulate_shots(circ, shots)
        times = []
        for _ in range(repeats):
            t0 = time.perf_counter()
            counts = sim.simulate_shots(circ, shots)
            times.append(time.perf_counter() - t0)
            if sum(counts.values()) != shots:
                raise RuntimeError(f"{name}: {sum(counts.values())} shots returned")
        out["seconds"][name] = {"median": float(np.median(times)), "samples": times}
        log(f"[pblock] {name}: median {np.median(times) * 1e3:.2f} ms")
    return out
