"""Benchmark: gate-amplitude updates/s and HBM GB/s of the dense statevector
engine on one MI355X.

Workload (one step): evolve a 32-qubit, depth-40 random circuit (Haar SU(2) on
every qubit plus CZ/CX on a random perfect matching per layer, 1,920 gates,
seed 1234) from |0...0> in complex64 with fused blocks of <= 5 qubits grouped
into LDS tile sweeps (12-qubit tiles, out of place), the block width and tile
rule chosen by the learned planner (shipped weights; --fusion greedy fuses as
far as the width allows); 3-qubit blocks run in split precision (FP32
operands as two FP16 parts on the FP16 matrix core, FP32 accumulation;
--block-arith 0: FP32 matrix cores). The same circuit also runs under the
greedy planner at that width and at 3 qubits, with FP32 matrix-core blocks,
and once on the plain complex128 path (one launch per compiled gate,
reference arithmetic) for the speed-up and the accuracy check.

Prints one JSON line. `value` = G * 2^n / seconds per step.
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "classical-quantum-sim-optimizer_amd"))
sys.path.insert(0, ROOT)

KIND_NAMES = {0: "one", 1: "x", 2: "swap", 3: "diag", 4: "dense", 5: "fused_valu", 6: "measure", 7: "reset",
              8: "fused_mfma", 9: "tile", 10: "tile128", 11: "xsweep", 12: "xsweep128"}
KERNEL_NAMES = {"one": "k_one", "x": "k_pairswap", "swap": "k_pairswap", "diag": "k_diag", "dense": "k_dense",
                "fused_valu": "k_fused", "fused_mfma": "k_fused_mfma", "measure": "k_collapse", "reset": "k_collapse",
                "tile": "k_tile", "tile128": "k_tile128", "xsweep": "k_xsweep", "xsweep128": "k_xsweep128"}
HBM_PEAK_GBS = 8000.0
MFMA_FP32_PEAK_TFLOPS = 157.3
MFMA_FP16_PEAK_TFLOPS = 2500.0  # dense FP16 matrix-core peak (split-precision blocks run there)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def per_kind(rows):
    agg = {}
    for r in rows:
        name = KIND_NAMES.get(r["kind"], str(r["kind"]))
        a = agg.setdefault(name, {"launches": 0, "ms": 0.0, "bytes": 0.0, "gates": 0})
        a["launches"] += 1
        a["ms"] += r["ms"]
        a["bytes"] += r["bytes"]
        a["gates"] += r["gates"]
    return agg


def profiled_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload` from the committed
    rocprofv3 PMC summaries (FETCH_SIZE and WRITE_SIZE, each scaled by the
    same counter's reading for a plain copy of known size); None if that
    kernel/workload pair has not been profiled."""
    import glob

    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        for line in open(path, encoding="utf-8"):
            if not line.startswith("traffic-per-launch:"):
                continue
            kv = dict(f.split("=", 1) for f in line.split(":", 1)[1].split())
            if kv.get("kernel") == kernel and kv.get("workload") == workload:
                return float(kv["bytes"])
    return None


def profiled_mfma_busy(kernel, workload):
    """Fraction of SIMD cycles with the matrix core busy for `kernel` on
    `workload` (SQ_VALU_MFMA_BUSY_CYCLES over GRBM_GUI_ACTIVE x SIMDs) from
    the committed rocprofv3 PMC summaries; None if not profiled."""
    import glob

    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        for line in open(path, encoding="utf-8"):
            if not line.startswith("mfma-busy:"):
                continue
            kv = dict(f.split("=", 1) for f in line.split(":", 1)[1].split())
            if kv.get("kernel") == kernel and kv.get("workload") == workload:
                return float(kv["frac"])
    return None


def mfma_flops_per_step(circ, precision, fusion, max_fused, tile, out_of_place):
    """Matrix-core flops of one step from the launch plan, as (tile sweeps,
    of which split precision, one-block launches): a k-qubit block pass is
    8 * 2^k * 2^n algorithmic flops (one complex MAC = 8 real flops); 1-qubit
    blocks run padded to 2 qubits in tile sweeps. A split-precision 3-qubit
    block issues 4x its algorithmic flops on the FP16 matrix core (two
    16x16x32 f16 MFMAs of K = 32 where the FP32 path has K = 16 in four
    16x16x4 f32 ones). A sweep's flops count the tiles it computes (under a
    zero start the leading sweeps store all-zero tiles without running their
    blocks: the plan's work_fraction). Also returns the plan's learned
    choices."""
    from dqsim import _core

    plan = _core.plan_json(circ, precision=precision, fusion=fusion, max_fused_qubits=max_fused, tile_qubits=tile,
                           out_of_place=out_of_place)
    n = plan["num_qubits"]
    tile_flops = split_flops = fused = 0.0
    for op in plan["launches"]:
        if op["kind"] == "tile" and not op.get("exact"):
            for b in op["blocks"]:
                f = 8.0 * (1 << max(2, len(b["qubits"]))) * float(1 << n) * op.get("work_fraction", 1.0)
                tile_flops += f
                if b.get("split"):
                    split_flops += f
        elif op["kind"] == "fused_mfma":
            fused += 8.0 * (1 << len(op["qubits"])) * float(1 << n)
    return tile_flops, split_flops, fused, plan["learned_choices"]


def device_copy_gbs(nbytes=4 << 30, reps=5):
    """Read+write rate of a plain device-to-device copy (the practical
    streaming ceiling beside the 8 TB/s peak)."""
    import torch

    src = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        dst.copy_(src)
    b.record()
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / reps
    del src, dst
    torch.cuda.empty_cache()
    return 2.0 * nbytes / (ms / 1e3) / 1e9


def host_cpu_info():
    """CPU model, logical CPUs (nproc), physical cores and the CPUs this
    process may run on."""
    model, cores = "unknown", set()
    phys, core = None, None
    try:
        with open("/proc/cpuinfo", encoding="utf-8") as f:
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name" and model == "unknown":
                    model = v
                elif k == "physical id":
                    phys = v
                elif k == "core id":
                    core = v
                elif not k and phys is not None and core is not None:
                    cores.add((phys, core))
                    phys, core = None, None
    except OSError:
        pass
    return {"model": model, "nproc": os.cpu_count() or 1, "physical_cores": len(cores) or None,
            "affinity_cpus": len(os.sched_getaffinity(0))}


def cpu_baseline(n, depth, repeats=5):
    """FP64 CPU restatement of the reference algorithm (the oracle) on the same
    circuit family: one warm-up, then the median of `repeats` runs, at every
    CPU this process's share allows (at most 16: the host share of one GPU)
    and at 4 threads (the reference's documented RAYON_NUM_THREADS), same n."""
    from dqsim import workloads
    from oracle import oracle

    info = host_cpu_info()
    circ = workloads.random_su2(n=n, depth=depth, seed=1234)
    g = workloads.gate_count(circ.data["instructions"])
    legs = {}
    for threads in sorted({4, min(16, info["affinity_cpus"])}, reverse=True):
        oracle.set_threads(threads)
        oracle.simulate(circ.data, 1)  # warm-up
        walls = []
        for _ in range(repeats):
            t0 = time.perf_counter()
            oracle.simulate(circ.data, 1)
            walls.append(time.perf_counter() - t0)
        walls.sort()
        med = walls[len(walls) // 2]
        legs[threads] = {"value": g * float(1 << n) / med, "median_s": med, "runs_s": walls}
        log(f"[bench] cpu oracle {threads} threads: {legs[threads]['value']:.3e} gate·amp/s (median {med:.2f} s)")
    top = max(legs)
    return {"value": legs[top]["value"], "unit": "gate_amp_updates/s", "cores": top, "kind": "port",
            "sample": f"random_su2 n={n} depth={depth} seed=1234 ({g} gates), FP64 oracle, median of {repeats} "
                      f"runs after a warm-up: {legs[top]['median_s']:.2f} s at {top} threads, "
                      f"{legs[4]['median_s']:.2f} s at 4 threads",
            "threads_4": legs[4]["value"], "host": info}


def brickwork_cpu_baseline(shots=1000, seed=42, repeats=5):
    """The oracle (FP64 C restatement of the reference algorithm) on the full
    20-qubit brickwork with 1000 shots, 4 threads and all threads: median of 5
    runs after a warm-up."""
    from dqsim import workloads
    from oracle import oracle

    circ = workloads.brickwork(n=20, layers=20, seed=seed)
    gates = workloads.gate_count(circ.data["instructions"])
    out = {}
    for threads in sorted({4, min(16, host_cpu_info()["affinity_cpus"])}):
        oracle.set_threads(threads)
        oracle.run_shots(circ.data, shots, seed)
        walls = []
        for _ in range(repeats):
            t0 = time.perf_counter()
            oracle.run_shots(circ.data, shots, seed)
            walls.append(time.perf_counter() - t0)
        walls.sort()
        med = walls[len(walls) // 2]
        out[f"threads_{threads}"] = {"wall_ms_median": med * 1e3, "gate_amp_per_s": gates * float(1 << 20) / med}
        log(f"[brickwork] cpu oracle, {threads} threads: {out[f'threads_{threads}']}")
    out["kind"] = "port"
    return out


def pblock_cpu_baseline():
    """The CPU P-block restatement (oracle/pblock.py: Python bookkeeping over
    the C FP64 kernels, OpenMP inside the kernels) on the same cases, one
    timed run each after a warm-up of the smallest case."""
    import os

    from dqsim import reports
    from oracle import oracle, pblock

    threads = min(16, os.cpu_count() or 1)
    oracle.set_threads(threads)
    out = {"unit": "s per case", "cores": threads, "kind": "port", "seconds": {}}
    cases = reports.pblock_cases()
    for name, opts, circ, shots in cases:
        data = pblock_transport_data(circ)
        kw = {"memory_mb": opts.get("max_memory_mb", 1024), "sample_terminal": opts.get("sample_terminal", True)}
        t0 = time.perf_counter()
        counts, strategy = pblock.simulate_shots(data, shots, opts["seed"], **kw)
        out["seconds"][name] = time.perf_counter() - t0
        log(f"[pblock-cpu] {name}: {out['seconds'][name] * 1e3:.1f} ms ({strategy})")
        if sum(counts.values()) != shots:
            raise RuntimeError(name)
    out["sample"] = "every case of the P-block report once"
    return out


def pblock_transport_data(circ) -> dict:
    """Plain circuit JSON of a report case (distributed inputs merged by the
    engine's transport, whose qubit labels are already 0..n-1 here)."""
    from dqsim import _pblock

    data = json.loads(_pblock._transport(circ, 1 << 40))
    n = sum(r["size"] for r in data["qregs"].values())
    data["qregs"] = {"q": {"name": "q", "base": 0, "size": n}}
    return data


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--qubits", type=int, default=32)
    ap.add_argument("--depth", type=int, default=40)
    ap.add_argument("--max-fused", type=int, default=5,
                    help="widest fused block (tile sweeps take up to 4 qubits)")
    ap.add_argument("--fusion", default="learned", choices=["none", "greedy", "learned"],
                    help="fusion planner (learned: the shipped weights choose block width and tile rule)")
    ap.add_argument("--no-compare", action="store_true",
                    help="skip the timing of the other fusion planners on the same circuit")
    ap.add_argument("--precision", default="complex64", choices=["complex64", "complex128"],
                    help="precision of the timed evolution (complex128 with --fusion none times the exact sweeps)")
    ap.add_argument("--tile", type=int, default=12, help="LDS tile qubits of fused sweeps (0 = one launch per block)")
    ap.add_argument("--pinned", action="store_true",
                    help="out-of-place sweeps keep qubits 0..3 in every tile (default: only in a run's first and "
                         "last tile)")
    ap.add_argument("--relabel", action="store_true",
                    help="relabelled tile sweeps (measured slower here; default: qubits 0..3 pinned in every tile)")
    ap.add_argument("--sweep-mode", type=int, default=None, choices=[0, 1, 2, 3, 4],
                    help="complex64 fused sweeps: 0 in place; out of place with 1 persistent 256-thread, "
                         "2 one-shot 256-thread, 3 one-shot 512-thread, 4 persistent 512-thread workgroups "
                         "(default: the library's)")
    ap.add_argument("--block-arith", type=int, default=None, choices=[0, 1],
                    help="3-qubit blocks of complex64 sweeps: 0 FP32 matrix cores, 1 split precision (FP16 parts, "
                         "FP32 accumulation) (default: the library's)")
    ap.add_argument("--wave-runs", type=int, default=None, choices=[0, 1],
                    help="256-thread out-of-place sweeps: 1 wave-owned block runs with a barrier after each run, "
                         "0 a barrier after every block (default: the library's)")
    ap.add_argument("--no-zero-start", action="store_true",
                    help="initialise the state and read it in the first sweeps (default: plans that open with "
                         "out-of-place sweeps start from a synthesised |0...0> and skip reading all-zero tiles)")
    ap.add_argument("--no-plain", action="store_true", help="skip the complex128 plain comparison run")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-qubits", type=int, default=24)
    ap.add_argument("--cpu-depth", type=int, default=10)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device copy of this many MiB (HBM counter calibration)")
    ap.add_argument("--report", choices=["precision", "batch", "planner", "brickwork", "batch-loop", "pblock"],
                    default=None,
                    help="run one secondary report instead of the headline benchmark")
    args = ap.parse_args()
    if args.report:
        from dqsim import reports

        fn = {"precision": reports.precision_report, "batch": reports.batch_report,
              "planner": reports.planner_report, "brickwork": reports.brickwork_report,
              "batch-loop": reports.batch_loop_report, "pblock": reports.pblock_report}[args.report]
        res = fn(log=log)
        if args.report == "brickwork":
            res["cpu_baseline"] = brickwork_cpu_baseline()
        if args.report == "pblock":
            res["cpu_baseline"] = pblock_cpu_baseline()
        print(json.dumps({"report": args.report, **res}), flush=True)
        return
    if args.copy_calibration:
        import torch

        src = torch.empty(args.copy_calibration << 20, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_mib": args.copy_calibration, "launches": 3}))
        return
    if args.gpus != 1:
        print(f"bench.py: this engine runs one state per GPU and this benchmark uses exactly one GPU "
              f"(--gpus {args.gpus} requested); multi-GPU sharding is not built.", file=sys.stderr)
        sys.exit(2)

    import dqsim
    from dqsim import workloads

    dqsim._core.set_sweep_relabel(args.relabel)
    dqsim._core.set_sweep_pinned(args.pinned)
    pinned = args.pinned
    if args.sweep_mode is not None:
        dqsim._core.set_sweep_mode(args.sweep_mode)
    if args.block_arith is not None:
        dqsim._core.set_block_arith(args.block_arith)
    block_arith = dqsim._core.set_block_arith(0)
    dqsim._core.set_block_arith(block_arith)
    if args.wave_runs is not None:
        dqsim._core.set_wave_runs(args.wave_runs)
    wave_runs = dqsim._core.set_wave_runs(1)
    dqsim._core.set_wave_runs(wave_runs)
    if args.no_zero_start:
        dqsim._core.set_zero_start(False)
    sweep_mode = dqsim._core.set_sweep_mode(1)
    dqsim._core.set_sweep_mode(sweep_mode)
    copy_gbs = device_copy_gbs()
    log(f"[bench] device copy (4 GiB, torch copy_): {copy_gbs:.0f} GB/s read+write")
    mfma_peak_tflops, _ = dqsim._core.mfma_fp32_peak(3)
    log(f"[bench] FP32 matrix-core microbenchmark: {mfma_peak_tflops:.1f} TFLOP/s")
    n, depth = args.qubits, args.depth
    circ = workloads.random_su2(n=n, depth=depth, seed=1234)
    gates = workloads.gate_count(circ.data["instructions"])
    mem_mb = (1 << n) * 16 * 2 // (1 << 20) + 4096
    fused = dqsim.StatevectorSimulator(seed=1, precision=args.precision, fusion=args.fusion,
                                       max_fused_qubits=args.max_fused, max_memory_mb=mem_mb, tile_qubits=args.tile)
    log(f"[bench] n={n} depth={depth} gates={gates} {args.precision} fusion={args.fusion} k<={args.max_fused} "
        f"tile={args.tile} relabel={args.relabel}")
    for w in range(args.warmup):
        t = time.perf_counter()
        fused.evolve_timed(circ, repeats=1, record=False)
        log(f"[bench] warmup {w}: {time.perf_counter() - t:.3f} s")
    # The timed steps run as a user's call does (one event pair around the
    # whole evolution); one more step afterwards records an event around
    # every launch for the per-kernel figures.
    dev_ms = []
    t0 = time.perf_counter()
    for s in range(args.steps):
        _, launches, g, total_ms, _ = fused.evolve_timed(circ, repeats=1, record=False)
        dev_ms.append(total_ms)
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    rows_all, launches, g, rec_ms, _ = fused.evolve_timed(circ, repeats=1, record=True)
    log(f"[bench] fused: {ms_per_step:.1f} ms/step wall, device {sum(dev_ms) / len(dev_ms):.1f} ms, "
        f"{launches} launches (recorded step: device {rec_ms:.1f} ms)")
    agg = per_kind(rows_all)
    dominant = max(agg.items(), key=lambda kv: kv[1]["ms"])
    dname, d = dominant
    total_bytes = sum(a["bytes"] for a in agg.values())
    value = gates * float(1 << n) / (ms_per_step / 1e3)

    tile_flops, split_flops, fused_flops, choices = mfma_flops_per_step(circ, args.precision, args.fusion, args.max_fused,
                                                           args.tile, sweep_mode != 0)
    mflops = tile_flops + fused_flops
    mfma_ms = sum(a["ms"] for k, a in agg.items() if k in ("tile", "xsweep", "fused_mfma"))
    mfma = None
    if mflops > 0 and mfma_ms > 0:
        ach = mflops / (mfma_ms / 1e3) / 1e12
        mfma = {"achieved_tflops": ach, "datasheet_peak_tflops": MFMA_FP32_PEAK_TFLOPS,
                "measured_peak_tflops": mfma_peak_tflops, "frac_of_datasheet": ach / MFMA_FP32_PEAK_TFLOPS,
                "frac_of_measured_peak": ach / mfma_peak_tflops, "flops_per_step": mflops,
                "split_precision_flops_per_step": split_flops, "busy_frac_profiled": None}
    # The same workload under the other fusion planners (greedy and learned
    # at this block width, and greedy at 3 qubits): all step times go into
    # the line.
    planner_cmp = None
    if args.fusion in ("greedy", "learned") and args.precision == "complex64" and not args.no_compare:
        per = {f"{args.fusion}_k{args.max_fused}": ms_per_step}
        for fusion, k in (("greedy", args.max_fused), ("learned", args.max_fused), ("greedy", 3)):
            key = f"{fusion}_k{k}"
            if key in per:
                continue
            alt = dqsim.StatevectorSimulator(seed=1, precision=args.precision, fusion=fusion, max_fused_qubits=k,
                                             max_memory_mb=mem_mb, tile_qubits=args.tile)
            for _ in range(max(1, args.warmup)):
                alt.evolve_timed(circ, repeats=1, record=False)
            ta = time.perf_counter()
            for _ in range(args.steps):
                alt.evolve_timed(circ, repeats=1, record=False)
            per[key] = (time.perf_counter() - ta) * 1e3 / args.steps
            del alt
        if block_arith == 1:
            # The same plan with 3-qubit blocks on the FP32 matrix cores.
            dqsim._core.set_block_arith(0)
            alt = dqsim.StatevectorSimulator(seed=1, precision=args.precision, fusion=args.fusion,
                                             max_fused_qubits=args.max_fused, max_memory_mb=mem_mb,
                                             tile_qubits=args.tile)
            for _ in range(max(1, args.warmup)):
                alt.evolve_timed(circ, repeats=1, record=False)
            ta = time.perf_counter()
            for _ in range(args.steps):
                alt.evolve_timed(circ, repeats=1, record=False)
            per[f"{args.fusion}_k{args.max_fused}_fp32_blocks"] = (time.perf_counter() - ta) * 1e3 / args.steps
            del alt
            dqsim._core.set_block_arith(1)
        lk, gk = f"learned_k{args.max_fused}", f"greedy_k{args.max_fused}"
        choice = dqsim._core.plan_json(circ, precision=args.precision, fusion="learned", max_fused_qubits=args.max_fused,
                                       tile_qubits=args.tile, out_of_place=sweep_mode != 0)["learned_choices"]
        planner_cmp = {"ms_per_step": per, "learned_speedup_vs_greedy_same_width": per[gk] / per[lk],
                       "learned_speedup_vs_greedy_k3": per["greedy_k3"] / per[lk], "learned_choices": choice,
                       "weights": "shipped planner_weights.npy (engine-recorded timings, trained on the GPU)"}
        fk = f"{args.fusion}_k{args.max_fused}_fp32_blocks"
        if fk in per:
            planner_cmp["split_blocks_speedup_vs_fp32_blocks"] = per[fk] / per[f"{args.fusion}_k{args.max_fused}"]
        log(f"[bench] planner: {per}, learned choice {choice}")
    extra = {"launches_per_step": launches, "kernels": {k: {kk: (round(vv, 4) if isinstance(vv, float) else vv)
                                                            for kk, vv in v.items()} for k, v in agg.items()},
             "device_ms_per_step": sum(dev_ms) / len(dev_ms),
             "hbm_gbs_whole_step": total_bytes / (sum(dev_ms) / len(dev_ms) / 1e3) / 1e9,
             "device_copy_gbs": copy_gbs, "planner": planner_cmp}

    if not args.no_plain:
        plain = dqsim.StatevectorSimulator(seed=1, precision="complex128", fusion="none", max_memory_mb=mem_mb,
                                           tile_qubits=0)
        t = time.perf_counter()
        prow, plaunch, _, pms, pres = plain.evolve_timed(circ, repeats=1, record=True, keep=True)
        pwall = time.perf_counter() - t
        pagg = per_kind(prow)
        pbytes = sum(a["bytes"] for a in pagg.values())
        log(f"[bench] plain complex128: {pwall:.2f} s wall, device {pms:.1f} ms, {plaunch} launches")
        _, _, _, _, fres = fused.evolve_timed(circ, repeats=1, record=False, keep=True)
        max_err, l2_err = fres.compare(pres)
        del fres, pres
        extra["plain_complex128"] = {
            "value": gates * float(1 << n) / (pms / 1e3), "ms": pms, "wall_s": pwall, "launches": plaunch,
            "hbm_gbs": pbytes / (pms / 1e3) / 1e9,
            "kernels": {k: {kk: (round(vv, 4) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                        for k, v in pagg.items()}}
        extra["speedup_vs_plain_complex128"] = pms / (sum(dev_ms) / len(dev_ms))
        extra["max_amp_err_vs_fp64"] = max_err
        extra["l2_err_vs_fp64"] = l2_err
        # The default complex128 path: the same unfused ops in exact LDS sweeps
        # (bit-identical to one launch per op).
        exact = dqsim.StatevectorSimulator(seed=1, precision="complex128", fusion="none", max_memory_mb=mem_mb)
        t = time.perf_counter()
        erow, elaunch, _, ems, eres = exact.evolve_timed(circ, repeats=1, record=True, keep=True)
        ewall = time.perf_counter() - t
        log(f"[bench] exact complex128 sweeps: {ewall:.2f} s wall, device {ems:.1f} ms, {elaunch} launches")
        _, _, _, _, pres = plain.evolve_timed(circ, repeats=1, record=False, keep=True)
        emax, _ = eres.compare(pres)
        del eres, pres
        extra["exact_complex128_sweeps"] = {
            "value": gates * float(1 << n) / (ems / 1e3), "ms": ems, "wall_s": ewall, "launches": elaunch,
            "speedup_vs_plain_complex128": pms / ems, "max_abs_diff_vs_plain": emax}

    # The plan that ran: tile rule and block width (the learned planner's
    # choice under --fusion learned).
    rule, kk = ("pinned" if pinned else "free"), args.max_fused
    if choices:
        rule, kk = ("pinned" if choices[0]["pinned"] else "free"), choices[0]["k"]
    if args.tile == 0:
        rule = "blocks"
    elif sweep_mode == 0:
        rule = "inplace"
    workload_key = (f"random_su2_n{n}_d{depth}_k{kk}_t{args.tile}_{rule}" + ("_split" if split_flops > 0 else "")
                    if args.precision == "complex64" and args.fusion in ("greedy", "learned")
                    else f"random_su2_n{n}_d{depth}_{args.precision}_{args.fusion}_t{args.tile}")
    if mfma is not None:
        mfma["busy_frac_profiled"] = profiled_mfma_busy(KERNEL_NAMES.get(dname, dname), workload_key)
        extra["mfma"] = mfma
    # Roofline of the dominant kernel per launch: algorithmic HBM bytes and
    # (for tile sweeps) algorithmic FP32 matrix-core flops over its measured
    # average launch time; the bound is the resource whose floor (bytes at
    # 8 TB/s, flops at 157.3 TFLOP/s) is higher.
    launch_ms = d["ms"] / d["launches"]
    bytes_pl = d["bytes"] / d["launches"]
    hbm = {"achieved": bytes_pl / (launch_ms / 1e3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "floor_ms": bytes_pl / (HBM_PEAK_GBS * 1e9) * 1e3}
    hbm["frac"] = hbm["achieved"] / HBM_PEAK_GBS
    hbm["frac_of_device_copy"] = hbm["achieved"] / copy_gbs
    kname = KERNEL_NAMES.get(dname, dname)
    roofline = {"bound": "hbm", "kernel": kname, "achieved": hbm["achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": hbm["frac"], "traffic": profiled_traffic(kname, workload_key),
                "algorithmic_bytes_per_launch": bytes_pl, "launch_ms": launch_ms, "hbm": hbm}
    if dname in ("xsweep", "tile") and tile_flops > 0:
        # Algorithmic flops per launch; the floor prices FP32 blocks at the
        # FP32 matrix-core peak and split-precision blocks' issued FP16 flops
        # (4x algorithmic) at the FP16 peak.
        flops_pl = tile_flops / d["launches"]
        split_pl = split_flops / d["launches"]
        floor_ms = ((flops_pl - split_pl) / (MFMA_FP32_PEAK_TFLOPS * 1e12)
                    + 4.0 * split_pl / (MFMA_FP16_PEAK_TFLOPS * 1e12)) * 1e3
        mf = {"achieved": flops_pl / (launch_ms / 1e3) / 1e12, "peak": MFMA_FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
              "floor_ms": floor_ms, "flops_per_launch": flops_pl, "split_precision_flops_per_launch": split_pl,
              "intensity_flop_per_byte": flops_pl / bytes_pl}
        mf["frac"] = mf["achieved"] / MFMA_FP32_PEAK_TFLOPS
        roofline["mfma"] = mf
        if mf["floor_ms"] > hbm["floor_ms"]:
            roofline.update({"bound": "mfma", "achieved": mf["achieved"], "peak": MFMA_FP32_PEAK_TFLOPS,
                             "unit": "TFLOP/s", "frac": mf["frac"]})
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args.cpu_qubits, args.cpu_depth)
        log(f"[bench] cpu baseline: {cpu['value']:.3e} gate·amp/s ({cpu['sample']})")

    line = {
        "metric": "gate·amp updates/s + HBM GB/s at 32q complex64-fused vs complex128-plain",
        "value": value,
        "unit": "gate_amp_updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64" if args.precision == "complex64" else "c128",
        "data": "synthetic",
        "config": {"workload": f"random_su2 n={n} depth={depth} seed=1234, {args.precision}, {args.fusion} fusion "
                               f"k<={args.max_fused}, LDS tile {args.tile} qubits", "qubits": n, "depth": depth, "gates": gates,
                   "relabelled_sweeps": args.relabel, "sweep_mode": sweep_mode, "block_arith": block_arith,
                   "wave_runs": wave_runs, "zero_start": not args.no_zero_start,
                   "tile_rule": rule, "block_width": kk, "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "detail": extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
