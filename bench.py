#!/usr/bin/env python3
"""Benchmark of the depth hot path on one MI355X.

Workload (default): 5 modules of 4160x3120 at f = 3300 px, the reference at
the centre of a 30 mm ring of four, 12-bit RGGB Bayer captures of a seeded
16-plane scene converted by luminance_half, 128 inverse-depth hypotheses
(near 800 mm, far 20000 mm), 8-path SGM, default options otherwise.

A step is one cost-field + SGM + winner-take-all pass over the capture with
every input already resident in HBM (the metric is "Mpix*hyp/s cost+SGM").
value = reference raster pixels x planes / step time / 1e6. The full path
(direct pass, filters and per-target decisions added) is timed as well and
reported beside it.

With --learned-scenes > 0 (the default) the line also carries
"learned_band": the learned band proposal trained on GPU-rendered half-scale
config 4 captures, its setting chosen on full config 4 captures and reported
on other, held-out full captures batched, for cost + SGM and end to end
(--workload config3 measures it on config 3).

Prints one JSON line. --gpus above 1 is rejected: one GPU is measured (the
path shards by capture, whole captures per GPU).
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = ("Mpix·hyp/s cost+SGM on 1 GPU; learned-band speedup at ≥98% "
          "±1-hyp agreement")
PEAK_FP32_TFLOPS = 157.3
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0
# Algorithmic work per (node, label, valid view) ZNCC of the cost volume:
# 81 taps x (3 bilinear samples' worth of f32 arithmetic + 2 weights) in
# f32 and the tap coordinates in f64 (the per-unit figures of the
# measurement plan: 1,620 FP32 + 540 FP64 flops).
FP32_PER_ZNCC = 1620
FP64_PER_ZNCC = 540
# HBM bytes per (node, label) of cost + SGM + WTA: 8 B cost-volume write,
# 8 paths x 12 B, 4 B WTA read.
BYTES_PER_UNIT = 108


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def host_cpu() -> dict:
    """The host's processor: model string, family/model/stepping numbers,
    logical CPUs, physical cores, the affinity mask and the cgroup CPU
    quota, and std::thread::available_parallelism's answer (the smaller of
    the affinity mask and the quota), which sizes the reference's row bands
    (depth.rs:729-731, 1204-1206)."""
    info = {"model": cpu_model(), "logical_cpus": os.cpu_count()}
    ids, cores, fields = set(), set(), {}
    phys = core = None
    try:
        for line in open("/proc/cpuinfo"):
            key, _, val = line.partition(":")
            key, val = key.strip(), val.strip()
            if key in ("cpu family", "model", "stepping", "vendor_id") and key not in fields:
                fields[key] = val
            if key == "physical id":
                phys = val
                ids.add(val)
            elif key == "core id":
                core = val
            elif not key and phys is not None and core is not None:
                cores.add((phys, core))
                phys = core = None
        if phys is not None and core is not None:
            cores.add((phys, core))
    except OSError:
        pass
    names = {"vendor_id": "vendor", "cpu family": "family", "model": "model_number",
             "stepping": "stepping"}
    info.update({names[k]: v for k, v in fields.items()})
    info["sockets"] = len(ids) or None
    info["physical_cores"] = len(cores) or None
    affinity = len(os.sched_getaffinity(0))
    info["affinity_cpus"] = affinity
    quota = None
    try:
        q, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            quota = -(-int(q) // int(period))
    except (OSError, ValueError):
        pass
    info["cgroup_cpu_quota"] = quota
    info["available_parallelism"] = min(affinity, quota) if quota else affinity
    return info


def cpu_threads(host: dict) -> int:
    """Threads of the CPU baseline: available_parallelism, as the reference
    sizes its row bands, within the host's CPU share for this GPU when the
    environment states one (OMP_NUM_THREADS on a shared multi-GPU host)."""
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    n = host["available_parallelism"]
    return max(1, min(n, share) if share > 0 else n)


def cpu_baseline(cap, rows: int, threads: int) -> dict:
    """The oracle's cost volume over a band of coarse rows with the
    reference's row-band threading, plus its single-threaded SGM over the
    same band, scaled to the whole capture."""
    import ctypes as C

    import numpy as np

    from chiaro_amd.depth import _ViewArrays
    from oracle import binding

    lib = binding.lib()
    lib.or_set_libm_mode(1)  # the platform libm, as the reference calls it
    ref = cap.modules[cap.reference_index]
    g = cap.options.grid(ref.width, ref.height)
    rows = min(rows, g.coarse_rows)
    first = (g.coarse_rows - rows) // 2
    nodes = rows * g.coarse_columns
    scores = np.empty((nodes, g.labels), np.float32)
    costs = np.empty((nodes, g.labels), np.float32)
    guidance = np.empty(nodes, np.float32)
    tested = np.empty(nodes, np.uint8)
    views = _ViewArrays(cap.modules)
    t0 = time.perf_counter()
    lib.or_cost_volume_rows(views.views, len(cap.modules), cap.reference_index,
                            C.byref(cap.options.to_c()), first, first + rows, threads,
                            scores.ctypes.data, costs.ctypes.data, guidance.ctypes.data,
                            tested.ctypes.data)
    t1 = time.perf_counter()
    out = np.empty_like(costs)
    lib.or_semi_global_costs(costs.ctypes.data, guidance.ctypes.data, g.coarse_columns, rows,
                             g.labels, out.ctypes.data)
    t2 = time.perf_counter()
    scale = g.coarse_rows / rows
    total = (t2 - t0) * scale
    pixels = ref.width * ref.height
    # the same cost volume on one thread over a few of those rows
    rows1 = min(8, rows)
    first1 = (g.coarse_rows - rows1) // 2
    t3 = time.perf_counter()
    lib.or_cost_volume_rows(views.views, len(cap.modules), cap.reference_index,
                            C.byref(cap.options.to_c()), first1, first1 + rows1, 1,
                            scores.ctypes.data, costs.ctypes.data, guidance.ctypes.data,
                            tested.ctypes.data)
    t4 = time.perf_counter()
    one_thread_cost = (t4 - t3) * g.coarse_rows / rows1
    return {
        "value": pixels * cap.options.planes / total / 1e6,
        "unit": "Mpix·hyp/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"oracle cost volume over coarse rows {first}..{first + rows} of {g.coarse_rows} "
                   f"({threads} threads, row bands) + its serial SGM over that band "
                   f"({t1 - t0:.2f} s + {t2 - t1:.2f} s), scaled by {scale:.2f} to the capture; "
                   f"host CPU: {cpu_model()}"),
        "seconds_sample": t2 - t0,
        "seconds_cost_sample": t1 - t0,
        "seconds_sgm_sample": t2 - t1,
        "seconds_estimated_full": total,
        "one_thread": {
            "sample": f"oracle cost volume over coarse rows {first1}..{first1 + rows1}, 1 thread",
            "seconds_sample": t4 - t3,
            "cost_value": pixels * cap.options.planes / one_thread_cost / 1e6,
            "unit": "Mpix·hyp/s (cost volume alone)",
        },
    }


def valu_roofline(zncc: int, ms: float, kernel: str, traffic) -> dict:
    """Roofline block of a ZNCC kernel. Its bound is the vector ALU: the
    published FP32 peak (157.3 TFLOP/s) counts a fused multiply-add as two
    flops and packed FP32 as two lanes, and bit-exact ZNCC arithmetic may use
    neither, so the kernel's own ceiling is one unfused, unpacked lane
    operation per lane per clock: 157.3 / 4 = 39.3 T lane-ops/s, which FP32
    and FP64 operations share (78.6 / 2)."""
    s = ms / 1000.0
    fp32 = zncc * FP32_PER_ZNCC / s / 1e12
    fp64 = zncc * FP64_PER_ZNCC / s / 1e12
    ceiling = PEAK_FP32_TFLOPS / 4.0
    return {
        "bound": "valu",
        "achieved": round(fp32, 3),
        "peak": PEAK_FP32_TFLOPS,
        "unit": "TFLOP/s",
        "frac": round(fp32 / PEAK_FP32_TFLOPS, 4),
        "traffic": traffic,
        "kernel": kernel,
        "compute": "vector FP32 + FP64 ALU (ZNCC gathers have no matrix form)",
        "avg_launch_ms": round(ms, 4),
        "algorithmic_fp32_per_zncc": FP32_PER_ZNCC,
        "algorithmic_fp64_per_zncc": FP64_PER_ZNCC,
        "zncc_per_launch": zncc,
        "fp64_achieved_tflops": round(fp64, 3),
        "issue_ceiling": {
            "definition": "unfused, unpacked lane operations: FP32 157.3/4 = FP64 78.6/2 "
                          "T ops/s, shared by both; achieved = (FP32 + FP64 algorithmic ops) / time",
            "peak_tops": round(ceiling, 2),
            "achieved_tops": round(fp32 + fp64, 3),
            "frac": round((fp32 + fp64) / ceiling, 4),
        },
    }


def load_traffic(kernel: str, workload: str):
    path = ROOT / "profiles" / "traffic.json"
    try:
        data = json.loads(path.read_text())
    except (OSError, ValueError):
        return None
    table = data.get(workload, {})
    # a templated kernel is listed under its instantiated name
    entry = table.get(kernel) or next((v for k, v in table.items() if k.startswith(kernel + "<")),
                                      None)
    return None if entry is None else entry.get("bytes_per_launch")


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=["config1", "config0", "config2", "config3", "config4",
                                           "config4-learned", "config4-align"],
                    default="config1")
    ap.add_argument("--train-scenes", type=int, default=512,
                    help="config3 / config4-learned: training scenes of the learned band")
    ap.add_argument("--train-all-modules", action="store_true",
                    help="config4-learned: training labels from all sixteen modules' full sweep "
                         "(default: the five wide modules)")
    ap.add_argument("--eval-scenes", type=int, default=24, help="config3: held-out scenes")
    ap.add_argument("--epochs", type=int, default=10,
                    help="config3 / config4-learned: training epochs")
    ap.add_argument("--learned-scenes", type=int, default=3200,
                    help="config1: training captures of the config 4 learned-band measurement "
                         "in the same line (0 skips it)")
    ap.add_argument("--learned-epochs", type=int, default=100)
    ap.add_argument("--learned-validation", type=int, default=8,
                    help="full config 4 captures that choose the band setting")
    ap.add_argument("--learned-test", type=int, default=8,
                    help="held-out full config 4 captures the band is reported on, in "
                         "batches of 4")
    ap.add_argument("--full-path", action="store_true",
                    help="config2/config4: time the whole path (direct pass, filters and "
                         "per-target decisions too) instead of cost + SGM")
    ap.add_argument("--captures", type=int, default=0,
                    help="config2/config4: captures per batch (default 8 and 4)")
    ap.add_argument("--tuning", choices=["default", "auto"], default="default",
                    help="launch shapes: the defaults, or chd_autotune on this capture")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--full-steps", type=int, default=2,
                    help="steps of the whole path (direct pass and targets) timed beside")
    ap.add_argument("--cpu-rows", type=int, default=256,
                    help="coarse rows of the CPU-baseline sample (0 skips it)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--copy-probe", type=float, default=0.0, metavar="GIB",
                    help="only copy GIB of device memory once (calibrates rocprofv3 byte counters)")
    ap.add_argument("--copy-bandwidth", type=float, default=0.0, metavar="GIB",
                    help="only time device-to-device copies of GIB (the practical ceiling of "
                         "kernels that read and write HBM in equal parts)")
    args = ap.parse_args()

    if args.copy_bandwidth > 0:
        import torch
        n = int(args.copy_bandwidth * (1 << 30)) // 4
        x = torch.ones(n, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        reps = 20
        start.record()
        for _ in range(reps):
            y.copy_(x)
        stop.record()
        torch.cuda.synchronize()
        ms = start.elapsed_time(stop) / reps
        print(json.dumps({"copy_bytes_each_way": n * 4, "ms_per_copy": round(ms, 4),
                          "read_plus_write_GBs": round(2 * n * 4 / (ms / 1000.0) / 1e9, 1)}))
        return 0

    if args.copy_probe > 0:
        import torch
        n = int(args.copy_probe * (1 << 30)) // 4
        x = torch.ones(n, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_each_way": n * 4}))
        return 0

    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported by this build "
              "(captures are independent, so the path shards by capture across GPUs)",
              file=sys.stderr)
        return 2

    if args.workload in ("config2", "config4"):
        return bench_batch(args)
    if args.workload == "config3":
        return bench_learned(args)
    if args.workload == "config4-learned":
        return bench_learned4(args)
    if args.workload == "config4-align":
        return bench_align(args)

    from chiaro_amd import _abi, scene
    from chiaro_amd.depth import Context, Session

    t = time.perf_counter()
    cap = scene.config1(seed=args.seed) if args.workload == "config1" else scene.config0(seed=args.seed)
    log(f"[bench] {args.workload} capture generated in {time.perf_counter() - t:.1f} s")
    ref = cap.modules[cap.reference_index]
    ctx = Context(0)
    sess = Session(ctx, cap.modules, cap.reference_index, cap.options)
    g = sess.grid
    units = g.coarse_columns * g.coarse_rows * g.labels
    stages = _abi.CHD_STAGE_COST | _abi.CHD_STAGE_SGM
    tuning = sess.autotune() if args.tuning == "auto" else sess.tuning()

    # Every kernel of the path runs on the context's stream, so a stream
    # synchronise on both sides brackets the timed region. (torch is not
    # used here: it carries its own HIP runtime, which cannot share the
    # process with the system runtime this library links.)
    for _ in range(args.warmup):
        sess.run(stages)
    ctx.synchronize()
    sess.reset_kernel_times()
    sess.enable_timing(True)
    log(f"[bench] timing {args.steps} steps")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sess.run(stages)
    ctx.synchronize()
    t1 = time.perf_counter()
    sess.enable_timing(False)
    kernels = sess.kernel_times()
    work = sess.work()
    ms_per_step = (t1 - t0) * 1000.0 / args.steps
    pixels = ref.width * ref.height
    value = pixels * cap.options.planes / (ms_per_step / 1000.0) / 1e6

    # the whole path, timed beside the headline step
    full_ms = None
    full_kernels = {}
    if args.full_steps > 0:
        sess.run(_abi.CHD_STAGE_ALL)
        ctx.synchronize()
        sess.reset_kernel_times()
        sess.enable_timing(True)
        f0 = time.perf_counter()
        for _ in range(args.full_steps):
            sess.run(_abi.CHD_STAGE_ALL)
        ctx.synchronize()
        full_ms = (time.perf_counter() - f0) * 1000.0 / args.full_steps
        sess.enable_timing(False)
        full_kernels = {k: round(v[1] / args.full_steps, 3) for k, v in sess.kernel_times().items()}
        work_full = sess.work()
    else:
        work_full = work

    # input preparation (luminance_half of the five u16 mosaics), host
    # buffers in and out: the kernel's own time is in the rocprof trace
    prep_ms = None
    if cap.raws is not None:
        from chiaro_amd import luminance_half
        shifted = [r << 4 for r in cap.raws]
        luminance_half(ctx, shifted[0], 64.0 * 16, 4095.0 * 16)
        p0 = time.perf_counter()
        for r in shifted:
            luminance_half(ctx, r, 64.0 * 16, 4095.0 * 16)
        prep_ms = (time.perf_counter() - p0) * 1000.0

    launches, total_ms = kernels["cost_volume"]
    cost_ms = total_ms / launches
    roofline = valu_roofline(work["cost_zncc"], cost_ms, "k_cost_volume",
                             load_traffic("cost_volume", args.workload))
    sgm_ms = sum(v[1] for k, v in kernels.items() if k.startswith("sgm")) / args.steps
    hbm = {
        "kernels": "sgm_paths + sgm_select",
        "ms_per_step": round(sgm_ms, 4),
        "algorithmic_bytes": units * (BYTES_PER_UNIT - 8),
        "achieved_GBs": round(units * (BYTES_PER_UNIT - 8) / (sgm_ms / 1000.0) / 1e9, 1),
        "peak_GBs": PEAK_HBM_GBS,
    }
    hbm["frac"] = round(hbm["achieved_GBs"] / PEAK_HBM_GBS, 4)
    # composite lower bound of the step: per stage the larger of its HBM
    # time and its FP32 + FP64 issue time at the published peaks
    t_cost = max(units * 8 / (PEAK_HBM_GBS * 1e9),
                 work["cost_zncc"] * (FP32_PER_ZNCC / (PEAK_FP32_TFLOPS * 1e12)
                                      + FP64_PER_ZNCC / (PEAK_FP64_TFLOPS * 1e12)))
    t_sgm = units * 96 / (PEAK_HBM_GBS * 1e9)
    t_select = units * 4 / (PEAK_HBM_GBS * 1e9)
    t_lb = (t_cost + t_sgm + t_select) * 1000.0
    # the same bound with the cost stage at the bit-exact issue ceiling:
    # unfused, unpacked lane operations (no FMA contraction, no packed FP32),
    # one per cycle per SIMD lane, FP32 and FP64 alike
    issue_tops = PEAK_FP32_TFLOPS / 4
    t_cost_issue = max(units * 8 / (PEAK_HBM_GBS * 1e9),
                       work["cost_zncc"] * (FP32_PER_ZNCC + FP64_PER_ZNCC) / (issue_tops * 1e12))
    t_lb_issue = (t_cost_issue + t_sgm + t_select) * 1000.0
    composite = {
        "definition": "sum over cost, SGM paths, select of max(bytes / 8 TB/s, "
                      "FP32 / 157.3 TFLOP/s + FP64 / 78.6 TFLOP/s); frac = lower bound / step",
        "lower_bound_ms": round(t_lb, 4),
        "step_ms": round(ms_per_step, 4),
        "frac": round(t_lb / ms_per_step, 4),
        "issue_ceiling": {
            "definition": "the cost stage at the unfused, unpacked issue ceiling "
                          f"({issue_tops:.2f} T lane-ops/s, FP32 and FP64 alike) instead of the "
                          "published peaks; SGM and select at 8 TB/s",
            "lower_bound_ms": round(t_lb_issue, 4),
            "frac": round(t_lb_issue / ms_per_step, 4),
        },
    }

    cpu = None
    if args.cpu_rows > 0:
        host = host_cpu()
        threads = args.cpu_threads or cpu_threads(host)
        log(f"[bench] CPU baseline: {args.cpu_rows} coarse rows on {threads} threads")
        cpu = cpu_baseline(cap, args.cpu_rows, threads)
        cpu["host"] = host

    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mpix·hyp/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (f64 coordinates)",
        "data": "synthetic",
        "config": {
            "workload": (f"{args.workload}: {len(cap.modules)} modules {ref.width}x{ref.height}, "
                         f"{cap.options.planes} hypotheses, 8-path SGM, seed {args.seed}"),
            "reference_raster": [ref.width, ref.height],
            "coarse_grid": [g.coarse_columns, g.coarse_rows, g.coarse_step],
            "labels": g.labels,
            "views": len(cap.modules),
            "parallelism": "1 GPU",
            "tuning": tuning,
        },
        "units_per_s": round(units / (ms_per_step / 1000.0), 1),
        "kernels_ms": {k: round(v[1] / args.steps, 4) for k, v in kernels.items()},
        "roofline": roofline,
        "sgm_hbm": hbm,
        "composite_roofline": composite,
        "full_path_ms": None if full_ms is None else round(full_ms, 2),
        "input_prep_ms_incl_pcie": None if prep_ms is None else round(prep_ms, 2),
        "full_path_kernels_ms": full_kernels,
        "work": work_full,
        "cpu_baseline": cpu,
    }
    if cpu:
        line["gpu_over_cpu"] = round(value / cpu["value"], 1)
    sess.close()
    if args.learned_scenes > 0 and args.workload == "config1":
        # the metric's second half: the learned band on config 4 (4 captures
        # of 16 modules batched, 256 hypotheses) against the plain full sweep
        line["learned_band"] = learned_config4(ctx, args.learned_scenes, args.learned_epochs,
                                               args.learned_validation, args.learned_test,
                                               args.seed)
    ctx.close()
    print(json.dumps(line))
    return 0


def bench_batch(args) -> int:
    """Config 2 (`captures` captures of ten modules of three sizes) or
    config 4 (captures of sixteen 4160x3120 wide, mid and tele modules at 256
    hypotheses) run as one batch (one SGM sweep launch over all captures,
    hipGraph replay); the same batch with the persistent cost-volume queue
    and the sessions run one after another are timed beside it."""
    from chiaro_amd import _abi, scene
    from chiaro_amd.depth import Batch, Context, Session

    count = args.captures or (8 if args.workload == "config2" else 4)
    rig_of = scene.rig_config2 if args.workload == "config2" else scene.rig_config4
    ctx = Context(0)
    t = time.perf_counter()
    caps, sessions = [], []
    for k in range(count):
        # rendered on the GPU; the sessions copy the device planes
        frames = scene.Frames(ctx, rig_of(seed=args.seed + k))
        frames.render(args.seed + k)
        caps.append(frames.capture())
        sessions.append(Session(ctx, caps[-1].modules, 0, caps[-1].options))
        frames.close()
    log(f"[bench] {count} {args.workload} captures rendered on the GPU in "
        f"{time.perf_counter() - t:.1f} s")
    tuning = sessions[0].autotune() if args.tuning == "auto" else sessions[0].tuning()
    if args.tuning == "auto":
        for s in sessions[1:]:
            s.set_tuning(**tuning)
    batch = Batch(ctx, sessions)
    one_sweep = Batch(ctx, sessions, one_sweep=True)
    queued = Batch(ctx, sessions, persistent=True)
    stages = _abi.CHD_STAGE_ALL if args.full_path else _abi.CHD_STAGE_COST | _abi.CHD_STAGE_SGM

    def timed(fn, steps):
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        ctx.synchronize()
        return (time.perf_counter() - t0) * 1000.0 / steps

    for _ in range(args.warmup):
        batch.run(stages)
    log(f"[bench] timing {args.steps} batched steps")
    ms = timed(lambda: batch.run(stages), args.steps)
    cost_ms = timed(lambda: batch.run(_abi.CHD_STAGE_COST), args.steps)
    zncc = sum(s.work()["cost_zncc"] for s in sessions)
    queued.run(stages)
    queued_ms = timed(lambda: queued.run(stages), max(1, args.steps // 2))
    one_sweep.run(stages)
    one_sweep_ms = timed(lambda: one_sweep.run(stages), max(1, args.steps // 2))

    def sequential():
        for s in sessions:
            s.run(stages)

    sequential()
    seq_ms = timed(sequential, max(1, args.steps // 2))
    ref = caps[0].modules[0]
    planes = caps[0].options.planes
    pixels = ref.width * ref.height * len(caps)
    if args.workload == "config2":
        workload = (f"config2: {len(caps)} captures x 10 modules (5 x 4160x3120 f=3300, "
                    f"3 x 2080x1560 f=4125, 2 x 1040x780 f=4425), {planes} hypotheses, "
                    f"8-path SGM, seeds {args.seed}..{args.seed + len(caps) - 1}, batched")
    else:
        workload = (f"config4: {len(caps)} captures x 16 modules 4160x3120 (5 wide f=3300, "
                    f"5 mid f=8250, 6 tele f=17700), 24-plane scenes, {planes} hypotheses, "
                    f"8-path SGM, seeds {args.seed}..{args.seed + len(caps) - 1}, batched")
    value = pixels * planes / (ms / 1000.0) / 1e6
    g = sessions[0].grid
    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mpix·hyp/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (f64 coordinates)",
        "data": "synthetic",
        "config": {
            "workload": workload,
            "coarse_grid": [g.coarse_columns, g.coarse_rows, g.coarse_step],
            "labels": g.labels,
            "parallelism": "1 GPU",
            "tuning": tuning,
        },
        "persistent_queue_ms_per_step": round(queued_ms, 3),
        "one_sweep_ms_per_step": round(one_sweep_ms, 3),
        "sequential_ms_per_step": round(seq_ms, 3),
        "batched_over_sequential": round(seq_ms / ms, 3),
        "cost_ms_per_step": round(cost_ms, 3),
        "roofline": valu_roofline(zncc, cost_ms, "k_cost_volume (one launch per capture)", None),
        "cpu_baseline": None,
    }
    batch.close()
    one_sweep.close()
    queued.close()
    for s in sessions:
        s.close()
    ctx.close()
    print(json.dumps(line))
    return 0


def learned_config3(ctx, train_scenes: int, epochs: int, eval_scenes: int, seed: int,
                    grid, ablation: bool = True) -> dict:
    """Trains the band proposal on the plain full sweep's label maps of
    `train_scenes` seeded config 3 scenes and measures, on `eval_scenes`
    held-out seeds, the full sweep (cost + SGM) against the band sweep
    (network + windows + cost + SGM) for each (width cap, tile, threshold) in
    `grid`; the fastest setting with at least 98 % +-1-hypothesis agreement
    is `best`."""
    import numpy as np

    from chiaro_amd import scene
    from chiaro_amd.proposal import (Proposal, agreement, build_dataset, constant_proposal,
                                     input_transform)

    t = time.perf_counter()
    offset, scale = input_transform([scene.config3(seed=s) for s in range(16)])
    x, y = build_dataset(ctx, range(train_scenes), lambda s: scene.config3(seed=s), offset, scale)
    log(f"[bench] {train_scenes} config3 scenes and labels in {time.perf_counter() - t:.1f} s")
    h, w = x.shape[1:3]
    net = Proposal(ctx, w, h, max_batch=32, seed=seed)
    net.set_input_transform(offset, scale)
    net.set_dataset(x, y)
    t = time.perf_counter()
    losses = net.train(epochs=epochs, batch=32, lr=1e-3, seed=seed, chunk=10,
                       log=lambda done, l: log(f"[bench] epoch {done}/{epochs}: loss {l[-1]:.4f}"))
    train_s = time.perf_counter() - t
    log(f"[bench] trained {epochs} epochs in {train_s:.1f} s, last losses {losses[-3:]}")
    evals = [scene.config3(seed=100000 + k) for k in range(eval_scenes)]
    rows = []
    for max_half, tile, thr in grid:
        stats = [agreement(ctx, cap, net, thr, tile=tile, max_half_width=max_half) for cap in evals]
        row = {k: float(np.mean([st[k] for st in stats])) for k in stats[0]}
        row["threshold"] = thr
        row["tile"] = tile
        row["max_half_width"] = max_half
        row["speedup"] = row["full_ms"] / row["band_ms"]
        rows.append(row)
        log(f"[bench] max half {max_half} tile {tile} threshold {thr}: "
            f"agreement {row['agreement']:.4f} fallback {row['fallback_fraction']:.3f} "
            f"speedup {row['speedup']:.3f} covers {row['window_covers_label']:.3f}")
    out = {}
    if ablation:
        # equal budget (+-8 labels everywhere, no fallback): the network's
        # centres against one centre, the training set's median label
        const = constant_proposal(ctx, w, h, float(np.median(y[y > 0])))
        for name, prop in (("network", net), ("constant_centre", const)):
            stats = [agreement(ctx, cap, prop, 0.0, tile=8) for cap in evals]
            out[name] = {k: round(float(np.mean([st[k] for st in stats])), 5)
                         for k in ("agreement", "window_covers_label", "band_ms")}
        const.close()
        log(f"[bench] ablation {out}")
    net.close()
    ok = [r for r in rows if r["agreement"] >= 0.98]
    ref = evals[0].modules[0]
    return {
        "rows": rows,
        "best": max(ok, key=lambda r: r["speedup"]) if ok else None,
        "ablation": out,
        "losses": losses,
        "train_s": train_s,
        "offset": offset,
        "scale": scale,
        "pixels_hyp": ref.width * ref.height * evals[0].options.planes,
        "full_ms": float(np.mean([r["full_ms"] for r in rows])),
    }


def learned_summary(res: dict, workload: str) -> dict:
    chosen = res["best"] or max(res["rows"], key=lambda r: r["agreement"])
    return {
        "workload": workload,
        "met_98_percent": res["best"] is not None,
        "threshold": chosen["threshold"],
        "tile": chosen["tile"],
        "max_half_width": chosen["max_half_width"],
        "speedup": round(chosen["speedup"], 3),
        "agreement": round(chosen["agreement"], 5),
        "tested_agreement": round(chosen["tested_agreement"], 5),
        "fallback_fraction": round(chosen["fallback_fraction"], 4),
        "full_ms": round(chosen["full_ms"], 4),
        "band_ms": round(chosen["band_ms"], 4),
        "ablation_equal_budget": res["ablation"],
        "input_transform": {"offset": res["offset"], "scale": res["scale"]},
        "train_seconds": round(res["train_s"], 2),
    }


def bench_learned(args) -> int:
    """Config 3: the learned band proposal over the full grid of settings
    (width caps 8..32, tiles 4 and 8, thresholds 0..0.9), with the
    equal-budget ablation; see learned_config3."""
    from chiaro_amd.depth import Context

    ctx = Context(0)
    grid = [(mh, tile, thr) for mh in (8, 16, 24, 32) for tile in (4, 8)
            for thr in (0.0, 0.2, 0.3, 0.4, 0.5, 0.6, 0.7, 0.8, 0.9)]
    res = learned_config3(ctx, args.train_scenes, args.epochs, args.eval_scenes, args.seed, grid)
    chosen = res["best"] or max(res["rows"], key=lambda r: r["agreement"])
    pixels_hyp, full_ms = res["pixels_hyp"], res["full_ms"]
    workload = (f"config3: 3 views 1040x780, 128 hypotheses, grid step 4; network trained on "
                f"{args.train_scenes} scenes x {args.epochs} epochs, evaluated on "
                f"{args.eval_scenes} held-out scenes")
    summary = learned_summary(res, workload)
    summary["thresholds"] = [{k: (round(v, 5) if isinstance(v, float) else v) for k, v in r.items()}
                             for r in res["rows"]]
    summary["losses"] = [round(l, 4) for l in res["losses"]]
    line = {
        "metric": METRIC,
        "value": round(pixels_hyp / (chosen["band_ms"] / 1000.0) / 1e6, 2),
        "unit": "Mpix·hyp/s",
        "n_gpus": 1,
        "steps": 3,
        "warmup": 1,
        "ms_per_step": round(chosen["band_ms"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 sweep, bf16 network",
        "data": "synthetic",
        "config": {"workload": workload, "parallelism": "1 GPU"},
        "full_sweep_ms": round(full_ms, 4),
        "full_sweep_value": round(pixels_hyp / (full_ms / 1000.0) / 1e6, 2),
        "learned_band": summary,
        "roofline": None,
        "cpu_baseline": None,
    }
    ctx.close()
    print(json.dumps(line))
    return 0


TRAIN_SEED, VALIDATION_SEED, TEST_SEED = 10_000, 200_000, 300_000
VALIDATION_MARGIN = 0.0015


def learned_config4(ctx, train_captures: int, epochs: int, validation_captures: int,
                    test_captures: int, seed: int, steps: int = 2, settings=None,
                    wide_only: bool = True) -> dict:
    """The learned band on config 4, every capture rendered on the GPU.

    Training: the full sweep's WTA label maps of `train_captures` half-scale
    config 4 captures of the five wide modules on a 4 px grid (the coarse
    grid, angles, disparities and 257 labels of the full capture at 8 px),
    rendered, converted and swept on the device (one session fed new planes
    per capture); the network inputs are built on the device from the same
    planes. Selection: the settings (confidence threshold, tile, width cap)
    are timed and scored on `validation_captures` full 16-module captures
    batched; the fastest with >= 98 % +-1-hypothesis agreement is chosen.
    Report: on `test_captures` other full captures in batches of four, that
    setting and the specification's (+-8 labels, tiles below confidence 0.5
    fall back) against the plain full sweep, for cost + SGM and end to end
    (cost, SGM, WTA, direct pass, filters, per-target consistency), with the
    agreement of the u8 consistency mask (provenance) and the per-target
    decision codes, averaged over the batches."""
    import numpy as np

    from chiaro_amd import _abi, scene
    from chiaro_amd.depth import Session
    from chiaro_amd.proposal import Proposal, batch_agreement_sessions, input_transform

    t_all = time.perf_counter()
    rig = scene.rig_config4(seed=TRAIN_SEED, scale=0.5, wide_only=wide_only)
    rig.options.grid_step = 4
    frames = scene.Frames(ctx, rig)
    # input standardisation from four captures' input planes
    probes = []
    for k in range(4):
        frames.render(TRAIN_SEED + k)
        mods = [scene.ModuleInput(name=str(i), luminance=frames.read(i)[1], width=0, height=0,
                                  camera=None, warp_points=np.zeros((1, 1, 2), np.float32),
                                  warp_step=32) for i in range(3)]
        probes.append(scene.Capture(modules=mods, reference_index=0, options=rig.options))
    offset, scale = input_transform(probes)
    frames.render(TRAIN_SEED)
    sess = Session(ctx, frames.capture().modules, 0, rig.options)
    g = sess.grid
    ph, pw = frames.planes()[0].shape
    h, w = ph // 2, pw // 2
    net = Proposal(ctx, w, h, max_batch=16, seed=seed)
    net.set_label_count(g.labels)
    net.set_input_transform(offset, scale)
    x = np.empty((train_captures, h, w, 16), np.uint16)
    y = np.empty((train_captures, h, w), np.uint16)
    stages = _abi.CHD_STAGE_COST | _abi.CHD_STAGE_SGM
    t = time.perf_counter()
    for k in range(train_captures):
        frames.render(TRAIN_SEED + k)
        sess.set_planes(frames.planes())
        sess.run(stages)
        _, _, best = sess.coarse()
        tested = sess.tested()
        labels = np.where(tested > 0, best, 0).astype(np.uint16)
        y[k] = labels.reshape(g.coarse_rows, g.coarse_columns)[:h, :w]
        x[k] = net.session_input(sess)
        if (k + 1) % 200 == 0:
            log(f"[bench] {k + 1} training captures on the device")
    data_s = time.perf_counter() - t
    sess.close()
    frames.close()
    log(f"[bench] {train_captures} half-scale config4 training captures in {data_s:.1f} s")
    net.set_dataset(x, y)
    del x, y
    t = time.perf_counter()
    losses = net.train(epochs=epochs, batch=16, lr=1e-3, seed=seed, chunk=10,
                       log=lambda done, l: log(f"[bench] epoch {done}/{epochs}: loss {l[-1]:.4f}"))
    train_s = time.perf_counter() - t
    log(f"[bench] trained {epochs} epochs in {train_s:.1f} s, losses {losses[0]:.3f} -> "
        f"{losses[-1]:.3f}")

    full_rig = scene.rig_config4(seed=TEST_SEED)
    frames = scene.Frames(ctx, full_rig)

    def sessions_for(first, count):
        out = []
        for k in range(count):
            frames.render(first + k)
            out.append(Session(ctx, frames.capture().modules, 0, full_rig.options))
        return out

    if settings is None:
        # fallback tiles sweep every label (0) or a +-48 / +-64 window
        settings = [(thr, tile, mh, 8, wide) for wide in (0, 48, 64)
                    for mh in (16, 20, 22, 24, 26, 28, 32) for tile in (4, 8)
                    for thr in (0.25, 0.3, 0.35, 0.4, 0.5)]
    validation = sessions_for(VALIDATION_SEED, validation_captures)
    val_rows = batch_agreement_sessions(ctx, validation, net, settings, steps)
    for s in validation:
        s.close()
    for r in val_rows:
        log(f"[bench] validation: max half {r['max_half_width']} fallback half "
            f"{r['fallback_half_width']} tile {r['tile']} threshold "
            f"{r['threshold']}: agreement {r['agreement']:.4f} fallback "
            f"{r['fallback_fraction']:.3f} speedup {r['speedup']:.3f}")
    # the choice keeps a margin over 98 % on the validation captures: with
    # eight of them the same setting's held-out agreement has come out up
    # to 0.1 points below its validation agreement
    ok = [r for r in val_rows if r["agreement"] >= 0.98 + VALIDATION_MARGIN]
    chosen = max(ok, key=lambda r: r["speedup"]) if ok else max(val_rows, key=lambda r: r["agreement"])
    chosen_setting = (chosen["threshold"], chosen["tile"], chosen["max_half_width"],
                      chosen["half_width"], chosen["fallback_half_width"])
    spec_setting = (0.5, 8, 8)
    # the held-out captures in batches of four (the configuration's captures
    # per launch); agreements averaged over the batches, times per batch
    test = sessions_for(TEST_SEED, test_captures)
    groups = [test[i:i + 4] for i in range(0, len(test), 4)]

    def over_groups(stages):
        per = [batch_agreement_sessions(ctx, g, net, [chosen_setting, spec_setting], steps,
                                        stages=stages) for g in groups]
        rows = []
        for j in range(2):
            r = dict(per[0][j])
            for key in ("agreement", "tested_agreement", "window_coverage", "mean_window",
                        "fallback_fraction", "full_ms", "band_ms",
                        "consistency_mask_agreement", "decision_agreement"):
                if key in r:
                    r[key] = sum(p[j][key] for p in per) / len(per)
            r["speedup"] = r["full_ms"] / r["band_ms"]
            rows.append(r)
        return rows

    cost_rows = over_groups(_abi.CHD_STAGE_COST | _abi.CHD_STAGE_SGM)
    e2e_rows = over_groups(_abi.CHD_STAGE_ALL)
    for s in test:
        s.close()
    frames.close()
    net.close()

    def pick(row, keys):
        return {k: (round(row[k], 5) if isinstance(row[k], float) else row[k]) for k in keys}

    keys = ("threshold", "tile", "max_half_width", "half_width", "fallback_half_width",
            "agreement", "tested_agreement",
            "window_coverage", "mean_window", "fallback_fraction", "full_ms", "band_ms",
            "speedup")
    e2e_keys = keys + ("consistency_mask_agreement", "decision_agreement")
    return {
        "workload": (f"config4: {test_captures} held-out captures x 16 modules 4160x3120, 256 "
                     f"hypotheses, in batches of 4; network trained on {train_captures} half-scale "
                     f"config4 captures ({'five wide' if wide_only else 'all sixteen'} modules, "
                     f"4 px grid) x {epochs} epochs; setting chosen on {validation_captures} "
                     f"other full captures (fastest with agreement >= "
                     f"{100 * (0.98 + VALIDATION_MARGIN):.2f} % there)"),
        "met_98_percent": cost_rows[0]["agreement"] >= 0.98,
        "speedup": round(cost_rows[0]["speedup"], 3),
        "agreement": round(cost_rows[0]["agreement"], 5),
        "end_to_end_speedup": round(e2e_rows[0]["speedup"], 3),
        "chosen": {"cost_sgm": pick(cost_rows[0], keys), "end_to_end": pick(e2e_rows[0], e2e_keys)},
        "specification_setting": {"cost_sgm": pick(cost_rows[1], keys),
                                  "end_to_end": pick(e2e_rows[1], e2e_keys)},
        "validation": [pick(r, keys) for r in val_rows],
        "seeds": {"train": [TRAIN_SEED, TRAIN_SEED + train_captures - 1],
                  "validation": [VALIDATION_SEED, VALIDATION_SEED + validation_captures - 1],
                  "test": [TEST_SEED, TEST_SEED + test_captures - 1]},
        "input_transform": {"offset": offset, "scale": scale},
        "training_data_seconds": round(data_s, 2),
        "train_seconds": round(train_s, 2),
        "total_seconds": round(time.perf_counter() - t_all, 2),
        "losses": [round(v, 4) for v in losses[:: max(1, len(losses) // 10)]],
    }


def _turn(degrees_x: float, degrees_y: float):
    """Rotation about x then y, degrees."""
    import numpy as np
    ax, ay = np.radians(degrees_x), np.radians(degrees_y)
    rx = np.array([[1, 0, 0], [0, np.cos(ax), -np.sin(ax)], [0, np.sin(ax), np.cos(ax)]])
    ry = np.array([[np.cos(ay), 0, np.sin(ay)], [0, 1, 0], [-np.sin(ay), 0, np.cos(ay)]])
    return ry @ rx


def bench_align(args) -> int:
    """config4-align: align_module (align.rs:270-519) and then
    refine_resolution_warp (resolution.rs:232-432) for the fifteen
    non-reference modules of one full config 4 capture rendered on the GPU,
    each module's factory model turned by up to 0.3 deg about x and y from
    its true orientation. The planes stay in HBM (the value); the same pass
    from host planes, as the reference's signature hands them over, is
    reported beside it. The kernel times are in the rocprof trace."""
    import ctypes as C

    import numpy as np

    from chiaro_amd import _abi, align, scene
    from chiaro_amd.depth import Context

    rig = scene.rig_config4(seed=args.seed)
    ctx = Context(0)
    frames = scene.Frames(ctx, rig)
    frames.render(args.seed)
    device = frames.planes()                                  # resident in HBM
    planes = [frames.read(i)[1] for i in range(len(rig.modules))]
    rng = np.random.default_rng(args.seed)
    modules, host_modules = [], []
    for i, m in enumerate(rig.modules):
        if i == rig.reference:
            factory = m
        else:
            turn = _turn(*rng.uniform(-0.3, 0.3, 2))
            rot = turn @ np.array(m.rotation_wc[:]).reshape(3, 3)
            factory = scene._module(m.focal_px, m.width, m.height, tuple(m.center[:]), rot)
        for out, lum in ((modules, device[i]), (host_modules, planes[i])):
            out.append(scene.ModuleInput(
                name=rig.names[i], luminance=lum, width=m.width, height=m.height,
                camera=scene.camera_for(factory),
                warp_points=scene.identity_warp(m.width, m.height), warp_step=32))
    ref = rig.reference
    w, h = rig.modules[ref].width, rig.modules[ref].height
    targets = [i for i in range(len(modules)) if i != ref]

    def one_pass(mods, lum):
        t0 = time.perf_counter()
        results = [align.align_module(ctx, mods, ref, t) for t in targets]
        t1 = time.perf_counter()
        warps = [align.refine_resolution_warp(ctx, lum[ref], lum[t], results[k].warp, w, h)
                 for k, t in enumerate(targets)]
        t2 = time.perf_counter()
        return results, warps, (t1 - t0) * 1000.0, (t2 - t1) * 1000.0

    align.align_module(ctx, modules, ref, targets[0])        # warm-up
    log(f"[bench] aligning {len(targets)} modules")
    for _ in range(args.warmup):
        one_pass(modules, device)
    align_ms, res_ms = [], []
    for _ in range(args.steps):
        results, warps, a, r = one_pass(modules, device)
        align_ms.append(a)
        res_ms.append(r)
    a_ms, r_ms = float(np.median(align_ms)), float(np.median(res_ms))
    per_module = (a_ms + r_ms) / len(targets)
    # the same pass from host planes (PCIe copies of both planes per call)
    host_results, host_warps, ha, hr = one_pass(host_modules, planes)
    same = all(np.array_equal(x.warp.points, y.warp.points, equal_nan=True)
               for x, y in zip(results, host_results)) and \
        all(np.array_equal(x.warp.points, y.warp.points, equal_nan=True)
            for x, y in zip(warps, host_warps))
    frames.close()
    statuses = {}
    for r in results:
        key = r.report["status"].split(":")[0]
        statuses[key] = statuses.get(key, 0) + 1
    modules_report = [{
        "module": modules[t].name, "status": results[k].report["status"],
        "inliers": results[k].report["inliers"], "patches": results[k].report["patches"],
        "residual_median_px": round(results[k].report["residual_median_px"], 3),
        "verified_fraction": round(warps[k].report["verified_fraction"], 4),
        "supported_fraction": round(warps[k].report["supported_fraction"], 4),
    } for k, t in enumerate(targets)]

    cpu = None
    if args.cpu_rows > 0:
        # the oracle on two modules (a wide and a tele), one thread
        from chiaro_amd.depth import _ViewArrays
        from oracle import binding
        lib = binding.lib()
        sample = [targets[0], targets[-1]]
        c0 = time.perf_counter()
        for t in sample:
            views = _ViewArrays([host_modules[ref], host_modules[t]])
            o = align.AlignOptions().ffi()
            wp = align.Warp.empty(w, h, 32)
            rep = _abi.AlignmentReport()
            lib.or_align_module(C.byref(views.views[0]), C.byref(views.views[1]), C.byref(o),
                                C.byref(wp.ffi()), C.byref(rep))
            out = align.Warp.empty(w, h, 32)
            rr = _abi.ResolutionReport()
            lib.or_refine_resolution_warp(planes[ref].ctypes.data, planes[ref].shape[1],
                                          planes[ref].shape[0], planes[t].ctypes.data,
                                          planes[t].shape[1], planes[t].shape[0],
                                          C.byref(wp.ffi()), w, h, C.byref(out.ffi()),
                                          C.byref(rr))
        cpu_ms = (time.perf_counter() - c0) * 1000.0 / len(sample)
        cpu = {"value": round(1000.0 / cpu_ms, 3), "unit": "modules/s", "cores": 1,
               "kind": "port", "sample": f"align_module + refine_resolution_warp of modules "
                                         f"{modules[sample[0]].name} and {modules[sample[1]].name}",
               "ms_per_module": round(cpu_ms, 1), "host": host_cpu()}
    line = {
        "metric": "modules/s: align_module + refine_resolution_warp on 1 GPU",
        "value": round(1000.0 / per_module, 3),
        "unit": "modules/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(a_ms + r_ms, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 (f64 geometry and fits)", "data": "synthetic",
        "config": {"workload": "config4-align: one 16-module config 4 capture (4160x3120 rasters, "
                               "2080x1560 planes), 15 modules aligned to the reference from factory "
                               "models turned by up to 0.3 deg, then refined at 32 px",
                   "seed": args.seed},
        "align_ms_per_step": round(a_ms, 3),
        "resolution_ms_per_step": round(r_ms, 3),
        "ms_per_module": round(per_module, 3),
        "host_planes": {"align_ms_per_step": round(ha, 3), "resolution_ms_per_step": round(hr, 3),
                        "ms_per_module": round((ha + hr) / len(targets), 3),
                        "same_warps_as_device_planes": same},
        "status_counts": statuses,
        "modules": modules_report,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    ctx.close()
    return 0


def bench_learned4(args) -> int:
    """Config 4 with the learned band as its own line (learned_config4); the
    value is the band's cost + SGM throughput on the held-out batch."""
    from chiaro_amd.depth import Context
    ctx = Context(0)
    res = learned_config4(ctx, args.train_scenes, args.epochs, args.learned_validation,
                          args.captures or args.learned_test, args.seed, steps=args.steps,
                          wide_only=not args.train_all_modules)
    ctx.close()
    row = res["chosen"]["cost_sgm"]
    pixels_hyp = 4160 * 3120 * 256 * (args.captures or args.learned_test)
    line = {
        "metric": METRIC,
        "value": round(pixels_hyp / (row["band_ms"] / 1000.0) / 1e6, 2),
        "unit": "Mpix·hyp/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": 1,
        "ms_per_step": round(row["band_ms"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 sweep, bf16 network",
        "data": "synthetic",
        "config": {"workload": "config4 learned band: " + res["workload"], "parallelism": "1 GPU"},
        "learned_band": res,
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
