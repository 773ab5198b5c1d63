"""Benchmark of the Field convolution path on one MI355X.

Workload (default): the FP32 CNN-layer triple, batch 16 images 64x64 with 16
planes, 32 filters 3x3x16 (He-normal, std sqrt(2/144)), BorderValid layer:
  ProjectFrame + blockReduceSum(16), BorderValid    y  = PF(x, W)        64x64 -> 62x62, 16x32 planes
  BackProjectFrame + blockReduceSum(32), BorderFull dx = BPF(dy, W)      62x62 -> 64x64, 16x16 planes
  FilterAdjoint + blockReduceSum(16), BorderValid   dW = FA(x, dy)       -> 3x3, 512 planes
dy is uniform[-1,1).  One step runs the three ops once; inputs are resident in
HBM before the timed region starts.

Prints one JSON line: metric/value (whole-job GFLOP/s of the triple), the
roofline of the dominant kernel (timed with HIP events on the stream the kernels
run on), and the CPU baseline (the oracle's FP32 restatement on a bounded sample
of the same workload, OpenMP on the host cores).
"""
import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "cct-core_amd"))

PEAK_FP32_TFLOPS = 157.3   # MI355X FP32 (vector and matrix)
PEAK_BF16_TFLOPS = 2500.0  # MI355X dense BF16/FP16 MFMA
PEAK_HBM_GBS = 8000.0      # MI355X HBM3E


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_inputs(seed=16, B=16, N=16, K=32, H=64):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, (B * N, H, H)).astype(np.float32)
    w = (rng.standard_normal((K * N, 3, 3)) * np.sqrt(2.0 / (N * 9))).astype(np.float32)
    dy = rng.uniform(-1, 1, (B * K, H - 2, H - 2)).astype(np.float32)
    return x, w, dy


def triple_descs(L, B=16, N=16, K=32, H=64, precision=0):
    f = lambda planes, s: (2, s, s, planes)  # noqa: E731
    pf = L.make_desc(L.PROJECT_FRAME_BLOCK_REDUCE_SUM, L.BORDER_VALID, L.CROSS_CORRELATION, L.NO_SAMPLING, 1, B,
                     f(B * N, H), f(K * N, 3), precision)
    bpf = L.make_desc(L.BACK_PROJECT_FRAME_BLOCK_REDUCE_SUM, L.BORDER_FULL, L.CONVOLUTION, L.NO_SAMPLING, 1, B,
                      f(B * K, H - 2), f(K * N, 3), precision)
    fa = L.make_desc(L.FILTER_ADJOINT_BLOCK_REDUCE_SUM, L.BORDER_VALID, L.CROSS_CORRELATION, L.NO_SAMPLING, 1, B,
                     f(B * N, H), f(B * K, H - 2), precision)
    return [("projectFrame+BRS", pf), ("backProjectFrame+BRS", bpf), ("filterAdjoint+BRS", fa)]


def cpu_baseline(sample_images=16, budget_s=10.0):
    """Oracle FP32 restatement (reference accumulation order), OpenMP over output
    planes, on `sample_images` of the batch; median GFLOP/s over repetitions."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    threads = max(1, min(cores, 16))
    B, N, K, H = sample_images, 16, 32, 64
    x, w, dy = make_inputs(B=B)
    ops = [(O.make_desc(O.PF_BRS, O.VALID, O.XCORR, O.NONE, 1, B, x.shape, w.shape), x, w),
           (O.make_desc(O.BPF_BRS, O.FULL, O.CONV, O.NONE, 1, B, dy.shape, w.shape), dy, w),
           (O.make_desc(O.FA_BRS, O.VALID, O.XCORR, O.NONE, 1, B, x.shape, dy.shape), x, dy)]
    flops = 2.0 * B * N * K * 9 * (62 * 62 + 64 * 64 + 62 * 62)
    for d, a, k in ops:  # warm-up
        O.conv(d, a, k, prec="f32", threads=threads)
    times = []
    t_all = time.perf_counter()
    while len(times) < 5 or (time.perf_counter() - t_all < budget_s and len(times) < 100):
        t0 = time.perf_counter()
        for d, a, k in ops:
            O.conv(d, a, k, prec="f32", threads=threads)
        times.append(time.perf_counter() - t0)
    med = statistics.median(times)
    return {"value": round(flops / med / 1e9, 3), "unit": "GFLOP/s", "cores": threads, "kind": "port",
            "sample": "%d of 16 images of the same triple (%.3g GFLOP), oracle FP32 restatement, median of %d reps"
                      % (B, flops / 1e9, len(times))}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-budget", type=float, default=10.0)
    ap.add_argument("--serial", action="store_true", help="run the three ops of a step back to back on one stream")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (--gpus %d requested)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    import cogconv._lib as L
    L.init(0)
    props = L.device_props(0)
    log("device:", props)

    x, w, dy = make_inputs()
    bx, bw, bdy = (L.DeviceBuffer.from_array(a) for a in (x, w, dy))
    ops = []
    for name, desc in triple_descs(L):
        plan = L.Plan(desc)
        out = L.DeviceBuffer(4 * int(np.prod(plan.out_shape)))
        ws = L.DeviceBuffer(plan.workspace_bytes) if plan.workspace_bytes else None
        ins = {"projectFrame+BRS": (bx, bw), "backProjectFrame+BRS": (bdy, bw), "filterAdjoint+BRS": (bx, bdy)}[name]
        ops.append({"name": name, "plan": plan, "in": ins, "out": out, "ws": ws})
        log("plan:", name, plan.description, "GFLOP %.4f" % (plan.flops / 1e9))
    flops_step = sum(op["plan"].flops for op in ops)
    streams = [L.Stream() for _ in ops]
    main_stream = streams[0]

    def launch(i, stream):
        op = ops[i]
        op["plan"].launch(op["in"][0], op["in"][1], op["out"], op["ws"], stream)

    # Phase 1: each op alone on one stream, timed with HIP events on that
    # stream (average launch duration; this feeds the roofline).
    iso_reps = max(20, args.steps // 4)
    per_op_ms = []
    for i in range(len(ops)):
        for _ in range(5):
            launch(i, main_stream)
        e0, e1 = L.Event(), L.Event()
        e0.record(main_stream)
        for _ in range(iso_reps):
            launch(i, main_stream)
        e1.record(main_stream)
        e1.synchronize()
        per_op_ms.append(e0.elapsed_ms(e1) / iso_reps)

    # Phase 2: the timed steps.  The three ops of a step are independent
    # (PF reads x, W; BPF reads dy, W; FA reads x, dy), so a step forks them
    # onto three streams and joins; --serial runs them back to back on one.
    fork = [L.Event() for _ in range(args.steps + args.warmup)]
    done = [[L.Event() for _ in ops] for _ in range(args.steps + args.warmup)]

    def run_step(k):
        if args.serial:
            for i in range(len(ops)):
                launch(i, main_stream)
            return
        fork[k].record(main_stream)
        for i, st in enumerate(streams):
            if i:
                L.check(L.lib().cogconv_stream_wait_event(st.handle, fork[k].handle))
            launch(i, st)
            if i:
                done[k][i].record(st)
        for i in range(1, len(ops)):
            L.check(L.lib().cogconv_stream_wait_event(main_stream.handle, done[k][i].handle))

    for k in range(args.warmup):
        run_step(k)
    L.synchronize()
    K = args.steps
    t_start, t_end = L.Event(), L.Event()
    wall0 = time.perf_counter()
    t_start.record(main_stream)
    for k in range(K):
        run_step(args.warmup + k)
    t_end.record(main_stream)
    t_end.synchronize()
    L.synchronize()
    wall = time.perf_counter() - wall0
    total_ms = t_start.elapsed_ms(t_end)
    ms_per_step = max(total_ms, wall * 1e3) / K
    value = flops_step / (ms_per_step * 1e-3) / 1e9

    dom = int(np.argmax(per_op_ms))
    dplan = ops[dom]["plan"]
    ai = dplan.flops / dplan.bytes
    ridge = PEAK_FP32_TFLOPS * 1e12 / (PEAK_HBM_GBS * 1e9)
    if ai >= ridge:
        achieved = dplan.flops / (per_op_ms[dom] * 1e-3) / 1e12
        roof = {"bound": "mfma", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / PEAK_FP32_TFLOPS, 4), "traffic": None}
    else:
        achieved = dplan.bytes / (per_op_ms[dom] * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": None}
    roof["kernel"] = ops[dom]["name"] + ": " + dplan.description
    roof["algorithmic_flops"] = dplan.flops
    roof["algorithmic_bytes"] = dplan.bytes
    roof["avg_launch_ms"] = round(per_op_ms[dom], 5)
    roof["timing"] = "HIP events around %d back-to-back launches of this op alone on its stream" % iso_reps

    cpu = None if args.no_cpu else cpu_baseline(budget_s=args.cpu_budget)
    line = {
        "metric": "conv GFLOP/s & % roofline (FP32 vs BF16-MFMA); p50/p99 request latency",
        "value": round(value, 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded uniform[-1,1) x and dy, He-normal filters)",
        "config": {"workload": "FP32 CNN-layer triple: PF+BRS Valid, BPF+BRS Full, FA+BRS Valid; batch 16 "
                               "images 64x64x16, 32 filters 3x3x16", "batch": 16, "image": [64, 64, 16],
                   "filters": [32, 3, 3, 16], "flops_per_step": flops_step,
                   "step": "serial on one stream" if args.serial else "three independent ops on three streams"},
        "per_op_isolated": {op["name"]: {"ms": round(t, 5), "gflops": round(op["plan"].flops / (t * 1e-3) / 1e9, 1),
                                         "kernel": op["plan"].description} for op, t in zip(ops, per_op_ms)},
        "roofline": roof,
        "cpu_baseline": cpu,
        "device": props["name"] + " " + props["arch"],
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
