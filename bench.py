"""Benchmark of the Field convolution path on one MI355X.

Default workload: large fields, FP32, with the BF16-MFMA form beside it.
Batch 8 images 256x256 with 32 planes; one step runs, back to back on one stream:
  ProjectFrame + blockReduceSum(32), BorderClamp, 32 filters 7x7x32     -> 256x256x32 per image
  ProjectFrame + blockReduceSum(32), BorderClamp, 32 filters 9x9x32     -> 256x256x32 per image
  ProjectFrame + blockReduceSum(32), UpsampleInputConvolution(2), BorderZero, the 7x7 bank
                                                                          -> 512x512x32 per image
  PlaneByPlane 1024x1024 ScalarField (*) 11x11 Gaussian-plus-noise filter, BorderClamp
(the reference rejects up-sampling with BorderClamp, so the up-sampled op runs
BorderZero).  Inputs are seeded and resident in HBM before the timed region.

Prints one JSON line: `value` = FP32 GFLOP/s of the timed steps; `bf16` = the
same step with the BF16-MFMA plans where they are offered (the two Clamp
ProjectFrames and the up-sampled ProjectFrame; PlaneByPlane stays FP32);
per-op medians of >= 20 launches, each
after a 512 MiB cache-flush write, timed with HIP events on the op's stream;
the roofline of the dominant FP32 kernel; the CPU baseline (the oracle's FP32
restatement, OpenMP, on a bounded sample of the same workload, on every host
core the process may use and on one core); and the CNN-layer triple's per-op
medians as extra keys.

Other workloads: --workload cnn-triple | precision | serving.
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
METRIC = "conv GFLOP/s & % roofline (FP32 vs BF16-MFMA); p50/p99 request latency"
FLUSH_BYTES = 512 << 20    # beyond the 256 MB last-level cache


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cores():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


# ---------------------------------------------------------------- inputs

def large_inputs(seed=4, B=8, N=32, K=32, H=256):
    """x: B*N planes of HxH uniform[-1,1); 7x7 and 9x9 He-normal banks of K
    filters over N planes; the 1024x1024 image and its 11x11 filter (a
    sigma-2 Gaussian plus N(0, 0.01) noise)."""
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, (B * N, H, H)).astype(np.float32)
    banks = {F: (rng.standard_normal((K * N, F, F)) * np.sqrt(2.0 / (N * F * F))).astype(np.float32) for F in (7, 9)}
    yy, xx = np.mgrid[-5:6, -5:6]
    gauss = np.exp(-(yy ** 2 + xx ** 2) / (2 * 2.0 ** 2))
    pbp_filter = (gauss / gauss.sum() + rng.normal(0, 0.01, gauss.shape)).astype(np.float32)[None]
    img = rng.uniform(0, 1, (1, 1024, 1024)).astype(np.float32)
    return x, banks, img, pbp_filter


def large_ops(L, precision, B=8, N=32, K=32, H=256):
    """(name, desc, operand-1 key, operand-2 key) of one step of the large-field
    workload.  BF16 is offered for the Clamp and the up-sampled ProjectFrames;
    PlaneByPlane runs FP32 in both steps."""
    pn = {L.FP32: "fp32", L.BF16: "bf16"}[precision]
    ops = []
    for F in (7, 9):
        ops.append(("projectFrame+BRS clamp %dx%d %s" % (F, F, pn),
                    L.make_desc(L.PROJECT_FRAME_BLOCK_REDUCE_SUM, L.BORDER_CLAMP, L.CROSS_CORRELATION,
                                L.NO_SAMPLING, 1, B, (2, H, H, B * N), (2, F, F, K * N), precision), "x", F))
    ops.append(("projectFrame+BRS up2 zero 7x7 %s" % pn,
                L.make_desc(L.PROJECT_FRAME_BLOCK_REDUCE_SUM, L.BORDER_ZERO, L.CROSS_CORRELATION,
                            L.UPSAMPLE_INPUT, 2, B, (2, H, H, B * N), (2, 7, 7, K * N), precision), "x", 7))
    ops.append(("planeByPlane clamp 1024x1024 11x11 fp32",
                L.make_desc(L.PLANE_BY_PLANE, L.BORDER_CLAMP, L.CONVOLUTION, L.NO_SAMPLING, 1, 1,
                            (2, 1024, 1024, 1), (2, 11, 11, 1)), "img", "pbp"))
    return ops


def triple_inputs(seed=16, B=16, N=16, K=32, H=64):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, (B * N, H, H)).astype(np.float32)
    w = (rng.standard_normal((K * N, 3, 3)) * np.sqrt(2.0 / (N * 9))).astype(np.float32)
    dy = rng.uniform(-1, 1, (B * K, H - 2, H - 2)).astype(np.float32)
    dyf = rng.uniform(-1, 1, (B * K, H, H)).astype(np.float32)
    return x, w, dy, dyf


def triple_descs(L, B=16, N=16, K=32, H=64, precision=0, extras=False):
    """The CNN-layer triple (Valid layer): PF+BRS Valid, BPF+BRS Full (its
    adjoint), FA+BRS Valid (the weight gradient).  With `extras`, also PF+BRS
    with BorderCyclic and with DownsampleOutputConvolution(2) (BorderZero)."""
    f = lambda planes, s: (2, s, s, planes)  # noqa: E731
    ops = [("projectFrame+BRS", L.make_desc(L.PROJECT_FRAME_BLOCK_REDUCE_SUM, L.BORDER_VALID, L.CROSS_CORRELATION,
                                            L.NO_SAMPLING, 1, B, f(B * N, H), f(K * N, 3), precision), "x", "w"),
           ("backProjectFrame+BRS", L.make_desc(L.BACK_PROJECT_FRAME_BLOCK_REDUCE_SUM, L.BORDER_FULL, L.CONVOLUTION,
                                                L.NO_SAMPLING, 1, B, f(B * K, H - 2), f(K * N, 3), precision),
            "dy", "w"),
           ("filterAdjoint+BRS", L.make_desc(L.FILTER_ADJOINT_BLOCK_REDUCE_SUM, L.BORDER_VALID, L.CROSS_CORRELATION,
                                             L.NO_SAMPLING, 1, B, f(B * N, H), f(B * K, H - 2), precision),
            "x", "dy")]
    if extras:
        ops.append(("projectFrame+BRS cyclic", L.make_desc(
            L.PROJECT_FRAME_BLOCK_REDUCE_SUM, L.BORDER_CYCLIC, L.CROSS_CORRELATION, L.NO_SAMPLING, 1, B,
            f(B * N, H), f(K * N, 3), precision), "x", "w"))
        ops.append(("projectFrame+BRS down2", L.make_desc(
            L.PROJECT_FRAME_BLOCK_REDUCE_SUM, L.BORDER_ZERO, L.CROSS_CORRELATION, L.DOWNSAMPLE_OUTPUT, 2, B,
            f(B * N, H), f(K * N, 3), precision), "x", "w"))
    return ops


# ---------------------------------------------------------------- timing

def tuned_plan(L, desc, args, flush_bytes=0):
    """The plan for `desc`, with its kernel configuration chosen by the
    autotuner (cogconv_autotune) unless --no-tune.  Candidates are timed with
    warm caches (flush_bytes=0), as in the timed steps, where the inputs are
    reused every step, or after a cache flush of flush_bytes per launch."""
    if not args.no_tune:
        if args.tune_cache:
            L.tune_cache_load(args.tune_cache)
        r = L.autotune(desc, warmup_steps=5, run_steps=60 if not flush_bytes else 20, flush_bytes=flush_bytes)
        log("tuned:", r["description"], "%.2f us" % r["avg_us"], "(cached)" if r["from_cache"] else
            "(%d candidates)" % r["variants"])
        if args.tune_cache:
            L.tune_cache_save(args.tune_cache)
    return L.Plan(desc)


class Op:
    """One planned op with its buffers."""

    def __init__(self, L, name, plan, in0, in1):
        self.name, self.plan, self.in0, self.in1 = name, plan, in0, in1
        self.out = L.DeviceBuffer(4 * int(np.prod(plan.out_shape)))
        self.ws = L.DeviceBuffer(plan.workspace_bytes) if plan.workspace_bytes else None

    def launch(self, stream):
        self.plan.launch(self.in0, self.in1, self.out, self.ws, stream)

    def close(self):
        self.out.close()
        if self.ws:
            self.ws.close()


def flushed_median_ms(L, op, stream, flush, reps=20, warm=2):
    """Median over `reps` launches of `op`, each timed alone with HIP events on
    `stream` after a write of the whole flush buffer (so no operand is left in
    the L2 or the last-level cache)."""
    for _ in range(warm):
        op.launch(stream)
    e0, e1 = L.Event(), L.Event()
    times = []
    for _ in range(reps):
        flush.zero(stream)
        e0.record(stream)
        op.launch(stream)
        e1.record(stream)
        e1.synchronize()
        times.append(e0.elapsed_ms(e1))
    e0.close()
    e1.close()
    return statistics.median(times), min(times), max(times)


def read_flushed_median_ms(L, op, stream, reps=20):
    """Median over `reps` launches of `op`, each timed alone after a read of
    FLUSH_BYTES (a torch sum over a resident buffer): the operands are gone
    from the L2 and the last-level cache as after a write flush, but the
    evicted lines are clean, so the op's own loads do not also pay for writing
    the flush's dirty lines back to HBM."""
    import torch
    buf = torch.ones(FLUSH_BYTES // 4, dtype=torch.float32, device="cuda")
    e0, e1 = L.Event(), L.Event()
    times = []
    for _ in range(reps):
        buf.sum()
        torch.cuda.synchronize()
        e0.record(stream)
        op.launch(stream)
        e1.record(stream)
        e1.synchronize()
        times.append(e0.elapsed_ms(e1))
    e0.close()
    e1.close()
    del buf
    return statistics.median(times)


def timed_steps(L, ops, stream, steps, warmup):
    """`warmup` untimed steps, then exactly `steps` steps (every op once, back
    to back on `stream`) bracketed by device synchronisations; returns ms per
    step (the larger of the event and wall-clock spans)."""
    for _ in range(warmup):
        for op in ops:
            op.launch(stream)
    L.synchronize()
    t0, t1 = L.Event(), L.Event()
    wall0 = time.perf_counter()
    t0.record(stream)
    for _ in range(steps):
        for op in ops:
            op.launch(stream)
    t1.record(stream)
    t1.synchronize()
    L.synchronize()
    wall = time.perf_counter() - wall0
    ms = max(t0.elapsed_ms(t1), wall * 1e3) / steps
    t0.close()
    t1.close()
    return ms


def roofline(plan, ms, peak_tflops, name, reps, traffic_file):
    """Roofline object for one launch of `plan` taking `ms`: compute-bound when
    its algorithmic FLOP/byte is above the ridge of `peak_tflops` and HBM."""
    ai = plan.flops / plan.bytes
    ridge = peak_tflops * 1e12 / (PEAK_HBM_GBS * 1e9)
    if ai >= ridge:
        achieved = plan.flops / (ms * 1e-3) / 1e12
        roof = {"bound": "mfma", "achieved": round(achieved, 3), "peak": peak_tflops, "unit": "TFLOP/s",
                "frac": round(achieved / peak_tflops, 4), "traffic": None}
    else:
        achieved = plan.bytes / (ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": None}
    roof["kernel"] = name + ": " + plan.description
    path = os.path.join(ROOT, "profiles", traffic_file)
    if os.path.exists(path):
        with open(path) as fh:
            t = json.load(fh).get(name)
        if t:
            roof["traffic"] = t["hbm_bytes_per_launch"]
            roof["traffic_source"] = ("profiles/%s (rocprofv3 FETCH_SIZE + WRITE_SIZE per dispatch, "
                                      "copy-calibrated)" % traffic_file)
    roof["algorithmic_flops"] = plan.flops
    roof["algorithmic_bytes"] = plan.bytes
    roof["arithmetic_intensity"] = round(ai, 1)
    roof["avg_launch_ms"] = round(ms, 5)
    roof["timing"] = ("median of %d launches of this op alone on its stream, HIP events, each after a "
                      "512 MiB cache-flush write" % reps)
    return roof


def copy_calibration(L):
    """Known-size reference for the HBM counters: fill 1 GiB, then copy its
    first 64 MiB (long evicted from the 256 MB last-level cache by the fill)
    three times."""
    src, dst = L.DeviceBuffer(1 << 30), L.DeviceBuffer(64 << 20)
    src.zero()
    L.synchronize()
    for _ in range(3):
        L.check(L.lib().cogconv_d2d_async(dst.ptr, src.ptr, 64 << 20, None))
    L.synchronize()
    src.close()
    dst.close()


# ---------------------------------------------------------------- CPU baseline

def cpu_baseline_large(budget_s=20.0):
    """The oracle's FP32 restatement (the reference's per-output accumulation
    order), OpenMP over output rows, on a bounded sample of the large-field
    step: image 0 of the batch for the 7x7 / 9x9 Clamp and the up-sampled
    ProjectFrames plus the whole PlaneByPlane, on every core the process may
    use; and on one core, image 0 with the first 2 of the 32 filters.  Median
    GFLOP/s (algorithmic FLOPs, as on the GPU) over >= 5 repetitions after one
    warm-up."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    x, banks, img, pbp = large_inputs()
    N, H = 32, 256
    x0 = x[:N]

    def sample(k):
        ops, flops = [], 0.0
        for F in (7, 9):
            w = banks[F][:k * N]
            ops.append((O.make_desc(O.PF_BRS, O.CLAMP, O.XCORR, O.NONE, 1, 1, x0.shape, w.shape), x0, w))
            flops += 2.0 * k * N * H * H * F * F
        w7 = banks[7][:k * N]
        ops.append((O.make_desc(O.PF_BRS, O.ZERO, O.XCORR, O.UP, 2, 1, x0.shape, w7.shape), x0, w7))
        flops += 2.0 * k * N * (2 * H) * (2 * H) * 49 / 4.0
        ops.append((O.make_desc(O.PBP, O.CLAMP, O.CONV, O.NONE, 1, 1, img.shape, pbp.shape), img, pbp))
        flops += 2.0 * 1024 * 1024 * 121
        return ops, flops

    def run(ops, flops, threads, budget):
        for d, a, k in ops:
            O.conv(d, a, k, prec="f32", threads=threads)
        times = []
        t_all = time.perf_counter()
        while len(times) < 5 or (time.perf_counter() - t_all < budget and len(times) < 20):
            t0 = time.perf_counter()
            for d, a, k in ops:
                O.conv(d, a, k, prec="f32", threads=threads)
            times.append(time.perf_counter() - t0)
        return flops / statistics.median(times) / 1e9, len(times)

    cores = host_cores()
    ops, flops = sample(32)
    v_all, n_all = run(ops, flops, cores, budget_s / 2)
    ops1, flops1 = sample(2)
    v_one, n_one = run(ops1, flops1, 1, budget_s / 2)
    return {"value": round(v_all, 3), "unit": "GFLOP/s", "cores": cores, "kind": "port",
            "sample": "image 0 of 8 of the FP32 step (7x7 and 9x9 Clamp PF+BRS, Up2 7x7 PF+BRS, and the whole "
                      "PlaneByPlane; %.3g GFLOP), oracle FP32 restatement, OpenMP over (output plane, 8-row band) on every core "
                      "in the process's affinity mask, median of %d reps" % (flops / 1e9, n_all),
            "one_core": {"value": round(v_one, 3), "unit": "GFLOP/s", "cores": 1,
                         "sample": "image 0, filters 0-1 of 32 of the same ops (%.3g GFLOP), median of %d reps"
                                   % (flops1 / 1e9, n_one)},
            "note": "the oracle materialises the zero-inserted operand of the up-sampled op (4x the multiply-adds "
                    "of the reference's kernel, which skips zero taps); GFLOP/s counts algorithmic FLOPs"}


# ---------------------------------------------------------------- PMC summaries

def pmc_traffic(fetch_csv, write_csv, calib_kernel="__amd_rocclr_copyBuffer", calib_bytes=64 << 20):
    """HBM bytes per dispatch per kernel from two rocprofv3 counter-collection
    CSVs (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch), each scaled by what
    the same counter reads for the 64 MiB copy of --copy-calibration."""
    import csv

    def read(path, counter):
        per = {}
        with open(path) as fh:
            for row in csv.DictReader(fh):
                if row.get("Counter_Name") != counter:
                    continue
                per.setdefault(row["Kernel_Name"], []).append(float(row["Counter_Value"]))
        return per
    fetch, write = read(fetch_csv, "FETCH_SIZE"), read(write_csv, "WRITE_SIZE")
    cal = [k for k in fetch if calib_kernel in k]
    if not cal:
        raise SystemExit("no copy kernel %r in %s (run the profiled command with --copy-calibration)"
                         % (calib_kernel, fetch_csv))
    # The calibration copy reads and writes calib_bytes; the median dispatch of the copy kernel.
    ff = calib_bytes / 1024.0 / statistics.median(fetch[cal[0]])
    wf = calib_bytes / 1024.0 / statistics.median(write[[k for k in write if calib_kernel in k][0]])
    out = {"_calibration": {"fetch_factor": round(ff, 4), "write_factor": round(wf, 4),
                            "method": "64 MiB device copy from a cold 1 GiB source, bench.py --copy-calibration"}}
    for k in fetch:
        if calib_kernel in k or k not in write:
            continue
        fk = statistics.median(fetch[k])
        wk = statistics.median(write[k])
        out[k] = {"fetch_kib": round(fk, 1), "write_kib": round(wk, 1),
                  "hbm_bytes_per_dispatch": int(round(1024.0 * (fk * ff + wk * wf)))}
    return out
def trace_summary(trace_csv):
    """Per-kernel dispatch durations from a rocprofv3 --kernel-trace CSV, split
    by grid size (a full-batch launch and a one-image check launch of one
    kernel are different work): calls, mean, median, min and max in us."""
    import csv
    import re
    groups = {}
    with open(trace_csv) as fh:
        for row in csv.DictReader(fh):
            name = row["Kernel_Name"]
            m = re.search(r"(\w+(?:<[^()]*>)?)\(", name)
            short = m.group(1) if m else name.split("(")[0]
            short = short.replace("cogconv::", "")
            grid = "x".join(row.get("Grid_Size_%s" % a, "?") for a in "XYZ")
            us = (int(row["End_Timestamp"]) - int(row["Start_Timestamp"])) / 1e3
            groups.setdefault((short, grid), []).append(us)
    out = []
    for (short, grid), v in sorted(groups.items(), key=lambda kv: -sum(kv[1])):
        out.append({"kernel": short, "grid": grid, "calls": len(v), "mean_us": round(statistics.mean(v), 2),
                    "median_us": round(statistics.median(v), 2), "min_us": round(min(v), 2),
                    "max_us": round(max(v), 2), "total_us": round(sum(v), 1)})
    return out


def precision_workload(args):
    """B32 128x128x64 planes, 64 filters 3x3x64: ProjectFrame+BRS and
    BackProjectFrame+BRS (BorderZero) and FilterAdjoint+BRS (BorderValid, dy
    126x126x64) in FP32, BF16, FP16 and split-BF16; per-op medians of
    --op-reps launches, each after a 512 MiB cache-flush write, HIP events;
    errors against the FP64 oracle on a 2-image slice (uniform[-1,1) and
    wide-range 10^U(-3,3) with random signs)."""
    import torch  # noqa: F401  (for the read flush; imported before the library loads the HIP runtime)
    import cogconv._lib as L
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    L.init(0)
    rng = np.random.default_rng(3)
    B, N, K, H = 32, 64, 64, 128
    x = rng.uniform(-1, 1, (B * N, H, H)).astype(np.float32)
    w = rng.uniform(-1, 1, (K * N, 3, 3)).astype(np.float32)
    dy = rng.uniform(-1, 1, (B * K, H, H)).astype(np.float32)
    dyv = rng.uniform(-1, 1, (B * K, H - 2, H - 2)).astype(np.float32)  # output gradient of a Valid layer
    bx, bw, bdy, bdyv = (L.DeviceBuffer.from_array(a) for a in (x, w, dy, dyv))
    f = lambda planes, side=H: (2, side, side, planes)  # noqa: E731
    stream = L.Stream()
    flush = L.DeviceBuffer(FLUSH_BYTES)
    names = {L.FP32: "fp32", L.BF16: "bf16", L.FP16: "fp16", L.SPLIT_BF16: "split_bf16"}
    res = {}
    for prec in (L.FP32, L.BF16, L.FP16, L.SPLIT_BF16):
        for opname, mode, border, orient, a, ab, t1, b1 in (
                ("projectFrame+BRS", L.PROJECT_FRAME_BLOCK_REDUCE_SUM, L.BORDER_ZERO, L.CROSS_CORRELATION, x, bx,
                 (2, 3, 3, K * N), bw),
                ("backProjectFrame+BRS", L.BACK_PROJECT_FRAME_BLOCK_REDUCE_SUM, L.BORDER_ZERO, L.CONVOLUTION, dy, bdy,
                 (2, 3, 3, K * N), bw),
                ("filterAdjoint+BRS", L.FILTER_ADJOINT_BLOCK_REDUCE_SUM, L.BORDER_VALID, L.CROSS_CORRELATION, x, bx,
                 f(B * K, H - 2), bdyv)):
            # Candidates timed after a cache flush: this workload compares the ops'
            # flushed medians.
            plan = tuned_plan(L, L.make_desc(mode, border, orient, L.NO_SAMPLING, 1, B, f(a.shape[0]), t1, prec),
                              args, flush_bytes=FLUSH_BYTES)
            op = Op(L, opname, plan, ab, b1)
            ms, lo, hi = flushed_median_ms(L, op, stream, flush, reps=args.op_reps)
            rms = read_flushed_median_ms(L, op, stream, reps=args.op_reps)
            warm = timed_steps(L, [op], stream, 20, 3)  # back to back, operands left in the caches
            res.setdefault(names[prec], {})[opname] = {
                "ms": round(ms, 4), "min_ms": round(lo, 4), "max_ms": round(hi, 4),
                "gflops": round(plan.flops / (ms * 1e-3) / 1e9, 1),
                "warm_ms": round(warm, 4), "warm_gflops": round(plan.flops / (warm * 1e-3) / 1e9, 1),
                "read_flushed_ms": round(rms, 4), "kernel": plan.description}
            log(names[prec], opname, plan.description, "%.4f ms flushed, %.4f ms read-flushed, %.4f ms warm" %
                (ms, rms, warm))
            op.close()
            plan.close()
    # Errors on a 2-image slice, both input sets.
    errs = {}
    for label, gen in (("uniform", lambda sh: rng.uniform(-1, 1, sh)),
                       ("wide", lambda sh: 10.0 ** rng.uniform(-3, 3, sh) * rng.choice([-1.0, 1.0], sh))):
        xs = gen((2 * N, H, H)).astype(np.float32)
        ws_ = gen((K * N, 3, 3)).astype(np.float32)
        dys = gen((2 * K, H - 2, H - 2)).astype(np.float32)
        od = O.make_desc(O.PF_BRS, O.ZERO, O.XCORR, O.NONE, 1, 2, xs.shape, ws_.shape)
        ref = O.conv(od, xs, ws_, threads=16)
        odf = O.make_desc(O.FA_BRS, O.VALID, O.XCORR, O.NONE, 1, 2, xs.shape, dys.shape)
        reff = O.conv(odf, xs, dys, threads=16)
        for prec in (L.FP32, L.BF16, L.FP16, L.SPLIT_BF16):
            for opname, mode, border, in1, r in (
                    ("projectFrame+BRS", L.PROJECT_FRAME_BLOCK_REDUCE_SUM, L.BORDER_ZERO, ws_, ref),
                    ("filterAdjoint+BRS", L.FILTER_ADJOINT_BLOCK_REDUCE_SUM, L.BORDER_VALID, dys, reff)):
                plan = L.Plan(L.make_desc(mode, border, L.CROSS_CORRELATION, L.NO_SAMPLING, 1, 2, f(2 * N),
                                          (2, in1.shape[1], in1.shape[2], in1.shape[0]), prec))
                a, k = L.DeviceBuffer.from_array(xs), L.DeviceBuffer.from_array(in1)
                out = L.DeviceBuffer(4 * int(np.prod(plan.out_shape)))
                wsb = L.DeviceBuffer(plan.workspace_bytes) if plan.workspace_bytes else None
                plan.launch(a, k, out, wsb, None)
                L.synchronize()
                got = out.download(plan.out_shape).astype(np.float64)
                d = np.abs(got - r)
                errs.setdefault(names[prec], {}).setdefault(label, {})[opname] = {
                    "max_abs": float(d.max()), "max_rel_normwise": float(d.max() / np.abs(r).max()),
                    "rel_l2": float(np.sqrt((d ** 2).sum() / (r ** 2).sum()))}
                for b_ in (a, k, out, wsb):
                    if b_ is not None:
                        b_.close()
    fp32 = res["fp32"]["projectFrame+BRS"]["gflops"]
    line = {"metric": "conv GFLOP/s & % roofline (FP32 vs BF16-MFMA); p50/p99 request latency",
            "workload": "precision variants: batch 32 images 128x128x64, 64 filters 3x3x64; PF/BPF BorderZero, "
                        "FA BorderValid (dy 126x126x64)",
            "results": res, "errors_vs_fp64": errs,
            "bf16_over_fp32_projectFrame": round(res["bf16"]["projectFrame+BRS"]["gflops"] / fp32, 2),
            "bf16_over_fp32_filterAdjoint": round(res["bf16"]["filterAdjoint+BRS"]["gflops"] /
                                                  res["fp32"]["filterAdjoint+BRS"]["gflops"], 2),
            "warm_bf16_over_fp32_projectFrame": round(res["bf16"]["projectFrame+BRS"]["warm_gflops"] /
                                                      res["fp32"]["projectFrame+BRS"]["warm_gflops"], 2),
            "warm_bf16_over_fp32_filterAdjoint": round(res["bf16"]["filterAdjoint+BRS"]["warm_gflops"] /
                                                       res["fp32"]["filterAdjoint+BRS"]["warm_gflops"], 2),
            "timing": "ms: median of --op-reps launches, each after a 512 MiB cache-flush write; read_flushed_ms: "
                      "the same after a 512 MiB read instead (caches emptied of the operands but left clean, so "
                      "the op does not pay for the flush's write-back); warm_ms: mean of 20 back-to-back launches "
                      "after 3 (operands resident in the last-level cache)",
            "peaks": {"fp32_tflops": PEAK_FP32_TFLOPS, "bf16_tflops": PEAK_BF16_TFLOPS, "hbm_gbs": PEAK_HBM_GBS}}
    print(json.dumps(line), flush=True)


def large_errors(L, x, banks, N=32, H=256):
    """Normwise and rel-L2 errors of the Clamp and the up-sampled ProjectFrames
    (FP32 and BF16) against the FP64 oracle on image 0."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    errs = {}
    xs = x[:N]
    a = L.DeviceBuffer.from_array(xs)
    for F in (7, 9):
        od = O.make_desc(O.PF_BRS, O.CLAMP, O.XCORR, O.NONE, 1, 1, xs.shape, banks[F].shape)
        ref = O.conv(od, xs, banks[F], threads=host_cores())
        wb = L.DeviceBuffer.from_array(banks[F])
        for prec, pn in ((L.FP32, "fp32"), (L.BF16, "bf16")):
            plan = L.Plan(L.make_desc(L.PROJECT_FRAME_BLOCK_REDUCE_SUM, L.BORDER_CLAMP, L.CROSS_CORRELATION,
                                      L.NO_SAMPLING, 1, 1, (2, H, H, N), (2, F, F, banks[F].shape[0]), prec))
            op = Op(L, "", plan, a, wb)
            op.launch(None)
            L.synchronize()
            got = op.out.download(plan.out_shape).astype(np.float64)
            d = np.abs(got - ref)
            errs["%dx%d %s" % (F, F, pn)] = {"max_abs": float(d.max()),
                                           "max_rel_normwise": float(d.max() / np.abs(ref).max()),
                                           "rel_l2": float(np.sqrt((d ** 2).sum() / (ref ** 2).sum()))}
            op.close()
            plan.close()
        wb.close()
    # The up-sampled 7x7 ProjectFrame (BorderZero) on image 0.
    od = O.make_desc(O.PF_BRS, O.ZERO, O.XCORR, O.UP, 2, 1, xs.shape, banks[7].shape)
    ref = O.conv(od, xs, banks[7], threads=host_cores())
    wb = L.DeviceBuffer.from_array(banks[7])
    for prec, pn in ((L.FP32, "fp32"), (L.BF16, "bf16")):
        plan = L.Plan(L.make_desc(L.PROJECT_FRAME_BLOCK_REDUCE_SUM, L.BORDER_ZERO, L.CROSS_CORRELATION,
                                  L.UPSAMPLE_INPUT, 2, 1, (2, H, H, N), (2, 7, 7, banks[7].shape[0]), prec))
        op = Op(L, "", plan, a, wb)
        op.launch(None)
        L.synchronize()
        got = op.out.download(plan.out_shape).astype(np.float64)
        d = np.abs(got - ref)
        errs["up2 7x7 %s" % pn] = {"max_abs": float(d.max()), "max_rel_normwise": float(d.max() / np.abs(ref).max()),
                                   "rel_l2": float(np.sqrt((d ** 2).sum() / (ref ** 2).sum()))}
        op.close()
        plan.close()
    wb.close()
    a.close()
    return errs


def triple_extras(L, args, stream, flush):
    """Per-op flushed medians of the FP32 CNN-layer triple (batch 16 images
    64x64x16, 32 filters 3x3x16) and of its BorderCyclic and Down2
    ProjectFrames."""
    x, w, dy, _ = triple_inputs()
    bufs = {"x": L.DeviceBuffer.from_array(x), "w": L.DeviceBuffer.from_array(w), "dy": L.DeviceBuffer.from_array(dy)}
    res = {}
    for name, desc, k0, k1 in triple_descs(L, extras=True):
        plan = tuned_plan(L, desc, args)
        op = Op(L, name, plan, bufs[k0], bufs[k1])
        med, lo, hi = flushed_median_ms(L, op, stream, flush)
        res[name] = {"ms": round(med, 5), "min_ms": round(lo, 5), "max_ms": round(hi, 5),
                     "gflops": round(plan.flops / (med * 1e-3) / 1e9, 1),
                     "frac_fp32_peak": round(plan.flops / (med * 1e-3) / 1e12 / PEAK_FP32_TFLOPS, 4),
                     "kernel": plan.description}
        log("cnn triple", name, plan.description, "%.4f ms" % med)
        op.close()
        plan.close()
    for b in bufs.values():
        b.close()
    return res


def large_workload(args):
    import cogconv._lib as L
    L.init(0)
    props = L.device_props(0)
    log("device:", props)
    if args.copy_calibration:
        copy_calibration(L)
    x, banks, img, pbp = large_inputs()
    bufs = {"x": L.DeviceBuffer.from_array(x), 7: L.DeviceBuffer.from_array(banks[7]),
            9: L.DeviceBuffer.from_array(banks[9]), "img": L.DeviceBuffer.from_array(img),
            "pbp": L.DeviceBuffer.from_array(pbp)}
    stream = L.Stream()
    flush = L.DeviceBuffer(FLUSH_BYTES)
    steps = {}
    for prec, pn in ((L.FP32, "fp32"), (L.BF16, "bf16")):
        ops = []
        for name, desc, k0, k1 in large_ops(L, prec):
            plan = tuned_plan(L, desc, args)
            ops.append(Op(L, name, plan, bufs[k0], bufs[k1]))
            log("plan:", name, plan.description, "GFLOP %.3f" % (plan.flops / 1e9))
        ms = timed_steps(L, ops, stream, args.steps, args.warmup)
        per_op = {}
        for op in ops:
            med, lo, hi = flushed_median_ms(L, op, stream, flush, reps=args.op_reps)
            per_op[op.name] = {"ms": round(med, 5), "min_ms": round(lo, 5), "max_ms": round(hi, 5),
                               "gflops": round(op.plan.flops / (med * 1e-3) / 1e9, 1),
                               "gflop": round(op.plan.flops / 1e9, 3), "kernel": op.plan.description}
            log(pn, op.name, op.plan.description, "%.4f ms" % med)
        steps[pn] = {"ops": ops, "ms_per_step": ms, "flops": sum(op.plan.flops for op in ops), "per_op": per_op}
    fp32, bf16 = steps["fp32"], steps["bf16"]
    value = fp32["flops"] / (fp32["ms_per_step"] * 1e-3) / 1e9
    bf16_value = bf16["flops"] / (bf16["ms_per_step"] * 1e-3) / 1e9
    # The dominant kernel: the FP32 op with the longest median launch.
    dom = max(fp32["ops"], key=lambda op: fp32["per_op"][op.name]["ms"])
    roof = roofline(dom.plan, fp32["per_op"][dom.name]["ms"], PEAK_FP32_TFLOPS, dom.name, args.op_reps,
                    "traffic_large.json")
    # BF16 beside it: the BF16 op with the longest median launch, against dense BF16 MFMA.
    bdom = max([op for op in bf16["ops"] if "bf16" in op.name], key=lambda op: bf16["per_op"][op.name]["ms"])
    broof = roofline(bdom.plan, bf16["per_op"][bdom.name]["ms"], PEAK_BF16_TFLOPS, bdom.name, args.op_reps,
                     "traffic_large.json")
    ratios = {}
    for F in (7, 9):
        a = fp32["per_op"]["projectFrame+BRS clamp %dx%d fp32" % (F, F)]["ms"]
        b = bf16["per_op"]["projectFrame+BRS clamp %dx%d bf16" % (F, F)]["ms"]
        ratios["clamp %dx%d" % (F, F)] = round(a / b, 2)
    ratios["up2 zero 7x7"] = round(fp32["per_op"]["projectFrame+BRS up2 zero 7x7 fp32"]["ms"] /
                                   bf16["per_op"]["projectFrame+BRS up2 zero 7x7 bf16"]["ms"], 2)
    errs = large_errors(L, x, banks)
    extras = None if args.no_extras else triple_extras(L, args, stream, flush)
    for st in steps.values():
        for op in st["ops"]:
            op.close()
            op.plan.close()
    cpu = None if args.no_cpu else cpu_baseline_large(budget_s=args.cpu_budget)
    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(fp32["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded uniform[-1,1) fields, He-normal filter banks, Gaussian-plus-noise 11x11 filter)",
        "kernel_choice": "heuristic (--no-tune)" if args.no_tune else "autotuned per op (cogconv_autotune)",
        "config": {"workload": "large fields: batch 8 images 256x256x32, 32 filters 7x7x32 and 9x9x32, "
                               "ProjectFrame+blockReduceSum BorderClamp; the 7x7 bank on the 2x up-sampled input "
                               "(BorderZero); PlaneByPlane 1024x1024 ScalarField with an 11x11 filter, BorderClamp",
                   "batch": 8, "image": [256, 256, 32], "filters": [[32, 7, 7, 32], [32, 9, 9, 32]],
                   "plane_by_plane": [1024, 1024, 11, 11], "flops_per_step": fp32["flops"],
                   "step": "the four ops back to back on one stream"},
        "per_op": fp32["per_op"],
        "bf16": {"value": round(bf16_value, 2), "unit": "GFLOP/s", "ms_per_step": round(bf16["ms_per_step"], 5),
                 "step": "the Clamp and the up-sampled ProjectFrames as BF16-MFMA plans (FP32 accumulation); "
                         "PlaneByPlane stays FP32",
                 "per_op": {k: v for k, v in bf16["per_op"].items() if "bf16" in k},
                 "bf16_over_fp32_per_op": ratios, "roofline": broof},
        "errors_vs_fp64_image0": errs,
        "roofline": roof,
        "cpu_baseline": cpu,
        "cnn_triple_per_op": extras,
        "peaks": {"fp32_tflops": PEAK_FP32_TFLOPS, "bf16_tflops": PEAK_BF16_TFLOPS, "hbm_gbs": PEAK_HBM_GBS},
        "device": props["name"] + " " + props["arch"],
    }
    flush.close()
    for b in bufs.values():
        b.close()
    print(json.dumps(line), flush=True)
RATE_SWEEP = (4000, 8000, 12000, 16000, 20000, 24000, 28000, 32000, 40000, 48000, 56000, 64000, 72000, 80000,
              96000, 112000, 128000, 144000, 160000)  # offered req/s, resident transport
P99_TARGETS_MS = (0.25, 0.5, 1.0, 2.0)


def serving_workload(args):
    """Poisson request stream (mean 2000 req/s, seeded) through the serving loop:
    field side uniform over {32, 48, ..., 256}, planes in {16, 32, 64}, 3x3 or 5x5
    filters with 16-64 outputs, 70/20/10 ProjectFrame / BackProjectFrame /
    FilterAdjoint (all with block reduction), 64 filter banks resident in HBM,
    an 8 GiB admission budget.  Run once with FP32 plans and once with BF16
    plans, open loop at the Poisson rate and closed loop (all requests queued
    at once) for capacity."""
    import cogconv._lib as L
    L.init(0)
    rng = np.random.default_rng(2000)
    n_req = args.requests
    sides = list(range(32, 257, 16))
    banks = []
    for _ in range(64):
        N = int(rng.choice([16, 32, 64]))
        K = int(rng.choice([16, 32, 48, 64]))
        F = int(rng.choice([3, 5]))
        banks.append((N, K, F, (rng.standard_normal((K * N, F, F)) * np.sqrt(2.0 / (N * F * F))).astype(np.float32)))
    kinds = rng.choice(["pf", "bpf", "fa"], size=n_req, p=[0.7, 0.2, 0.1])
    bank_of = rng.integers(0, 64, size=n_req)
    side_of = rng.choice(sides, size=n_req)
    gaps = rng.exponential(1.0 / 2000.0, size=n_req)
    arrivals = np.cumsum(gaps)
    SLOTS = 48
    maxe = 64 * 256 * 256
    pool = [(L.HostBuffer((maxe,)), L.HostBuffer((maxe,)), L.HostBuffer((maxe,))) for _ in range(SLOTS)]
    for a, b, _ in pool:
        a.array[:] = rng.uniform(-1, 1, maxe).astype(np.float32)
        b.array[:] = rng.uniform(-1, 1, maxe).astype(np.float32)

    def desc_for(i, precision):
        N, K, F, _ = banks[bank_of[i]]
        s = int(side_of[i])
        k = kinds[i]
        if k == "pf":
            return L.make_desc(L.PROJECT_FRAME_BLOCK_REDUCE_SUM, L.BORDER_ZERO, L.CROSS_CORRELATION, L.NO_SAMPLING, 1,
                               1, (2, s, s, N), (2, F, F, K * N), precision), True, N * s * s, K * s * s
        if k == "bpf":
            return L.make_desc(L.BACK_PROJECT_FRAME_BLOCK_REDUCE_SUM, L.BORDER_ZERO, L.CONVOLUTION, L.NO_SAMPLING, 1,
                               1, (2, s, s, K), (2, F, F, K * N), precision), True, K * s * s, N * s * s
        d = s - F + 1
        return L.make_desc(L.FILTER_ADJOINT_BLOCK_REDUCE_SUM, L.BORDER_VALID, L.CROSS_CORRELATION, L.NO_SAMPLING, 1,
                           1, (2, s, s, N), (2, d, d, K), precision), False, N * s * s, K * N * F * F

    # Resident slots: field state already in HBM (graph state), so requests move no data over PCIe.
    dpool = [(L.DeviceBuffer(4 * maxe), L.DeviceBuffer(4 * maxe), L.DeviceBuffer(4 * maxe)) for _ in range(SLOTS)]
    for (a, b, _), (da, db, _) in zip(pool, dpool):
        da.upload(a.array)
        db.upload(b.array)

    tuned = None
    if args.serving_tune:
        # Time every distinct descriptor of the trace (both precisions) once, so the
        # server's plans come from the autotuning cache (a short sample per candidate).
        t0 = time.perf_counter()
        if args.tune_cache and os.path.exists(args.tune_cache):
            L.tune_cache_load(args.tune_cache)
        seen = {}
        for precision in (L.FP32, L.BF16):
            for i in range(n_req):
                d = desc_for(i, precision)[0]
                key = bytes(d)
                if key not in seen:
                    seen[key] = d
        for d in seen.values():
            L.autotune(d, warmup_steps=1, run_steps=3, flush_bytes=0)
        if args.tune_cache:
            L.tune_cache_save(args.tune_cache)
        tuned = {"descriptors": len(seen), "seconds": round(time.perf_counter() - t0, 1)}
        log("serving: tuned", tuned)
    results = {}
    only_prec, only_run = (args.serving_only.split(":") + [""])[:2] if args.serving_only else (None, None)
    for transport in ("resident", "host"):
        if only_prec and transport != "resident":
            continue
        for label, precision in (("fp32", L.FP32), ("bf16", L.BF16)):
            if only_prec and label != only_prec:
                continue
            srv = L.Server(budget_bytes=8 << 30, max_inflight=16, streams=4, max_stack=8, graph_after=args.graph_after)
            for N, K, F, w in banks:
                srv.register_bank(w)
            reqs = []
            for i in range(n_req):
                desc, banked, n0, nout = desc_for(i, precision)
                sl = i % SLOTS
                if transport == "host":
                    a, b, o = (h.ptr for h in pool[sl])
                else:
                    a, b, o = (d.ptr for d in dpool[sl])
                reqs.append(L.Server.request(desc, a, o, bank=int(bank_of[i]) if banked else -1,
                                             in1=None if banked else b, resident=transport == "resident"))
            runs = [("open_loop_2000_per_s", arrivals), ("closed_loop", np.zeros(n_req))]
            if transport == "resident":
                # Offered-load sweep: the same request sequence at higher Poisson rates.
                for rate in RATE_SWEEP:
                    runs.append(("open_loop_%d_per_s" % rate, arrivals * (2000.0 / rate)))
            if only_run:
                # A closed-loop pass first, unrecorded: plans made, banks packed, as in the full run.
                srv.submit_timed(reqs, np.zeros(n_req))
                srv.drain()
                want = "closed_loop" if only_run == "closed" else "open_loop_%s_per_s" % only_run
                runs = [(k, a) for k, a in runs if k == want]
            for mode_label, arr in runs:
                srv.reset_stats()
                # Native submission at the trace's arrival times (all at once for the closed loop).
                # Requests reuse SLOTS buffer sets; concurrent requests on one slot race on
                # its output, which this throughput measurement does not read.
                srv.submit_timed(reqs, arr)
                srv.drain()
                st = srv.stats()
                r = {"p50_ms": round(st["p50_ms"], 3), "p99_ms": round(st["p99_ms"], 3),
                     "mean_ms": round(st["mean_ms"], 3), "requests_per_s": round(st["requests_per_s"], 1),
                     "completed": st["completed"],
                     "gflops": round(st["gflop"] / max(1e-9, st["last_complete_s"] - st["first_submit_s"]), 1),
                     "launches": st["launches"], "stacked_requests": st["stacked"],
                     "graph_launches": st["graph_launches"],
                     "peak_device_bytes": st["peak_bytes"]}
                results.setdefault(transport, {}).setdefault(label, {})[mode_label] = r
                log(transport, label, mode_label, r)
            srv.close()
    if only_prec:
        print(json.dumps({"workload": "serving (one run)", "run": args.serving_only, "results": results}), flush=True)
        return
    res = results["resident"]
    cap = res["bf16"]["closed_loop"]["requests_per_s"] / max(1e-9, res["fp32"]["closed_loop"]["requests_per_s"])

    def sustained(label, p99_ms):
        """Highest offered rate whose p99 stays within p99_ms while keeping up (>= 95% of the offered rate)."""
        best = 0
        for rate in RATE_SWEEP:
            r = res[label]["open_loop_%d_per_s" % rate]
            if r["p99_ms"] <= p99_ms and r["requests_per_s"] >= 0.95 * rate:
                best = rate
        return best
    at_equal_p99 = {}
    for p99_ms in P99_TARGETS_MS:
        f, b = sustained("fp32", p99_ms), sustained("bf16", p99_ms)
        at_equal_p99["p99<=%gms" % p99_ms] = {"fp32_req_per_s": f, "bf16_req_per_s": b,
                                              "ratio": round(b / f, 2) if f else None}
    line = {"metric": "conv GFLOP/s & % roofline (FP32 vs BF16-MFMA); p50/p99 request latency",
            "workload": "serving: Poisson 2000 req/s, sides 32..256, planes 16/32/64, 3x3/5x5 filters with 16-64 "
                        "outputs, 70/20/10 PF/BPF/FA, 64 resident banks, 8 GiB admission budget",
            "requests": n_req, "results": results, "bf16_over_fp32_capacity_resident": round(cap, 2),
            "plans": "autotuned per descriptor before the runs (%s)" % tuned if tuned else "heuristic",
            "sustained_rate_at_equal_p99_resident": at_equal_p99,
            "transports": {"resident": "request fields already in HBM (graph state); no PCIe copies",
                           "host": "request fields in pinned host memory, copied in and out per request"}}
    print(json.dumps(line), flush=True)


def cpu_baseline_triple(sample_images=16, budget_s=10.0):
    """Oracle FP32 restatement, OpenMP over output planes on every core the
    process may use, on `sample_images` of the CNN-layer triple's batch."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    threads = host_cores()
    B, N, K = sample_images, 16, 32
    x, w, dy, _ = triple_inputs(B=B)
    ops = [(O.make_desc(O.PF_BRS, O.VALID, O.XCORR, O.NONE, 1, B, x.shape, w.shape), x, w),
           (O.make_desc(O.BPF_BRS, O.FULL, O.CONV, O.NONE, 1, B, dy.shape, w.shape), dy, w),
           (O.make_desc(O.FA_BRS, O.VALID, O.XCORR, O.NONE, 1, B, x.shape, dy.shape), x, dy)]
    flops = 2.0 * B * N * K * 9 * (62 * 62 + 64 * 64 + 62 * 62)
    for d, a, k in ops:
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


def cnn_triple_workload(args):
    """The FP32 CNN-layer triple: batch 16 images 64x64x16, 32 filters 3x3x16
    (He-normal), PF+BRS Valid, BPF+BRS Full, FA+BRS Valid; one step runs the
    three ops once."""
    import cogconv._lib as L
    L.init(0)
    props = L.device_props(0)
    log("device:", props)
    if args.copy_calibration:
        copy_calibration(L)
    x, w, dy, _ = triple_inputs()
    bufs = {"x": L.DeviceBuffer.from_array(x), "w": L.DeviceBuffer.from_array(w), "dy": L.DeviceBuffer.from_array(dy)}
    ops = []
    for name, desc, k0, k1 in triple_descs(L):
        plan = tuned_plan(L, desc, args)
        ops.append(Op(L, name, plan, bufs[k0], bufs[k1]))
        log("plan:", name, plan.description, "GFLOP %.4f" % (plan.flops / 1e9))
    flops_step = sum(op.plan.flops for op in ops)
    stream = L.Stream()
    flush = L.DeviceBuffer(FLUSH_BYTES)
    schedule = "concurrent" if args.concurrent else args.schedule
    if schedule == "serial":
        ms_per_step = timed_steps(L, ops, stream, args.steps, args.warmup)
    else:
        # "layer": the forward ProjectFrame first, then the two backward ops
        # (independent of each other) on two streams; "concurrent": all three forked.
        streams = [stream] + [L.Stream() for _ in ops[1:]]
        n = args.steps + args.warmup
        fork = [L.Event() for _ in range(n)]
        done = [[L.Event() for _ in ops] for _ in range(n)]

        def run_step(k):
            first = 1 if schedule == "layer" else 0
            for i in range(first):
                ops[i].launch(stream)
            fork[k].record(stream)
            for i in range(first, len(ops)):
                if i:
                    L.check(L.lib().cogconv_stream_wait_event(streams[i].handle, fork[k].handle))
                ops[i].launch(streams[i])
                if i:
                    done[k][i].record(streams[i])
            for i in range(max(1, first), len(ops)):
                L.check(L.lib().cogconv_stream_wait_event(stream.handle, done[k][i].handle))
        for k in range(args.warmup):
            run_step(k)
        L.synchronize()
        t0, t1 = L.Event(), L.Event()
        wall0 = time.perf_counter()
        t0.record(stream)
        for k in range(args.steps):
            run_step(args.warmup + k)
        t1.record(stream)
        t1.synchronize()
        L.synchronize()
        ms_per_step = max(t0.elapsed_ms(t1), (time.perf_counter() - wall0) * 1e3) / args.steps
    per_op = {}
    for op in ops:
        med, lo, hi = flushed_median_ms(L, op, stream, flush, reps=args.op_reps)
        per_op[op.name] = {"ms": round(med, 5), "min_ms": round(lo, 5), "max_ms": round(hi, 5),
                           "gflops": round(op.plan.flops / (med * 1e-3) / 1e9, 1), "kernel": op.plan.description}
    value = flops_step / (ms_per_step * 1e-3) / 1e9
    dom = max(ops, key=lambda op: per_op[op.name]["ms"])
    roof = roofline(dom.plan, per_op[dom.name]["ms"], PEAK_FP32_TFLOPS, dom.name, args.op_reps,
                    "traffic_cnn_triple.json")
    cpu = None if args.no_cpu else cpu_baseline_triple(budget_s=args.cpu_budget)
    line = {
        "metric": METRIC, "value": round(value, 2), "unit": "GFLOP/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 5), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded uniform[-1,1) x and dy, He-normal filters)",
        "kernel_choice": "heuristic (--no-tune)" if args.no_tune else "autotuned per op (cogconv_autotune)",
        "config": {"workload": "FP32 CNN-layer triple: PF+BRS Valid, BPF+BRS Full, FA+BRS Valid; batch 16 "
                               "images 64x64x16, 32 filters 3x3x16", "batch": 16, "image": [64, 64, 16],
                   "filters": [32, 3, 3, 16], "flops_per_step": flops_step,
                   "step": {"layer": "ProjectFrame, then BackProjectFrame and FilterAdjoint on two streams",
                            "serial": "ops back to back on one stream",
                            "concurrent": "three ops on three streams"}[schedule]},
        "per_op": per_op, "roofline": roof, "cpu_baseline": cpu, "device": props["name"] + " " + props["arch"],
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--op-reps", type=int, default=20, help="flushed launches per op for the per-op medians")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-extras", action="store_true", help="large workload: skip the CNN-triple per-op keys")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--schedule", default="serial", choices=("serial", "layer", "concurrent"),
                    help="cnn-triple: serial (default) runs the three ops back to back on one stream; layer runs "
                         "ProjectFrame, then BackProjectFrame and FilterAdjoint on two streams; concurrent forks all "
                         "three")
    ap.add_argument("--concurrent", action="store_true", help="same as --schedule concurrent")
    ap.add_argument("--workload", default="large", choices=("large", "cnn-triple", "precision", "serving"))
    ap.add_argument("--requests", type=int, default=3000, help="requests of the serving workload")
    ap.add_argument("--serving-tune", action="store_true",
                    help="serving: autotune every distinct request descriptor first (the plans then use the cache)")
    ap.add_argument("--serving-only", default=None, metavar="PRECISION:RUN",
                    help="serving workload: only this precision (fp32 / bf16) and run (closed / a rate), resident "
                         "transport, e.g. bf16:closed (for profiling one run)")
    ap.add_argument("--graph-after", type=int, default=4,
                    help="serving: launches of one (plan, bank, stream) before it is captured into a hipGraph "
                         "(0 = never)")
    ap.add_argument("--no-tune", action="store_true",
                    help="use the heuristic kernel configuration instead of autotuning each op's descriptor")
    ap.add_argument("--tune-cache", default=None,
                    help="autotuning cache file: loaded before planning, saved after (cached samples are not re-timed)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run three 64 MiB device-to-device copies (a known-size reference for HBM counters)")
    ap.add_argument("--pmc-traffic", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="no GPU work: print HBM bytes per dispatch per kernel from two rocprofv3 counter CSVs "
                         "(FETCH_SIZE, WRITE_SIZE) of a run made with --copy-calibration")
    ap.add_argument("--trace-summary", metavar="KERNEL_TRACE_CSV",
                    help="no GPU work: print per-kernel, per-grid dispatch durations of a rocprofv3 --kernel-trace CSV")
    args = ap.parse_args()
    if args.trace_summary:
        for r in trace_summary(args.trace_summary):
            print(json.dumps(r))
        return
    if args.pmc_traffic:
        print(json.dumps(pmc_traffic(*args.pmc_traffic), indent=1))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (--gpus %d requested)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    {"large": large_workload, "cnn-triple": cnn_triple_workload, "precision": precision_workload,
     "serving": serving_workload}[args.workload](args)


if __name__ == "__main__":
    main()
