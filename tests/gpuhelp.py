"""Helpers for the GPU parity tests: run one ConvolveOp through the C ABI and
compare it with the oracle.

Tolerance (FP32 plans): for every output element
    |gpu - oracle_f64| <= 1e-5 * sum_taps |tap * sample|
(the magnitude sum is the oracle run on |in0|, |in1|), and normwise
    max |gpu - oracle_f64| <= 1e-5 * max |oracle_f64|.
Index arithmetic errors (wrong border, phase, plane or flip) show up as O(1)
relative errors and fail both.
"""
import numpy as np

import cogconv._lib as L
import oracle as O

TOL_FP32 = 1e-5


def desc_pair(mode, border, orient, sampling, step, batch, a_shape, k_shape, precision=L.FP32, algo=L.ALGO_AUTO):
    od = O.make_desc(mode, border, orient, sampling, step, batch, a_shape, k_shape)

    def ft(s):
        return (2, s[1], s[2], s[0]) if len(s) == 3 else (1, 1, s[1], s[0])

    cd = L.make_desc(mode, border, orient, sampling, step, batch, ft(a_shape), ft(k_shape), precision, algo)
    return od, cd


def run_gpu(cdesc, a, k, stream=None, variant=None):
    plan = L.Plan(cdesc, variant)
    ba = L.DeviceBuffer.from_array(a.astype(np.float32))
    bk = L.DeviceBuffer.from_array(k.astype(np.float32))
    bo = L.DeviceBuffer(4 * int(np.prod(plan.out_shape)))
    ws = L.DeviceBuffer(plan.workspace_bytes) if plan.workspace_bytes else None
    plan.launch(ba, bk, bo, ws, stream)
    L.synchronize()
    out = bo.download(plan.out_shape)
    for b in (ba, bk, bo, ws):
        if b is not None:
            b.close()
    plan.close()
    return out


def oracle_ref(odesc, a, k, threads=0):
    ref = O.conv(odesc, a, k, threads=threads)
    mag = O.conv(odesc, np.abs(a), np.abs(k), threads=threads)
    return ref, mag


def assert_parity(got, ref, mag, tol=TOL_FP32, what=""):
    assert got.shape == ref.shape, (what, got.shape, ref.shape)
    assert np.all(np.isfinite(got)), what
    err = np.abs(got.astype(np.float64) - ref)
    bound = tol * mag + 1e-30
    bad = err > bound
    assert not bad.any(), "%s: %d elements out of bound, worst err %.3g vs bound %.3g at %s" % (
        what, int(bad.sum()), float(err[bad].max()), float(bound[bad][np.argmax(err[bad])]),
        np.unravel_index(np.argmax(np.where(bad, err, 0)), err.shape))
    scale = max(float(np.abs(ref).max()), 1e-30)
    assert err.max() <= tol * scale, "%s: normwise %.3g" % (what, err.max() / scale)
    return float(err.max() / scale)
