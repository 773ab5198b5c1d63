"""Precision variants (BF16 / FP16 MFMA, split-BF16) against the FP64 oracle.

Tolerances, normwise (max |gpu - oracle_f64| / max |oracle_f64|):
  BF16 and FP16 implicit GEMM: <= 1e-2
  split-BF16 (three-term):     <= 1e-5
on uniform[-1,1) inputs and on wide-range inputs |v| = 10^U(-3,3) with random
signs.  All borders, both block-reduce frame modes, 1/3/5/7 filters.
"""
import zlib

import numpy as np
import pytest

import cogconv._lib as L
import oracle as O
from gpuhelp import desc_pair, run_gpu

pytestmark = pytest.mark.gpu

TOL = {L.BF16: 1e-2, L.FP16: 1e-2, L.SPLIT_BF16: 1e-5}
NAMES = {L.BF16: "bf16", L.FP16: "fp16", L.SPLIT_BF16: "split_bf16"}


@pytest.fixture(scope="module", autouse=True)
def device():
    L.init(0)
    yield
    L.synchronize()


def inputs(rng, shape, wide):
    if not wide:
        return rng.uniform(-1, 1, shape).astype(np.float32)
    mag = 10.0 ** rng.uniform(-3, 3, shape)
    return (mag * rng.choice([-1.0, 1.0], shape)).astype(np.float32)


def normwise(got, ref):
    return float(np.abs(got.astype(np.float64) - ref).max() / np.abs(ref).max())


CASES = []
for mode in (O.PF_BRS, O.BPF_BRS):
    for border in (O.ZERO, O.CLAMP, O.CYCLIC, O.VALID, O.FULL):
        for F in (3, 1, 5):
            CASES.append((mode, border, F))


@pytest.mark.parametrize("precision", (L.BF16, L.FP16, L.SPLIT_BF16))
@pytest.mark.parametrize("mode,border,F", CASES)
def test_precision_variant_accuracy(precision, mode, border, F):
    rng = np.random.default_rng(1000 * precision + 10 * mode + border + F)
    B, N, K, H, W = 2, 40, 24, 19, 37  # N not a multiple of 32, ragged tiles
    for wide in (False, True):
        a = inputs(rng, (B * N, H, W), wide)
        k = inputs(rng, (K * N, F, F), wide)
        orient = O.CONV if mode == O.BPF_BRS else O.XCORR
        od, cd = desc_pair(mode, border, orient, O.NONE, 1, B, a.shape, k.shape, precision=precision)
        print("CASE precision", NAMES[precision], mode, border, F, "wide" if wide else "uniform", flush=True)
        got = run_gpu(cd, a, k)
        ref = O.conv(od, a, k)
        err = normwise(got, ref)
        print("  normwise error %.3g" % err)
        assert got.shape == ref.shape
        assert err <= TOL[precision], (NAMES[precision], wide, err)


VEC4_CASES = []
for _mode in (O.PF_BRS, O.BPF_BRS):
    for _border in (O.ZERO, O.VALID, O.FULL):
        for _F in (1, 3, 5, 7, 9):
            VEC4_CASES.append((_mode, _border, _F))


@pytest.mark.parametrize("precision", (L.BF16, L.FP16, L.SPLIT_BF16))
@pytest.mark.parametrize("mode,border,F", VEC4_CASES)
def test_precision_vec4_staging(precision, mode, border, F):
    """Field widths that are multiples of 4 with the borders that read 0 outside
    the field: the kernel stages 4-column groups with 16-byte loads.  Ragged
    column and row tiles; the plane count leaves a partial 8-plane group."""
    rng = np.random.default_rng(700 + 100 * precision + 10 * border + F + mode)
    B, N, K, H, W = 2, 36, 20, 23, 68
    a = inputs(rng, (B * N, H, W), False)
    k = inputs(rng, (K * N, F, F), False)
    orient = O.CONV if mode == O.BPF_BRS else O.XCORR
    od, cd = desc_pair(mode, border, orient, O.NONE, 1, B, a.shape, k.shape, precision=precision)
    try:
        plan = L.Plan(cd)
    except L.CogconvError as e:
        assert precision == L.SPLIT_BF16 and e.status == L.ERR_UNSUPPORTED
        return
    desc = plan.description
    plan.close()
    assert "ld16" in desc, desc
    print("CASE vec4", NAMES[precision], mode, border, F, desc, flush=True)
    got = run_gpu(cd, a, k)
    ref = O.conv(od, a, k)
    err = normwise(got, ref)
    print("  normwise error %.3g" % err)
    assert got.shape == ref.shape
    assert err <= TOL[precision]


def test_precision_c3_shape_uniform():
    """B32 128x128x64 planes, 64 filters 3x3x64, BorderZero on a 4-image slice."""
    rng = np.random.default_rng(33)
    B, N, K = 4, 64, 64
    x = rng.uniform(-1, 1, (B * N, 128, 128)).astype(np.float32)
    w = rng.uniform(-1, 1, (K * N, 3, 3)).astype(np.float32)
    od = O.make_desc(O.PF_BRS, O.ZERO, O.XCORR, O.NONE, 1, B, x.shape, w.shape)
    ref = O.conv(od, x, w, threads=16)
    for precision in (L.BF16, L.FP16, L.SPLIT_BF16):
        _, cd = desc_pair(O.PF_BRS, O.ZERO, O.XCORR, O.NONE, 1, B, x.shape, w.shape, precision=precision)
        print("CASE c3 slice", NAMES[precision], flush=True)
        got = run_gpu(cd, x, w)
        err = normwise(got, ref)
        print("  normwise error %.3g" % err)
        assert err <= TOL[precision]


def test_precision_rejects_unsupported_modes():
    """Precision variants exist for the contraction modes only; others fail at plan time."""
    _, cd = desc_pair(O.PBP, O.ZERO, O.XCORR, O.NONE, 1, 1, (1, 16, 16), (1, 3, 3), precision=L.BF16)
    with pytest.raises(L.CogconvError) as e:
        L.Plan(cd)
    assert e.value.status == L.ERR_UNSUPPORTED


FA_CASES = [
    # (orientation, sampling, step, (B, N, H, W), (K, dh, dw))
    (O.XCORR, O.NONE, 1, (4, 16, 34, 34), (32, 32, 32)),      # 3x3 dW, c2-like planes
    (O.CONV, O.NONE, 1, (2, 5, 30, 70), (3, 26, 64)),        # 5x7 dW, ragged planes
    (O.XCORR, O.DOWN, 2, (2, 6, 37, 41), (20, 18, 20)),      # down-sampled x taps: dW 10x11
    (O.XCORR, O.UP, 2, (2, 4, 41, 181), (12, 18, 88)),       # up-sampled dy, two tap-column chunks
    (O.XCORR, O.NONE, 1, (1, 3, 20, 300), (40, 18, 294)),    # three chunks, MB 3
]


@pytest.mark.parametrize("precision", (L.BF16, L.FP16, L.SPLIT_BF16))
@pytest.mark.parametrize("orient,sampling,step,xs,ys", FA_CASES)
def test_filter_adjoint_precision_accuracy(precision, orient, sampling, step, xs, ys):
    """FilterAdjoint + blockReduceSum(batch), BorderValid, in the precision
    variants: same tolerances as the frame modes."""
    B, N, H, W = xs
    K, dh, dw = ys
    rng = np.random.default_rng(7000 + 100 * precision + dh + dw)
    for wide in (False, True):
        x = inputs(rng, (B * N, H, W), wide)
        dy = inputs(rng, (B * K, dh, dw), wide)
        od, cd = desc_pair(O.FA_BRS, O.VALID, orient, sampling, step, B, x.shape, dy.shape, precision=precision)
        plan = L.Plan(cd)
        print("CASE fa precision", NAMES[precision], xs, ys, "wide" if wide else "uniform", plan.description,
              flush=True)
        assert plan.description.startswith(("fa_igemm_" + NAMES[precision], "fa_rows_" + NAMES[precision]))
        plan.close()
        got = run_gpu(cd, x, dy)
        ref = O.conv(od, x, dy, threads=16)
        err = normwise(got, ref)
        print("  normwise error %.3g" % err)
        assert got.shape == ref.shape
        assert err <= TOL[precision], (NAMES[precision], wide, err)


@pytest.mark.parametrize("precision", (L.BF16, L.SPLIT_BF16))
@pytest.mark.parametrize("K,dw", [(56, 36), (40, 36), (20, 28)])
def test_filter_adjoint_precision_every_candidate(precision, K, dw):
    """Every candidate configuration of a FilterAdjoint with four, three and two
    row blocks (the last one padded) against the FP64 oracle: the 4-wave kernel
    and the 8-wave ones (wave quads owning two row blocks each for K = 56;
    otherwise waves splitting tap rows and 32-column slices, 8 x 1 for the
    one-slice dy rows of dw = 28)."""
    rng = np.random.default_rng(5600 + precision + K)
    B, N, H, W, dh = 2, 6, 30, dw + 4, 27
    x = inputs(rng, (B * N, H, W), False)
    dy = inputs(rng, (B * K, dh, dw), False)
    od, cd = desc_pair(O.FA_BRS, O.VALID, O.XCORR, O.NONE, 1, B, x.shape, dy.shape, precision=precision)
    ref = O.conv(od, x, dy, threads=16)
    waves = set()
    for v in range(L.variant_count(cd)):
        plan = L.Plan(cd, v)
        desc = plan.description
        plan.close()
        print("CASE fa candidate", NAMES[precision], v, desc, flush=True)
        waves.add(desc.split("waves=")[1].split()[0])
        got = run_gpu(cd, x, dy, variant=v)
        err = normwise(got, ref)
        assert err <= TOL[precision], (v, desc, err)
    assert waves == {"4", "8"}, waves


def test_filter_adjoint_precision_c3_slice():
    """B32 128x128x64 / dy 126x126x64 on a 4-image slice, BF16 / FP16 / split."""
    rng = np.random.default_rng(34)
    B, N, K = 4, 64, 64
    x = rng.uniform(-1, 1, (B * N, 128, 128)).astype(np.float32)
    dy = rng.uniform(-1, 1, (B * K, 126, 126)).astype(np.float32)
    od = O.make_desc(O.FA_BRS, O.VALID, O.XCORR, O.NONE, 1, B, x.shape, dy.shape)
    ref = O.conv(od, x, dy, threads=16)
    for precision in (L.BF16, L.FP16, L.SPLIT_BF16):
        _, cd = desc_pair(O.FA_BRS, O.VALID, O.XCORR, O.NONE, 1, B, x.shape, dy.shape, precision=precision)
        print("CASE c3 FA slice", NAMES[precision], flush=True)
        got = run_gpu(cd, x, dy)
        err = normwise(got, ref)
        print("  normwise error %.3g" % err)
        assert err <= TOL[precision]


@pytest.mark.parametrize("precision", (L.BF16, L.FP16, L.SPLIT_BF16))
@pytest.mark.parametrize("mode,border,F", [(O.PF_BRS, O.CLAMP, 7), (O.PF_BRS, O.CLAMP, 9), (O.BPF_BRS, O.ZERO, 9),
                                           (O.PF_BRS, O.CYCLIC, 7)])
def test_precision_large_filters(precision, mode, border, F):
    """7x7 and 9x9 filters (the large-field filter banks); split-BF16 is not
    offered where its two-term staging does not fit LDS."""
    rng = np.random.default_rng(500 + 10 * F + border + precision)
    B, N, K, H, W = 2, 24, 20, 30, 41
    a = inputs(rng, (B * N, H, W), False)
    k = inputs(rng, (K * N, F, F), False)
    orient = O.CONV if mode == O.BPF_BRS else O.XCORR
    od, cd = desc_pair(mode, border, orient, O.NONE, 1, B, a.shape, k.shape, precision=precision)
    try:
        plan = L.Plan(cd)
    except L.CogconvError as e:
        assert precision == L.SPLIT_BF16 and e.status == L.ERR_UNSUPPORTED
        return
    print("CASE large filter", NAMES[precision], mode, border, F, plan.description, flush=True)
    plan.close()
    got = run_gpu(cd, a, k)
    ref = O.conv(od, a, k)
    err = normwise(got, ref)
    print("  normwise error %.3g" % err)
    assert err <= TOL[precision]


EVERY_CASES = []
for _mode in (O.PF_BRS, O.BPF_BRS):
    for _border in (O.ZERO, O.CLAMP, O.CYCLIC, O.VALID, O.FULL):
        for _F in (3, 5, 7, 9):
            EVERY_CASES.append((_mode, _border, _F))


@pytest.mark.parametrize("precision", (L.BF16, L.FP16))
@pytest.mark.parametrize("mode,border,F", EVERY_CASES)
def test_precision_frame_every_candidate(precision, mode, border, F):
    """Every candidate configuration of the BF16 / FP16 frame kernel (output
    block sizes, whole-filter or tap-row groups, the double-buffered one-row
    ring, 16-byte staging with the Clamp / Cyclic halo pass) adds the same
    products in the same order, so all agree bit for bit; the first meets the
    oracle bound.  Width 68 (a multiple of 4: 16-byte staging) and 37 (per-column
    staging); 36 planes (a partial 8-plane group), ragged row and column tiles."""
    rng = np.random.default_rng(900 + 100 * precision + 10 * border + F + 3 * mode)
    B, N, K = 2, 36, 20
    orient = O.CONV if mode == O.BPF_BRS else O.XCORR
    for H, W in ((23, 68), (21, 37)):
        a = inputs(rng, (B * N, H, W), False)
        k = inputs(rng, (K * N, F, F), False)
        od, cd = desc_pair(mode, border, orient, O.NONE, 1, B, a.shape, k.shape, precision=precision)
        ref = O.conv(od, a, k)
        first = None
        names = []
        for v in range(L.variant_count(cd)):
            p = L.Plan(cd, v)
            name = p.description
            p.close()
            names.append(name)
            print("CASE every", NAMES[precision], mode, border, F, W, v, name, flush=True)
            got = run_gpu(cd, a, k, variant=v)
            if first is None:
                first = got
                err = normwise(got, ref)
                assert err <= TOL[precision], (name, err)
            else:
                assert np.array_equal(got, first), "%s differs from %s" % (name, names[0])
        if W % 4 == 0:
            assert any("ld16" in n for n in names), names
        if F > 1:
            assert any("ring" in n for n in names), names


def rounded(a, precision):
    """`a` rounded to the MFMA input type (round to nearest even), as float32:
    products of such values are exact in FP32, so a kernel that adds them in
    FP32 meets an FP64 oracle run on the rounded inputs to FP32 summation error."""
    a = np.ascontiguousarray(a, np.float32)
    if precision == L.FP16:
        return a.astype(np.float16).astype(np.float32)
    bits = a.view(np.uint32).astype(np.uint64)
    bits = (bits + 0x7FFF + ((bits >> 16) & 1)) & 0xFFFF0000
    return bits.astype(np.uint32).view(np.float32)


ROWS_CASES = [
    # (B, N, H, W), K: dW side H - dh + 1 with dh = H - OD + 1
    ((2, 64, 34, 36), 64, 3),   # c3 planes; dy rows 34 wide: a full and a 2-pixel chunk, the last
                                # 16-byte piece of the last row running past the image's buffer
    ((3, 20, 23, 70), 40, 3),   # partial x- and dy-plane tiles; x rows 70 wide (rows not 16-byte aligned), dy rows 68
    ((2, 16, 19, 37), 24, 1),   # 1x1 dW; odd widths
    ((2, 12, 21, 45), 20, 5),   # 5x5 dW; dy rows 41 wide
]


@pytest.mark.parametrize("precision", (L.BF16, L.FP16))
@pytest.mark.parametrize("xs,K,OD", ROWS_CASES)
def test_filter_adjoint_rows_every_candidate(precision, xs, K, OD):
    """Every configuration of the row-streaming FilterAdjoint kernel (tiles,
    row splits, load widths): against the FP64 oracle on the inputs rounded to
    the MFMA type (normwise <= 2e-6: only FP32 summation differs), against the
    oracle on the FP32 inputs (the precision bound), and bit for bit between
    tile shapes with the same row splits (same partial sums, same order)."""
    B, N, H, W = xs
    dh, dw = H - OD + 1, W - OD + 1
    rng = np.random.default_rng(8800 + 10 * precision + OD + K)
    x = inputs(rng, (B * N, H, W), False)
    dy = inputs(rng, (B * K, dh, dw), False)
    od, cd = desc_pair(O.FA_BRS, O.VALID, O.XCORR, O.NONE, 1, B, x.shape, dy.shape, precision=precision)
    ref = O.conv(od, x, dy, threads=16)
    ref_r = O.conv(od, rounded(x, precision), rounded(dy, precision), threads=16)
    by_split = {}
    seen = 0
    for v in range(L.variant_count(cd)):
        plan = L.Plan(cd, v)
        desc = plan.description
        plan.close()
        if not desc.startswith("fa_rows_"):
            continue
        seen += 1
        print("CASE fa rows", NAMES[precision], xs, K, OD, v, desc, flush=True)
        got = run_gpu(cd, x, dy, variant=v)
        assert got.shape == ref.shape
        err_r = normwise(got, ref_r)
        assert err_r <= 2e-6, (desc, err_r)
        assert normwise(got, ref) <= TOL[precision], desc
        key = desc.split("rowsplits=")[1].split()[0]
        if key in by_split:
            assert np.array_equal(got, by_split[key][1]), (desc, by_split[key][0])
        else:
            by_split[key] = (desc, got)
    assert seen >= 2, "row-streaming kernel not offered"


# (mode, border, F, N, K, H, W): widths 44 and 74 output columns (last 8-column
# group past the row end, 4-byte-aligned rows), 40 output planes (two blocks),
# 36 planes (a partial 32-plane chunk), BorderFull / Valid halos.
UP2_LP_CASES = [("pf_brs", O.ZERO, 7, 36, 20, 13, 22), ("bpf_brs", O.FULL, 5, 16, 40, 9, 17),
                ("pf_brs", O.CYCLIC, 9, 8, 16, 11, 37), ("pf_brs", O.VALID, 3, 5, 3, 12, 9),
                ("bpf_brs", O.ZERO, 3, 40, 33, 6, 70)]
_UP2_MODES = {"pf_brs": O.PF_BRS, "bpf_brs": O.BPF_BRS}


@pytest.mark.parametrize("precision", (L.BF16, L.FP16))
@pytest.mark.parametrize("mode,border,F,N,K,H,W", UP2_LP_CASES)
def test_upsampled_lp_every_candidate(precision, mode, border, F, N, K, H, W):
    """UpsampleInputConvolution(2) in BF16 / FP16: the fused four-phase kernel at
    each output block, against the FP64 oracle on the inputs rounded to the MFMA
    type (normwise <= 2e-6: only FP32 summation differs), against the oracle on
    the FP32 inputs (the precision bound), bit for bit between output blocks,
    and with an epilogue equal to the same ops applied afterwards."""
    rng = np.random.default_rng(zlib.crc32(repr(("uplp", precision, mode, border, F, N, K)).encode()))
    B = 2
    a = inputs(rng, (B * N, H, W), False)
    k = inputs(rng, (K * N, F, F), False)
    orient = O.CONV if mode == "bpf_brs" else O.XCORR
    od, cd = desc_pair(_UP2_MODES[mode], border, orient, O.UP, 2, B, a.shape, k.shape, precision=precision)
    ref = O.conv(od, a, k, threads=16)
    ref_r = O.conv(od, rounded(a, precision), rounded(k, precision), threads=16)
    first = None
    seen = 0
    for v in range(L.variant_count(cd)):
        plan = L.Plan(cd, v)
        desc = plan.description
        plan.close()
        print("CASE up2 lp", NAMES[precision], mode, border, F, N, K, H, W, v, desc, flush=True)
        assert "_up2 " in desc and "fused" in desc, desc
        seen += 1
        got = run_gpu(cd, a, k, variant=v)
        assert got.shape == ref.shape
        assert normwise(got, ref_r) <= 2e-6, (desc, normwise(got, ref_r))
        assert normwise(got, ref) <= TOL[precision], desc
        if first is None:
            first = got
            plan = L.Plan(cd, v)
            plan.set_epilogue(L.make_epilogue([(L.EPI_MUL_CONST, 3.0, 0), (L.EPI_MAX_CONST, 0.5, 0)]))
            ba, bk = L.DeviceBuffer.from_array(a), L.DeviceBuffer.from_array(k)
            bo = L.DeviceBuffer(4 * int(np.prod(plan.out_shape)))
            ws = L.DeviceBuffer(plan.workspace_bytes)
            plan.launch(ba, bk, bo, ws)
            L.synchronize()
            fused = bo.download(plan.out_shape)
            assert np.array_equal(fused, np.maximum(got * np.float32(3.0), np.float32(0.5)))
            for buf in (ba, bk, bo, ws):
                buf.close()
            plan.close()
        else:
            assert np.array_equal(got, first), desc
    assert seen >= 1
