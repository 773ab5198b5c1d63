"""Parity of the HIP path (through the C ABI) with the oracle and the reference's golden vectors.

All tests here run on device 0 and are marked `gpu`.  Each prints the case it
runs before touching the device.
"""
import sys
import zlib

import numpy as np
import pytest

import cogconv
import cogconv._lib as L
import oracle as O
from conftest import load_case, load_orphan
from gpuhelp import assert_parity, desc_pair, oracle_ref, run_gpu
from test_oracle import ADJ_BORDER, ADJ_SAMPLING, case_params, ref_field_equal

pytestmark = pytest.mark.gpu

MODES = {"pbp": O.PBP, "pf": O.PF, "pf_brs": O.PF_BRS, "bpf": O.BPF, "bpf_brs": O.BPF_BRS, "fa": O.FA,
         "fa_brs": O.FA_BRS}


def announce(*what):
    print("CASE", *what, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def device():
    L.init(0)
    yield
    L.synchronize()


@pytest.mark.parametrize("i", range(25))
def test_golden_wx_gpu(manifest, i):
    """DynamicConvolutionSpec 'convolution' test: wx == convolve(x, k, border, sampling)."""
    c = manifest["dynamic"][i]
    border, sampling, step = case_params(c)
    if sampling == O.UP and border == O.CLAMP:
        pytest.skip("the reference rejects Upsample + BorderClamp (DynamicConvolutionGenerator.scala:531-536)")
    z = load_case(i)
    announce("golden wx", i, c["border"], c["sampling"], step)
    od, cd = desc_pair(O.PBP, border, O.CONV, sampling, step, 1, (1,) + z["x"].shape, (1,) + z["k"].shape)
    got = run_gpu(cd, z["x"], z["k"])[0]
    assert ref_field_equal(got, z["wx"])
    ref, mag = oracle_ref(od, z["x"], z["k"])
    assert_parity(got, ref[0], mag[0], what="wx%d" % i)


@pytest.mark.parametrize("i", range(25))
def test_golden_wty_gpu(manifest, i):
    c = manifest["dynamic"][i]
    border, sampling, step = case_params(c)
    b2, s2 = ADJ_BORDER.get(border, border), ADJ_SAMPLING[sampling]
    if s2 == O.UP and b2 == O.CLAMP:
        pytest.skip("adjoint is Upsample + BorderClamp, rejected by the reference")
    z = load_case(i)
    announce("golden wty", i, b2, s2, step)
    od, cd = desc_pair(O.PBP, b2, O.XCORR, s2, step, 1, (1,) + z["y"].shape, (1,) + z["k"].shape)
    got = run_gpu(cd, z["y"], z["k"])[0]
    assert ref_field_equal(got, z["wty"])
    ref, mag = oracle_ref(od, z["y"], z["k"])
    assert_parity(got, ref[0], mag[0], what="wty%d" % i)


@pytest.mark.parametrize("t", (0, 3))
@pytest.mark.parametrize("border", ("Zero", "Clamp", "Cyclic"))
def test_golden_orphan_gpu(manifest, t, border):
    z = load_orphan(t)
    announce("orphan", t, border)
    meta = [m for m in manifest["orphan"] if m["case"] == t][0]
    b = O.BORDER_NAMES["Border" + border]
    od, cd = desc_pair(O.PBP, b, O.CONV, O.NONE, 1, 1, (1,) + z["target"].shape, (1,) + z["filter"].shape)
    got = run_gpu(cd, z["target"], z["filter"])[0]
    band = meta["band"]
    r0, c0 = meta["centre_origin"]
    for part, sl in (("top", np.s_[:band]), ("bottom", np.s_[-band:]), ("left", np.s_[:, :band]),
                     ("right", np.s_[:, -band:]), ("centre", np.s_[r0:r0 + 32, c0:c0 + 32])):
        assert ref_field_equal(got[sl], z["%s_%s" % (border, part)]), part


IMAGE = np.array([[0, 1, 2, 3], [2, 3, 4, 5], [8, 7, 6, 4], [5, 4, 9, 8]], np.float32)
GAUSS = np.array([[0.25, 0.5, 0.25], [0.5, 1.0, 0.5], [0.25, 0.5, 0.25]], np.float32)


def test_closed_forms_gpu():
    """StaticConvolutionSpec.scala:62-233 through the Field API."""
    announce("closed forms")
    with cogconv.ComputeGraph() as cg:
        src = cogconv.ScalarField(IMAGE)
        k = cogconv.ScalarField(GAUSS)
        clamp = cogconv.convolve(src, k, cogconv.BorderClamp)
        cyclic = cogconv.convolve(src, k, cogconv.BorderCyclic)
        zero = cogconv.convolve(src, k, cogconv.BorderZero)
        xk = cogconv.ScalarField(np.array([[1, 2, 0], [0, 3, 0], [0, 0, 0]], np.float32))
        xc = cogconv.crossCorrelate(src, xk, cogconv.BorderClamp)
        img1 = cogconv.ScalarField(np.array([1, 2, 3, 4], np.float32))
        k1 = cogconv.ScalarField(np.array([0.25, 0.5, 0.25], np.float32))
        one_d = [cogconv.convolve(img1, k1, b) for b in (cogconv.BorderClamp, cogconv.BorderCyclic,
                                                         cogconv.BorderZero)]
        cogconv.probe(clamp, cyclic, zero, xc, *one_d)
    cg.step()
    assert np.array_equal(cg.read(clamp), [[3.0, 6.0, 10.0, 13.0], [12.5, 14.0, 15.75, 16.75],
                                          [22.5, 22.5, 23.0, 22.0], [22.0, 23.5, 28.25, 29.25]])
    assert np.array_equal(cg.read(cyclic), [[10.5, 10.5, 15.5, 15.5], [13.75, 14.0, 15.75, 15.5],
                                           [22.0, 22.5, 23.0, 22.5], [18.75, 19.0, 22.75, 22.5]])
    assert np.array_equal(cg.read(zero), [[2.25, 5.0, 8.0, 7.5], [9.5, 14.0, 15.75, 12.5],
                                         [16.75, 22.5, 23.0, 16.75], [12.75, 18.0, 20.75, 16.0]])
    assert np.array_equal(cg.read(xc), [[0, 5, 11, 17], [6, 11, 17, 23], [30, 29, 29, 26], [39, 34, 46, 38]])
    for f, want in zip(one_d, ([1.25, 2.0, 3.0, 3.75], [2.0, 2.0, 3.0, 3.0], [1.0, 2.0, 3.0, 2.75])):
        assert np.array_equal(cg.read(f), want)
    cg.release()


FRAME_CASES = []
for mode in ("pf", "pf_brs", "bpf", "bpf_brs", "fa", "fa_brs", "pbp"):
    for border in (O.ZERO, O.CLAMP, O.CYCLIC, O.VALID, O.FULL):
        for sampling, step in ((O.NONE, 1), (O.DOWN, 2), (O.UP, 2)):
            FRAME_CASES.append((mode, border, sampling, step))


def frame_inputs(mode, border, sampling, step, rng):
    B, N, K = 2, 3, 2
    H, W, F = 10, 12, 3
    if sampling == O.DOWN and border == O.VALID:
        H, W = 10, 12
    if sampling == O.DOWN and border == O.FULL:
        H, W = 10, 12
    if mode in ("fa", "fa_brs"):
        if border == O.VALID:
            dh, dw = (H - 3 + 1, W - 3 + 1) if sampling != O.UP else ((H - 3 + 1) // 2, (W - 3 + 1) // 2)
        else:
            dh, dw = 5, 5
        a = rng.uniform(-1, 1, (B * N, H, W)).astype(np.float32)
        k = rng.uniform(-1, 1, (B * K, dh, dw)).astype(np.float32)
        return B, a, k
    if mode == "pbp":
        a = rng.uniform(-1, 1, (N, H, W)).astype(np.float32)
        k = rng.uniform(-1, 1, (N, F, F + 2)).astype(np.float32)
        return 1, a, k
    a = rng.uniform(-1, 1, (B * N, H, W)).astype(np.float32)
    k = rng.uniform(-1, 1, (K * N, F, F + 2)).astype(np.float32)
    return B, a, k


@pytest.mark.parametrize("mode,border,sampling,step", FRAME_CASES)
@pytest.mark.parametrize("orient", (O.XCORR, O.CONV))
def test_frame_modes_gpu(mode, border, sampling, step, orient):
    rng = np.random.default_rng(zlib.crc32(repr((mode, border, sampling, step, orient)).encode()))
    batch, a, k = frame_inputs(mode, border, sampling, step, rng)
    od, cd = desc_pair(MODES[mode], border, orient, sampling, step, batch, a.shape, k.shape)
    try:
        O.output_type(od)
    except O.OracleError:
        with pytest.raises(L.CogconvError):
            L.Plan(cd)
        return
    announce("frame", mode, border, sampling, step, orient, a.shape, k.shape)
    got = run_gpu(cd, a, k)
    ref, mag = oracle_ref(od, a, k)
    assert_parity(got, ref, mag, what="%s b%d s%d" % (mode, border, sampling))


ALGO_CASES = []
for mode in ("pf_brs", "bpf_brs", "fa_brs"):
    for border in (O.ZERO, O.CLAMP, O.CYCLIC, O.VALID, O.FULL):
        for sampling, step in ((O.NONE, 1), (O.DOWN, 2)):
            ALGO_CASES.append((mode, border, sampling, step))


@pytest.mark.parametrize("mode,border,sampling,step", ALGO_CASES)
@pytest.mark.parametrize("algo", (L.ALGO_DIRECT, L.ALGO_IMPLICIT_GEMM, L.ALGO_SMALL_OUTPUT))
def test_every_kernel_family_gpu(mode, border, sampling, step, algo):
    """Each kernel family, forced through the algorithm hint, on square-filter frame cases."""
    rng = np.random.default_rng(zlib.crc32(repr((mode, border, sampling, step, algo)).encode()))
    B, N, K, H, W = 2, 5, 3, 12, 20
    if mode == "fa_brs":
        a = rng.uniform(-1, 1, (B * N, H, W)).astype(np.float32)
        k = rng.uniform(-1, 1, (B * K, H - 2, W - 2) if border == O.VALID else (B * K, 5, 5)).astype(np.float32)
    else:
        a = rng.uniform(-1, 1, (B * N, H, W)).astype(np.float32)
        k = rng.uniform(-1, 1, (K * N, 3, 3)).astype(np.float32)
    od, cd = desc_pair(MODES[mode], border, O.CONV if mode == "bpf_brs" else O.XCORR, sampling, step, B,
                       a.shape, k.shape, algo=algo)
    try:
        O.output_type(od)
    except O.OracleError:
        return
    try:
        plan = L.Plan(cd)
    except L.CogconvError as e:
        assert e.status == L.ERR_UNSUPPORTED
        return
    announce("algo", algo, plan.description)
    plan.close()
    got = run_gpu(cd, a, k)
    ref, mag = oracle_ref(od, a, k)
    assert_parity(got, ref, mag, what="%s algo %d" % (mode, algo))


RT_CASES = []
for _fh, _fw in ((1, 1), (2, 4), (5, 5), (3, 8), (11, 11), (16, 15), (7, 13), (15, 15)):
    for _border in (O.ZERO, O.CLAMP, O.CYCLIC, O.VALID, O.FULL):
        if (_fh % 2 and _fw % 2) or _border in (O.VALID, O.FULL):  # even filters need Valid or Full
            RT_CASES.append((_fh, _fw, _border))


@pytest.mark.parametrize("fh,fw,border", RT_CASES)
@pytest.mark.parametrize("mode", ("pbp", "pf"))
def test_register_tiled_direct_gpu(fh, fw, border, mode):
    """The register-tiled direct kernel (no sampling, filters up to 16x15) on planes that leave ragged
    32x128 output tiles, for every border, both orientations."""
    rng = np.random.default_rng(zlib.crc32(repr((fh, fw, border, mode)).encode()))
    H, W = 45 + fh, 150 + 2 * fw
    if mode == "pbp":
        B, a = 1, rng.uniform(-1, 1, (3, H, W)).astype(np.float32)
        k = rng.uniform(-1, 1, (3, fh, fw)).astype(np.float32)
    else:
        B, N, K = 2, 2, 3
        a = rng.uniform(-1, 1, (B * N, H, W)).astype(np.float32)
        k = rng.uniform(-1, 1, (K * N, fh, fw)).astype(np.float32)
    for orient in (O.XCORR, O.CONV):
        od, cd = desc_pair(MODES[mode], border, orient, O.NONE, 1, B, a.shape, k.shape, algo=L.ALGO_DIRECT)
        plan = L.Plan(cd)
        desc = plan.description
        plan.close()
        assert desc.startswith("direct_rt"), desc
        announce("direct_rt", mode, border, orient, a.shape, k.shape)
        got = run_gpu(cd, a, k)
        ref, mag = oracle_ref(od, a, k)
        assert_parity(got, ref, mag, what="direct_rt %s b%d o%d %dx%d" % (mode, border, orient, fh, fw))


def test_c1_project_frame_block_reduce_sum():
    """x 32x32x3 uniform[-1,1), 8 filters 5x5x3, BorderZero: projectFrame + blockReduceSum(3) -> 32x32x8."""
    announce("c1")
    rng = np.random.default_rng(2024)
    x = rng.uniform(-1, 1, (3, 32, 32)).astype(np.float32)
    w = rng.uniform(-1, 1, (24, 5, 5)).astype(np.float32)
    with cogconv.ComputeGraph() as cg:
        y = cogconv.blockReduceSum(cogconv.projectFrame(cogconv.VectorField(x), cogconv.VectorField(w),
                                                        cogconv.BorderZero), 3)
        cogconv.probe(y)
    cg.step()
    assert [p.description.split()[1] for p in cg.plans()] == ["ProjectFrameBlockReduceSum"]
    print(cg.plans()[0].description)
    got = cg.read(y)
    cg.release()
    od = O.make_desc(O.PF_BRS, O.ZERO, O.XCORR, O.NONE, 1, 1, x.shape, w.shape)
    ref, mag = oracle_ref(od, x, w)
    assert got.shape == (8, 32, 32)
    assert_parity(got, ref, mag, what="c1")


def c2_inputs():
    rng = np.random.default_rng(16)
    B, N, K = 16, 16, 32
    x = rng.uniform(-1, 1, (B * N, 64, 64)).astype(np.float32)
    w = (rng.standard_normal((K * N, 3, 3)) * np.sqrt(2.0 / 144.0)).astype(np.float32)
    return B, N, K, x, w, rng


@pytest.mark.parametrize("border,sampling,step", [
    (O.VALID, O.NONE, 1), (O.CYCLIC, O.NONE, 1), (O.ZERO, O.NONE, 1), (O.CYCLIC, O.DOWN, 2)])
def test_c2_project_frame_full_size(border, sampling, step):
    B, N, K, x, w, _ = c2_inputs()
    announce("c2 PF", border, sampling, step)
    od, cd = desc_pair(O.PF_BRS, border, O.XCORR, sampling, step, B, x.shape, w.shape)
    got = run_gpu(cd, x, w)
    ref, mag = oracle_ref(od, x, w, threads=16)
    assert_parity(got, ref, mag, what="c2 PF")


@pytest.mark.parametrize("border", (O.FULL, O.CYCLIC))
def test_c2_back_project_frame_full_size(border):
    B, N, K, x, w, rng = c2_inputs()
    s = 62 if border == O.FULL else 64
    dy = rng.uniform(-1, 1, (B * K, s, s)).astype(np.float32)
    announce("c2 BPF", border)
    # The ProjectFrame bank W[k*N + n] read as BackProjectFrame's W[m*L + l] (m = k, l = n).
    od, cd = desc_pair(O.BPF_BRS, border, O.CONV, O.NONE, 1, B, dy.shape, w.shape)
    got = run_gpu(cd, dy, w)
    ref, mag = oracle_ref(od, dy, w, threads=16)
    assert got.shape == (B * N, 64, 64)
    assert_parity(got, ref, mag, what="c2 BPF")


def test_c2_filter_adjoint_full_size():
    B, N, K, x, w, rng = c2_inputs()
    dy = rng.uniform(-1, 1, (B * K, 62, 62)).astype(np.float32)
    announce("c2 FA")
    od, cd = desc_pair(O.FA_BRS, O.VALID, O.XCORR, O.NONE, 1, B, x.shape, dy.shape)
    got = run_gpu(cd, x, dy)
    ref, mag = oracle_ref(od, x, dy, threads=16)
    assert got.shape == (K * N, 3, 3)
    assert_parity(got, ref, mag, what="c2 FA")


@pytest.mark.parametrize("shape,orient,sampling,step", [
    (((2, 3, 40, 150), (2, 4, 38, 148)), O.XCORR, O.NONE, 1),   # two tap-column chunks
    (((1, 2, 30, 300), (1, 5, 26, 292)), O.CONV, O.NONE, 1),    # three chunks, 5x9 output
    (((2, 2, 41, 181), (2, 3, 18, 88)), O.XCORR, O.UP, 2),      # up-sampled dy, 175 tap columns
])
def test_filter_adjoint_column_chunks_gpu(shape, orient, sampling, step):
    """FilterAdjoint+BRS, BorderValid, with dy wider than one staged row: the
    tap columns are split into chunks whose partial sums meet in the split
    reduction."""
    (B, N, H, W), (_, K, dh, dw) = shape
    rng = np.random.default_rng(zlib.crc32(repr(shape).encode()))
    x = rng.uniform(-1, 1, (B * N, H, W)).astype(np.float32)
    dy = rng.uniform(-1, 1, (B * K, dh, dw)).astype(np.float32)
    od, cd = desc_pair(O.FA_BRS, O.VALID, orient, sampling, step, B, x.shape, dy.shape, algo=L.ALGO_IMPLICIT_GEMM)
    plan = L.Plan(cd)
    announce("FA chunks", shape, plan.description)
    assert plan.description.startswith("fa_igemm_f32")
    plan.close()
    got = run_gpu(cd, x, dy)
    ref, mag = oracle_ref(od, x, dy, threads=16)
    assert_parity(got, ref, mag, what="FA chunks")


def test_c3_filter_adjoint_fp32_full_size():
    """Batch 32, x 128x128x64, dy 126x126x64 (BorderValid): dW 64x64x3x3, FP32."""
    rng = np.random.default_rng(33)
    B, N, K = 32, 64, 64
    x = rng.uniform(-1, 1, (B * N, 128, 128)).astype(np.float32)
    dy = rng.uniform(-1, 1, (B * K, 126, 126)).astype(np.float32)
    od, cd = desc_pair(O.FA_BRS, O.VALID, O.XCORR, O.NONE, 1, B, x.shape, dy.shape)
    plan = L.Plan(cd)
    announce("c3 FA fp32", plan.description)
    plan.close()
    got = run_gpu(cd, x, dy)
    ref, mag = oracle_ref(od, x, dy, threads=16)
    assert got.shape == (K * N, 3, 3)
    assert_parity(got, ref, mag, what="c3 FA")


UP_CASES = [(mode, border, F) for mode in ("pf_brs", "bpf_brs") for border in (O.ZERO, O.CYCLIC, O.VALID, O.FULL)
            for F in (3, 5, 7, 9)]


@pytest.mark.parametrize("mode,border,F", UP_CASES)
def test_upsampled_implicit_gemm_gpu(mode, border, F):
    """UpsampleInputConvolution(2) on the frame block-reduce modes through the
    four-phase implicit GEMM (each phase a dense sub-filter convolution)."""
    rng = np.random.default_rng(zlib.crc32(repr(("up", mode, border, F)).encode()))
    B, N, K, H, W = 2, 5, 3, 13, 22
    a = rng.uniform(-1, 1, (B * N, H, W)).astype(np.float32)
    k = rng.uniform(-1, 1, (K * N, F, F)).astype(np.float32)
    orient = O.CONV if mode == "bpf_brs" else O.XCORR
    od, cd = desc_pair(MODES[mode], border, orient, O.UP, 2, B, a.shape, k.shape, algo=L.ALGO_IMPLICIT_GEMM)
    try:
        O.output_type(od)
    except O.OracleError:
        return
    plan = L.Plan(cd)
    announce("up2 igemm", mode, border, F, plan.description)
    assert plan.description.startswith("igemm_f32_up2")
    plan.close()
    got = run_gpu(cd, a, k)
    ref, mag = oracle_ref(od, a, k)
    assert_parity(got, ref, mag, what="up2 %s b%d F%d" % (mode, border, F))


# (mode, border, F, N, K, H, W): output widths 44 (a last 8-column group past
# the row end), 2W + F - 1 odd (BorderFull), output planes 20 and 40 (two
# output blocks), an input narrower than one 32-column tile.
UPV_CASES = [("pf_brs", O.ZERO, 7, 12, 20, 13, 22), ("bpf_brs", O.FULL, 5, 6, 40, 9, 17),
             ("pf_brs", O.CYCLIC, 9, 8, 16, 11, 37), ("pf_brs", O.VALID, 3, 5, 3, 12, 9),
             ("bpf_brs", O.ZERO, 3, 16, 33, 6, 70)]


@pytest.mark.parametrize("mode,border,F,N,K,H,W", UPV_CASES)
def test_upsampled_every_variant_gpu(mode, border, F, N, K, H, W):
    """Every up-sampled implicit-GEMM candidate (the fused four-phase launch at
    each output block and plane chunk, and the four phase launches) against the
    oracle; the fused form with an epilogue equals its plain form with the same
    ops applied afterwards, bit for bit."""
    rng = np.random.default_rng(zlib.crc32(repr(("upv", mode, border, F, N, K)).encode()))
    B = 2
    if mode == "pf_brs":
        a = rng.uniform(-1, 1, (B * N, H, W)).astype(np.float32)
        k = rng.uniform(-1, 1, (K * N, F, F)).astype(np.float32)
    else:  # K = output planes per image: dy carries N planes per image
        a = rng.uniform(-1, 1, (B * N, H, W)).astype(np.float32)
        k = rng.uniform(-1, 1, (N * K, F, F)).astype(np.float32)
    orient = O.CONV if mode == "bpf_brs" else O.XCORR
    od, cd = desc_pair(MODES[mode], border, orient, O.UP, 2, B, a.shape, k.shape)
    ref, mag = oracle_ref(od, a, k)
    seen = 0
    for v in range(L.variant_count(cd)):
        plan = L.Plan(cd, v)
        desc = plan.description
        plan.close()
        if not desc.startswith("igemm_f32_up2"):
            continue
        announce("up2 variant", mode, border, F, N, K, H, W, v, desc)
        got = run_gpu(cd, a, k, variant=v)
        assert_parity(got, ref, mag, what="up2 %s b%d F%d v%d %s" % (mode, border, F, v, desc))
        if "fused" not in desc:
            continue
        seen += 1
        plan = L.Plan(cd, v)
        plan.set_epilogue(L.make_epilogue([(L.EPI_MUL_CONST, 3.0, 0), (L.EPI_MAX_CONST, 0.5, 0)]))
        ba, bk = L.DeviceBuffer.from_array(a), L.DeviceBuffer.from_array(k)
        bo = L.DeviceBuffer(4 * int(np.prod(plan.out_shape)))
        plan.launch(ba, bk, bo, None)
        L.synchronize()
        fused = bo.download(plan.out_shape)
        assert np.array_equal(fused, np.maximum(got * np.float32(3.0), np.float32(0.5)))
        for buf in (ba, bk, bo):
            buf.close()
        plan.close()
    assert seen >= 1, "no fused up-sampled launch"


def test_c4_plane_by_plane_full_size():
    """PlaneByPlane 1024x1024 scalar field, 11x11 Gaussian-plus-noise filter, BorderClamp."""
    rng = np.random.default_rng(44)
    img = rng.uniform(0, 1, (1, 1024, 1024)).astype(np.float32)
    yy, xx = np.mgrid[-5:6, -5:6]
    gauss = np.exp(-(yy ** 2 + xx ** 2) / 8.0)
    k = (gauss / gauss.sum() + rng.normal(0, 0.01, gauss.shape)).astype(np.float32)[None]
    announce("c4 PBP")
    od, cd = desc_pair(O.PBP, O.CLAMP, O.CONV, O.NONE, 1, 1, img.shape, k.shape)
    got = run_gpu(cd, img, k)
    ref, mag = oracle_ref(od, img, k, threads=16)
    assert_parity(got, ref, mag, what="c4 PBP")


@pytest.mark.parametrize("F", (7, 9))
def test_c4_upsampled_project_frame_slice(F):
    """UpsampleInputConvolution(2), BorderZero, ProjectFrame+BRS with the
    large-field filter sizes, on a one-image 96x96x8 slice."""
    rng = np.random.default_rng(45 + F)
    x = rng.uniform(-1, 1, (8, 96, 96)).astype(np.float32)
    w = rng.uniform(-1, 1, (6 * 8, F, F)).astype(np.float32)
    od, cd = desc_pair(O.PF_BRS, O.ZERO, O.XCORR, O.UP, 2, 1, x.shape, w.shape)
    announce("c4 up2", F)
    got = run_gpu(cd, x, w)
    ref, mag = oracle_ref(od, x, w, threads=16)
    assert_parity(got, ref, mag, what="c4 up2")


@pytest.mark.parametrize("F", (7, 9))
def test_c4_clamp_project_frame_slice(F):
    """BorderClamp ProjectFrame+BRS, 7x7 / 9x9 banks of 32 filters on a one-image 256x256x32 slice."""
    rng = np.random.default_rng(47 + F)
    x = rng.uniform(-1, 1, (32, 256, 256)).astype(np.float32)
    w = (rng.standard_normal((32 * 32, F, F)) * np.sqrt(2.0 / (32 * F * F))).astype(np.float32)
    od, cd = desc_pair(O.PF_BRS, O.CLAMP, O.XCORR, O.NONE, 1, 1, x.shape, w.shape)
    announce("c4 clamp", F, L.Plan(cd).description)
    got = run_gpu(cd, x, w)
    ref, mag = oracle_ref(od, x, w, threads=16)
    assert_parity(got, ref, mag, what="c4 clamp")


def test_fused_equals_unfused():
    """The reduce fusion of ProjectFrameTensorReduceSumOptimizer changes nothing
    beyond rounding (StaticConvolutionSpec.scala:1742-1768)."""
    announce("fused vs unfused")
    rng = np.random.default_rng(77)
    B, N, K = 3, 4, 5
    x = rng.uniform(-1, 1, (B * N, 20, 20)).astype(np.float32)
    w = rng.uniform(-1, 1, (K * N, 3, 3)).astype(np.float32)
    dy = rng.uniform(-1, 1, (B * K, 18, 18)).astype(np.float32)
    outs = []
    for optimize in (True, False):
        with cogconv.ComputeGraph(optimize=optimize) as cg:
            xf, wf, dyf = cogconv.VectorField(x), cogconv.VectorField(w), cogconv.VectorField(dy)
            y = cogconv.blockReduceSum(cogconv.projectFrame(xf, wf, cogconv.BorderZero, batchSize=B), N)
            dx = cogconv.blockReduceSum(cogconv.backProjectFrame(dyf, wf, cogconv.BorderFull, batchSize=B), K)
            dw = cogconv.blockReduceSum(cogconv.crossCorrelateFilterAdjoint(xf, dyf, cogconv.BorderValid,
                                                                            batchSize=B), B)
            cogconv.probe(y, dx, dw)
        cg.step()
        n_plans = len(cg.plans())
        outs.append((cg.read(y), cg.read(dx), cg.read(dw), n_plans))
        cg.release()
    assert outs[0][3] == 3 and outs[1][3] == 3
    for a, b in zip(outs[0][:3], outs[1][:3]):
        assert a.shape == b.shape
        assert np.abs(a - b).max() <= 1e-5 * np.abs(b).max()


def test_cog_merging_flags_gpu():
    """Cog.projectFrameMerging / fastProjectFrame (Cog.scala:64-72) switch the
    fusion and the kernel family; the values do not change."""
    announce("Cog flags")
    rng = np.random.default_rng(78)
    x = rng.uniform(-1, 1, (2 * 4, 18, 18)).astype(np.float32)
    w = rng.uniform(-1, 1, (3 * 4, 3, 3)).astype(np.float32)
    results = []
    saved = (cogconv.Cog.projectFrameMerging, cogconv.Cog.fastProjectFrame)
    try:
        for merging, fast in ((True, True), (False, True), (True, False)):
            cogconv.Cog.projectFrameMerging, cogconv.Cog.fastProjectFrame = merging, fast
            with cogconv.ComputeGraph() as cg:
                y = cogconv.blockReduceSum(cogconv.projectFrame(cogconv.VectorField(x), cogconv.VectorField(w),
                                                                cogconv.BorderZero, batchSize=2), 4)
                cogconv.probe(y)
            cg.step()
            kinds = [p.description.split()[0] for p in cg.plans()]
            results.append((cg.read(y), kinds))
            cg.release()
    finally:
        cogconv.Cog.projectFrameMerging, cogconv.Cog.fastProjectFrame = saved
    assert results[0][1] == ["igemm_f32"]
    # The direct family (LDS-tiled or register-tiled) for the other two.
    assert [k.startswith("direct") for k in results[1][1]] == [True]   # unfused: ProjectFrame then blockReduceSum
    assert [k.startswith("direct") for k in results[2][1]] == [True]   # fused, one output plane per block
    for r, _ in results[1:]:
        assert np.abs(r - results[0][0]).max() <= 1e-5 * np.abs(results[0][0]).max()


def test_block_reduce_sum_gpu():
    announce("blockReduceSum")
    rng = np.random.default_rng(3)
    x = rng.uniform(-1, 1, (24, 17, 19)).astype(np.float32)
    bi = L.DeviceBuffer.from_array(x)
    bo = L.DeviceBuffer(4 * 6 * 17 * 19)
    L.block_reduce_sum(bi, bo, 17 * 19, 24, 4)
    L.synchronize()
    got = bo.download((6, 17, 19))
    assert np.abs(got - O.block_reduce_sum(x, 4)).max() <= 1e-6


@pytest.mark.parametrize("case", [
    # (mode, border, sampling, step, batch, a_shape, k_shape)
    ("pbp", O.VALID, O.NONE, 1, 1, (1, 9, 9), (1, 9, 9)),        # filter == input: 1x1 output
    ("pbp", O.ZERO, O.NONE, 1, 1, (1, 1, 1), (1, 1, 1)),         # single pixel
    ("pbp", O.CYCLIC, O.NONE, 1, 1, (1, 3, 3), (1, 7, 7)),       # filter wider than the field wraps twice
    ("pbp", O.CLAMP, O.NONE, 1, 1, (1, 2, 70), (1, 5, 9)),       # thin field
    ("pbp", O.ZERO, O.NONE, 1, 1, (1, 200), (1, 15)),            # 1-D field
    ("pbp", O.CYCLIC, O.DOWN, 2, 1, (1, 40), (1, 5)),            # 1-D downsample
    ("pbp", O.ZERO, O.UP, 3, 1, (1, 21), (1, 7)),                # 1-D upsample
    ("pbp", O.ZERO, O.NONE, 1, 1, (4, 33, 65), (1, 5, 5)),       # vector image, scalar filter
    ("pbp", O.ZERO, O.NONE, 1, 1, (1, 33, 65), (4, 5, 5)),       # scalar image, vector filter
    ("fa", O.ZERO, O.NONE, 1, 2, (4, 9, 9), (6, 9, 9)),          # FilterAdjoint, full-size output
    ("fa", O.VALID, O.UP, 2, 2, (4, 64, 72), (6, 24, 28)),      # StaticConvolutionSpec.scala:1426
    ("fa_brs", O.VALID, O.UP, 4, 3, (6, 70, 70), (6, 15, 15)),   # up-sampled dy, summed over batch
    ("fa_brs", O.VALID, O.DOWN, 2, 2, (4, 22, 22), (6, 11, 11)),
    ("pf", O.VALID, O.NONE, 1, 1, (3, 70, 70), (192, 7, 7)),     # DynamicConvolutionSpec.scala:115-129 shape
    ("pf_brs", O.ZERO, O.NONE, 1, 4, (4, 16, 16), (5, 1, 1)),    # 1x1 filters, N = 1
    ("bpf_brs", O.CYCLIC, O.DOWN, 3, 2, (6, 27, 27), (12, 5, 5)),
    ("pf_brs", O.FULL, O.UP, 2, 2, (6, 13, 11), (9, 3, 5)),
])
def test_edge_cases_gpu(case):
    mode, border, sampling, step, batch, a_shape, k_shape = case
    announce("edge", case)
    rng = np.random.default_rng(zlib.crc32(repr(case).encode()))
    a = rng.uniform(-1, 1, a_shape).astype(np.float32)
    k = rng.uniform(-1, 1, k_shape).astype(np.float32)
    od, cd = desc_pair(MODES[mode], border, O.CONV, sampling, step, batch, a_shape, k_shape)
    got = run_gpu(cd, a, k)
    ref, mag = oracle_ref(od, a, k)
    assert_parity(got, ref, mag, what=str(case))


def test_split_reduction_is_deterministic():
    """The small-output kernel sums partials in a fixed order: two launches agree bitwise."""
    announce("determinism")
    rng = np.random.default_rng(8)
    x = rng.uniform(-1, 1, (32, 40, 40)).astype(np.float32)
    dy = rng.uniform(-1, 1, (64, 38, 38)).astype(np.float32)
    _, cd = desc_pair(O.FA_BRS, O.VALID, O.XCORR, O.NONE, 1, 8, x.shape, dy.shape)
    a = run_gpu(cd, x, dy)
    b = run_gpu(cd, x, dy)
    assert np.array_equal(a, b)
