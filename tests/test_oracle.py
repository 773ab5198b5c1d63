"""Pins the CPU oracle (oracle/cogconv_oracle.c) before anything is checked against it.

Every oracle function is checked against something outside itself:
  - the reference's golden vectors (DynamicConvolutionSpec.scala:131-197 data):
    wx = convolve(x, k, border, sampling); wty = crossCorrelate(y, k, adjoint border,
    adjoint sampling); orphan convolve{Zero,Clamp,Cyclic} vectors;
  - the closed-form matrices of StaticConvolutionSpec.scala:36-233;
  - torch's CPU conv2d / conv_transpose2d / weight gradient for the frame modes;
  - adjoint identities <A x, y> = <x, A^T y>;
  - the reference's plane-decomposition and blockReduceSum identities
    (StaticConvolutionSpec.scala:1612-1797, 491-582, 1258-1314).
"""
import numpy as np
import pytest

import oracle as O
from conftest import load_case, load_orphan

ADJ_BORDER = {O.VALID: O.FULL, O.FULL: O.VALID}
ADJ_SAMPLING = {O.NONE: O.NONE, O.UP: O.DOWN, O.DOWN: O.UP}


def ref_field_equal(a, b):
    """The reference's field comparison RefFieldMemory.approxEqual
    (src/test/scala/cogx/reference/memory/RefFieldMemory.scala:453-484):
    |a-b| <= 0.001, or |a/b - 1| <= 1e-5 with b != 0, elementwise."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    close = np.abs(a - b) <= 0.001
    with np.errstate(divide="ignore", invalid="ignore"):
        rel = (b != 0) & (np.abs(a / np.where(b == 0, 1, b) - 1) <= 1e-5)
    return bool(np.all(close | rel))


def case_params(c):
    return O.BORDER_NAMES[c["border"]], O.SAMPLING_NAMES[c["sampling"]], max(1, c["step"])


def scalar_desc(mode_orient, border, sampling, step, a, k):
    return O.make_desc(O.PBP, border, mode_orient, sampling, step, 1, (1,) + a.shape, (1,) + k.shape)


@pytest.mark.parametrize("i", range(25))
def test_golden_wx(manifest, i):
    c = manifest["dynamic"][i]
    z = load_case(i)
    border, sampling, step = case_params(c)
    d = scalar_desc(O.CONV, border, sampling, step, z["x"], z["k"])
    if sampling == O.UP and border == O.CLAMP:
        # The reference rejects integrated upsampling with BorderClamp
        # (DynamicConvolutionGenerator.scala:531-536) and excludes these cases
        # from the golden test (DynamicConvolutionSpec.scala:156-161).
        with pytest.raises(O.OracleError) as e:
            O.output_type(d)
        assert e.value.code == 6
        return
    assert O.output_type(d)[1:3] == z["wx"].shape
    got = O.conv(d, z["x"], z["k"])[0]
    assert ref_field_equal(got, z["wx"])
    assert np.abs(got - z["wx"]).max() <= 1e-6 * np.abs(z["wx"]).max()
    got32 = O.conv(d, z["x"], z["k"], prec="f32")[0]
    assert np.abs(got32 - z["wx"]).max() <= 1e-6 * np.abs(z["wx"]).max()


@pytest.mark.parametrize("i", range(25))
def test_golden_wty(manifest, i):
    """wty = W^T y: cross-correlation with the adjoint border and sampling."""
    c = manifest["dynamic"][i]
    z = load_case(i)
    border, sampling, step = case_params(c)
    b2, s2 = ADJ_BORDER.get(border, border), ADJ_SAMPLING[sampling]
    d = scalar_desc(O.XCORR, b2, s2, step, z["y"], z["k"])
    if s2 == O.UP and b2 == O.CLAMP:
        with pytest.raises(O.OracleError):
            O.output_type(d)
        return
    assert O.output_type(d)[1:3] == z["wty"].shape
    got = O.conv(d, z["y"], z["k"])[0]
    assert ref_field_equal(got, z["wty"])
    assert np.abs(got - z["wty"]).max() <= 1e-6 * np.abs(z["wty"]).max()


@pytest.mark.parametrize("t", (0, 3))
@pytest.mark.parametrize("border", ("Zero", "Clamp", "Cyclic"))
def test_golden_orphan(manifest, t, border):
    z = load_orphan(t)
    meta = [m for m in manifest["orphan"] if m["case"] == t][0]
    b = O.BORDER_NAMES["Border" + border]
    d = scalar_desc(O.CONV, b, O.NONE, 1, z["target"], z["filter"])
    got = O.conv(d, z["target"], z["filter"])[0]
    band = meta["band"]
    r0, c0 = meta["centre_origin"]
    for part, sl in (("top", np.s_[:band]), ("bottom", np.s_[-band:]), ("left", np.s_[:, :band]),
                     ("right", np.s_[:, -band:]), ("centre", np.s_[r0:r0 + 32, c0:c0 + 32])):
        want = z["%s_%s" % (border, part)]
        assert ref_field_equal(got[sl], want), part


IMAGE = np.array([[0, 1, 2, 3], [2, 3, 4, 5], [8, 7, 6, 4], [5, 4, 9, 8]], np.float32)
GAUSS = np.array([[0.25, 0.5, 0.25], [0.5, 1.0, 0.5], [0.25, 0.5, 0.25]], np.float32)


@pytest.mark.parametrize("border,expected", [
    (O.CLAMP, [[3.0, 6.0, 10.0, 13.0], [12.5, 14.0, 15.75, 16.75], [22.5, 22.5, 23.0, 22.0], [22.0, 23.5, 28.25, 29.25]]),
    (O.CYCLIC, [[10.5, 10.5, 15.5, 15.5], [13.75, 14.0, 15.75, 15.5], [22.0, 22.5, 23.0, 22.5], [18.75, 19.0, 22.75, 22.5]]),
    (O.ZERO, [[2.25, 5.0, 8.0, 7.5], [9.5, 14.0, 15.75, 12.5], [16.75, 22.5, 23.0, 16.75], [12.75, 18.0, 20.75, 16.0]]),
])
def test_closed_form_2d(border, expected):
    """StaticConvolutionSpec.scala:89-124, 182-202."""
    d = scalar_desc(O.CONV, border, O.NONE, 1, IMAGE, GAUSS)
    assert np.array_equal(O.conv(d, IMAGE, GAUSS)[0], np.array(expected))


def test_closed_form_cross_correlate_clamp():
    """StaticConvolutionSpec.scala:206-233: crossCorrelate does not flip the filter."""
    k = np.array([[1, 2, 0], [0, 3, 0], [0, 0, 0]], np.float32)
    want = [[0, 5, 11, 17], [6, 11, 17, 23], [30, 29, 29, 26], [39, 34, 46, 38]]
    d = scalar_desc(O.XCORR, O.CLAMP, O.NONE, 1, IMAGE, k)
    assert np.array_equal(O.conv(d, IMAGE, k)[0], np.array(want, np.float64))


@pytest.mark.parametrize("border,expected", [
    (O.CLAMP, [1.25, 2.0, 3.0, 3.75]), (O.CYCLIC, [2.0, 2.0, 3.0, 3.0]), (O.ZERO, [1.0, 2.0, 3.0, 2.75])])
def test_closed_form_1d(border, expected):
    """StaticConvolutionSpec.scala:62-85 (1-D fields)."""
    img = np.array([1, 2, 3, 4], np.float32)
    k = np.array([0.25, 0.5, 0.25], np.float32)
    d = O.make_desc(O.PBP, border, O.CONV, O.NONE, 1, 1, (1, 4), (1, 3))
    assert O.output_type(d) == (1, 1, 4, 1, 0)
    assert np.array_equal(O.conv(d, img, k)[0, 0], np.array(expected))


@pytest.mark.parametrize("n", (2, 3, 5))
def test_filter_adjoint_flip_ordering(n):
    """StaticConvolutionSpec.scala:1160-1200: 1x1 filter upsampled by the input size."""
    x = np.array([[n * r + c + 1 for c in range(n)] for r in range(n)], np.float32)
    f = np.ones((1, 1), np.float32)
    for orient, want in ((O.CONV, n * n), (O.XCORR, 1.0)):
        d = O.make_desc(O.FA, O.VALID, orient, O.UP, n, 1, (1, n, n), (1, 1, 1))
        out = O.conv(d, x, f)
        assert out.shape == (1, 1, 1)
        assert out[0, 0, 0] == want


def test_filter_adjoint_upsample_decomposition():
    """StaticConvolutionSpec.scala:1318-1430: FA with UpsampleInputConvolution(s)
    equals a Valid convolution with the zero-inserted 2nd operand."""
    rng = np.random.default_rng(3)
    B, N, Kf, H, W, s = 2, 3, 2, 16, 18, 2
    x = rng.random((B * N, H, W), dtype=np.float32)
    kr, kc = (H - 5 + 1) // s, (W - 5 + 1) // s
    dy = rng.random((B * Kf, kr, kc), dtype=np.float32)
    d = O.make_desc(O.FA, O.VALID, O.CONV, O.UP, s, B, x.shape, dy.shape)
    out = O.conv(d, x, dy)
    for i in range(out.shape[0]):
        b, w = i % B, i // B
        up = O.upsample(dy[b * Kf + w // N], s, s)
        ds = scalar_desc(O.CONV, O.VALID, O.NONE, 1, x[0], up)
        assert np.allclose(out[i], O.conv(ds, x[b * N + w % N], up)[0], rtol=0, atol=1e-9)


def _torch():
    return pytest.importorskip("torch")


@pytest.mark.parametrize("border,sampling,step", [
    (O.ZERO, O.NONE, 1), (O.VALID, O.NONE, 1), (O.FULL, O.NONE, 1), (O.CYCLIC, O.NONE, 1), (O.CLAMP, O.NONE, 1),
    (O.ZERO, O.DOWN, 2), (O.VALID, O.DOWN, 2), (O.CYCLIC, O.DOWN, 2)])
def test_project_frame_brs_vs_torch(border, sampling, step):
    """ProjectFrameBlockReduceSum == torch conv2d (cross-correlation) after the border's padding."""
    torch = _torch()
    F = torch.nn.functional
    rng = np.random.default_rng(11)
    B, N, K, H, Fs = 3, 4, 5, 12, 3
    x = rng.uniform(-1, 1, (B * N, H, H)).astype(np.float32)
    w = rng.uniform(-1, 1, (K * N, Fs, Fs)).astype(np.float32)
    d = O.make_desc(O.PF_BRS, border, O.XCORR, sampling, step, B, x.shape, w.shape)
    got = O.conv(d, x, w)
    xt = torch.from_numpy(x.astype(np.float64)).reshape(B, N, H, H)
    wt = torch.from_numpy(w.astype(np.float64)).reshape(K, N, Fs, Fs)
    h = (Fs - 1) // 2
    if border in (O.CYCLIC, O.CLAMP):
        xt = F.pad(xt, (h, h, h, h), mode="circular" if border == O.CYCLIC else "replicate")
        pad = 0
    else:
        pad = {O.ZERO: h, O.VALID: 0, O.FULL: Fs - 1}[border]
    ref = F.conv2d(xt, wt, stride=step, padding=pad).reshape(B * K, *got.shape[1:]).numpy()
    assert np.abs(got - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("border", (O.ZERO, O.FULL))
def test_back_project_frame_brs_vs_torch(border):
    """BackProjectFrameBlockReduceSum == conv_transpose2d with the filter bank as (M, L, F, F)."""
    torch = _torch()
    rng = np.random.default_rng(12)
    B, M, L, H, Fs = 2, 4, 3, 10, 3
    dy = rng.uniform(-1, 1, (B * M, H, H)).astype(np.float32)
    w = rng.uniform(-1, 1, (M * L, Fs, Fs)).astype(np.float32)
    d = O.make_desc(O.BPF_BRS, border, O.CONV, O.NONE, 1, B, dy.shape, w.shape)
    got = O.conv(d, dy, w)
    pad = (Fs - 1) // 2 if border == O.ZERO else 0
    ref = torch.nn.functional.conv_transpose2d(
        torch.from_numpy(dy.astype(np.float64)).reshape(B, M, H, H),
        torch.from_numpy(w.astype(np.float64)).reshape(M, L, Fs, Fs), padding=pad)
    assert np.abs(got - ref.reshape(B * L, *got.shape[1:]).numpy()).max() <= 1e-12


def test_filter_adjoint_brs_vs_torch():
    """FilterAdjoint(Valid) + blockReduceSum(batch) == the conv2d weight gradient."""
    torch = _torch()
    rng = np.random.default_rng(13)
    B, N, K, H, Fs = 3, 4, 5, 11, 3
    x = rng.uniform(-1, 1, (B * N, H, H)).astype(np.float32)
    dy = rng.uniform(-1, 1, (B * K, H - Fs + 1, H - Fs + 1)).astype(np.float32)
    d = O.make_desc(O.FA_BRS, O.VALID, O.XCORR, O.NONE, 1, B, x.shape, dy.shape)
    got = O.conv(d, x, dy)
    ref = torch.nn.grad.conv2d_weight(torch.from_numpy(x.astype(np.float64)).reshape(B, N, H, H),
                                      (K, N, Fs, Fs),
                                      torch.from_numpy(dy.astype(np.float64)).reshape(B, K, H - Fs + 1, H - Fs + 1))
    assert np.abs(got - ref.reshape(K * N, Fs, Fs).numpy()).max() <= 1e-12


def test_block_reduce_identities():
    """blockReduceSum(PF, N) == PF_BRS, blockReduceSum(BPF, N) == BPF_BRS,
    blockReduceSum(FA, B) == FA_BRS (StaticConvolutionSpec.scala:1742-1768, 2169-2241)."""
    rng = np.random.default_rng(5)
    B, N, K, H = 2, 3, 4, 9
    x = rng.uniform(-1, 1, (B * N, H, H)).astype(np.float32)
    w = rng.uniform(-1, 1, (K * N, 3, 3)).astype(np.float32)
    for plain, fused in ((O.PF, O.PF_BRS), (O.BPF, O.BPF_BRS)):
        p = O.conv(O.make_desc(plain, O.ZERO, O.XCORR, O.NONE, 1, B, x.shape, w.shape), x, w)
        f = O.conv(O.make_desc(fused, O.ZERO, O.XCORR, O.NONE, 1, B, x.shape, w.shape), x, w)
        assert np.allclose(O.block_reduce_sum(p.astype(np.float32), N), f, atol=1e-5)
        assert np.allclose(p.reshape(-1, N, H, H).sum(1), f, atol=1e-12)
    dy = rng.uniform(-1, 1, (B * K, 7, 7)).astype(np.float32)
    p = O.conv(O.make_desc(O.FA, O.VALID, O.XCORR, O.NONE, 1, B, x.shape, dy.shape), x, dy)
    f = O.conv(O.make_desc(O.FA_BRS, O.VALID, O.XCORR, O.NONE, 1, B, x.shape, dy.shape), x, dy)
    assert np.allclose(p.reshape(-1, B, 3, 3).sum(1), f, atol=1e-12)


def test_project_frame_plane_decomposition():
    """StaticConvolutionSpec.scala:1653-1698: plane b*T+w == crossCorrelate(x[b*N + w%N], W[w])."""
    rng = np.random.default_rng(6)
    B, N, T, H = 3, 2, 6, 8
    x = rng.random((B * N, H, H), dtype=np.float32)
    w = rng.random((T, 3, 3), dtype=np.float32)
    out = O.conv(O.make_desc(O.PF, O.CYCLIC, O.XCORR, O.NONE, 1, B, x.shape, w.shape), x, w)
    for b in range(B):
        for t in range(T):
            ref = O.conv(scalar_desc(O.XCORR, O.CYCLIC, O.NONE, 1, x[0], w[0]), x[b * N + t % N], w[t])[0]
            assert np.array_equal(out[b * T + t], ref)


def test_back_project_frame_plane_decomposition():
    """StaticConvolutionSpec.scala:536-582: plane p uses image p%N and filter K*(p%N) + p/N."""
    rng = np.random.default_rng(8)
    B, N, T, H = 2, 3, 6, 7
    K = T // N
    x = rng.random((B * N, H, H), dtype=np.float32)
    w = rng.random((T, 3, 3), dtype=np.float32)
    out = O.conv(O.make_desc(O.BPF, O.ZERO, O.CONV, O.NONE, 1, B, x.shape, w.shape), x, w)
    for b in range(B):
        for p in range(T):
            ref = O.conv(scalar_desc(O.CONV, O.ZERO, O.NONE, 1, x[0], w[0]), x[b * N + p % N], w[K * (p % N) + p // N])[0]
            assert np.array_equal(out[b * T + p], ref)


def test_filter_adjoint_plane_decomposition():
    """StaticConvolutionSpec.scala:1258-1314: output i uses image i%B, in1 = b*N + (i/B)%N, in2 = b*K + (i/B)/N."""
    rng = np.random.default_rng(9)
    B, N, K, H, O_ = 3, 2, 4, 15, 5
    x = rng.random((B * N, H, H + 2), dtype=np.float32)
    dy = rng.random((B * K, H - O_ + 1, H + 2 - O_ + 1), dtype=np.float32)
    out = O.conv(O.make_desc(O.FA, O.VALID, O.CONV, O.NONE, 1, B, x.shape, dy.shape), x, dy)
    for i in range(N * K * B):
        b, w = i % B, i // B
        ref = O.conv(scalar_desc(O.CONV, O.VALID, O.NONE, 1, x[0], dy[0]), x[b * N + w % N], dy[b * K + w // N])[0]
        assert np.array_equal(out[i], ref)


@pytest.mark.parametrize("border,sampling,step", [
    (O.ZERO, O.NONE, 1), (O.CYCLIC, O.NONE, 1), (O.VALID, O.NONE, 1), (O.FULL, O.NONE, 1),
    (O.ZERO, O.DOWN, 2), (O.VALID, O.DOWN, 3), (O.CYCLIC, O.DOWN, 2), (O.FULL, O.DOWN, 2)])
def test_adjoint_identity(border, sampling, step):
    """<PF(x, W), y> == <x, BPF_xcorr-adjoint(y, W)> == <W, FA(x, y)>: the layer's
    adjoint triple with Valid<->Full and Down(s)<->Up(s) (FilterAdjointToeplitzHyperKernel.scala:92-125)."""
    rng = np.random.default_rng(21)
    B, N, K, Fs = 2, 3, 2, 3
    H = {O.VALID: 14, O.FULL: 10}.get(border, 12)
    x = rng.uniform(-1, 1, (B * N, H, H)).astype(np.float32)
    w = rng.uniform(-1, 1, (K * N, Fs, Fs)).astype(np.float32)
    y = O.conv(O.make_desc(O.PF_BRS, border, O.XCORR, sampling, step, B, x.shape, w.shape), x, w)
    dy = rng.uniform(-1, 1, y.shape).astype(np.float32)
    lhs = float((y * dy).sum())
    # dX = BPF_BRS(dy, W) with the adjoint border and sampling: filter bank read as (M=K, L=N)
    wt = w.reshape(K, N, Fs, Fs).reshape(K * N, Fs, Fs)
    adjb = ADJ_BORDER.get(border, border)
    adjs = ADJ_SAMPLING[sampling]
    if border == O.CLAMP:
        pytest.skip("BorderClamp has no adjoint")
    dx = O.conv(O.make_desc(O.BPF_BRS, adjb, O.CONV, adjs, step, B, dy.shape, wt.shape), dy, wt)
    if dx.shape != x.shape:
        pytest.skip("adjoint shape differs for this combination")
    assert abs(lhs - float((x * dx).sum())) <= 1e-9 * max(1.0, abs(lhs))
    if border == O.VALID and sampling == O.NONE:
        dw = O.conv(O.make_desc(O.FA_BRS, O.VALID, O.XCORR, O.NONE, 1, B, x.shape, dy.shape), x, dy)
        assert abs(lhs - float((w * dw).sum())) <= 1e-9 * max(1.0, abs(lhs))


def test_block_reduce_sum_brute_force():
    rng = np.random.default_rng(2)
    x = rng.random((12, 5, 7), dtype=np.float32)
    got = O.block_reduce_sum(x, 4)
    for p in range(3):
        assert np.allclose(got[p], sum(x[4 * p + j].astype(np.float64) for j in range(4)), rtol=0, atol=1e-12)


def test_brute_force_tiny_all_borders():
    """Tiny inputs against a direct sum written from the definition of
    up-sampling as skipped taps (not materialised zeros)."""
    rng = np.random.default_rng(4)
    for border in (O.ZERO, O.CLAMP, O.CYCLIC, O.VALID, O.FULL):
        for sampling, step in ((O.NONE, 1), (O.UP, 2), (O.DOWN, 2)):
            if sampling == O.UP and border == O.CLAMP:
                continue
            H, W, FH, FW = 6, 8, 3, 3
            if sampling == O.DOWN and border == O.VALID:
                W, H = 8, 6
            a = rng.uniform(-1, 1, (H, W)).astype(np.float32)
            k = rng.uniform(-1, 1, (FH, FW)).astype(np.float32)
            for orient in (O.XCORR, O.CONV):
                d = scalar_desc(orient, border, sampling, step, a, k)
                try:
                    _, oh, ow, _, _ = O.output_type(d)
                except O.OracleError:
                    continue
                got = O.conv(d, a, k)[0]
                up = step if sampling == O.UP else 1
                dn = step if sampling == O.DOWN else 1
                U, V = H * up, W * up
                h = {O.FULL: FH - 1, O.VALID: 0}.get(border, (FH - 1) // 2)
                hc = {O.FULL: FW - 1, O.VALID: 0}.get(border, (FW - 1) // 2)
                want = np.zeros((oh, ow))
                for R in range(oh):
                    for C in range(ow):
                        s = 0.0
                        for i in range(FH):
                            for j in range(FW):
                                t = k[FH - 1 - i, FW - 1 - j] if orient == O.CONV else k[i, j]
                                yy, xx = R * dn - h + i, C * dn - hc + j
                                if border == O.CYCLIC:
                                    yy, xx = yy % U, xx % V
                                elif border == O.CLAMP:
                                    yy, xx = min(max(yy, 0), U - 1), min(max(xx, 0), V - 1)
                                elif not (0 <= yy < U and 0 <= xx < V):
                                    continue
                                if yy % up or xx % up:
                                    continue
                                s += float(t) * float(a[yy // up, xx // up])
                        want[R, C] = s
                assert np.abs(got - want).max() <= 1e-12, (border, sampling, orient)


# ---- separable convolution (CogFunctionAPI.scala:1277-1356) ----------------------------------

SEP_ROW = np.array([0.1, 0.3, 0.4, 0.2, 0.0], np.float32)   # DynamicConvolutionSpec.scala:75-77
SEP_COL = np.array([0.0, 0.5, -0.4, -0.2, -0.1], np.float32)


@pytest.mark.parametrize("border", [O.ZERO, O.CLAMP, O.CYCLIC])
@pytest.mark.parametrize("orient", [O.XCORR, O.CONV])
def test_separable_equals_outer_product_convolution(border, orient):
    """The reference's own check (DynamicConvolutionSpec.scala:74-110, there with
    BorderZero): separable(a, row, col) == convolve(a, col outer row); the identity
    holds for every border whose padding is per coordinate."""
    rng = np.random.default_rng(40 + border * 2 + orient)
    for size, rf, cf in ((10, SEP_ROW, SEP_COL), (17, rng.uniform(-1, 1, 7), rng.uniform(-1, 1, 3)),
                         (23, rng.uniform(-1, 1, 3), rng.uniform(-1, 1, 9))):
        a = rng.uniform(-1, 1, (size, size + 3)).astype(np.float32)
        rf = np.asarray(rf, np.float32)
        cf = np.asarray(cf, np.float32)
        square = np.outer(cf, rf)  # float32 taps, as the reference's outer product (Matrix of Float)
        got = O.separable(a, rf, cf, border, orient)
        # 2-D reference in float64 from the definition (outer-product taps, per-coordinate border)
        d = O.make_desc(O.PBP, border, orient, O.NONE, 1, 1, (1,) + a.shape, (1,) + square.shape)
        ref2d = O.conv(d, a, square.astype(np.float32))[0]
        # the float32 outer product rounds the taps: compare in the rounding of the taps
        assert np.abs(got - ref2d).max() <= 1e-6 * max(1.0, np.abs(ref2d).max())


@pytest.mark.parametrize("border,pad", [(O.ZERO, "constant"), (O.CLAMP, "replicate"), (O.CYCLIC, "circular")])
def test_separable_matches_torch(border, pad):
    """torch conv2d on the padded plane with the outer-product weight (xcorr),
    and the two passes as separate 1-D convolutions."""
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(7)
    a = rng.uniform(-1, 1, (19, 26)).astype(np.float32)
    rf = rng.uniform(-1, 1, 5).astype(np.float32)
    cf = rng.uniform(-1, 1, 7).astype(np.float32)
    x = torch.from_numpy(a.astype(np.float64))[None, None]
    xp = torch.nn.functional.pad(x, (2, 2, 3, 3), mode=pad)
    w = torch.from_numpy(np.outer(cf.astype(np.float64), rf.astype(np.float64)))[None, None]
    ref = torch.nn.functional.conv2d(xp, w)[0, 0].numpy()
    got = O.separable(a, rf, cf, border, O.XCORR)
    assert np.abs(got - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())
    # convolution orientation = cross-correlation with both filters reversed
    got_c = O.separable(a, rf[::-1], cf[::-1], border, O.CONV)
    assert np.abs(got_c - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())


def test_separable_brute_force_and_errors():
    """Tiny case written out tap by tap; even filters are rejected."""
    a = np.arange(12, dtype=np.float32).reshape(3, 4)
    rf = np.array([1.0, 2.0, 3.0], np.float32)
    cf = np.array([0.0, 1.0, -1.0], np.float32)
    mid = np.zeros((3, 4))
    for r in range(3):
        for c in range(4):
            mid[r, c] = sum(rf[i] * (a[r, c - 1 + i] if 0 <= c - 1 + i < 4 else 0.0) for i in range(3))
    out = np.zeros((3, 4))
    for r in range(3):
        for c in range(4):
            out[r, c] = sum(cf[j] * (mid[r - 1 + j, c] if 0 <= r - 1 + j < 3 else 0.0) for j in range(3))
    assert np.array_equal(O.separable(a, rf, cf, O.ZERO, O.XCORR), out)
    with pytest.raises(O.OracleError):
        O.separable(a, np.ones(4, np.float32), cf, O.ZERO, O.XCORR)
