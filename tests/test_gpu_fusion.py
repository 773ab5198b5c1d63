"""Elementwise sinks merged into the convolution's stores (the reference's
forward merge of a kernel into its elementwise sink, HyperKernel.scala:478-530,
under Cog.localMemoryMerging), and the separable route
(CogFunctionAPI.scala:1277-1356).

A merged graph must give bit for bit what the unmerged graph gives (the same
float ops, in the same order, with contraction off), run fewer launch steps,
and agree with the FP64 oracle followed by the same ops in float64."""
import numpy as np
import pytest

import cogconv
import cogconv._lib as L
import oracle as O
from cogconv import api

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def device():
    L.init(0)
    yield
    L.synchronize()


def run_graph(build, merge, precision=L.FP32):
    """build() creates the fields and returns the result field; `merge` sets
    Cog.localMemoryMerging (the other optimizers stay on, so both graphs sum
    in the same order)."""
    cg = cogconv.ComputeGraph(precision=precision)
    with cg:
        out = build()
        cogconv.probe(out)
    try:
        cogconv.Cog.localMemoryMerging = merge
        cg.step()
        steps = cg.kernel_steps()
    finally:
        cogconv.Cog.localMemoryMerging = True
    got = cg.read(out)
    cg.release()
    return got, steps


def check_merge(build, ref, precision=L.FP32, tol=1e-5, what=""):
    print("CASE merge %s" % what, flush=True)
    fused, n_fused = run_graph(build, True, precision)
    plain, n_plain = run_graph(build, False, precision)
    assert n_fused < n_plain, (what, n_fused, n_plain)
    assert np.array_equal(fused, plain), "%s: merged differs from unmerged by %.3g" % (
        what, np.abs(fused - plain).max())
    if ref is not None:
        ref = np.asarray(ref, np.float64).reshape(fused.shape)
        err = np.abs(fused - ref).max() / max(1e-30, np.abs(ref).max())
        assert err <= tol, "%s: normwise %.3g against the oracle" % (what, err)
    return n_fused, n_plain


def conv_ref(mode, border, orient, sampling, step, batch, a, k):
    return O.conv(O.make_desc(mode, border, orient, sampling, step, batch, a.shape, k.shape), a, k)


@pytest.mark.parametrize("H", [32, 13])
def test_project_frame_bias_relu(H):
    """PF + blockReduceSum + per-plane bias (0-D VectorField) + max(., 0): one
    launch.  H = 32 takes the 16-byte store path of the MFMA kernel, 13 the LDS tile."""
    rng = np.random.default_rng(H)
    B, N, K = 2, 8, 16
    x = rng.uniform(-1, 1, (B * N, H, H)).astype(np.float32)
    w = rng.uniform(-1, 1, (K * N, 3, 3)).astype(np.float32)
    bias = rng.uniform(-1, 1, B * K).astype(np.float32)

    def build():
        y = cogconv.blockReduceSum(cogconv.projectFrame(cogconv.VectorField(x), cogconv.VectorField(w),
                                                        cogconv.BorderZero, batchSize=B), N)
        return cogconv.max(y + cogconv.VectorField(bias), 0.0)
    ref = conv_ref(O.PF_BRS, O.ZERO, O.XCORR, O.NONE, 1, B, x, w) + bias[:, None, None]
    n, m = check_merge(build, np.maximum(ref, 0.0), what="pf bias relu H=%d" % H)
    assert n == 1


def test_back_project_frame_scalar_field_bias_tanh():
    rng = np.random.default_rng(2)
    B, N, K, H = 2, 8, 16, 24
    dy = rng.uniform(-1, 1, (B * K, H, H)).astype(np.float32)
    w = rng.uniform(-1, 1, (K * N, 3, 3)).astype(np.float32)
    s = rng.uniform(-1, 1, (H, H)).astype(np.float32)  # ScalarField: broadcast over the planes

    def build():
        y = cogconv.blockReduceSum(cogconv.backProjectFrame(cogconv.VectorField(dy), cogconv.VectorField(w),
                                                            cogconv.BorderZero, batchSize=B), K)
        return cogconv.tanh(y * 0.25 + cogconv.ScalarField(s))
    ref = conv_ref(O.BPF_BRS, O.ZERO, O.CONV, O.NONE, 1, B, dy, w)
    check_merge(build, np.tanh(ref * 0.25 + s[None]), what="bpf scalar bias tanh")


def test_filter_adjoint_scale_and_shift():
    """The FilterAdjoint reduction kernel applies the epilogue."""
    rng = np.random.default_rng(3)
    B, N, K, H = 4, 8, 16, 20
    x = rng.uniform(-1, 1, (B * N, H, H)).astype(np.float32)
    dy = rng.uniform(-1, 1, (B * K, H - 2, H - 2)).astype(np.float32)

    def build():
        g = cogconv.blockReduceSum(cogconv.crossCorrelateFilterAdjoint(cogconv.VectorField(x), cogconv.VectorField(dy),
                                                                       cogconv.BorderValid, batchSize=B), B)
        return 1.0 - g * 0.5
    ref = conv_ref(O.FA_BRS, O.VALID, O.XCORR, O.NONE, 1, B, x, dy)
    check_merge(build, 1.0 - ref * 0.5, what="fa scale shift")


@pytest.mark.parametrize("F,border", [(5, O.CLAMP), (17, O.CYCLIC)])
def test_plane_by_plane_unary_chain(F, border):
    """Direct kernels (register-tiled for 5x5, the LDS-tiled one for 17x17):
    abs, sqrt, exp and log merged."""
    rng = np.random.default_rng(F)
    x = rng.uniform(-1, 1, (40, 45)).astype(np.float32)
    k = rng.uniform(-1, 1, (F, F)).astype(np.float32)
    pol = {O.CLAMP: cogconv.BorderClamp, O.CYCLIC: cogconv.BorderCyclic}[border]

    def build():
        c = cogconv.convolve(cogconv.ScalarField(x), cogconv.ScalarField(k), pol)
        return cogconv.log(cogconv.exp(cogconv.sqrt(cogconv.abs(c)) * 0.5) + 1.0)
    ref = conv_ref(O.PBP, border, O.CONV, O.NONE, 1, 1, x[None], k[None])[0]
    check_merge(build, np.log(np.exp(np.sqrt(np.abs(ref)) * 0.5) + 1.0), what="pbp unary F=%d" % F)


def test_upsampled_phases_and_field_operand():
    """Up-sampled ProjectFrame (four phase launches with strided stores) with a
    same-type field added and the result divided by a 0-D scalar."""
    rng = np.random.default_rng(5)
    B, N, K, H = 1, 8, 16, 12
    x = rng.uniform(-1, 1, (B * N, H, H)).astype(np.float32)
    w = rng.uniform(-1, 1, (K * N, 5, 5)).astype(np.float32)
    other = rng.uniform(-1, 1, (K, 2 * H, 2 * H)).astype(np.float32)

    def build():
        y = cogconv.blockReduceSum(cogconv.projectFrame(cogconv.VectorField(x), cogconv.VectorField(w),
                                                        cogconv.BorderZero, cogconv.UpsampleInputConvolution(2),
                                                        batchSize=B), N)
        return (y + cogconv.VectorField(other)) / cogconv.ScalarField(4.0)
    ref = conv_ref(O.PF_BRS, O.ZERO, O.XCORR, O.UP, 2, B, x, w)
    check_merge(build, (ref + other) / 4.0, what="up2 field operand")


@pytest.mark.parametrize("precision", [L.BF16, L.SPLIT_BF16])
def test_low_precision_frame_kernel(precision):
    rng = np.random.default_rng(6)
    B, N, K, H = 2, 32, 16, 32
    x = rng.uniform(-1, 1, (B * N, H, H)).astype(np.float32)
    w = rng.uniform(-1, 1, (K * N, 3, 3)).astype(np.float32)
    bias = rng.uniform(-1, 1, B * K).astype(np.float32)

    def build():
        y = cogconv.blockReduceSum(cogconv.projectFrame(cogconv.VectorField(x), cogconv.VectorField(w),
                                                        cogconv.BorderZero, batchSize=B), N)
        return cogconv.max(y - cogconv.VectorField(bias), 0.0)
    ref = np.maximum(conv_ref(O.PF_BRS, O.ZERO, O.XCORR, O.NONE, 1, B, x, w) - bias[:, None, None], 0.0)
    check_merge(build, ref, precision=precision, tol=1e-2 if precision == L.BF16 else 1e-5,
                what="lp frame %d" % precision)


def test_reduce_of_a_field_with_epilogue_and_chain():
    """blockReduceSum(blockReduceSum(v, 2), 3) merges into one reduce, which
    takes the signum / min epilogue."""
    rng = np.random.default_rng(7)
    v = rng.uniform(-1, 1, (12, 9, 11)).astype(np.float32)

    def build():
        r = cogconv.blockReduceSum(cogconv.blockReduceSum(cogconv.VectorField(v), 2), 3)
        return cogconv.min(cogconv.signum(r) * 2.0, 1.5)
    ref = v.astype(np.float64).reshape(2, 6, 9, 11).sum(axis=1)
    n, m = check_merge(build, np.minimum(np.sign(ref) * 2.0, 1.5), what="reduce chain")
    assert n == 1 and m == 4


# ---- separable route ------------------------------------------------------------------------

@pytest.mark.parametrize("border", [O.ZERO, O.CLAMP, O.CYCLIC])
@pytest.mark.parametrize("orient", [O.XCORR, O.CONV])
@pytest.mark.parametrize("shape,fr,fc", [((10, 10), 5, 5), ((70, 133), 7, 3), ((33, 65), 1, 63), ((200, 129), 31, 9)])
def test_separable_against_oracle(border, orient, shape, fr, fc):
    print("CASE separable %s %s %s %d %d" % (border, orient, shape, fr, fc), flush=True)
    rng = np.random.default_rng(fr * 100 + fc)
    x = rng.uniform(-1, 1, shape).astype(np.float32)
    rf = rng.uniform(-1, 1, fr).astype(np.float32)
    cf = rng.uniform(-1, 1, fc).astype(np.float32)
    ref = O.separable(x, rf, cf, border, orient)
    mag = O.separable(np.abs(x), np.abs(rf), np.abs(cf), border, orient)
    bx, br, bc = (L.DeviceBuffer.from_array(a) for a in (x, rf, cf))
    bo = L.DeviceBuffer(x.nbytes)
    sp = L.SeparablePlan((2, shape[0], shape[1], 1, 0), fr, fc, border, orient)
    sp.launch(bx, br, bc, bo)
    L.synchronize()
    got = bo.download(shape).astype(np.float64)
    assert np.all(np.abs(got - ref) <= 1e-5 * mag + 1e-30)
    # float-order restatement of the two passes: the kernel sums in the same order
    ref32 = O.separable(x, rf, cf, border, orient, prec="f32")
    assert np.abs(got - ref32).max() <= 1e-6 * max(1.0, np.abs(ref).max())
    for b in (bx, br, bc, bo):
        b.close()


def test_separable_equals_square_convolution_on_gpu():
    """The reference's own check (DynamicConvolutionSpec.scala:74-110) on the device:
    convolve(a, col outer row, BorderZero) == convolveSeparable(a, row, col, BorderZero),
    and the same for cross-correlation; plus a merged epilogue on the separable step."""
    rng = np.random.default_rng(10)
    a = rng.uniform(0, 1, (10, 10)).astype(np.float32)
    row = np.array([0.1, 0.3, 0.4, 0.2, 0.0], np.float32)
    col = np.array([0.0, 0.5, -0.4, -0.2, -0.1], np.float32)
    square = np.outer(col, row)
    cg = cogconv.ComputeGraph()
    with cg:
        af, rf, cf, sf = (cogconv.ScalarField(v) for v in (a, row, col, square))
        smooth = cogconv.convolve(af, sf, cogconv.BorderZero)
        smooth_sep = cogconv.convolveSeparable(af, rf, cf, cogconv.BorderZero)
        cross = cogconv.crossCorrelate(af, sf, cogconv.BorderZero)
        cross_sep = af.crossCorrelateSeparable(rf, cf, cogconv.BorderZero)
        relu_sep = cogconv.max(cogconv.convolveSeparable(af, rf, cf, cogconv.BorderZero), 0.0)
        cogconv.probe(smooth, smooth_sep, cross, cross_sep, relu_sep)
    cg.step()
    for p, q in ((smooth, smooth_sep), (cross, cross_sep)):
        u, v = cg.read(p), cg.read(q)
        assert np.abs(u - v).max() <= 1e-5 * max(1.0, np.abs(u).max())
    assert np.array_equal(cg.read(relu_sep), np.maximum(cg.read(smooth_sep), 0.0))
    cg.release()


def test_capi_epilogue_with_field_operand():
    """cogconv_launch_epilogue with a field operand; cogconv_launch refuses such a plan."""
    rng = np.random.default_rng(11)
    x = rng.uniform(-1, 1, (1, 30, 34)).astype(np.float32)
    k = rng.uniform(-1, 1, (1, 3, 3)).astype(np.float32)
    other = rng.uniform(1, 2, (1, 30, 34)).astype(np.float32)
    desc = L.make_desc(L.PLANE_BY_PLANE, L.BORDER_ZERO, L.CROSS_CORRELATION, L.NO_SAMPLING, 1, 1, (2, 30, 34, 1, 0),
                       (2, 3, 3, 1, 0))
    plan = L.Plan(desc)
    plan.set_epilogue(L.make_epilogue([(L.EPI_RSUB_FIELD, 0.0, 0), (L.EPI_MUL_CONST, 3.0, 0)], [(2, 30, 34, 1, 0)]))
    bx, bk, bo, bf = (L.DeviceBuffer.from_array(a) for a in (x, k, np.zeros_like(x), other))
    with pytest.raises(L.CogconvError):
        plan.launch(bx, bk, bo)
    plan.launch_epilogue(bx, bk, bo, None, [bf])
    L.synchronize()
    got = bo.download(x.shape).astype(np.float64)
    ref = conv_ref(O.PBP, O.ZERO, O.XCORR, O.NONE, 1, 1, x, k)
    want = (other - ref) * 3.0
    assert np.abs(got - want).max() <= 1e-5 * np.abs(want).max()
    for b in (bx, bk, bo, bf):
        b.close()
    plan.close()
