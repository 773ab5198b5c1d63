"""FFT route of PlaneByPlane convolution (DynamicConvolutionGenerator.convolveBy2DFFT,
DynamicConvolutionGenerator.scala:322-476): border expansion, filter layout,
spectrum product and trim around hipFFT transforms, for BorderZero / Clamp /
Cyclic, both orientations, and image / filter plane broadcasting.

Tolerance: an FFT spreads rounding over the whole plane, so the bound is
normwise (max |gpu - oracle_f64| <= 1e-5 * max |oracle_f64|, and rel-L2 <= 1e-6),
not the per-element bound of the direct kernels."""
import numpy as np
import pytest

import cogconv._lib as L
import oracle as O
from gpuhelp import desc_pair, run_gpu

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def device():
    L.init(0)
    yield
    L.synchronize()


def fft_variant(cd):
    for i in range(L.variant_count(cd)):
        p = L.Plan(cd, i)
        name = p.description
        p.close()
        if name.startswith("fft"):
            return i, name
    raise AssertionError("no FFT variant offered")


def check(got, ref, what):
    err = np.abs(got.astype(np.float64) - ref)
    scale = np.abs(ref).max()
    assert got.shape == ref.shape, what
    assert err.max() <= 1e-5 * scale, "%s: normwise %.3g" % (what, err.max() / scale)
    assert np.sqrt((err ** 2).sum() / (ref ** 2).sum()) <= 1e-6, what


@pytest.mark.parametrize("border", [O.ZERO, O.CLAMP, O.CYCLIC])
@pytest.mark.parametrize("orient", [O.XCORR, O.CONV])
@pytest.mark.parametrize("shape,fshape", [((100, 73), (31, 31)), ((64, 64), (9, 5)), ((57, 90), (33, 17))])
def test_fft_route_matches_oracle(border, orient, shape, fshape):
    rng = np.random.default_rng(shape[0] + fshape[0] + 7 * border + orient)
    x = rng.uniform(-1, 1, (1,) + shape).astype(np.float32)
    k = rng.uniform(-1, 1, (1,) + fshape).astype(np.float32)
    od, cd = desc_pair(O.PBP, border, orient, O.NONE, 1, 1, x.shape, k.shape)
    vi, name = fft_variant(cd)
    print("CASE fft %s border=%d orient=%d %s" % (name, border, orient, shape), flush=True)
    got = run_gpu(cd, x, k, variant=vi)
    check(got, O.conv(od, x, k), name)


@pytest.mark.parametrize("pa,pk", [(3, 3), (1, 4), (4, 1)])
def test_fft_route_plane_broadcast(pa, pk):
    """PlaneByPlane pairs plane t of each operand, a single plane broadcasting
    (FieldDefines.scala:101-107)."""
    rng = np.random.default_rng(pa * 10 + pk)
    x = rng.uniform(-1, 1, (pa, 48, 40)).astype(np.float32)
    k = rng.uniform(-1, 1, (pk, 21, 21)).astype(np.float32)
    od, cd = desc_pair(O.PBP, O.CLAMP, O.CONV, O.NONE, 1, 1, x.shape, k.shape)
    vi, name = fft_variant(cd)
    print("CASE fft broadcast %s" % name, flush=True)
    check(run_gpu(cd, x, k, variant=vi), O.conv(od, x, k), name)


def test_fft_route_is_the_default_for_large_filters_and_repeats():
    """A 41 x 41 filter on a 256 x 256 field plans the FFT route by default; a
    second launch of the same plan reuses its transforms and agrees bit for bit."""
    rng = np.random.default_rng(12)
    x = rng.uniform(-1, 1, (2, 256, 256)).astype(np.float32)
    k = rng.uniform(-1, 1, (2, 41, 41)).astype(np.float32)
    od, cd = desc_pair(O.PBP, O.ZERO, O.XCORR, O.NONE, 1, 1, x.shape, k.shape)
    plan = L.Plan(cd)
    print("CASE fft default %s" % plan.description, flush=True)
    assert plan.description.startswith("fft")
    ba, bk = L.DeviceBuffer.from_array(x), L.DeviceBuffer.from_array(k)
    o1, o2 = L.DeviceBuffer(x.nbytes), L.DeviceBuffer(x.nbytes)
    ws = L.DeviceBuffer(plan.workspace_bytes)
    plan.launch(ba, bk, o1, ws)
    plan.launch(ba, bk, o2, ws)
    L.synchronize()
    a, b = o1.download(x.shape), o2.download(x.shape)
    assert np.array_equal(a, b)
    check(a, O.conv(od, x, k, threads=16), plan.description)
    for buf in (ba, bk, o1, o2, ws):
        buf.close()
    plan.close()
