"""cogconv_launch_noalloc: the plan's workspace comes from the library's
per-device scratch array; results must equal cogconv_launch's (with a caller
workspace) bit for bit, on two calls in a row, the second with a smaller input."""
import ctypes

import numpy as np
import pytest

import cogconv._lib as L
import oracle as O
from gpuhelp import desc_pair


def _plans(kind, big):
    """(descriptor, operand 1, operand 2) of a plan that needs a workspace:
    FP32 FilterAdjoint+BRS (split partials) or a BF16 ProjectFrame+BRS (packed bank)."""
    rng = np.random.default_rng(11 if big else 12)
    if kind == "fa":
        B, N, K, H, F = (8, 16, 32, 40, 3) if big else (2, 8, 16, 20, 5)
        x = rng.uniform(-1, 1, (B * N, H, H)).astype(np.float32)
        dy = rng.uniform(-1, 1, (B * K, H - F + 1, H - F + 1)).astype(np.float32)
        _, cd = desc_pair(O.FA_BRS, O.VALID, O.XCORR, O.NONE, 1, B, x.shape, dy.shape)
        return cd, x, dy
    B, N, K, H, F = (4, 32, 32, 48, 3) if big else (1, 32, 16, 24, 3)
    x = rng.uniform(-1, 1, (B * N, H, H)).astype(np.float32)
    w = rng.uniform(-1, 1, (K * N, F, F)).astype(np.float32)
    _, cd = desc_pair(O.PF_BRS, O.ZERO, O.XCORR, O.NONE, 1, B, x.shape, w.shape, precision=L.BF16)
    return cd, x, w


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["fa", "bf16_frame"])
def test_noalloc_matches_launch_on_two_calls(kind):
    L.init(0)
    for big in (True, False):
        cd, a, k = _plans(kind, big)
        plan = L.Plan(cd)
        print("CASE noalloc %s big=%s %s workspace=%d" % (kind, big, plan.description, plan.workspace_bytes),
              flush=True)
        assert plan.workspace_bytes > 0
        ba, bk = L.DeviceBuffer.from_array(a), L.DeviceBuffer.from_array(k)
        o1 = L.DeviceBuffer(4 * int(np.prod(plan.out_shape)))
        o2 = L.DeviceBuffer(4 * int(np.prod(plan.out_shape)))
        ws = L.DeviceBuffer(plan.workspace_bytes)
        o2.upload(np.full(plan.out_shape, np.nan, np.float32))
        plan.launch(ba, bk, o1, ws, None)
        plan.launch_noalloc(ba, bk, o2, None)
        L.synchronize()
        ref = o1.download(plan.out_shape)
        got = o2.download(plan.out_shape)
        assert np.all(np.isfinite(ref))
        assert np.array_equal(got, ref), "noalloc differs from launch: max %.3g" % np.abs(got - ref).max()
        for b in (ba, bk, o1, o2, ws):
            b.close()
        plan.close()


def test_noalloc_rejects_oversized_workspace():
    """A plan whose workspace exceeds the scratch array is refused before any device work."""
    size = ctypes.c_size_t()
    L.check(L.lib().cogconv_noalloc_scratch_bytes(ctypes.byref(size)))
    assert size.value == 64 << 20
    # A BF16 ProjectFrame+BRS with a 512 x 1024-plane bank of 9x9 filters: an 81 MiB packed bank.
    _, cd = desc_pair(O.PF_BRS, O.ZERO, O.XCORR, O.NONE, 1, 1, (1024, 64, 64), (512 * 1024, 9, 9),
                      precision=L.BF16)
    plan = L.Plan(cd)
    assert plan.workspace_bytes > size.value
    fake = ctypes.c_void_p(256)  # never dereferenced: the size check comes first
    st = L.lib().cogconv_launch_noalloc(plan.handle, fake, fake, fake, None)
    assert st == L.ERR_UNSUPPORTED
    assert "at most" in L.last_error()
    plan.close()
