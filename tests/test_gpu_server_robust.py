"""Serving-loop liveness: admission blocked by launch slots or memory while
another worker issues, a request that can never fit once packed banks hold
part of the budget, and drain after a failed request.  Each case runs in a
daemon thread joined with a host-side time limit, so a deadlocked server
fails the test instead of hanging it."""
import threading

import numpy as np
import pytest

import cogconv._lib as L
import oracle as O
from gpuhelp import TOL_FP32, desc_pair

pytestmark = pytest.mark.gpu

LIMIT_S = 60


@pytest.fixture(scope="module", autouse=True)
def device():
    L.init(0)
    yield
    L.synchronize()


def within_limit(fn):
    box = {}

    def body():
        try:
            box["value"] = fn()
        except BaseException as e:  # re-raised in the test thread
            box["error"] = e
    t = threading.Thread(target=body, daemon=True)
    t.start()
    t.join(LIMIT_S)
    assert not t.is_alive(), "server did not finish within %d s (deadlock)" % LIMIT_S
    if "error" in box:
        raise box["error"]
    return box.get("value")


def pf_requests(rng, w, n, side):
    K = w.shape[0] // 16
    xs = [rng.uniform(-1, 1, (16, side, side)).astype(np.float32) for _ in range(n)]
    outs = [np.zeros((K, side, side), np.float32) for _ in xs]
    return xs, outs


def check_outputs(w, xs, outs):
    od = O.make_desc(O.PF_BRS, O.ZERO, O.XCORR, O.NONE, 1, 1, xs[0].shape, w.shape)
    for x, o in zip(xs, outs):
        ref = O.conv(od, x, w)
        assert np.abs(o - ref).max() <= TOL_FP32 * np.abs(ref).max()


def run_backlog(srv, w, bid, xs, outs):
    _, cd = desc_pair(O.PF_BRS, O.ZERO, O.XCORR, O.NONE, 1, 1, xs[0].shape, w.shape)
    rq = [L.Server.request(cd, x, o, bank=bid) for x, o in zip(xs, outs)]
    tickets = srv.submit_timed(rq, np.zeros(len(rq)))
    assert all(tickets)
    srv.drain()
    return srv.stats()


@pytest.mark.parametrize("max_inflight,streams,stack", [(1, 2, 1), (1, 3, 8), (2, 3, 1)])
def test_one_launch_slot_many_workers(max_inflight, streams, stack):
    print("CASE server slots max_inflight=%d streams=%d stack=%d" % (max_inflight, streams, stack), flush=True)
    rng = np.random.default_rng(11)
    w = rng.uniform(-1, 1, (16 * 16, 3, 3)).astype(np.float32)
    srv = L.Server(budget_bytes=64 << 20, max_inflight=max_inflight, streams=streams, max_stack=stack)
    bid = srv.register_bank(w)
    xs, outs = pf_requests(rng, w, 40, 32)
    st = within_limit(lambda: run_backlog(srv, w, bid, xs, outs))
    assert st["completed"] == 40
    check_outputs(w, xs, outs)
    srv.close()


def test_budget_for_two_with_three_workers():
    print("CASE server budget for two, three workers", flush=True)
    rng = np.random.default_rng(12)
    w = rng.uniform(-1, 1, (16 * 16, 3, 3)).astype(np.float32)
    side = 48
    one = 4 * 16 * side * side * 2  # in0 + out of one request (16 planes each)
    srv = L.Server(budget_bytes=2 * one + 4096, max_inflight=8, streams=3, max_stack=1)
    bid = srv.register_bank(w)
    xs, outs = pf_requests(rng, w, 30, side)
    st = within_limit(lambda: run_backlog(srv, w, bid, xs, outs))
    assert st["completed"] == 30
    assert st["peak_bytes"] <= 2 * one + 4096
    check_outputs(w, xs, outs)
    srv.close()


def test_request_that_cannot_fit_after_packing_fails_and_drain_returns():
    """A BF16 resident request packs its bank into the top of the arena; a host
    request sized between (budget - pack) and the budget then never fits.  It
    is refused at submit (if the pack already exists) or fails at admission,
    and drain returns either way."""
    print("CASE server request that never fits", flush=True)
    rng = np.random.default_rng(13)
    N, K, side = 16, 16, 64
    w = rng.uniform(-1, 1, (K * N, 3, 3)).astype(np.float32)
    _, cd16 = desc_pair(O.PF_BRS, O.ZERO, O.XCORR, O.NONE, 1, 1, (N, side, side), w.shape, precision=L.BF16)
    pack = L.Plan(cd16).workspace_bytes
    assert pack > 4096
    need = 2 * 4 * N * side * side  # host FP32 request: in0 + out
    budget = need + pack // 2
    srv = L.Server(budget_bytes=budget, max_inflight=4, streams=2, max_stack=1)
    bid = srv.register_bank(w)
    x = rng.uniform(-1, 1, (N, side, side)).astype(np.float32)
    dx, dout = L.DeviceBuffer.from_array(x), L.DeviceBuffer(4 * K * side * side)
    _, cd32 = desc_pair(O.PF_BRS, O.ZERO, O.XCORR, O.NONE, 1, 1, x.shape, w.shape)
    big_out = np.zeros((K, side, side), np.float32)
    rq = [L.Server.request(cd16, dx, dout, bank=bid, resident=True), L.Server.request(cd32, x, big_out, bank=bid)]

    def body():
        tickets = srv.submit_timed(rq, np.zeros(2))
        assert tickets[0]
        srv.wait(tickets[0])
        if tickets[1]:
            with pytest.raises(L.CogconvError) as e:
                srv.wait(tickets[1])
            assert e.value.status == L.ERR_BUDGET
        srv.drain()
    within_limit(body)
    got = dout.download((K, side, side)).astype(np.float64)
    ref = O.conv(O.make_desc(O.PF_BRS, O.ZERO, O.XCORR, O.NONE, 1, 1, x.shape, w.shape), x, w)
    assert np.abs(got - ref).max() <= 1e-2 * np.abs(ref).max()
    dx.close()
    dout.close()
    srv.close()
