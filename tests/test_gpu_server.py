"""The serving loop (cogconv_server_*): continuous admission, resident banks,
shape-bucket stacking, and results that are the same alone and in the mix."""
import numpy as np
import pytest

import cogconv._lib as L
import oracle as O
from gpuhelp import TOL_FP32, desc_pair

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def device():
    L.init(0)
    yield
    L.synchronize()


def make_requests(rng, banks, n):
    """A mix of ProjectFrame / BackProjectFrame (with banks) and FilterAdjoint (dy sent along)."""
    reqs = []
    for i in range(n):
        bid = int(rng.integers(len(banks)))
        w, N, K, F = banks[bid]
        side = int(rng.choice([12, 16, 24]))
        kind = ("pf", "bpf", "fa")[i % 3]
        if kind == "pf":
            x = rng.uniform(-1, 1, (N, side, side)).astype(np.float32)
            od, cd = desc_pair(O.PF_BRS, O.ZERO, O.XCORR, O.NONE, 1, 1, x.shape, w.shape)
            reqs.append((cd, od, x, None, bid, w))
        elif kind == "bpf":
            dy = rng.uniform(-1, 1, (K, side, side)).astype(np.float32)
            od, cd = desc_pair(O.BPF_BRS, O.ZERO, O.CONV, O.NONE, 1, 1, dy.shape, w.shape)
            reqs.append((cd, od, dy, None, bid, w))
        else:
            x = rng.uniform(-1, 1, (N, side, side)).astype(np.float32)
            dy = rng.uniform(-1, 1, (K, side - F + 1, side - F + 1)).astype(np.float32)
            od, cd = desc_pair(O.FA_BRS, O.VALID, O.XCORR, O.NONE, 1, 1, x.shape, dy.shape)
            reqs.append((cd, od, x, dy, -1, None))
    return reqs


def run(server, reqs, one_at_a_time):
    """Alone: one request at a time.  In the mix: all requests queued natively at
    once (cogconv_server_submit_timed, zero arrival times), so that admission
    finds a backlog and stacks requests of one bucket into shared launches."""
    outs = [np.zeros(O.output_type(od)[3:4] + O.output_type(od)[1:3], np.float32) for _, od, *_ in reqs]
    if one_at_a_time:
        for (cd, od, a, b, bid, _), out in zip(reqs, outs):
            server.wait(server.submit(cd, a, out, bank=bid, in1=b))
        return outs
    rq = [L.Server.request(cd, a, out, bank=bid, in1=b) for (cd, od, a, b, bid, _), out in zip(reqs, outs)]
    tickets = server.submit_timed(rq, np.zeros(len(rq)))
    assert all(tickets)
    for t in tickets:
        server.wait(t)
    return outs


def test_alone_and_in_the_mix_agree():
    print("CASE server alone vs mix", flush=True)
    rng = np.random.default_rng(5)
    banks_np = []
    srv = L.Server(budget_bytes=64 << 20, max_inflight=6, streams=3, max_stack=8)
    for _ in range(3):
        N, K, F = 8, 16, 3
        w = rng.uniform(-1, 1, (K * N, F, F)).astype(np.float32)
        banks_np.append((w, N, K, F))
    for w, *_ in banks_np:
        srv.register_bank(w)
    reqs = make_requests(rng, banks_np, 48)
    alone = run(srv, reqs, one_at_a_time=True)
    srv.reset_stats()
    mixed = run(srv, reqs, one_at_a_time=False)
    st = srv.stats()
    print("stats", st)
    assert st["completed"] == len(reqs)
    assert st["stacked"] > 0, "no request shared a launch"
    for (cd, od, a, b, bid, w), x, y in zip(reqs, alone, mixed):
        assert np.array_equal(x, y)
        ref = O.conv(od, a, w if bid >= 0 else b)
        assert np.abs(x - ref).max() <= TOL_FP32 * max(1.0, np.abs(ref).max())
    srv.close()


def test_budget_admission_and_rejection():
    print("CASE server budget", flush=True)
    rng = np.random.default_rng(6)
    N, K, F = 16, 16, 3
    w = rng.uniform(-1, 1, (K * N, F, F)).astype(np.float32)
    # Room for about two requests at a time: the rest wait for memory.
    one = 4 * N * 64 * 64 + 4 * K * 64 * 64
    srv = L.Server(budget_bytes=2 * one + 4096, max_inflight=8, streams=2, max_stack=1)
    bid = srv.register_bank(w)
    xs = [rng.uniform(-1, 1, (N, 64, 64)).astype(np.float32) for _ in range(10)]
    outs = [np.zeros((K, 64, 64), np.float32) for _ in xs]
    for x, o in zip(xs, outs):
        _, cd = desc_pair(O.PF_BRS, O.ZERO, O.XCORR, O.NONE, 1, 1, x.shape, w.shape)
        srv.submit(cd, x, o, bank=bid)
    srv.drain()
    st = srv.stats()
    assert st["completed"] == 10
    assert st["peak_bytes"] <= 2 * one + 4096
    od = O.make_desc(O.PF_BRS, O.ZERO, O.XCORR, O.NONE, 1, 1, xs[0].shape, w.shape)
    for x, o in zip(xs, outs):
        ref = O.conv(od, x, w)
        assert np.abs(o - ref).max() <= TOL_FP32 * np.abs(ref).max()
    # A request bigger than the whole budget is refused at submit.
    big = rng.uniform(-1, 1, (N, 256, 256)).astype(np.float32)
    _, cd = desc_pair(O.PF_BRS, O.ZERO, O.XCORR, O.NONE, 1, 1, big.shape, w.shape)
    with pytest.raises(L.CogconvError) as e:
        srv.submit(cd, big, np.zeros((K, 256, 256), np.float32), bank=bid)
    assert e.value.status == L.ERR_BUDGET
    srv.close()


def test_illegal_request_fails_at_submit():
    srv = L.Server(budget_bytes=16 << 20)
    w = np.zeros((4, 4, 4), np.float32)
    bid = srv.register_bank(w)
    x = np.zeros((2, 16, 16), np.float32)
    _, cd = desc_pair(O.PF_BRS, O.ZERO, O.XCORR, O.NONE, 1, 1, x.shape, w.shape)
    with pytest.raises(L.CogconvError) as e:
        srv.submit(cd, x, np.zeros((2, 16, 16), np.float32), bank=bid)
    assert e.value.status == L.ERR_ILLEGAL_OP  # even filter with BorderZero
    srv.close()
