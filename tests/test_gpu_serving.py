"""The serving loop at the serving workload's own shapes, and its request-table
launches.

* cogconv_launch_table equals cogconv_launch image for image, bit for bit.
* Resident requests of one bucket stack into one table launch; hot launches
  are replayed from captured hipGraphs; each result equals the same request run
  alone, bit for bit.
* A seeded sample of 64 requests drawn the way the serving benchmark draws its
  trace (sides 32..256 step 16, 16/32/64 planes, 3x3 / 5x5 filters with 16..64
  outputs, 70/20/10 ProjectFrame / BackProjectFrame / FilterAdjoint, 64 banks)
  runs through the server in FP32 and BF16, resident and in host memory, with
  stacking on; FP32 results meet the per-element bound against the FP64 oracle,
  BF16 results 1e-2 normwise, and every result is the same alone and in the mix.
"""
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


@pytest.mark.parametrize("precision", [L.FP32, L.BF16])
@pytest.mark.parametrize("mode", [O.PF_BRS, O.BPF_BRS])
def test_table_launch_matches_launch(precision, mode):
    rng = np.random.default_rng(40 + precision * 2 + mode)
    B, N, K, H, W, F = 3, 32, 16, 20, 36, 3
    orient = O.CONV if mode == O.BPF_BRS else O.XCORR
    a = rng.uniform(-1, 1, (B * N, H, W)).astype(np.float32)
    k = rng.uniform(-1, 1, (K * N, F, F)).astype(np.float32)
    _, cd = desc_pair(mode, O.ZERO, orient, O.NONE, 1, B, a.shape, k.shape, precision=precision)
    plan = L.Plan(cd)
    print("CASE table", precision, mode, plan.description, flush=True)
    assert plan.table_supported()
    per_out = int(np.prod(plan.out_shape)) // B
    bk, ba = L.DeviceBuffer.from_array(k), L.DeviceBuffer.from_array(a)
    ref = L.DeviceBuffer(4 * per_out * B)
    ws = L.DeviceBuffer(plan.workspace_bytes) if plan.workspace_bytes else None
    plan.launch(ba, bk, ref, ws)
    # each image in its own buffer, in reverse order in memory
    imgs = [L.DeviceBuffer.from_array(a[b * N:(b + 1) * N]) for b in range(B)]
    outs = [L.DeviceBuffer(4 * per_out) for _ in range(B)]
    tab = L.DeviceBuffer(16 * B)
    plan.launch_table(list(zip(imgs, outs)), tab, in1=bk, workspace=ws)
    L.synchronize()
    want = ref.download(plan.out_shape)
    got = np.concatenate([o.download((plan.out_shape[0] // B,) + plan.out_shape[1:]) for o in outs])
    assert np.array_equal(got, want)
    if ws is not None:  # on a bank packed once
        wsp = L.DeviceBuffer(plan.workspace_bytes)
        assert plan.pack_operand(bk, wsp)
        outs2 = [L.DeviceBuffer(4 * per_out) for _ in range(B)]
        plan.launch_table(list(zip(imgs, outs2)), tab, workspace=wsp, packed=True)
        L.synchronize()
        got2 = np.concatenate([o.download((plan.out_shape[0] // B,) + plan.out_shape[1:]) for o in outs2])
        assert np.array_equal(got2, want)
    plan.close()


def test_resident_stacking_and_graph_replay():
    """Resident requests of a few buckets, queued at once: they stack into table
    launches and, once a (plan, bank, stream) is hot, replay captured graphs."""
    print("CASE resident stacking", flush=True)
    rng = np.random.default_rng(50)
    N, K, F = 32, 16, 3
    banks = [rng.uniform(-1, 1, (K * N, F, F)).astype(np.float32) for _ in range(2)]
    for precision in (L.FP32, L.BF16):
        srv = L.Server(budget_bytes=256 << 20, max_inflight=8, streams=2, max_stack=4, graph_after=2)
        bids = [srv.register_bank(w) for w in banks]
        reqs = []
        for i in range(48):
            side = (24, 40)[i % 2]
            bank = i % 2
            x = rng.uniform(-1, 1, (N, side, side)).astype(np.float32)
            od, cd = desc_pair(O.PF_BRS, O.ZERO, O.XCORR, O.NONE, 1, 1, x.shape, banks[bank].shape,
                               precision=precision)
            reqs.append((cd, od, x, bank, L.DeviceBuffer.from_array(x), L.DeviceBuffer(4 * K * side * side)))
        # alone
        alone = []
        for cd, od, x, bank, bx, bo in reqs:
            srv.wait(srv.submit(cd, bx, bo.ptr, bank=bids[bank]))
            alone.append(bo.download((K,) + x.shape[1:]))
        srv.reset_stats()
        rq = [L.Server.request(cd, bx, bo, bank=bids[bank], resident=True) for cd, od, x, bank, bx, bo in reqs]
        tickets = srv.submit_timed(rq, np.zeros(len(rq)))
        for t in tickets:
            srv.wait(t)
        st = srv.stats()
        print("stats", precision, st, flush=True)
        assert st["completed"] == len(reqs)
        assert st["stacked"] > 0, "no resident request shared a launch"
        assert st["graph_launches"] > 0, "no launch was replayed from a graph"
        for (cd, od, x, bank, bx, bo), a in zip(reqs, alone):
            got = bo.download(a.shape)
            assert np.array_equal(got, a)
            ref = O.conv(od, x, banks[bank])
            tol = TOL_FP32 if precision == L.FP32 else 1e-2
            assert np.abs(got - ref).max() <= tol * np.abs(ref).max()
        srv.close()
        for r in reqs:
            r[4].close()
            r[5].close()


def c5_sample(n, seed=2000):
    """n requests drawn as the serving benchmark draws its trace."""
    rng = np.random.default_rng(seed)
    banks = []
    for _ in range(64):
        N = int(rng.choice([16, 32, 64]))
        K = int(rng.choice([16, 32, 48, 64]))
        F = int(rng.choice([3, 5]))
        banks.append((N, K, F, (rng.standard_normal((K * N, F, F)) * np.sqrt(2.0 / (N * F * F))).astype(np.float32)))
    kinds = rng.choice(["pf", "bpf", "fa"], size=n, p=[0.7, 0.2, 0.1])
    bank_of = rng.integers(0, 64, size=n)
    side_of = rng.choice(list(range(32, 257, 16)), size=n)
    reqs = []
    for i in range(n):
        N, K, F, w = banks[bank_of[i]]
        s = int(side_of[i])
        if kinds[i] == "pf":
            x = rng.uniform(-1, 1, (N, s, s)).astype(np.float32)
            reqs.append(("pf", int(bank_of[i]), x, None, (O.PF_BRS, O.ZERO, O.XCORR), w))
        elif kinds[i] == "bpf":
            dy = rng.uniform(-1, 1, (K, s, s)).astype(np.float32)
            reqs.append(("bpf", int(bank_of[i]), dy, None, (O.BPF_BRS, O.ZERO, O.CONV), w))
        else:
            x = rng.uniform(-1, 1, (N, s, s)).astype(np.float32)
            dy = rng.uniform(-1, 1, (K, s - F + 1, s - F + 1)).astype(np.float32)
            reqs.append(("fa", -1, x, dy, (O.FA_BRS, O.VALID, O.XCORR), None))
    return banks, reqs


def test_serving_sample_against_oracle():
    banks, reqs = c5_sample(64)
    # Oracle once per request (FP64), with the magnitude run for the FP32 bound.
    refs = []
    for kind, bank, a, b, (mode, border, orient), w in reqs:
        k = w if b is None else b
        od = O.make_desc(mode, border, orient, O.NONE, 1, 1, a.shape, k.shape)
        refs.append((od, O.conv(od, a, k), O.conv(od, np.abs(a), np.abs(k))))
    for precision in (L.FP32, L.BF16):
        for resident in (True, False):
            print("CASE c5 sample precision=%d resident=%s" % (precision, resident), flush=True)
            srv = L.Server(budget_bytes=4 << 30, max_inflight=16, streams=4, max_stack=8, graph_after=2)
            bids = [srv.register_bank(w) for (_, _, _, w) in banks]
            bufs, descs = [], []
            for (kind, bank, a, b, (mode, border, orient), w), (od, ref, _) in zip(reqs, refs):
                k = w if b is None else b
                _, cd = desc_pair(mode, border, orient, O.NONE, 1, 1, a.shape, k.shape, precision=precision)
                descs.append(cd)
                if resident:
                    bufs.append((L.DeviceBuffer.from_array(a), L.DeviceBuffer.from_array(b) if b is not None else None,
                                 L.DeviceBuffer(ref.size * 4)))
                else:
                    ha, ho = L.HostBuffer(a.shape), L.HostBuffer(ref.shape)
                    ha.array[:] = a
                    hb = None
                    if b is not None:
                        hb = L.HostBuffer(b.shape)
                        hb.array[:] = b
                    bufs.append((ha, hb, ho))

            def addr(x):
                if x is None:
                    return None
                return x.ptr if isinstance(x, L.DeviceBuffer) else x.array

            def result(i):
                o = bufs[i][2]
                shape = refs[i][1].shape
                return o.download(shape) if isinstance(o, L.DeviceBuffer) else o.array.reshape(shape).copy()

            def submit_all():
                rq = [L.Server.request(cd, addr(ba), addr(bo), bank=bids[r[1]] if r[1] >= 0 else -1, in1=addr(bb),
                                       resident=resident) for cd, (ba, bb, bo), r in zip(descs, bufs, reqs)]
                tk = srv.submit_timed(rq, np.zeros(len(rq)))
                assert all(tk)
                for t in tk:
                    srv.wait(t)
            # alone: one at a time
            alone = []
            for i, (cd, (ba, bb, bo), r) in enumerate(zip(descs, bufs, reqs)):
                rq = [L.Server.request(cd, addr(ba), addr(bo), bank=bids[r[1]] if r[1] >= 0 else -1, in1=addr(bb),
                                       resident=resident)]
                srv.wait(srv.submit_timed(rq, [0.0])[0])
                alone.append(result(i))
            srv.reset_stats()
            submit_all()
            st = srv.stats()
            print("stats", st, flush=True)
            assert st["completed"] == len(reqs)
            for i, ((od, ref, mag), a) in enumerate(zip(refs, alone)):
                got = result(i)
                assert np.array_equal(got, a), "request %d differs alone and in the mix" % i
                d = np.abs(got.astype(np.float64) - ref)
                if precision == L.FP32:
                    assert np.all(d <= TOL_FP32 * mag + 1e-30), "request %d (%s)" % (i, reqs[i][0])
                else:
                    assert d.max() <= 1e-2 * np.abs(ref).max(), "request %d (%s)" % (i, reqs[i][0])
            srv.close()
            for t in bufs:
                for x in t:
                    if x is not None:
                        x.close()
