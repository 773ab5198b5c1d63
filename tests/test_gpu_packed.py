"""Resident operand 2: cogconv_plan_pack_operand + cogconv_launch_packed give the
same bits as cogconv_launch, and the serving loop's packed resident banks give
the same bits as a plain launch of the same plan."""
import numpy as np
import pytest

import cogconv._lib as L
import oracle as O
from gpuhelp import desc_pair

pytestmark = pytest.mark.gpu

NAMES = {L.BF16: "bf16", L.FP16: "fp16", L.SPLIT_BF16: "split_bf16"}


@pytest.fixture(scope="module", autouse=True)
def device():
    L.init(0)
    yield
    L.synchronize()


@pytest.mark.parametrize("precision", (L.BF16, L.FP16, L.SPLIT_BF16))
@pytest.mark.parametrize("mode", (O.PF_BRS, O.BPF_BRS))
@pytest.mark.parametrize("W", (68, 37))  # 16-byte and 4-byte staging
def test_launch_packed_matches_launch(precision, mode, W):
    rng = np.random.default_rng(31 + precision + mode + W)
    B, N, K, H, F = 2, 40, 24, 21, 3
    orient = O.CONV if mode == O.BPF_BRS else O.XCORR
    a1 = rng.uniform(-1, 1, (B * N, H, W)).astype(np.float32)
    a2 = rng.uniform(-1, 1, (B * N, H, W)).astype(np.float32)
    k = rng.uniform(-1, 1, (K * N, F, F)).astype(np.float32)
    od, cd = desc_pair(mode, O.ZERO, orient, O.NONE, 1, B, a1.shape, k.shape, precision=precision)
    plan = L.Plan(cd)
    print("CASE packed", NAMES[precision], mode, W, plan.description, flush=True)
    assert plan.workspace_bytes > 0
    bk = L.DeviceBuffer.from_array(k)
    ws_a = L.DeviceBuffer(plan.workspace_bytes)
    ws_p = L.DeviceBuffer(plan.workspace_bytes)
    assert plan.pack_operand(bk, ws_p)
    for a in (a1, a2):  # the packed form serves several launches
        ba = L.DeviceBuffer.from_array(a)
        o1 = L.DeviceBuffer(4 * int(np.prod(plan.out_shape)))
        o2 = L.DeviceBuffer(4 * int(np.prod(plan.out_shape)))
        plan.launch(ba, bk, o1, ws_a)
        plan.launch_packed(ba, o2, ws_p)
        L.synchronize()
        r1, r2 = o1.download(plan.out_shape), o2.download(plan.out_shape)
        assert np.array_equal(r1, r2)
        for b in (ba, o1, o2):
            b.close()
    for b in (bk, ws_a, ws_p):
        b.close()
    plan.close()


def test_pack_operand_is_a_no_op_for_fp32():
    print("CASE packed fp32", flush=True)
    od, cd = desc_pair(O.PF_BRS, O.ZERO, O.XCORR, O.NONE, 1, 2, (2 * 8, 16, 16), (4 * 8, 3, 3))
    plan = L.Plan(cd)
    bk = L.DeviceBuffer(4 * 4 * 8 * 9)
    assert not plan.pack_operand(bk, None)
    with pytest.raises(L.CogconvError):
        plan.launch_packed(bk, bk, bk)
    bk.close()
    plan.close()


def test_server_resident_packed_bank_matches_launch():
    """Resident BF16 requests on a registered bank: the server packs the bank once
    per plan and runs every request on the packed form; each result equals a plain
    launch of the same plan bit for bit."""
    print("CASE server packed bank", flush=True)
    rng = np.random.default_rng(77)
    N, K, F, H, W = 16, 32, 3, 24, 32
    w = rng.uniform(-1, 1, (K * N, F, F)).astype(np.float32)
    srv = L.Server(budget_bytes=64 << 20, max_inflight=6, streams=3, max_stack=1)
    bid = srv.register_bank(w)
    bw = L.DeviceBuffer.from_array(w)
    od, cd = desc_pair(O.PF_BRS, O.ZERO, O.XCORR, O.NONE, 1, 1, (N, H, W), w.shape, precision=L.BF16)
    plan = L.Plan(cd)
    ws = L.DeviceBuffer(plan.workspace_bytes) if plan.workspace_bytes else None
    ins, outs, refs = [], [], []
    for i in range(12):
        x = rng.uniform(-1, 1, (N, H, W)).astype(np.float32)
        bx = L.DeviceBuffer.from_array(x)
        bo = L.DeviceBuffer(4 * int(np.prod(plan.out_shape)))
        br = L.DeviceBuffer(4 * int(np.prod(plan.out_shape)))
        plan.launch(bx, bw, br, ws)
        ins.append(bx)
        outs.append(bo)
        refs.append(br)
    L.synchronize()
    tickets = [srv.submit(cd, bx, bo, bank=bid) for bx, bo in zip(ins, outs)]
    for t in tickets:
        srv.wait(t)
    assert srv.stats()["completed"] == len(tickets)
    for bo, br in zip(outs, refs):
        assert np.array_equal(bo.download(plan.out_shape), br.download(plan.out_shape))
    ref = O.conv(od, ins[0].download((N, H, W)), w)
    got = outs[0].download(plan.out_shape)
    assert np.abs(got - ref).max() <= 1e-2 * np.abs(ref).max()
    srv.close()
    for b in ins + outs + refs + [bw] + ([ws] if ws else []):
        b.close()
    plan.close()


@pytest.mark.parametrize("precision", (L.BF16, L.FP16))
def test_server_packed_bank_shared_across_field_sizes(precision):
    """One registered bank serves resident requests of several field sizes and
    both frame modes in one mix: plans that pack the bank identically share one
    packed copy (the packed layout does not depend on the field size), and every
    result equals a plain launch of its own plan bit for bit."""
    print("CASE server shared packed bank", NAMES[precision], flush=True)
    rng = np.random.default_rng(88 + precision)
    N, K, F = 32, 32, 5
    w = rng.uniform(-1, 1, (K * N, F, F)).astype(np.float32)
    srv = L.Server(budget_bytes=256 << 20, max_inflight=8, streams=4, max_stack=1)
    bid = srv.register_bank(w)
    bw = L.DeviceBuffer.from_array(w)
    cases = []
    for i, (mode, side) in enumerate([(m, s) for s in (24, 40, 72, 37) for m in (O.PF_BRS, O.BPF_BRS)] * 2):
        orient = O.CONV if mode == O.BPF_BRS else O.XCORR
        x = rng.uniform(-1, 1, (N if mode == O.PF_BRS else K, side, side)).astype(np.float32)
        _, cd = desc_pair(mode, O.ZERO, orient, O.NONE, 1, 1, x.shape, w.shape, precision=precision)
        plan = L.Plan(cd)
        ws = L.DeviceBuffer(plan.workspace_bytes) if plan.workspace_bytes else None
        bx = L.DeviceBuffer.from_array(x)
        bo = L.DeviceBuffer(4 * int(np.prod(plan.out_shape)))
        br = L.DeviceBuffer(4 * int(np.prod(plan.out_shape)))
        plan.launch(bx, bw, br, ws)
        cases.append((cd, plan, ws, bx, bo, br))
    L.synchronize()
    tickets = [srv.submit(cd, bx, bo, bank=bid) for cd, _, _, bx, bo, _ in cases]
    for t in tickets:
        srv.wait(t)
    assert srv.stats()["completed"] == len(cases)
    for cd, plan, ws, bx, bo, br in cases:
        assert np.array_equal(bo.download(plan.out_shape), br.download(plan.out_shape)), plan.description
    srv.close()
    for cd, plan, ws, bx, bo, br in cases:
        for b in (ws, bx, bo, br):
            if b is not None:
                b.close()
        plan.close()
    bw.close()
