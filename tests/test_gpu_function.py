"""Checkpoint / restore of compiled graphs and the reentrant evaluate()
Function Interface (cogconv_function_*).

Reference: ComputeGraph.write / readFromFile (ComputeGraph.scala:523-666), whose
restored kernels re-run saved OpenCL source (OpenCLRestoredDeviceKernel.scala:28-60);
here restoring re-plans each saved step.  The Function Interface follows
CogJNIWrapper.java:85-161 (create, inputs / outputs, evaluate, shutdown)."""
import threading

import numpy as np
import pytest

import cogconv
import cogconv._lib as L
import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def device():
    L.init(0)
    yield
    L.synchronize()


def layer_graph(x, w, bias, dy, precision=L.FP32):
    """y = max(PF+BRS(x, W) + bias, 0); dW = FA+BRS(x, dy); s = separable(x plane 0)."""
    B, N = 2, x.shape[0] // 2
    cg = cogconv.ComputeGraph(precision=precision)
    with cg:
        xf, wf, bf, dyf = (cogconv.VectorField(a) for a in (x, w, bias, dy))
        y = cogconv.max(cogconv.blockReduceSum(cogconv.projectFrame(xf, wf, cogconv.BorderZero, batchSize=B), N) + bf,
                        0.0)
        dw = cogconv.blockReduceSum(cogconv.crossCorrelateFilterAdjoint(xf, dyf, cogconv.BorderValid, batchSize=B), B)
        x0 = cogconv.ScalarField(x[0])
        s = cogconv.tanh(cogconv.convolveSeparable(x0, cogconv.ScalarField(np.array([0.25, 0.5, 0.25], np.float32)),
                                                   cogconv.ScalarField(np.array([1.0, 0.0, -1.0], np.float32)),
                                                   cogconv.BorderClamp))
        cogconv.probe(y, dw, s)
    return cg, (xf, dyf, x0), (y, dw, s)


def make_inputs(seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, (2 * 8, 20, 20)).astype(np.float32)
    dy = rng.uniform(-1, 1, (2 * 16, 18, 18)).astype(np.float32)
    return x, dy


def test_write_restore_evaluate_matches_the_graph(tmp_path):
    print("CASE function restore", flush=True)
    rng = np.random.default_rng(1)
    x, dy = make_inputs(2)
    w = rng.uniform(-1, 1, (16 * 8, 3, 3)).astype(np.float32)
    bias = rng.uniform(-1, 1, 2 * 16).astype(np.float32)
    cg, (xf, dyf, x0), outs = layer_graph(x, w, bias, dy)
    path = cg.write(str(tmp_path / "layer.cgraph"), inputs=[xf, dyf, x0], name="layer", description="test graph")
    cg.step()
    want = [cg.read(o) for o in outs]
    cg.release()
    fn = cogconv.readFromFile(path)
    assert fn.name == "layer"
    assert [t for _, t in fn.inputs] == [xf.ftype, dyf.ftype, x0.ftype]
    assert [t for _, t in fn.outputs] == [o.ftype for o in outs]
    assert fn.optimal_thread_count() >= 1
    got = fn.evaluate([x, dy, x[0]])
    for g, w_ in zip(got, want):
        assert np.array_equal(g, w_)
    # and against the oracle on new inputs
    x2, dy2 = make_inputs(3)
    y2, dw2, _ = fn.evaluate([x2, dy2, x2[0]])
    ref_y = np.maximum(O.conv(O.make_desc(O.PF_BRS, O.ZERO, O.XCORR, O.NONE, 1, 2, x2.shape, w.shape), x2, w) +
                       bias[:, None, None], 0)
    ref_dw = O.conv(O.make_desc(O.FA_BRS, O.VALID, O.XCORR, O.NONE, 1, 2, x2.shape, dy2.shape), x2, dy2)
    assert np.abs(y2 - ref_y).max() <= 1e-5 * np.abs(ref_y).max()
    assert np.abs(dw2 - ref_dw).max() <= 1e-5 * np.abs(ref_dw).max()
    fn.close()


def test_evaluate_is_reentrant(tmp_path):
    """8 host threads evaluate at once, 4 calls each on their own inputs; every
    result equals a serial evaluate of the same inputs bit for bit."""
    print("CASE function threads", flush=True)
    rng = np.random.default_rng(4)
    x, dy = make_inputs(5)
    w = rng.uniform(-1, 1, (16 * 8, 3, 3)).astype(np.float32)
    bias = rng.uniform(-1, 1, 2 * 16).astype(np.float32)
    cg, ins, _ = layer_graph(x, w, bias, dy, precision=L.BF16)
    path = cg.write(str(tmp_path / "layer_bf16.cgraph"), inputs=list(ins))
    fn = cogconv.readFromFile(path)
    jobs = [make_inputs(100 + i) for i in range(32)]
    serial = [fn.evaluate([a, b, a[0]]) for a, b in jobs]
    results = [None] * len(jobs)
    errors = []

    def worker(t):
        try:
            for i in range(t, len(jobs), 8):
                a, b = jobs[i]
                results[i] = fn.evaluate([a, b, a[0]])
        except Exception as e:  # surfaced below
            errors.append(e)
    threads = [threading.Thread(target=worker, args=(t,)) for t in range(8)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not errors, errors
    for s, r in zip(serial, results):
        for u, v in zip(s, r):
            assert np.array_equal(u, v)
    fn.close()
