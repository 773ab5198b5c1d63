"""Host-side graph compilation (no device): which fields merge into which
launch steps.  Mirrors the optimizer conditions of the reference:
  TensorReduceOptimizer.scala:59-80 (reduce chains),
  ProjectFrameTensorReduceSumOptimizer.scala:108-131 (conv + reduce),
  HyperKernel.canMergeWithSink, HyperKernel.scala:478-530 (elementwise sinks)."""
import numpy as np
import pytest

import cogconv
import cogconv._lib as L
from cogconv import api


def _frame_graph(**kw):
    rng = np.random.default_rng(0)
    cg = cogconv.ComputeGraph(**kw)
    with cg:
        x = cogconv.VectorField(rng.uniform(-1, 1, (2 * 4, 12, 12)))
        w = cogconv.VectorField(rng.uniform(-1, 1, (3 * 4, 3, 3)))
        bias = cogconv.VectorField(np.arange(6, dtype=np.float32))  # 0-D, one value per output plane
    return cg, x, w, bias


def kinds(cg):
    return [(st["kind"], len(st["epilogue"])) for st in cg.schedule().steps]


def test_conv_reduce_bias_relu_is_one_step():
    cg, x, w, bias = _frame_graph()
    with cg:
        y = cogconv.blockReduceSum(cogconv.projectFrame(x, w, cogconv.BorderZero, batchSize=2), 4)
        z = api.max(y + bias, 0.0)
        cogconv.probe(z)
    st = cg.schedule().steps
    assert kinds(cg) == [("conv", 2)]
    assert st[0]["mode"] == L.PROJECT_FRAME_BLOCK_REDUCE_SUM
    assert st[0]["result"] is z
    assert [op for op, _, _ in st[0]["epilogue"]] == [L.EPI_ADD_FIELD, L.EPI_MAX_CONST]
    assert st[0]["operands"] == [bias]


def test_probed_or_shared_intermediates_stop_merging():
    cg, x, w, bias = _frame_graph()
    with cg:
        y = cogconv.blockReduceSum(cogconv.projectFrame(x, w, cogconv.BorderZero, batchSize=2), 4)
        a = y * 2.0
        cogconv.probe(a)  # a probed link keeps its own buffer: merging stops after it
        b = cogconv.tanh(a)
        c = -b
        d = c + 1.0  # c has two sinks: d and e
        e = c * 3.0
        cogconv.probe(d, e)
    assert kinds(cg) == [("conv", 1), ("elementwise", 2), ("elementwise", 1), ("elementwise", 1)]


def test_merging_can_be_switched_off():
    cg, x, w, bias = _frame_graph()
    with cg:
        y = cogconv.blockReduceSum(cogconv.projectFrame(x, w, cogconv.BorderZero, batchSize=2), 4)
        z = cogconv.abs(y - 0.5)
        cogconv.probe(z)
    try:
        cogconv.Cog.localMemoryMerging = False
        assert kinds(cg) == [("conv", 0), ("elementwise", 1), ("elementwise", 1)]
    finally:
        cogconv.Cog.localMemoryMerging = True
    assert kinds(cg) == [("conv", 2)]
    cg2, x2, w2, _ = _frame_graph(optimize=False)
    with cg2:
        y2 = cogconv.blockReduceSum(cogconv.projectFrame(x2, w2, cogconv.BorderZero, batchSize=2), 4)
        cogconv.probe(cogconv.abs(y2 - 0.5))
    assert kinds(cg2) == [("conv", 0), ("reduce", 0), ("elementwise", 1), ("elementwise", 1)]


def test_reduce_chain_merges_then_fuses_with_the_convolution():
    """blockReduceSum(blockReduceSum(conv, 2), 2) == blockReduceSum(conv, 4), which
    then fuses into ProjectFrameBlockReduceSum."""
    cg, x, w, _ = _frame_graph()
    with cg:
        pf = cogconv.projectFrame(x, w, cogconv.BorderZero, batchSize=2)
        y = cogconv.blockReduceSum(cogconv.blockReduceSum(pf, 2), 2)
        cogconv.probe(y)
    st = cg.schedule().steps
    assert kinds(cg) == [("conv", 0)]
    assert st[0]["mode"] == L.PROJECT_FRAME_BLOCK_REDUCE_SUM
    # a probed inner reduce is kept
    cg2, x2, w2, _ = _frame_graph()
    with cg2:
        inner = cogconv.blockReduceSum(cogconv.projectFrame(x2, w2, cogconv.BorderZero, batchSize=2), 2)
        cogconv.probe(inner, cogconv.blockReduceSum(inner, 2))
    assert kinds(cg2) == [("conv", 0), ("reduce", 0), ("reduce", 0)]


def test_reduce_chain_of_a_plain_field():
    rng = np.random.default_rng(1)
    cg = cogconv.ComputeGraph()
    with cg:
        v = cogconv.VectorField(rng.uniform(-1, 1, (12, 5, 5)))
        r = cogconv.blockReduceSum(cogconv.blockReduceSum(cogconv.blockReduceSum(v, 2), 3), 2)
        cogconv.probe(r)
    st = cg.schedule().steps
    assert [(s["kind"], s["factor"]) for s in st] == [("reduce", 12)]
    assert st[0]["src"] is v


def test_elementwise_types_follow_the_reference_broadcast_rule():
    """ResultType.scala:29-58: same type, 0-D of the same tensor shape, scalar
    field of the same field shape, or a 0-D scalar."""
    rng = np.random.default_rng(2)
    with cogconv.ComputeGraph():
        v = cogconv.VectorField(rng.uniform(-1, 1, (3, 4, 5)))
        s = cogconv.ScalarField(rng.uniform(-1, 1, (4, 5)))
        z3 = cogconv.VectorField(np.ones(3, np.float32))
        z1 = cogconv.ScalarField(2.0)
        assert (v + s).ftype == v.ftype
        assert (s + v).ftype == v.ftype
        assert (v * z3).ftype == v.ftype
        assert (z1 - v).ftype == v.ftype
        other = cogconv.VectorField(rng.uniform(-1, 1, (2, 4, 5)))
        with pytest.raises(L.CogconvError):
            v + other
        # the operand order is kept for non-commutative ops: s - v runs as RSUB on v
        n = s - v
        assert n.inputs[0] is v and n.params["ops"] == [(L.EPI_RSUB_FIELD, 0.0)]


def test_merged_step_runs_after_its_operands():
    """The bias is computed by another convolution created after the first one:
    the merged step must run after it."""
    rng = np.random.default_rng(3)
    cg = cogconv.ComputeGraph()
    with cg:
        x = cogconv.ScalarField(rng.uniform(-1, 1, (16, 16)))
        k = cogconv.ScalarField(rng.uniform(-1, 1, (3, 3)))
        a = cogconv.convolve(x, k, cogconv.BorderZero)
        b = cogconv.crossCorrelate(x, k, cogconv.BorderClamp)
        c = a + b
        cogconv.probe(c, b)
    st = cg.schedule().steps
    assert [s["result"] for s in st] == [b, c]
    assert st[1]["operands"] == [b]


def test_separable_node_checks():
    with cogconv.ComputeGraph():
        x = cogconv.ScalarField(np.zeros((8, 8), np.float32))
        r = cogconv.ScalarField(np.zeros(5, np.float32))
        c = cogconv.ScalarField(np.zeros(3, np.float32))
        assert cogconv.convolveSeparable(x, r, c, cogconv.BorderZero).ftype == x.ftype
        assert x.crossCorrelateSeparable(r, c, cogconv.BorderClamp).ftype == x.ftype
        with pytest.raises(L.CogconvError):
            cogconv.convolveSeparable(x, cogconv.ScalarField(np.zeros(4, np.float32)), c, cogconv.BorderZero)
        with pytest.raises(L.CogconvError):
            cogconv.convolveSeparable(x, r, c, cogconv.BorderValid)
        v = cogconv.VectorField(np.zeros((2, 8, 8), np.float32))
        with pytest.raises(L.CogconvError):
            cogconv.convolveSeparable(v, r, c, cogconv.BorderZero)


def test_field_method_forms_build_the_same_ops():
    """Field.scala:1467-1633 forward to the CogFunctionAPI functions."""
    rng = np.random.default_rng(4)
    with cogconv.ComputeGraph():
        x = cogconv.VectorField(rng.uniform(-1, 1, (4, 10, 10)))
        w = cogconv.VectorField(rng.uniform(-1, 1, (8, 3, 3)))
        for meth, fn in (("projectFrame", cogconv.projectFrame), ("backProjectFrame", cogconv.backProjectFrame)):
            a = getattr(x, meth)(w, cogconv.BorderZero, batchSize=2)
            b = fn(x, w, cogconv.BorderZero, batchSize=2)
            assert a.ftype == b.ftype and a.params == b.params
        s = cogconv.ScalarField(rng.uniform(-1, 1, (10, 10)))
        k = cogconv.ScalarField(rng.uniform(-1, 1, (5, 5)))
        assert s.convolve(k, cogconv.BorderCyclic).params["orientation"] == L.CONVOLUTION
        assert s.crossCorrelate(k, cogconv.BorderClamp).params["orientation"] == L.CROSS_CORRELATION
        dy = cogconv.VectorField(rng.uniform(-1, 1, (4, 8, 8)))
        fa = x.crossCorrelateFilterAdjoint(dy, cogconv.BorderValid, batchSize=2)
        assert fa.params["mode"] == L.FILTER_ADJOINT
        pfb = x._projectFrameBlockReduceSum(w, cogconv.BorderZero, batchSize=2)
        assert pfb.params["mode"] == L.PROJECT_FRAME_BLOCK_REDUCE_SUM
        assert pfb.ftype == cogconv.blockReduceSum(cogconv.projectFrame(x, w, cogconv.BorderZero, batchSize=2),
                                                   2).ftype
