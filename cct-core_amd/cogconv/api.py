"""Field / ComputeGraph operator API for the convolution path.

Mirrors the reference's user API for this path so that graph code reads the
same way:
  convolve / crossCorrelate               src/main/scala/cogx/api/CogFunctionAPI.scala:1245-1280
  projectFrame / backProjectFrame         CogFunctionAPI.scala:1370-1377, 1425-1432
  crossCorrelateFilterAdjoint / convolveFilterAdjoint   CogFunctionAPI.scala:1480-1510
  blockReduceSum                          CogFunctionAPI.scala:755-756
  ComputeGraph step / read / release      src/main/scala/cogx/runtime/ComputeGraph.scala:228-457
Border and sampling policies are the reference's objects: BorderZero,
BorderClamp, BorderCyclic, BorderValid, BorderFull, NoSamplingConvolution,
UpsampleInputConvolution(step), DownsampleOutputConvolution(step).

Graph building raises immediately on an illegal op (as the reference raises at
graph construction): the legality checks and result types come from
libcogconv's cogconv_output_type.  On `step`, each op runs through libcogconv;
a convolution whose only consumer is a matching blockReduceSum is fused into
the *BlockReduceSum vector mode, as ProjectFrameTensorReduceSumOptimizer does
(src/main/scala/cogx/compiler/optimizer/ProjectFrameTensorReduceSumOptimizer.scala:64-179).

Field data are numpy arrays in planar layout: a ScalarField is (rows, cols) or
(cols,) for 1-D; a VectorField is (planes, rows, cols) or (planes, cols).
"""
import threading

import numpy as np

from . import _lib as L


class _Policy:
    def __init__(self, name, code):
        self.name, self.code = name, code

    def __repr__(self):
        return self.name


BorderZero = _Policy("BorderZero", L.BORDER_ZERO)
BorderClamp = _Policy("BorderClamp", L.BORDER_CLAMP)
BorderCyclic = _Policy("BorderCyclic", L.BORDER_CYCLIC)
BorderValid = _Policy("BorderValid", L.BORDER_VALID)
BorderFull = _Policy("BorderFull", L.BORDER_FULL)


class _Sampling:
    def __init__(self, name, code, step):
        self.name, self.code, self.step = name, code, step

    def __repr__(self):
        return "%s(%d)" % (self.name, self.step) if self.code != L.NO_SAMPLING else self.name

    def __eq__(self, other):
        return isinstance(other, _Sampling) and (self.code, self.step) == (other.code, other.step)

    def __hash__(self):
        return hash((self.code, self.step))


NoSamplingConvolution = _Sampling("NoSamplingConvolution", L.NO_SAMPLING, 1)


def UpsampleInputConvolution(step):
    return _Sampling("UpsampleInputConvolution", L.UPSAMPLE_INPUT, int(step))


def DownsampleOutputConvolution(step):
    return _Sampling("DownsampleOutputConvolution", L.DOWNSAMPLE_OUTPUT, int(step))


_tls = threading.local()


def _current_graph():
    g = getattr(_tls, "graph", None)
    if g is None:
        raise RuntimeError("fields must be created inside `with ComputeGraph() as cg:`")
    return g


class Field:
    """A node of a ComputeGraph: an input with data, or the result of an op."""

    def __init__(self, graph, ftype, op=None, inputs=(), params=None, data=None):
        self.graph = graph
        self.ftype = tuple(ftype)  # (dims, rows, cols, planes, tensor_order)
        self.op = op
        self.inputs = list(inputs)
        self.params = params or {}
        self.data = data
        self.probed = False
        self.sinks = []
        for i in self.inputs:
            i.sinks.append(self)
        graph._fields.append(self)

    @property
    def fieldShape(self):
        d, r, c, _, _ = self.ftype
        return (r, c) if d == 2 else (c,)

    @property
    def tensorPoints(self):
        return self.ftype[3]

    @property
    def shape(self):
        """numpy shape of the planar data of this field."""
        d, r, c, p, order = self.ftype
        spatial = (r, c) if d == 2 else (c,)
        return spatial if (order == 0 and p == 1) else (p,) + spatial

    def __repr__(self):
        return "Field(%s, shape=%s)" % (self.op or "input", self.shape)


def _input_field(arr, vector):
    arr = np.ascontiguousarray(arr, dtype=np.float32)
    if vector:
        if arr.ndim not in (2, 3):
            raise ValueError("VectorField data must be (planes, rows, cols) or (planes, cols)")
        planes = arr.shape[0]
        dims = arr.ndim - 1
        rows, cols = (arr.shape[1], arr.shape[2]) if dims == 2 else (1, arr.shape[1])
        order = 1
    else:
        if arr.ndim not in (1, 2):
            raise ValueError("ScalarField data must be (rows, cols) or (cols,)")
        planes, dims, order = 1, arr.ndim, 0
        rows, cols = arr.shape if dims == 2 else (1, arr.shape[0])
    return Field(_current_graph(), (dims, rows, cols, planes, order), data=arr)


def ScalarField(data):
    return _input_field(data, vector=False)


def VectorField(data):
    """data: planar (planes, rows, cols) or (planes, cols) array."""
    return _input_field(data, vector=True)


def probe(*fields):
    for f in fields:
        f.probed = True


def _convolve(field, filt, border, orientation, sampling, mode, batch):
    if sampling.code != L.NO_SAMPLING and sampling.step == 1:
        sampling = NoSamplingConvolution  # CogFunctionAPI.scala:1140-1144
    desc = L.make_desc(mode, border.code, orientation, sampling.code, sampling.step, batch, field.ftype, filt.ftype)
    out = L.output_type(desc)  # raises CogconvError on an illegal op, as graph building does
    return Field(field.graph, out, op="convolve", inputs=(field, filt),
                 params={"mode": mode, "border": border, "orientation": orientation, "sampling": sampling,
                         "batch": batch})


def crossCorrelate(field, filt, border, sampling=NoSamplingConvolution, batchSize=1):
    return _convolve(field, filt, border, L.CROSS_CORRELATION, sampling, L.PLANE_BY_PLANE, batchSize)


def convolve(field, filt, border, sampling=NoSamplingConvolution, batchSize=1):
    return _convolve(field, filt, border, L.CONVOLUTION, sampling, L.PLANE_BY_PLANE, batchSize)


def projectFrame(field, filt, border, sampling=NoSamplingConvolution, batchSize=1):
    return _convolve(field, filt, border, L.CROSS_CORRELATION, sampling, L.PROJECT_FRAME, batchSize)


def backProjectFrame(field, filt, border, sampling=NoSamplingConvolution, batchSize=1):
    return _convolve(field, filt, border, L.CONVOLUTION, sampling, L.BACK_PROJECT_FRAME, batchSize)


def crossCorrelateFilterAdjoint(field, filt, border, sampling=NoSamplingConvolution, batchSize=1):
    return _convolve(field, filt, border, L.CROSS_CORRELATION, sampling, L.FILTER_ADJOINT, batchSize)


def convolveFilterAdjoint(field, filt, border, sampling=NoSamplingConvolution, batchSize=1):
    return _convolve(field, filt, border, L.CONVOLUTION, sampling, L.FILTER_ADJOINT, batchSize)


def blockReduceSum(field, factor):
    d, r, c, p, _ = field.ftype
    if factor < 1 or p % factor:
        raise L.CogconvError(L.ERR_ILLEGAL_OP,
                             "tensor length %d is not a multiple of the reduction factor %d" % (p, factor))
    q = p // factor
    return Field(field.graph, (d, r, c, q, 0 if q == 1 else 1), op="blockReduceSum", inputs=(field,),
                 params={"factor": factor})


class Cog:
    """Global switches of the convolution path, mirroring the reference's
    `Cog` flags (src/main/scala/cogx/parameters/Cog.scala:64-72), with the same
    defaults.  The merging flags gate the reduce fusion per vector mode;
    `fastProjectFrame` lets the planner use multi-plane blocks for the fused
    ProjectFrame (the MFMA implicit GEMM) rather than one output plane per
    block; `localMemoryMerging` has no counterpart here (no elementwise
    epilogue fusion yet) and is kept for source compatibility."""
    fastProjectFrame = True
    projectFrameMerging = True
    backProjectFrameMerging = True
    filterAdjointMerging = True
    localMemoryMerging = True


_FUSED = {L.PROJECT_FRAME: L.PROJECT_FRAME_BLOCK_REDUCE_SUM,
          L.BACK_PROJECT_FRAME: L.BACK_PROJECT_FRAME_BLOCK_REDUCE_SUM,
          L.FILTER_ADJOINT: L.FILTER_ADJOINT_BLOCK_REDUCE_SUM}


class ComputeGraph:
    """Builds, compiles (fusing conv + blockReduceSum) and steps a graph of conv ops.

    `optimize` enables the reduce fusion; `precision` selects the arithmetic of
    the contraction modes (L.FP32, L.BF16, L.FP16, L.SPLIT_BF16).
    """

    def __init__(self, optimize=True, device=0, precision=L.FP32):
        self.optimize = optimize
        self.device = device
        self.precision = precision
        self._fields = []
        self._compiled = False
        self._released = False

    def __enter__(self):
        self._prev = getattr(_tls, "graph", None)
        _tls.graph = self
        return self

    def __exit__(self, *exc):
        _tls.graph = self._prev
        return False

    # -- compile -------------------------------------------------------------
    def _fusable(self, red):
        conv = red.inputs[0]
        if conv.op != "convolve" or conv.probed or len(conv.sinks) != 1:
            return None
        mode = conv.params["mode"]
        if mode not in _FUSED:
            return None
        enabled = {L.PROJECT_FRAME: Cog.projectFrameMerging, L.BACK_PROJECT_FRAME: Cog.backProjectFrameMerging,
                   L.FILTER_ADJOINT: Cog.filterAdjointMerging}[mode]
        if not enabled:
            return None
        batch = conv.params["batch"]
        p0 = conv.inputs[0].ftype[3]
        factor = red.params["factor"]
        # reduce factor that the fused mode sums over
        want = batch if mode == L.FILTER_ADJOINT else p0 // batch
        return _FUSED[mode] if factor == want else None

    def _compile(self):
        if self._compiled:
            return
        L.init(self.device)
        self._stream = L.Stream()
        self._buffers = {}
        self._steps = []
        order = list(self._fields)  # creation order is a topological order
        fused = {}  # id(blockReduceSum field) -> fused vector mode
        skip = set()  # convolutions absorbed into a fused reduce
        if self.optimize:
            for f in order:
                if f.op == "blockReduceSum":
                    mode = self._fusable(f)
                    if mode is not None:
                        fused[id(f)] = mode
                        skip.add(id(f.inputs[0]))
        for f in order:
            if f.op is None:
                self._buffers[id(f)] = L.DeviceBuffer.from_array(f.data)
        for f in order:
            if f.op is None or id(f) in skip:
                continue
            if id(f) in fused:
                self._add_conv(f, f.inputs[0], fused[id(f)])
                continue
            if f.op == "convolve":
                self._add_conv(f, f, f.params["mode"])
            else:
                self._add_reduce(f)
        self._compiled = True

    def _buffer_for(self, f):
        b = self._buffers.get(id(f))
        if b is None:
            _, r, c, p, _ = f.ftype
            b = L.DeviceBuffer(4 * r * c * p)
            self._buffers[id(f)] = b
        return b

    def _add_conv(self, result, conv, mode):
        a, k = conv.inputs
        pr = conv.params
        precision = self.precision if mode in _FUSED.values() else L.FP32
        algo = L.ALGO_AUTO if Cog.fastProjectFrame else L.ALGO_DIRECT
        desc = L.make_desc(mode, pr["border"].code, pr["orientation"], pr["sampling"].code, pr["sampling"].step,
                           pr["batch"], a.ftype, k.ftype, precision=precision,
                           algo=algo if precision == L.FP32 else L.ALGO_AUTO)
        plan = L.Plan(desc)
        _, r, c, p, _ = result.ftype
        if plan.out_shape != (p, r, c):
            raise RuntimeError("internal: fused result type %s != %s" % (plan.out_shape, (p, r, c)))
        ws = L.DeviceBuffer(plan.workspace_bytes) if plan.workspace_bytes else None
        self._steps.append(("conv", plan, a, k, result, ws))

    def _add_reduce(self, f):
        (src,) = f.inputs
        _, r, c, p, _ = src.ftype
        self._steps.append(("reduce", f.params["factor"], src, f, r * c, p))

    # -- run -----------------------------------------------------------------
    def step(self, count=1):
        if self._released:
            raise RuntimeError("ComputeGraph has been released")
        self._compile()
        for _ in range(count):
            for st in self._steps:
                if st[0] == "conv":
                    _, plan, a, k, out, ws = st
                    plan.launch(self._buffer_for(a), self._buffer_for(k), self._buffer_for(out), ws, self._stream)
                else:
                    _, factor, src, out, points, planes = st
                    L.block_reduce_sum(self._buffer_for(src), self._buffer_for(out), points, planes, factor,
                                       self._stream)
        self._stream.synchronize()

    def read(self, field):
        self._compile()
        if id(field) not in self._buffers:
            raise RuntimeError("field %r was fused away; probe() it to keep it readable" % (field,))
        return self._buffers[id(field)].download(field.shape)

    def plans(self):
        self._compile()
        return [st[1] for st in self._steps if st[0] == "conv"]

    def release(self):
        if getattr(self, "_buffers", None):
            for b in self._buffers.values():
                b.close()
        self._buffers = {}
        self._steps = []
        if getattr(self, "_stream", None):
            self._stream.close()
        self._released = True
