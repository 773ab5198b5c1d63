"""Field / ComputeGraph operator API for the convolution path.

Mirrors the reference's user API for this path so that graph code reads the
same way:
  convolve / crossCorrelate               src/main/scala/cogx/api/CogFunctionAPI.scala:1245-1280
  projectFrame / backProjectFrame         CogFunctionAPI.scala:1370-1377, 1425-1432
  crossCorrelateFilterAdjoint / convolveFilterAdjoint   CogFunctionAPI.scala:1480-1510
  blockReduceSum                          CogFunctionAPI.scala:755-756
  Field method forms of the above         src/main/scala/cogx/compiler/parser/syntaxtree/Field.scala:1467-1633
  elementwise sinks: Field + - * / with a Float or a Field (Field.scala:180-390),
  max / min / abs / exp / log / sq / sqrt / tanh / signum (CogFunctionAPI.scala:109-303, 1574-1590)
  ComputeGraph step / read / release      src/main/scala/cogx/runtime/ComputeGraph.scala:228-457
Border and sampling policies are the reference's objects: BorderZero,
BorderClamp, BorderCyclic, BorderValid, BorderFull, NoSamplingConvolution,
UpsampleInputConvolution(step), DownsampleOutputConvolution(step).

Graph building raises immediately on an illegal op (as the reference raises at
graph construction): the legality checks and result types come from
libcogconv's cogconv_output_type.  Compiling a graph applies the reference's
optimizers that touch the convolution:
  * chained blockReduceSums merge into one (TensorReduceOptimizer.scala:38-80);
  * a convolution whose only consumer is a matching blockReduceSum becomes the
    *BlockReduceSum vector mode (ProjectFrameTensorReduceSumOptimizer.scala:64-179);
  * a chain of elementwise sinks merges into the convolution's (or reduce's)
    store as an epilogue, under Cog.localMemoryMerging, when each link has a
    single, unprobed sink of the same field type (HyperKernel.canMergeWithSink,
    src/main/scala/cogx/compiler/codegenerator/opencl/fragments/HyperKernel.scala:478-530).

Field data are numpy arrays in planar layout: a ScalarField is (rows, cols) or
(cols,) for 1-D, or a number for a 0-D field; a VectorField is (planes, rows,
cols) or (planes, cols), or (planes,) for a 0-D field.
"""
import builtins
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
_BORDERS = {p.code: p for p in (BorderZero, BorderClamp, BorderCyclic, BorderValid, BorderFull)}


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


def _sampling_of(code, step):
    if code == L.UPSAMPLE_INPUT:
        return UpsampleInputConvolution(step)
    if code == L.DOWNSAMPLE_OUTPUT:
        return DownsampleOutputConvolution(step)
    return NoSamplingConvolution


_tls = threading.local()


def _current_graph():
    g = getattr(_tls, "graph", None)
    if g is None:
        raise RuntimeError("fields must be created inside `with ComputeGraph() as cg:`")
    return g


def _result_type(t1, t2):
    """Result FieldType of an elementwise binary op, or None
    (src/main/scala/cogx/compiler/parser/semantics/ResultType.scala:29-58)."""
    def field_shape(t):
        return (t[0], t[1], t[2])

    def tensor_shape(t):
        return (t[3], t[4])

    def zero_d(t):
        return t[0] == 0

    def tensor0(t):
        return t[4] == 0
    if t1 == t2:
        return t1
    if tensor_shape(t1) == tensor_shape(t2) and zero_d(t1):
        return t2
    if tensor_shape(t1) == tensor_shape(t2) and zero_d(t2):
        return t1
    if field_shape(t1) == field_shape(t2) and tensor0(t1):
        return t2
    if field_shape(t1) == field_shape(t2) and tensor0(t2):
        return t1
    if zero_d(t1) and tensor0(t1):
        return t2
    if zero_d(t2) and tensor0(t2):
        return t1
    return None


# (op with a constant on the right, op with a field on the right, the same with the operands swapped)
_BINARY = {
    "+": (L.EPI_ADD_CONST, L.EPI_ADD_FIELD, L.EPI_ADD_CONST, L.EPI_ADD_FIELD),
    "-": (L.EPI_SUB_CONST, L.EPI_SUB_FIELD, L.EPI_RSUB_CONST, L.EPI_RSUB_FIELD),
    "*": (L.EPI_MUL_CONST, L.EPI_MUL_FIELD, L.EPI_MUL_CONST, L.EPI_MUL_FIELD),
    "/": (L.EPI_DIV_CONST, L.EPI_DIV_FIELD, L.EPI_RDIV_CONST, L.EPI_RDIV_FIELD),
    "max": (L.EPI_MAX_CONST, L.EPI_MAX_FIELD, L.EPI_MAX_CONST, L.EPI_MAX_FIELD),
    "min": (L.EPI_MIN_CONST, L.EPI_MIN_FIELD, L.EPI_MIN_CONST, L.EPI_MIN_FIELD),
}


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
        return (r, c) if d == 2 else ((c,) if d == 1 else ())

    @property
    def tensorPoints(self):
        return self.ftype[3]

    @property
    def shape(self):
        """numpy shape of the planar data of this field."""
        d, r, c, p, order = self.ftype
        spatial = (r, c) if d == 2 else ((c,) if d == 1 else ())
        if order == 0 and p == 1:
            return spatial
        return (p,) + spatial

    def __repr__(self):
        return "Field(%s, shape=%s)" % (self.op or "input", self.shape)

    # -- convolution method forms (Field.scala:1467-1633) ---------------------
    def crossCorrelate(self, filt, borderPolicy, samplingPolicy=NoSamplingConvolution):
        return crossCorrelate(self, filt, borderPolicy, samplingPolicy)

    def convolve(self, filt, borderPolicy, samplingPolicy=NoSamplingConvolution):
        return convolve(self, filt, borderPolicy, samplingPolicy)

    def crossCorrelateSeparable(self, rowFilter, columnFilter, borderPolicy):
        return crossCorrelateSeparable(self, rowFilter, columnFilter, borderPolicy)

    def convolveSeparable(self, rowFilter, columnFilter, borderPolicy):
        return convolveSeparable(self, rowFilter, columnFilter, borderPolicy)

    def projectFrame(self, filt, borderPolicy, samplingPolicy=NoSamplingConvolution, batchSize=1):
        return projectFrame(self, filt, borderPolicy, samplingPolicy, batchSize)

    def backProjectFrame(self, filt, borderPolicy, samplingPolicy=NoSamplingConvolution, batchSize=1):
        return backProjectFrame(self, filt, borderPolicy, samplingPolicy, batchSize)

    def crossCorrelateFilterAdjoint(self, filt, borderPolicy, samplingPolicy=NoSamplingConvolution, batchSize=1):
        return crossCorrelateFilterAdjoint(self, filt, borderPolicy, samplingPolicy, batchSize)

    def convolveFilterAdjoint(self, filt, borderPolicy, samplingPolicy=NoSamplingConvolution, batchSize=1):
        return convolveFilterAdjoint(self, filt, borderPolicy, samplingPolicy, batchSize)

    def _projectFrameBlockReduceSum(self, filt, borderPolicy, samplingPolicy=NoSamplingConvolution, batchSize=1):
        """Field.scala:1553-1559 (private to the reference's compiler)."""
        return _convolve(self, filt, borderPolicy, L.CROSS_CORRELATION, samplingPolicy,
                         L.PROJECT_FRAME_BLOCK_REDUCE_SUM, batchSize)

    def _backProjectFrameBlockReduceSum(self, filt, borderPolicy, samplingPolicy=NoSamplingConvolution, batchSize=1):
        """Field.scala:1592-1598 (private to the reference's compiler)."""
        return _convolve(self, filt, borderPolicy, L.CONVOLUTION, samplingPolicy,
                         L.BACK_PROJECT_FRAME_BLOCK_REDUCE_SUM, batchSize)

    def blockReduceSum(self, factor):
        return blockReduceSum(self, factor)

    # -- elementwise operators (Field.scala:180-390, 478) ---------------------
    def __add__(self, other):
        return _binary("+", self, other)

    def __radd__(self, other):
        return _binary("+", other, self)

    def __sub__(self, other):
        return _binary("-", self, other)

    def __rsub__(self, other):
        return _binary("-", other, self)

    def __mul__(self, other):
        return _binary("*", self, other)

    def __rmul__(self, other):
        return _binary("*", other, self)

    def __truediv__(self, other):
        return _binary("/", self, other)

    def __rtruediv__(self, other):
        return _binary("/", other, self)

    def __neg__(self):
        return _unary(L.EPI_NEG, self)


def _input_field(arr, vector):
    if not vector and np.ndim(arr) == 0:  # 0-D ScalarField
        arr = np.asarray(arr, dtype=np.float32).reshape(())
        return Field(_current_graph(), (0, 1, 1, 1, 0), data=arr)
    arr = np.ascontiguousarray(arr, dtype=np.float32)
    if vector:
        if arr.ndim == 1:  # 0-D VectorField
            return Field(_current_graph(), (0, 1, 1, arr.shape[0], 1), data=arr)
        if arr.ndim not in (2, 3):
            raise ValueError("VectorField data must be (planes, rows, cols), (planes, cols) or (planes,)")
        planes = arr.shape[0]
        dims = arr.ndim - 1
        rows, cols = (arr.shape[1], arr.shape[2]) if dims == 2 else (1, arr.shape[1])
        order = 1
    else:
        if arr.ndim not in (1, 2):
            raise ValueError("ScalarField data must be (rows, cols), (cols,) or a number")
        planes, dims, order = 1, arr.ndim, 0
        rows, cols = arr.shape if dims == 2 else (1, arr.shape[0])
    return Field(_current_graph(), (dims, rows, cols, planes, order), data=arr)


def ScalarField(data):
    return _input_field(data, vector=False)


def VectorField(data):
    """data: planar (planes, rows, cols) or (planes, cols) array, or (planes,) for a 0-D field."""
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


def _separable(field, row_filter, column_filter, border, orientation):
    """convolveSeparableInternal (CogFunctionAPI.scala:1277-1316): a 2-D
    ScalarField, odd 1-D scalar row and column filters, BorderClamp / Zero /
    Cyclic; rows first, then columns."""
    def check(cond, msg):
        if not cond:
            raise L.CogconvError(L.ERR_ILLEGAL_OP, msg)
    ft = field.ftype
    check(ft[0] == 2, "field must be 2D")
    check(ft[4] == 0, "field must be scalar")
    for name, f in (("row", row_filter), ("column", column_filter)):
        check(f.ftype[0] == 1, "%s filter must be 1D" % name)
        check(f.ftype[4] == 0, "%s filter must be scalar" % name)
        check(f.ftype[2] % 2 == 1, "%s filter must be an odd size" % name)
    check(border in (BorderClamp, BorderZero, BorderCyclic), "unsupported border policy %s" % border)
    return Field(field.graph, ft, op="separable", inputs=(field, row_filter, column_filter),
                 params={"border": border, "orientation": orientation})


def crossCorrelateSeparable(field, rowFilter, columnFilter, borderPolicy):
    return _separable(field, rowFilter, columnFilter, borderPolicy, L.CROSS_CORRELATION)


def convolveSeparable(field, rowFilter, columnFilter, borderPolicy):
    return _separable(field, rowFilter, columnFilter, borderPolicy, L.CONVOLUTION)


def blockReduceSum(field, factor):
    d, r, c, p, _ = field.ftype
    if factor < 1 or p % factor:
        raise L.CogconvError(L.ERR_ILLEGAL_OP,
                             "tensor length %d is not a multiple of the reduction factor %d" % (p, factor))
    q = p // factor
    return Field(field.graph, (d, r, c, q, 0 if q == 1 else 1), op="blockReduceSum", inputs=(field,),
                 params={"factor": factor})


def _is_number(x):
    return isinstance(x, (int, float, np.floating, np.integer)) and not isinstance(x, bool)


def _binary(name, a, b):
    """An elementwise binary op as a one-op epilogue node: `src` is the operand
    whose type is the result type, `operand` the other field (if any)."""
    c_op, f_op, rc_op, rf_op = _BINARY[name]
    if isinstance(a, Field) and _is_number(b):
        return Field(a.graph, a.ftype, op="elementwise", inputs=(a,), params={"ops": [(c_op, float(b))]})
    if _is_number(a) and isinstance(b, Field):
        return Field(b.graph, b.ftype, op="elementwise", inputs=(b,), params={"ops": [(rc_op, float(a))]})
    if isinstance(a, Field) and isinstance(b, Field):
        rt = _result_type(a.ftype, b.ftype)
        if rt is None:
            raise L.CogconvError(L.ERR_ILLEGAL_OP, "incompatible field types for %s: %s and %s" %
                                 (name, a.ftype, b.ftype))
        if rt == a.ftype:
            return Field(a.graph, rt, op="elementwise", inputs=(a, b), params={"ops": [(f_op, 0.0)]})
        return Field(b.graph, rt, op="elementwise", inputs=(b, a), params={"ops": [(rf_op, 0.0)]})
    raise TypeError("unsupported operands for %s: %r and %r" % (name, a, b))


def _unary(op, field):
    return Field(field.graph, field.ftype, op="elementwise", inputs=(field,), params={"ops": [(op, 0.0)]})


def max(field, value):  # noqa: A001 - the reference's name
    """max(field, Float) or max(field, Field) (CogFunctionAPI.scala:109, 1574)."""
    if not isinstance(field, Field):
        return builtins.max(field, value)
    return _binary("max", field, value)


def min(field, value):  # noqa: A001 - the reference's name
    """min(field, Float) or min(field, Field) (CogFunctionAPI.scala:121, 1584)."""
    if not isinstance(field, Field):
        return builtins.min(field, value)
    return _binary("min", field, value)


def abs(field):  # noqa: A001 - the reference's name
    return _unary(L.EPI_ABS, field) if isinstance(field, Field) else builtins.abs(field)


def exp(field):
    return _unary(L.EPI_EXP, field)


def log(field):
    return _unary(L.EPI_LOG, field)


def sq(field):
    return _unary(L.EPI_SQ, field)


def sqrt(field):
    return _unary(L.EPI_SQRT, field)


def tanh(field):
    return _unary(L.EPI_TANH, field)


def signum(field):
    return _unary(L.EPI_SIGNUM, field)


class Cog:
    """Global switches of the convolution path, mirroring the reference's
    `Cog` flags (src/main/scala/cogx/parameters/Cog.scala:64-72), with the same
    defaults.  The merging flags gate the reduce fusion per vector mode;
    `fastProjectFrame` lets the planner use multi-plane blocks for the fused
    ProjectFrame (the MFMA implicit GEMM) rather than one output plane per
    block; `localMemoryMerging` gates the merging of elementwise sinks into the
    convolution's stores (the reference gates forward merging of kernels that
    use local memory with it, HyperKernel.scala:512)."""
    fastProjectFrame = True
    projectFrameMerging = True
    backProjectFrameMerging = True
    filterAdjointMerging = True
    localMemoryMerging = True


_FUSED = {L.PROJECT_FRAME: L.PROJECT_FRAME_BLOCK_REDUCE_SUM,
          L.BACK_PROJECT_FRAME: L.BACK_PROJECT_FRAME_BLOCK_REDUCE_SUM,
          L.FILTER_ADJOINT: L.FILTER_ADJOINT_BLOCK_REDUCE_SUM}


def _buffer_ftype(ftype):
    """FieldTypeC of a field for the C ABI (0-D fields are 1x1)."""
    d, r, c, p, o = ftype
    return L.field_type(d, r, c, p, o)


class _Schedule:
    """The compiled form of a graph: which fields run as which steps, with the
    optimizers applied.  Pure host logic (no device work), so its decisions can
    be checked without a GPU.

    steps: list of dicts, in an order that respects every data dependence:
      {"kind": "conv", "result": F, "conv": C, "mode": m, "epilogue": [(op, value, slot)], "operands": [fields]}
      {"kind": "reduce", "result": F, "src": S, "factor": f, "epilogue": [...], "operands": [...]}
      {"kind": "separable", "result": F, "node": N, "epilogue": [...], "operands": [...]}
      {"kind": "elementwise", "result": F, "src": S, "epilogue": [...], "operands": [...]}
    Fields absorbed into a step have no buffer of their own (`absorbed`).
    """

    def __init__(self, fields, optimize=True):
        self.absorbed = set()
        order = list(fields)  # creation order is a topological order
        # 1. Reduce chains: blockReduceSum(blockReduceSum(x, a), b) -> blockReduceSum(x, a*b)
        #    when the inner reduce has one sink and is unprobed (TensorReduceOptimizer.scala:59-80).
        red_src, red_factor = {}, {}
        for f in order:
            if f.op == "blockReduceSum":
                src, factor = f.inputs[0], f.params["factor"]
                if optimize:
                    while src.op == "blockReduceSum" and len(src.sinks) == 1 and not src.probed:
                        self.absorbed.add(id(src))
                        factor *= red_factor[id(src)]
                        src = red_src[id(src)]
                red_src[id(f)], red_factor[id(f)] = src, factor
        # 2. Convolution + blockReduceSum -> *BlockReduceSum mode.
        fused_conv = {}
        if optimize:
            for f in order:
                if f.op == "blockReduceSum" and id(f) not in self.absorbed:
                    mode = self._fusable(red_src[id(f)], red_factor[id(f)])
                    if mode is not None:
                        fused_conv[id(f)] = (red_src[id(f)], mode)
                        self.absorbed.add(id(red_src[id(f)]))
        # 3. Steps, with chains of elementwise sinks merged into their producer.
        self.steps = []
        pending = {}  # id(step head field) -> step under construction
        for f in order:
            if f.op is None or id(f) in self.absorbed:
                continue
            if f.op == "elementwise" and id(f) in pending:
                continue  # merged into the step that produced its source
            if id(f) in fused_conv:
                conv, mode = fused_conv[id(f)]
                st = {"kind": "conv", "result": f, "conv": conv, "mode": mode}
            elif f.op == "convolve":
                st = {"kind": "conv", "result": f, "conv": f, "mode": f.params["mode"]}
            elif f.op == "blockReduceSum":
                st = {"kind": "reduce", "result": f, "src": red_src[id(f)], "factor": red_factor[id(f)]}
            elif f.op == "separable":
                st = {"kind": "separable", "result": f, "node": f}
            else:
                st = {"kind": "elementwise", "result": f, "src": f.inputs[0]}
            st["epilogue"], st["operands"] = [], []
            if st["kind"] == "elementwise":
                self._append_ops(st, f)
            if optimize and Cog.localMemoryMerging:
                self._merge_sinks(st, pending)
            self.steps.append(st)
        # A merged step runs where its last sink was created, so that every field
        # operand its epilogue reads is computed first.
        pos = {id(f): i for i, f in enumerate(order)}
        self.steps.sort(key=lambda st: pos[id(st["result"])])

    @staticmethod
    def _fusable(conv, factor):
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
        want = batch if mode == L.FILTER_ADJOINT else p0 // batch
        return _FUSED[mode] if factor == want else None

    @staticmethod
    def _append_ops(st, node):
        """Adds the ops of elementwise `node` (whose source is the step's current
        result) to the step's epilogue; False if they do not fit."""
        operands = list(st["operands"])
        slot = 0
        for x in node.inputs[1:]:
            hit = [i for i, o in enumerate(operands) if o is x]
            if hit:
                slot = hit[0]
            else:
                operands.append(x)
                slot = len(operands) - 1
        if len(st["epilogue"]) + len(node.params["ops"]) > L.EPI_MAX_OPS or len(operands) > L.EPI_MAX_OPERANDS:
            return False
        for (op, value) in node.params["ops"]:
            st["epilogue"].append((op, value, slot if op >= L.EPI_ADD_FIELD else 0))
        st["operands"] = operands
        return True

    def _merge_sinks(self, st, pending):
        """Absorbs the chain of single elementwise sinks that follows the step's
        result (HyperKernel.canMergeWithSink: a pure edge, the same field shape
        and tensor shape, the source unprobed)."""
        cur = st["result"]
        while True:
            if cur.probed or len(cur.sinks) != 1:
                return
            s = cur.sinks[0]
            if s.op != "elementwise" or s.inputs[0] is not cur or s.ftype != cur.ftype:
                return
            if any(x is cur for x in s.inputs[1:]):
                return  # the source feeds the sink twice: not a pure edge
            if not self._append_ops(st, s):
                return
            self.absorbed.add(id(cur))
            pending[id(s)] = st
            st["result"] = s
            cur = s


def readFromFile(filename, device=0):
    """Restore a graph saved by ComputeGraph.write (ComputeGraph.readFromFile,
    ComputeGraph.scala:646-666): every step is re-planned for `device`; the
    result is a Function whose evaluate(inputs) runs it (reentrant)."""
    return L.Function(filename, device)


class ComputeGraph:
    """Builds, compiles (with the reference's conv optimizers) and steps a graph
    of conv ops.

    `optimize` enables the reduce-chain merge, the reduce fusion and the
    elementwise-sink merge; `precision` selects the arithmetic of the
    contraction modes (L.FP32, L.BF16, L.FP16, L.SPLIT_BF16).
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
    def schedule(self):
        """The compiled schedule (host only): steps and absorbed fields."""
        return _Schedule(self._fields, self.optimize)

    def _compile(self):
        if self._compiled:
            return
        L.init(self.device)
        self._stream = L.Stream()
        self._buffers = {}
        self._steps = []
        sched = self.schedule()
        self._absorbed = sched.absorbed
        for f in self._fields:
            if f.op is None:
                self._buffers[id(f)] = L.DeviceBuffer.from_array(np.ascontiguousarray(f.data, dtype=np.float32))
        for st in sched.steps:
            epi = None
            if st["epilogue"]:
                epi = L.make_epilogue(st["epilogue"], [_buffer_ftype(o.ftype) for o in st["operands"]])
            if st["kind"] == "conv":
                self._add_conv(st, epi)
            elif st["kind"] == "separable":
                self._add_separable(st, epi)
            else:
                st["epi"] = epi
                self._steps.append(st)
        self._compiled = True

    def _buffer_for(self, f):
        b = self._buffers.get(id(f))
        if b is None:
            _, r, c, p, _ = f.ftype
            b = L.DeviceBuffer(4 * r * c * p)
            self._buffers[id(f)] = b
        return b

    def _add_conv(self, st, epi):
        conv, mode, result = st["conv"], st["mode"], st["result"]
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
        if epi is not None:
            plan.set_epilogue(epi)
        st["plan"] = plan
        st["ws"] = L.DeviceBuffer(plan.workspace_bytes) if plan.workspace_bytes else None
        self._steps.append(st)

    def _add_separable(self, st, epi):
        node = st["node"]
        x, rf, cf = node.inputs
        st["sep"] = L.SeparablePlan(x.ftype, rf.ftype[2], cf.ftype[2], node.params["border"].code,
                                    node.params["orientation"])
        if epi is not None:
            st["sep"].set_epilogue(epi)
        self._steps.append(st)

    # -- run -----------------------------------------------------------------
    def _run_step(self, st):
        out = self._buffer_for(st["result"])
        operands = [self._buffer_for(o) for o in st["operands"]]
        if st["kind"] == "conv":
            a, k = st["conv"].inputs
            if operands:
                st["plan"].launch_epilogue(self._buffer_for(a), self._buffer_for(k), out, st["ws"], operands,
                                           self._stream)
            else:
                st["plan"].launch(self._buffer_for(a), self._buffer_for(k), out, st["ws"], self._stream)
        elif st["kind"] == "reduce":
            src = st["src"]
            _, r, c, p, _ = src.ftype
            L.block_reduce_sum(self._buffer_for(src), out, r * c, p, st["factor"], self._stream,
                               epilogue=st["epi"], operands=operands, in_type=_buffer_ftype(src.ftype))
        elif st["kind"] == "separable":
            x, rf, cf = st["node"].inputs
            st["sep"].launch(self._buffer_for(x), self._buffer_for(rf), self._buffer_for(cf), out, operands,
                             self._stream)
        else:
            L.elementwise(self._buffer_for(st["src"]), out, _buffer_ftype(st["result"].ftype), st["epi"], operands,
                          self._stream)

    def step(self, count=1):
        if self._released:
            raise RuntimeError("ComputeGraph has been released")
        self._compile()
        for _ in range(count):
            for st in self._steps:
                self._run_step(st)
        self._stream.synchronize()

    def read(self, field):
        self._compile()
        if id(field) not in self._buffers or id(field) in self._absorbed:
            raise RuntimeError("field %r was fused away; probe() it to keep it readable" % (field,))
        return self._buffers[id(field)].download(field.shape)

    def plans(self):
        self._compile()
        return [st["plan"] for st in self._steps if st["kind"] == "conv"]

    def kernel_steps(self):
        """Number of launch steps one graph step runs (after merging)."""
        self._compile()
        return len(self._steps)

    # -- checkpoint ----------------------------------------------------------
    def write(self, filename, inputs=(), outputs=None, name="ComputeGraph", description=None):
        """Save the compiled graph (ComputeGraph.write, ComputeGraph.scala:523-612):
        per launch step its descriptor, the field types, and the data of every
        data field not listed in `inputs`; `inputs` become the inputs of the
        restored Function, `outputs` (default: the probed fields) its outputs.
        readFromFile() restores it, re-planning each step for its device."""
        sched = self.schedule()
        inputs = list(inputs)
        if outputs is None:
            outputs = [f for f in self._fields if f.probed]
        ids, lines = {}, []

        def fid(f):
            if id(f) not in ids:
                if id(f) in sched.absorbed:
                    raise RuntimeError("field %r was merged away; probe() it to save it as an output" % (f,))
                ids[id(f)] = len(ids)
                if f.op is None:
                    kind = "input" if any(f is x for x in inputs) else "const"
                else:
                    kind = "computed"
                d, r, c, p, o = f.ftype
                lines.append("field %d %d %d %d %d %d %s %s" % (ids[id(f)], d, r, c, p, o, kind,
                                                               getattr(f, "name", None) or "f%d" % ids[id(f)]))
            return ids[id(f)]
        for f in inputs:
            fid(f)
        for st in sched.steps:
            if st["kind"] == "conv":
                conv = st["conv"]
                a, k = conv.inputs
                pr = conv.params
                mode = st["mode"]
                precision = self.precision if mode in _FUSED.values() else L.FP32
                ia, ik = fid(a), fid(k)
                line = "step conv %d %d %d %d %d %d %d %d %d %d" % (
                    fid(st["result"]), ia, ik, mode, pr["border"].code, pr["orientation"], pr["sampling"].code,
                    pr["sampling"].step, pr["batch"], precision)
            elif st["kind"] == "reduce":
                isrc = fid(st["src"])
                line = "step reduce %d %d %d" % (fid(st["result"]), isrc, st["factor"])
            elif st["kind"] == "separable":
                x, rf, cf = st["node"].inputs
                ix, ir, ic = fid(x), fid(rf), fid(cf)
                line = "step separable %d %d %d %d %d %d" % (fid(st["result"]), ix, ir, ic,
                                                            st["node"].params["border"].code,
                                                            st["node"].params["orientation"])
            else:
                isrc = fid(st["src"])
                line = "step elementwise %d %d" % (fid(st["result"]), isrc)
            opids = [fid(o) for o in st["operands"]]
            lines.append(line)
            for op, value, slot in st["epilogue"]:
                lines.append("epi %d %s %d" % (op, float(np.float32(value)).hex(), slot))
            for o in opids:
                lines.append("operand %d" % o)
        for f in outputs:
            lines.append("output %d %s" % (fid(f), getattr(f, "name", None) or "out%d" % ids[id(f)]))
        order = sorted(ids.items(), key=lambda kv: kv[1])
        by_id = {id(f): f for f in self._fields}
        header = ["COGCONV-GRAPH 1 0", "name " + name]
        if description:
            header.append("description " + description.replace("\n", " "))
        with open(filename, "wb") as fh:
            fh.write(("\n".join(header + sorted([x for x in lines if x.startswith("field ")],
                                                key=lambda x: int(x.split()[1])) +
                               [x for x in lines if not x.startswith("field ")] + ["end"]) + "\n").encode())
            for key, _ in order:
                f = by_id[key]
                if f.op is None and not any(f is x for x in inputs):
                    fh.write(np.ascontiguousarray(f.data, dtype="<f4").tobytes())
        return filename

    def release(self):
        if getattr(self, "_buffers", None):
            for b in self._buffers.values():
                b.close()
        self._buffers = {}
        self._steps = []
        if getattr(self, "_stream", None):
            self._stream.close()
        self._released = True
