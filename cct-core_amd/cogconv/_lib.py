"""ctypes binding of libcogconv.so (include/cogconv.h).

Loads the in-tree library and fails loudly if it has not been built: there is
no fallback implementation behind this package.
"""
import ctypes
import os

import numpy as np

_PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG_ROOT, "libcogconv.so")

# Status codes
OK, ERR_INVALID_ARG, ERR_ILLEGAL_OP, ERR_RUNTIME, ERR_OOM, ERR_NOT_READY, ERR_UNSUPPORTED, ERR_BUDGET = range(8)
# VectorMode
(PLANE_BY_PLANE, PROJECT_FRAME, PROJECT_FRAME_BLOCK_REDUCE_SUM, BACK_PROJECT_FRAME,
 BACK_PROJECT_FRAME_BLOCK_REDUCE_SUM, FILTER_ADJOINT, FILTER_ADJOINT_BLOCK_REDUCE_SUM) = range(7)
# BorderPolicy
BORDER_ZERO, BORDER_CLAMP, BORDER_CYCLIC, BORDER_VALID, BORDER_FULL = range(5)
# Sampling
NO_SAMPLING, UPSAMPLE_INPUT, DOWNSAMPLE_OUTPUT = range(3)
# Orientation
CROSS_CORRELATION, CONVOLUTION = range(2)
# Precision
FP32, BF16, FP16, SPLIT_BF16 = range(4)
# Algorithm hint
ALGO_AUTO, ALGO_DIRECT, ALGO_SMALL_OUTPUT, ALGO_IMPLICIT_GEMM = range(4)


class CogconvError(RuntimeError):
    """A non-zero status from libcogconv; `status` holds the code."""

    def __init__(self, status, message):
        super().__init__("cogconv error %d: %s" % (status, message))
        self.status = status


class FieldTypeC(ctypes.Structure):
    _fields_ = [("dims", ctypes.c_int32), ("rows", ctypes.c_int32), ("cols", ctypes.c_int32),
                ("planes", ctypes.c_int32), ("tensor_order", ctypes.c_int32)]

    def as_tuple(self):
        return (self.dims, self.rows, self.cols, self.planes, self.tensor_order)


class DescC(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int32), ("border", ctypes.c_int32), ("orientation", ctypes.c_int32),
                ("sampling", ctypes.c_int32), ("step", ctypes.c_int32), ("batch", ctypes.c_int32),
                ("in0", FieldTypeC), ("in1", FieldTypeC), ("precision", ctypes.c_int32),
                ("algo", ctypes.c_int32)]


class DevicePropsC(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 128), ("arch", ctypes.c_char * 64),
                ("compute_units", ctypes.c_int32), ("wavefront_size", ctypes.c_int32),
                ("lds_bytes_per_block", ctypes.c_int32), ("max_threads_per_block", ctypes.c_int32),
                ("clock_khz", ctypes.c_int32), ("memory_clock_khz", ctypes.c_int32),
                ("memory_bus_width", ctypes.c_int32), ("l2_bytes", ctypes.c_int32),
                ("hbm_bytes", ctypes.c_uint64)]


class ServerConfigC(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int32), ("hbm_budget_bytes", ctypes.c_uint64), ("max_inflight", ctypes.c_int32),
                ("streams", ctypes.c_int32), ("max_stack", ctypes.c_int32), ("graph_after", ctypes.c_int32)]


class RequestC(ctypes.Structure):
    _fields_ = [("desc", DescC), ("bank", ctypes.c_int32), ("flags", ctypes.c_int32), ("host_in0", ctypes.c_void_p),
                ("host_in1", ctypes.c_void_p), ("host_out", ctypes.c_void_p)]


REQUEST_RESIDENT = 1


class ServerStatsC(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64), ("rejected", ctypes.c_uint64),
                ("launches", ctypes.c_uint64), ("stacked", ctypes.c_uint64), ("graph_launches", ctypes.c_uint64),
                ("p50_ms", ctypes.c_double),
                ("p99_ms", ctypes.c_double), ("mean_ms", ctypes.c_double), ("max_ms", ctypes.c_double),
                ("first_submit_s", ctypes.c_double), ("last_complete_s", ctypes.c_double),
                ("requests_per_s", ctypes.c_double), ("gflop", ctypes.c_double), ("peak_bytes", ctypes.c_uint64)]


class TuneOptionsC(ctypes.Structure):
    _fields_ = [("warmup_steps", ctypes.c_int32), ("run_steps", ctypes.c_int32), ("flush_bytes", ctypes.c_int64),
                ("force", ctypes.c_int32)]


class TuneResultC(ctypes.Structure):
    _fields_ = [("variant", ctypes.c_int32), ("variants", ctypes.c_int32), ("avg_us", ctypes.c_double),
                ("from_cache", ctypes.c_int32), ("description", ctypes.c_char * 192)]


# Elementwise epilogue ops (cogconv.h COGCONV_EPI_*)
(EPI_ADD_CONST, EPI_SUB_CONST, EPI_RSUB_CONST, EPI_MUL_CONST, EPI_DIV_CONST, EPI_RDIV_CONST, EPI_MAX_CONST,
 EPI_MIN_CONST) = range(1, 9)
EPI_NEG, EPI_ABS, EPI_SQ, EPI_SQRT, EPI_EXP, EPI_LOG, EPI_TANH, EPI_SIGNUM = range(16, 24)
(EPI_ADD_FIELD, EPI_SUB_FIELD, EPI_RSUB_FIELD, EPI_MUL_FIELD, EPI_DIV_FIELD, EPI_RDIV_FIELD, EPI_MAX_FIELD,
 EPI_MIN_FIELD) = range(32, 40)
EPI_MAX_OPS, EPI_MAX_OPERANDS = 8, 4


class EpilogueOpC(ctypes.Structure):
    _fields_ = [("op", ctypes.c_int32), ("value", ctypes.c_float), ("operand", ctypes.c_int32)]


class EpilogueC(ctypes.Structure):
    _fields_ = [("n_ops", ctypes.c_int32), ("ops", EpilogueOpC * EPI_MAX_OPS), ("n_operands", ctypes.c_int32),
                ("operands", FieldTypeC * EPI_MAX_OPERANDS)]


def make_epilogue(ops, operand_types=()):
    """ops: [(op, value, operand slot)]; operand_types: FieldTypeC or tuples, one per slot."""
    if len(ops) > EPI_MAX_OPS or len(operand_types) > EPI_MAX_OPERANDS:
        raise ValueError("epilogue too long")
    e = EpilogueC()
    e.n_ops = len(ops)
    for i, (op, value, slot) in enumerate(ops):
        e.ops[i] = EpilogueOpC(op, value, slot)
    e.n_operands = len(operand_types)
    for i, t in enumerate(operand_types):
        e.operands[i] = t if isinstance(t, FieldTypeC) else field_type(*t)
    return e


_P = ctypes.c_void_p
_I = ctypes.c_int
_SZ = ctypes.c_size_t

# name -> argtypes (restype int unless listed in _RESTYPES)
SIGNATURES = {
    "cogconv_last_error": [],
    "cogconv_abi_version": [],
    "cogconv_init": [ctypes.c_int32],
    "cogconv_shutdown": [],
    "cogconv_device_count": [ctypes.POINTER(ctypes.c_int32)],
    "cogconv_device_props_get": [ctypes.c_int32, ctypes.POINTER(DevicePropsC)],
    "cogconv_device_synchronize": [],
    "cogconv_buf_alloc": [_SZ, ctypes.POINTER(_P)],
    "cogconv_buf_free": [_P],
    "cogconv_h2d": [_P, _P, _SZ],
    "cogconv_d2h": [_P, _P, _SZ],
    "cogconv_h2d_async": [_P, _P, _SZ, _P],
    "cogconv_d2h_async": [_P, _P, _SZ, _P],
    "cogconv_d2d_async": [_P, _P, _SZ, _P],
    "cogconv_fill_zero": [_P, _SZ, _P],
    "cogconv_host_alloc": [_SZ, ctypes.POINTER(_P)],
    "cogconv_host_free": [_P],
    "cogconv_stream_create": [ctypes.POINTER(_P)],
    "cogconv_stream_destroy": [_P],
    "cogconv_stream_synchronize": [_P],
    "cogconv_event_create": [ctypes.POINTER(_P)],
    "cogconv_event_destroy": [_P],
    "cogconv_event_record": [_P, _P],
    "cogconv_stream_wait_event": [_P, _P],
    "cogconv_event_query": [_P],
    "cogconv_event_synchronize": [_P],
    "cogconv_event_elapsed_ms": [_P, _P, ctypes.POINTER(ctypes.c_float)],
    "cogconv_output_type": [ctypes.POINTER(DescC), ctypes.POINTER(FieldTypeC)],
    "cogconv_plan_create": [ctypes.POINTER(DescC), ctypes.POINTER(_P)],
    "cogconv_plan_destroy": [_P],
    "cogconv_plan_output_type": [_P, ctypes.POINTER(FieldTypeC)],
    "cogconv_plan_workspace_bytes": [_P, ctypes.POINTER(_SZ)],
    "cogconv_plan_flops": [_P, ctypes.POINTER(ctypes.c_double)],
    "cogconv_plan_bytes": [_P, ctypes.POINTER(ctypes.c_double)],
    "cogconv_plan_describe": [_P, ctypes.c_char_p, _SZ],
    "cogconv_launch": [_P, _P, _P, _P, _P, _P],
    "cogconv_plan_pack_operand": [_P, _P, _P, _P, _P],
    "cogconv_launch_packed": [_P, _P, _P, _P, _P],
    "cogconv_noalloc_scratch_bytes": [ctypes.POINTER(_SZ)],
    "cogconv_plan_table_supported": [_P, ctypes.POINTER(ctypes.c_int32)],
    "cogconv_launch_table": [_P, _P, _P, _P, _P, ctypes.c_int32, _P],
    "cogconv_plan_set_epilogue": [_P, ctypes.POINTER(EpilogueC)],
    "cogconv_launch_epilogue": [_P, _P, _P, _P, _P, ctypes.POINTER(_P), _P],
    "cogconv_elementwise": [_P, _P, ctypes.POINTER(FieldTypeC), ctypes.POINTER(EpilogueC), ctypes.POINTER(_P), _P],
    "cogconv_block_reduce_sum_epilogue": [_P, _P, ctypes.POINTER(FieldTypeC), ctypes.c_int32,
                                          ctypes.POINTER(EpilogueC), ctypes.POINTER(_P), _P],
    "cogconv_function_create": [ctypes.c_char_p, ctypes.c_int32, ctypes.POINTER(_P)],
    "cogconv_function_destroy": [_P],
    "cogconv_function_name": [_P, ctypes.c_char_p, _SZ],
    "cogconv_function_optimal_thread_count": [_P, ctypes.POINTER(ctypes.c_int32)],
    "cogconv_function_input_count": [_P, ctypes.POINTER(ctypes.c_int32)],
    "cogconv_function_output_count": [_P, ctypes.POINTER(ctypes.c_int32)],
    "cogconv_function_io": [_P, ctypes.c_int32, ctypes.c_int32, ctypes.c_char_p, _SZ, ctypes.POINTER(FieldTypeC)],
    "cogconv_function_evaluate": [_P, ctypes.POINTER(_P), ctypes.POINTER(_P)],
    "cogconv_separable": [_P, ctypes.POINTER(FieldTypeC), _P, ctypes.c_int32, _P, ctypes.c_int32, ctypes.c_int32,
                          ctypes.c_int32, _P, ctypes.POINTER(EpilogueC), ctypes.POINTER(_P), _P],
    "cogconv_launch_noalloc": [_P, _P, _P, _P, _P],
    "cogconv_block_reduce_sum": [_P, _P, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, _P],
    "cogconv_plan_variant_count": [ctypes.POINTER(DescC), ctypes.POINTER(ctypes.c_int32)],
    "cogconv_plan_create_variant": [ctypes.POINTER(DescC), ctypes.c_int32, ctypes.POINTER(_P)],
    "cogconv_plan_variant": [_P, ctypes.POINTER(ctypes.c_int32)],
    "cogconv_autotune": [ctypes.POINTER(DescC), ctypes.POINTER(TuneOptionsC), _P, ctypes.POINTER(TuneResultC)],
    "cogconv_tune_cache_clear": [],
    "cogconv_tune_cache_save": [ctypes.c_char_p],
    "cogconv_tune_cache_load": [ctypes.c_char_p],
    "cogconv_server_create": [ctypes.POINTER(ServerConfigC), ctypes.POINTER(_P)],
    "cogconv_server_destroy": [_P],
    "cogconv_server_register_bank": [_P, _P, ctypes.POINTER(FieldTypeC), ctypes.POINTER(ctypes.c_int32)],
    "cogconv_server_submit": [_P, ctypes.POINTER(RequestC), ctypes.POINTER(ctypes.c_uint64)],
    "cogconv_server_query": [_P, ctypes.c_uint64],
    "cogconv_server_submit_timed": [_P, ctypes.POINTER(RequestC), ctypes.POINTER(ctypes.c_double), ctypes.c_int64,
                                    ctypes.POINTER(ctypes.c_uint64)],
    "cogconv_server_wait": [_P, ctypes.c_uint64],
    "cogconv_server_drain": [_P],
    "cogconv_server_stats_get": [_P, ctypes.POINTER(ServerStatsC)],
    "cogconv_server_stats_reset": [_P],
}
_RESTYPES = {"cogconv_last_error": ctypes.c_char_p}

_lib = None


def lib():
    """The loaded libcogconv (raises if the library has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libcogconv.so is not built (expected at %s); run __graft_entry__.build()" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = _RESTYPES.get(name, ctypes.c_int)
        _lib = L
    return _lib


def last_error():
    return lib().cogconv_last_error().decode("utf-8", "replace")


def check(status):
    if status != OK:
        raise CogconvError(status, last_error())
    return status


def device_count():
    n = ctypes.c_int32(0)
    check(lib().cogconv_device_count(ctypes.byref(n)))
    return n.value


def device_props(device=0):
    p = DevicePropsC()
    check(lib().cogconv_device_props_get(device, ctypes.byref(p)))
    return {"name": p.name.decode(), "arch": p.arch.decode(), "compute_units": p.compute_units,
            "wavefront_size": p.wavefront_size, "lds_bytes_per_block": p.lds_bytes_per_block,
            "max_threads_per_block": p.max_threads_per_block, "clock_khz": p.clock_khz,
            "memory_clock_khz": p.memory_clock_khz, "memory_bus_width": p.memory_bus_width,
            "l2_bytes": p.l2_bytes, "hbm_bytes": p.hbm_bytes}


def init(device=0):
    check(lib().cogconv_init(device))


def synchronize():
    check(lib().cogconv_device_synchronize())


class DeviceBuffer:
    """A device allocation owned by Python (freed on close / garbage collection)."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        check(lib().cogconv_buf_alloc(self.nbytes, ctypes.byref(p)))
        self.ptr = p.value

    @classmethod
    def from_array(cls, arr):
        arr = np.ascontiguousarray(arr)
        b = cls(arr.nbytes)
        b.upload(arr)
        return b

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        check(lib().cogconv_h2d(self.ptr, arr.ctypes.data, arr.nbytes))

    def download(self, shape, dtype=np.float32):
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes <= self.nbytes
        check(lib().cogconv_d2h(out.ctypes.data, self.ptr, out.nbytes))
        return out

    def zero(self, stream=None):
        s = stream.handle if isinstance(stream, Stream) else stream
        check(lib().cogconv_fill_zero(self.ptr, self.nbytes, s))

    def close(self):
        if getattr(self, "ptr", None):
            lib().cogconv_buf_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Stream:
    def __init__(self):
        p = ctypes.c_void_p()
        check(lib().cogconv_stream_create(ctypes.byref(p)))
        self.handle = p.value

    def synchronize(self):
        check(lib().cogconv_stream_synchronize(self.handle))

    def close(self):
        if getattr(self, "handle", None):
            lib().cogconv_stream_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Event:
    def __init__(self):
        p = ctypes.c_void_p()
        check(lib().cogconv_event_create(ctypes.byref(p)))
        self.handle = p.value

    def record(self, stream=None):
        check(lib().cogconv_event_record(self.handle, stream.handle if isinstance(stream, Stream) else stream))

    def synchronize(self):
        check(lib().cogconv_event_synchronize(self.handle))

    def query(self):
        s = lib().cogconv_event_query(self.handle)
        if s == ERR_NOT_READY:
            return False
        check(s)
        return True

    def elapsed_ms(self, end):
        ms = ctypes.c_float()
        check(lib().cogconv_event_elapsed_ms(self.handle, end.handle, ctypes.byref(ms)))
        return ms.value

    def close(self):
        if getattr(self, "handle", None):
            lib().cogconv_event_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def field_type(dims, rows, cols, planes, order=None):
    if order is None:
        order = 0 if planes == 1 else 1
    return FieldTypeC(dims, rows, cols, planes, order)


def make_desc(mode, border, orientation, sampling, step, batch, in0, in1, precision=FP32, algo=ALGO_AUTO):
    """in0 / in1: FieldTypeC or (dims, rows, cols, planes[, order]) tuples."""
    def ft(x):
        return x if isinstance(x, FieldTypeC) else field_type(*x)
    return DescC(mode, border, orientation, sampling, step, batch, ft(in0), ft(in1), precision, algo)


def output_type(desc):
    out = FieldTypeC()
    check(lib().cogconv_output_type(ctypes.byref(desc), ctypes.byref(out)))
    return out.as_tuple()


class Plan:
    """One ConvolveOp execution plan (cogconv_plan_create / cogconv_launch);
    `variant` fixes one of the plan's candidate configurations."""

    def __init__(self, desc, variant=None):
        self.desc = desc
        p = ctypes.c_void_p()
        if variant is None:
            check(lib().cogconv_plan_create(ctypes.byref(desc), ctypes.byref(p)))
        else:
            check(lib().cogconv_plan_create_variant(ctypes.byref(desc), variant, ctypes.byref(p)))
        self.handle = p.value
        v = ctypes.c_int32()
        check(lib().cogconv_plan_variant(self.handle, ctypes.byref(v)))
        self.variant = v.value
        out = FieldTypeC()
        check(lib().cogconv_plan_output_type(self.handle, ctypes.byref(out)))
        self.out_type = out.as_tuple()
        ws = ctypes.c_size_t()
        check(lib().cogconv_plan_workspace_bytes(self.handle, ctypes.byref(ws)))
        self.workspace_bytes = ws.value
        fl = ctypes.c_double()
        check(lib().cogconv_plan_flops(self.handle, ctypes.byref(fl)))
        self.flops = fl.value
        by = ctypes.c_double()
        check(lib().cogconv_plan_bytes(self.handle, ctypes.byref(by)))
        self.bytes = by.value
        buf = ctypes.create_string_buffer(256)
        check(lib().cogconv_plan_describe(self.handle, buf, 256))
        self.description = buf.value.decode()

    @property
    def out_shape(self):
        _, rows, cols, planes, _ = self.out_type
        return (planes, rows, cols)

    def launch(self, in0, in1, out, workspace=None, stream=None):
        def ptr(x):
            return x.ptr if isinstance(x, DeviceBuffer) else x
        s = stream.handle if isinstance(stream, Stream) else stream
        check(lib().cogconv_launch(self.handle, ptr(in0), ptr(in1), ptr(out), ptr(workspace), s))

    def pack_operand(self, in1, workspace, stream=None):
        """cogconv_plan_pack_operand: True when the plan packed operand 2 into workspace."""
        def ptr(x):
            return x.ptr if isinstance(x, DeviceBuffer) else x
        s = stream.handle if isinstance(stream, Stream) else stream
        packed = ctypes.c_int32()
        check(lib().cogconv_plan_pack_operand(self.handle, ptr(in1), ptr(workspace), s, ctypes.byref(packed)))
        return bool(packed.value)

    def launch_packed(self, in0, out, workspace, stream=None):
        """cogconv_launch_packed: run on the operand 2 packed earlier into workspace."""
        def ptr(x):
            return x.ptr if isinstance(x, DeviceBuffer) else x
        s = stream.handle if isinstance(stream, Stream) else stream
        check(lib().cogconv_launch_packed(self.handle, ptr(in0), ptr(out), ptr(workspace), s))

    def set_epilogue(self, epilogue):
        """cogconv_plan_set_epilogue (an EpilogueC, or None to remove it)."""
        check(lib().cogconv_plan_set_epilogue(self.handle, ctypes.byref(epilogue) if epilogue is not None else None))
        self.epilogue_operands = epilogue.n_operands if epilogue is not None else 0

    def launch_epilogue(self, in0, in1, out, workspace, operands, stream=None):
        """cogconv_launch_epilogue: operands are the device buffers of the epilogue's field slots."""
        def ptr(x):
            return x.ptr if isinstance(x, DeviceBuffer) else x
        s = stream.handle if isinstance(stream, Stream) else stream
        arr = (_P * max(1, len(operands)))(*[ptr(o) for o in operands])
        check(lib().cogconv_launch_epilogue(self.handle, ptr(in0), ptr(in1), ptr(out), ptr(workspace), arr, s))

    def table_supported(self):
        v = ctypes.c_int32()
        check(lib().cogconv_plan_table_supported(self.handle, ctypes.byref(v)))
        return bool(v.value)

    def launch_table(self, pairs, device_table, in1=None, workspace=None, packed=False, stream=None):
        """cogconv_launch_table: pairs = [(in0, out)] one per image of the plan's batch
        (DeviceBuffers or device addresses); device_table: a DeviceBuffer of 16 bytes a pair."""
        def ptr(x):
            return x.ptr if isinstance(x, DeviceBuffer) else x
        host = (ctypes.c_void_p * (2 * len(pairs)))(*[ptr(v) for pr in pairs for v in pr])
        s = stream.handle if isinstance(stream, Stream) else stream
        check(lib().cogconv_launch_table(self.handle, host, ptr(device_table), ptr(in1), ptr(workspace),
                                         1 if packed else 0, s))
        if stream is None:
            synchronize()  # the pageable host table must outlive the copy

    def launch_noalloc(self, in0, in1, out, stream=None):
        """cogconv_launch_noalloc: the workspace comes from the library's per-device scratch array."""
        def ptr(x):
            return x.ptr if isinstance(x, DeviceBuffer) else x
        s = stream.handle if isinstance(stream, Stream) else stream
        check(lib().cogconv_launch_noalloc(self.handle, ptr(in0), ptr(in1), ptr(out), s))

    def close(self):
        if getattr(self, "handle", None):
            lib().cogconv_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def variant_count(desc):
    n = ctypes.c_int32()
    check(lib().cogconv_plan_variant_count(ctypes.byref(desc), ctypes.byref(n)))
    return n.value


def autotune(desc, warmup_steps=1, run_steps=5, flush_bytes=512 << 20, force=False, stream=None):
    """Times the candidate configurations of `desc` (cogconv_autotune); returns
    dict(variant, variants, avg_us, from_cache, description)."""
    o = TuneOptionsC(warmup_steps, run_steps, flush_bytes, 1 if force else 0)
    r = TuneResultC()
    s = stream.handle if isinstance(stream, Stream) else stream
    check(lib().cogconv_autotune(ctypes.byref(desc), ctypes.byref(o), s, ctypes.byref(r)))
    return {"variant": r.variant, "variants": r.variants, "avg_us": r.avg_us, "from_cache": bool(r.from_cache),
            "description": r.description.decode()}


def tune_cache_clear():
    check(lib().cogconv_tune_cache_clear())


def tune_cache_save(path):
    check(lib().cogconv_tune_cache_save(os.fsencode(path)))


def tune_cache_load(path):
    check(lib().cogconv_tune_cache_load(os.fsencode(path)))


def _ptr(x):
    return x.ptr if isinstance(x, DeviceBuffer) else x


def _stream(stream):
    return stream.handle if isinstance(stream, Stream) else stream


def _ptr_array(operands):
    return (_P * max(1, len(operands)))(*[_ptr(o) for o in operands])


def block_reduce_sum(inp, out, points, in_planes, factor, stream=None, epilogue=None, operands=(), in_type=None):
    """cogconv_block_reduce_sum, or cogconv_block_reduce_sum_epilogue with an EpilogueC
    (`in_type`: the input's FieldType, whose field shape the epilogue's operands follow)."""
    if epilogue is None:
        check(lib().cogconv_block_reduce_sum(_ptr(inp), _ptr(out), points, in_planes, factor, _stream(stream)))
        return
    t = in_type if isinstance(in_type, FieldTypeC) else field_type(*(in_type or (2, 1, int(points), int(in_planes))))
    check(lib().cogconv_block_reduce_sum_epilogue(_ptr(inp), _ptr(out), ctypes.byref(t), factor,
                                                  ctypes.byref(epilogue), _ptr_array(operands), _stream(stream)))


class SeparablePlan:
    """cogconv_separable on a fixed field type, filter lengths, border and orientation."""

    def __init__(self, ftype, row_taps, column_taps, border, orientation):
        self.ftype = ftype if isinstance(ftype, FieldTypeC) else field_type(*ftype)
        self.row_taps, self.column_taps = int(row_taps), int(column_taps)
        self.border, self.orientation = border, orientation
        self.epilogue = None

    def set_epilogue(self, epilogue):
        self.epilogue = epilogue

    def launch(self, inp, row_filter, column_filter, out, operands=(), stream=None):
        check(lib().cogconv_separable(_ptr(inp), ctypes.byref(self.ftype), _ptr(row_filter), self.row_taps,
                                      _ptr(column_filter), self.column_taps, self.border, self.orientation, _ptr(out),
                                      ctypes.byref(self.epilogue) if self.epilogue is not None else None,
                                      _ptr_array(operands), _stream(stream)))


def elementwise(inp, out, ftype, epilogue, operands=(), stream=None):
    """cogconv_elementwise: out = epilogue(inp) over a field of type `ftype`."""
    def ptr(x):
        return x.ptr if isinstance(x, DeviceBuffer) else x
    s = stream.handle if isinstance(stream, Stream) else stream
    t = ftype if isinstance(ftype, FieldTypeC) else field_type(*ftype)
    arr = (_P * max(1, len(operands)))(*[ptr(o) for o in operands])
    check(lib().cogconv_elementwise(ptr(inp), ptr(out), ctypes.byref(t), ctypes.byref(epilogue), arr, s))


class Function:
    """A compiled graph restored from a file and run through the reentrant
    evaluate() Function Interface (cogconv_function_*)."""

    def __init__(self, path, device=0):
        p = ctypes.c_void_p()
        check(lib().cogconv_function_create(str(path).encode(), device, ctypes.byref(p)))
        self.handle = p.value
        buf = ctypes.create_string_buffer(256)
        check(lib().cogconv_function_name(self.handle, buf, 256))
        self.name = buf.value.decode()
        self.inputs = self._io(0)
        self.outputs = self._io(1)

    def _io(self, output):
        n = ctypes.c_int32()
        fn = lib().cogconv_function_output_count if output else lib().cogconv_function_input_count
        check(fn(self.handle, ctypes.byref(n)))
        res = []
        for i in range(n.value):
            buf = ctypes.create_string_buffer(256)
            t = FieldTypeC()
            check(lib().cogconv_function_io(self.handle, output, i, buf, 256, ctypes.byref(t)))
            res.append((buf.value.decode(), t.as_tuple()))
        return res

    @staticmethod
    def _shape(t):
        d, r, c, p, o = t
        spatial = (r, c) if d == 2 else ((c,) if d == 1 else ())
        return spatial if (o == 0 and p == 1) else (p,) + spatial

    def optimal_thread_count(self):
        n = ctypes.c_int32()
        check(lib().cogconv_function_optimal_thread_count(self.handle, ctypes.byref(n)))
        return n.value

    def evaluate(self, inputs):
        """inputs: arrays in input order (planar float32); returns the outputs in order."""
        if len(inputs) != len(self.inputs):
            raise ValueError("expected %d inputs, got %d" % (len(self.inputs), len(inputs)))
        ins = []
        for a, (name, t) in zip(inputs, self.inputs):
            a = np.ascontiguousarray(a, dtype=np.float32)
            if a.size != t[1] * t[2] * t[3]:
                raise ValueError("input %s has %d values, its field type holds %d" % (name, a.size, t[1] * t[2] * t[3]))
            ins.append(a)
        outs = [np.empty(self._shape(t), np.float32) for _, t in self.outputs]
        ip = (_P * max(1, len(ins)))(*[a.ctypes.data for a in ins])
        op = (_P * max(1, len(outs)))(*[o.ctypes.data for o in outs])
        check(lib().cogconv_function_evaluate(self.handle, ip, op))
        return outs

    def close(self):
        if getattr(self, "handle", None):
            lib().cogconv_function_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class HostBuffer:
    """Pinned host memory viewed as a float32 numpy array."""

    def __init__(self, shape):
        self.shape = tuple(shape)
        n = int(np.prod(self.shape)) * 4
        p = ctypes.c_void_p()
        check(lib().cogconv_host_alloc(n, ctypes.byref(p)))
        self.ptr = p.value
        self.array = np.ctypeslib.as_array((ctypes.c_float * (n // 4)).from_address(self.ptr)).reshape(self.shape)

    def close(self):
        if getattr(self, "ptr", None):
            self.array = None
            lib().cogconv_host_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """The serving loop of libcogconv (cogconv_server_*)."""

    def __init__(self, budget_bytes=1 << 30, max_inflight=8, streams=4, max_stack=8, device=0, graph_after=4):
        cfg = ServerConfigC(device, budget_bytes, max_inflight, streams, max_stack, graph_after)
        p = ctypes.c_void_p()
        check(lib().cogconv_server_create(ctypes.byref(cfg), ctypes.byref(p)))
        self.handle = p.value
        self._keep = {}

    def register_bank(self, weights):
        """weights: planar (planes, rows, cols) float32; returns the bank id."""
        w = np.ascontiguousarray(weights, dtype=np.float32)
        ft = FieldTypeC(2, w.shape[1], w.shape[2], w.shape[0], 0 if w.shape[0] == 1 else 1)
        bid = ctypes.c_int32()
        check(lib().cogconv_server_register_bank(self.handle, w.ctypes.data, ctypes.byref(ft), ctypes.byref(bid)))
        return bid.value

    def submit(self, desc, in0, out, bank=-1, in1=None):
        """in0 / in1 / out: numpy float32 arrays (pinned HostBuffer arrays for overlap),
        or DeviceBuffers / device addresses (then no copies: the request's state is
        resident in HBM); they must stay alive until the ticket completes."""
        resident = isinstance(in0, (DeviceBuffer, int))

        def addr(x):
            if x is None:
                return None
            if isinstance(x, DeviceBuffer):
                return x.ptr
            if isinstance(x, int):
                return x
            return x.ctypes.data
        req = RequestC(desc, bank, REQUEST_RESIDENT if resident else 0, addr(in0), addr(in1), addr(out))
        t = ctypes.c_uint64()
        check(lib().cogconv_server_submit(self.handle, ctypes.byref(req), ctypes.byref(t)))
        self._keep[t.value] = (in0, in1, out)
        return t.value

    def submit_timed(self, requests, arrivals):
        """requests: list of RequestC (built by request()); arrivals: seconds from now.
        Submits natively at those times; returns the tickets (0 = refused)."""
        n = len(requests)
        arr = (RequestC * n)(*requests)
        ts = (ctypes.c_double * n)(*[float(a) for a in arrivals])
        tk = (ctypes.c_uint64 * n)()
        st = lib().cogconv_server_submit_timed(self.handle, arr, ts, n, tk)
        if st not in (OK, ERR_BUDGET):
            check(st)
        return list(tk)

    @staticmethod
    def request(desc, in0, out, bank=-1, in1=None, resident=False):
        """A RequestC for submit_timed; buffers are addresses (ints) or numpy arrays."""
        def addr(x):
            if x is None:
                return None
            if isinstance(x, DeviceBuffer):
                return x.ptr
            if isinstance(x, int):
                return x
            return x.ctypes.data
        return RequestC(desc, bank, REQUEST_RESIDENT if resident else 0, addr(in0), addr(in1), addr(out))

    def done(self, ticket):
        s = lib().cogconv_server_query(self.handle, ticket)
        if s == ERR_NOT_READY:
            return False
        check(s)
        self._keep.pop(ticket, None)
        return True

    def wait(self, ticket):
        check(lib().cogconv_server_wait(self.handle, ticket))
        self._keep.pop(ticket, None)

    def drain(self):
        check(lib().cogconv_server_drain(self.handle))
        self._keep.clear()

    def stats(self):
        st = ServerStatsC()
        check(lib().cogconv_server_stats_get(self.handle, ctypes.byref(st)))
        return {n: getattr(st, n) for n, _ in ServerStatsC._fields_}

    def reset_stats(self):
        check(lib().cogconv_server_stats_reset(self.handle))

    def close(self):
        if getattr(self, "handle", None):
            lib().cogconv_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
