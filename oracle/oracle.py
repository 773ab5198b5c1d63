"""ctypes wrapper of the CPU oracle (liboracle.so, built from cogconv_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
bench.py's cpu_baseline leg, always as the checker.  The product path
(libcogconv and the cogconv package) never imports this module.

Codes match include/cogconv.h:
  modes:    PBP=0 PF=1 PF_BRS=2 BPF=3 BPF_BRS=4 FA=5 FA_BRS=6
  borders:  ZERO=0 CLAMP=1 CYCLIC=2 VALID=3 FULL=4
  sampling: NONE=0 UP=1 DOWN=2
  orient:   XCORR=0 CONV=1
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

PBP, PF, PF_BRS, BPF, BPF_BRS, FA, FA_BRS = range(7)
ZERO, CLAMP, CYCLIC, VALID, FULL = range(5)
NONE, UP, DOWN = range(3)
XCORR, CONV = range(2)

BORDER_NAMES = {"BorderZero": ZERO, "BorderClamp": CLAMP, "BorderCyclic": CYCLIC,
                "BorderValid": VALID, "BorderFull": FULL}
SAMPLING_NAMES = {"NoSamplingConvolution": NONE, "UpsampleInputConvolution": UP,
                  "DownsampleOutputConvolution": DOWN}

ERRORS = {1: "field dimensions", 2: "batch", 3: "filter depth", 4: "even filter",
          5: "downsample divisibility", 6: "upsample with BorderClamp",
          7: "PlaneByPlane shapes", 8: "FilterAdjoint upsample border", 9: "argument"}


class Desc(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in (
        "mode", "border", "orient", "sampling", "step", "batch",
        "in0_dims", "in0_rows", "in0_cols", "in0_planes", "in0_order",
        "in1_dims", "in1_rows", "in1_cols", "in1_planes", "in1_order")]


class OutType(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in ("dims", "rows", "cols", "planes", "order")]


class OracleError(ValueError):
    def __init__(self, code):
        super().__init__("oracle rejects ConvolveOp: %s (code %d)" % (ERRORS.get(code, "?"), code))
        self.code = code


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = ctypes.CDLL(path)
        L.oracle_output_type.argtypes = [ctypes.POINTER(Desc), ctypes.POINTER(OutType)]
        L.oracle_conv.argtypes = [ctypes.POINTER(Desc), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                  ctypes.c_int, ctypes.c_int]
        L.oracle_scalar.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                    ctypes.c_int] + [ctypes.c_int] * 11 + [ctypes.c_void_p]
        L.oracle_block_reduce_sum.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                              ctypes.c_void_p]
        L.oracle_upsample.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_void_p]
        L.oracle_plane_pair.argtypes = [ctypes.POINTER(Desc), ctypes.c_int, ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        L.oracle_max_threads.restype = ctypes.c_int
        L.oracle_separable.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                       ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_void_p]
        _LIB = L
    return _LIB


def make_desc(mode, border, orient, sampling, step, batch, in0_shape, in1_shape, in0_order=None, in1_order=None):
    """in*_shape = (planes, rows, cols) for 2-D fields or (planes, cols) for 1-D fields."""
    def unpack(s):
        if len(s) == 3:
            return 2, s[1], s[2], s[0]
        return 1, 1, s[1], s[0]
    d0, r0, c0, p0 = unpack(in0_shape)
    d1, r1, c1, p1 = unpack(in1_shape)
    if in0_order is None:
        in0_order = 0 if p0 == 1 else 1
    if in1_order is None:
        in1_order = 0 if p1 == 1 else 1
    return Desc(mode, border, orient, sampling, step, batch, d0, r0, c0, p0, in0_order, d1, r1, c1, p1, in1_order)


def output_type(desc):
    out = OutType()
    e = lib().oracle_output_type(ctypes.byref(desc), ctypes.byref(out))
    if e:
        raise OracleError(e)
    return out.dims, out.rows, out.cols, out.planes, out.order


def conv(desc, in0, in1, prec="f64", threads=0):
    """Full ConvolveOp on planar float32 arrays; returns (planes, rows, cols)."""
    _, rows, cols, planes, _ = output_type(desc)
    a = np.ascontiguousarray(in0, dtype=np.float32)
    f = np.ascontiguousarray(in1, dtype=np.float32)
    assert a.size == desc.in0_planes * desc.in0_rows * desc.in0_cols
    assert f.size == desc.in1_planes * desc.in1_rows * desc.in1_cols
    out = np.empty((planes, rows, cols), dtype=np.float64 if prec == "f64" else np.float32)
    e = lib().oracle_conv(ctypes.byref(desc), a.ctypes.data, f.ctypes.data, out.ctypes.data,
                          0 if prec == "f64" else 1, int(threads))
    if e:
        raise OracleError(e)
    return out


def scalar(a, f, border, orient, upA=(1, 1), upF=(1, 1), dn=(1, 1), out_shape=None, prec="f64"):
    a = np.ascontiguousarray(a, dtype=np.float32)
    f = np.ascontiguousarray(f, dtype=np.float32)
    oh, ow = out_shape
    out = np.empty((oh, ow), dtype=np.float64 if prec == "f64" else np.float32)
    lib().oracle_scalar(a.ctypes.data, a.shape[0], a.shape[1], f.ctypes.data, f.shape[0], f.shape[1],
                        border, orient, upA[0], upA[1], upF[0], upF[1], dn[0], dn[1], oh, ow,
                        0 if prec == "f64" else 1, out.ctypes.data)
    return out


def plane_pair(desc, o, r):
    pa, pf = ctypes.c_int(), ctypes.c_int()
    lib().oracle_plane_pair(ctypes.byref(desc), o, r, ctypes.byref(pa), ctypes.byref(pf))
    return pa.value, pf.value


def block_reduce_sum(x, factor):
    """x: (planes, rows, cols) -> (planes/factor, rows, cols) in float64."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    planes = x.shape[0]
    points = x[0].size
    out = np.empty((planes // factor,) + x.shape[1:], dtype=np.float64)
    e = lib().oracle_block_reduce_sum(x.ctypes.data, points, planes, factor, out.ctypes.data)
    if e:
        raise OracleError(e)
    return out


def upsample(p, ur, uc):
    p = np.ascontiguousarray(p, dtype=np.float32)
    out = np.empty((p.shape[0] * ur, p.shape[1] * uc), dtype=np.float32)
    lib().oracle_upsample(p.ctypes.data, p.shape[0], p.shape[1], ur, uc, out.ctypes.data)
    return out


def separable(x, row_filter, column_filter, border, orient, prec="f64"):
    """Separable convolution of a 2-D plane: the row pass, then the column pass
    (float64 result; with prec="f32" each pass is summed in float in tap order)."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    rf = np.ascontiguousarray(row_filter, dtype=np.float32).ravel()
    cf = np.ascontiguousarray(column_filter, dtype=np.float32).ravel()
    out = np.empty(x.shape, dtype=np.float64)
    e = lib().oracle_separable(x.ctypes.data, x.shape[0], x.shape[1], rf.ctypes.data, rf.size, cf.ctypes.data,
                               cf.size, border, orient, 0 if prec == "f64" else 1, out.ctypes.data)
    if e:
        raise OracleError(e)
    return out


def max_threads():
    return lib().oracle_max_threads()
