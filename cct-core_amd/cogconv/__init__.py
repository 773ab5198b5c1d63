"""cogconv: MI355X-native Field convolution path of cct-core (libcogconv.so + host API).

`cogconv._lib` binds the C ABI of include/cogconv.h; `cogconv.api` mirrors the
reference's Field / ComputeGraph operator API for this path (including its
elementwise sinks, which merge into the convolution's stores).
"""
from . import _lib as lib  # noqa: F401
from ._lib import CogconvError  # noqa: F401
from .api import (  # noqa: F401
    BorderClamp, BorderCyclic, BorderFull, BorderValid, BorderZero, Cog, ComputeGraph,
    DownsampleOutputConvolution, NoSamplingConvolution, ScalarField, UpsampleInputConvolution, VectorField, abs,
    backProjectFrame, blockReduceSum, convolve, convolveFilterAdjoint, convolveSeparable, crossCorrelate,
    crossCorrelateFilterAdjoint, crossCorrelateSeparable, exp, log, max, min, probe, projectFrame, readFromFile, signum,
    sq, sqrt, tanh)
