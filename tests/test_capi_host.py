"""Host-side checks of libcogconv (no GPU needed).

- the library loads and exports every function include/cogconv.h declares;
- cogconv_output_type (the reference's shape function and legality checks)
  agrees with the oracle's on a seeded sweep of ConvolveOps, legal and illegal;
- plans build on the host and report the algorithmic FLOPs of the configs;
- the Field / ComputeGraph API raises at graph construction on illegal ops.
"""
import ctypes
import itertools
import os
import re

import numpy as np
import pytest

import cogconv
import cogconv._lib as L
import oracle as O
from conftest import ROOT


def declared_functions():
    text = open(os.path.join(ROOT, "include", "cogconv.h")).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+char\s*\*|int)\s+(cogconv_\w+)\s*\(", text, re.M)))


def test_header_declares_what_the_binding_binds():
    assert declared_functions() == sorted(L.SIGNATURES)


def test_library_exports_every_declared_symbol():
    lib = L.lib()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing


def test_abi_version_and_error_convention():
    lib = L.lib()
    assert lib.cogconv_abi_version() == 1
    # Every failure is a status code with a message, never an exception.
    status = lib.cogconv_plan_create(None, None)
    assert status == L.ERR_INVALID_ARG
    assert "null" in L.last_error()


def _sweep():
    rng = np.random.default_rng(1234)
    modes = range(7)
    borders = range(5)
    samplings = ((O.NONE, 1), (O.UP, 2), (O.UP, 3), (O.DOWN, 2), (O.DOWN, 3), (O.UP, 1), (O.DOWN, 1))
    cases = []
    for mode, border, (samp, step) in itertools.product(modes, borders, samplings):
        for _ in range(3):
            dims = int(rng.integers(1, 3))
            batch = int(rng.choice([1, 2, 3]))
            n = int(rng.choice([1, 2, 3, 4]))
            k = int(rng.choice([1, 2, 5]))
            rows = int(rng.integers(4, 20)) if dims == 2 else 1
            cols = int(rng.integers(4, 20))
            frows = int(rng.choice([1, 2, 3, 4, 5])) if dims == 2 else 1
            fcols = int(rng.choice([1, 2, 3, 5]))
            p0 = n * batch
            if mode in (O.FA, O.FA_BRS):
                p1 = k * batch
            elif mode == O.PBP:
                p1 = int(rng.choice([1, p0]))
            else:
                p1 = n * k
            if rng.random() < 0.1:
                p1 += 1  # provoke depth errors
            in0 = (p0, rows, cols) if dims == 2 else (p0, cols)
            in1 = (p1, frows, fcols) if dims == 2 else (p1, fcols)
            cases.append((mode, border, samp, step, batch, in0, in1))
    return cases


@pytest.mark.parametrize("case", _sweep())
def test_output_type_matches_oracle(case):
    mode, border, samp, step, batch, in0, in1 = case
    od = O.make_desc(mode, border, O.XCORR, samp, step, batch, in0, in1)
    dims = 2 if len(in0) == 3 else 1

    def ft(s):
        return (dims, s[1] if dims == 2 else 1, s[-1], s[0])

    cd = L.make_desc(mode, border, L.CROSS_CORRELATION, samp, step, batch, ft(in0), ft(in1))
    try:
        want = O.output_type(od)
    except O.OracleError:
        with pytest.raises(L.CogconvError) as e:
            L.output_type(cd)
        assert e.value.status in (L.ERR_ILLEGAL_OP, L.ERR_INVALID_ARG)
        assert L.last_error()
        return
    assert L.output_type(cd) == want


def c2_plan(mode, border, sampling=L.NO_SAMPLING, step=1):
    B, N, K = 16, 16, 32
    if mode == L.PROJECT_FRAME_BLOCK_REDUCE_SUM:
        in0, in1 = (2, 64, 64, B * N), (2, 3, 3, K * N)
    elif mode == L.BACK_PROJECT_FRAME_BLOCK_REDUCE_SUM:
        s = 62 if border == L.BORDER_FULL else 64
        in0, in1 = (2, s, s, B * K), (2, 3, 3, K * N)
    else:
        in0, in1 = (2, 64, 64, B * N), (2, 62, 62, B * K)
    return L.Plan(L.make_desc(mode, border, L.CROSS_CORRELATION if mode != L.BACK_PROJECT_FRAME_BLOCK_REDUCE_SUM
                              else L.CONVOLUTION, sampling, step, B, in0, in1))


def test_plan_flops_match_the_algorithmic_counts():
    """Algorithmic FLOPs of the CNN-layer triple (B16 64x64x16, 32 filters 3x3x16)."""
    assert c2_plan(L.PROJECT_FRAME_BLOCK_REDUCE_SUM, L.BORDER_ZERO).flops == 603979776
    assert c2_plan(L.PROJECT_FRAME_BLOCK_REDUCE_SUM, L.BORDER_CYCLIC).flops == 603979776
    assert c2_plan(L.PROJECT_FRAME_BLOCK_REDUCE_SUM, L.BORDER_VALID).flops == 566820864
    assert c2_plan(L.PROJECT_FRAME_BLOCK_REDUCE_SUM, L.BORDER_ZERO, L.DOWNSAMPLE_OUTPUT, 2).flops == 150994944
    assert c2_plan(L.BACK_PROJECT_FRAME_BLOCK_REDUCE_SUM, L.BORDER_ZERO).flops == 603979776
    assert c2_plan(L.BACK_PROJECT_FRAME_BLOCK_REDUCE_SUM, L.BORDER_FULL).flops == 603979776
    p = c2_plan(L.FILTER_ADJOINT_BLOCK_REDUCE_SUM, L.BORDER_VALID)
    assert p.flops == 566820864
    assert p.out_shape == (512, 3, 3)


def test_plan_is_host_only_and_describes_its_kernel():
    p = c2_plan(L.FILTER_ADJOINT_BLOCK_REDUCE_SUM, L.BORDER_VALID)
    assert p.description.startswith(("small_output", "fa_igemm", "igemm"))
    assert p.workspace_bytes >= 0
    q = c2_plan(L.PROJECT_FRAME_BLOCK_REDUCE_SUM, L.BORDER_CYCLIC)
    assert q.out_shape == (512, 64, 64)


def test_illegal_ops_raise_at_graph_construction():
    x = np.zeros((3, 8, 8), np.float32)
    with cogconv.ComputeGraph():
        xf = cogconv.VectorField(x)
        even = cogconv.VectorField(np.zeros((6, 4, 4), np.float32))
        with pytest.raises(cogconv.CogconvError, match="odd shape"):
            cogconv.projectFrame(xf, even, cogconv.BorderZero)
        odd = cogconv.VectorField(np.zeros((6, 3, 3), np.float32))
        with pytest.raises(cogconv.CogconvError, match="BorderClamp"):
            cogconv.projectFrame(xf, odd, cogconv.BorderClamp, cogconv.UpsampleInputConvolution(2))
        with pytest.raises(cogconv.CogconvError, match="divisible"):
            cogconv.projectFrame(cogconv.VectorField(np.zeros((3, 9, 9), np.float32)), odd, cogconv.BorderZero,
                                 cogconv.DownsampleOutputConvolution(2))
        bad = cogconv.VectorField(np.zeros((4, 3, 3), np.float32))
        with pytest.raises(cogconv.CogconvError, match="multiple"):
            cogconv.projectFrame(xf, bad, cogconv.BorderZero)
        y = cogconv.projectFrame(xf, odd, cogconv.BorderZero)
        assert y.shape == (6, 8, 8)
        assert cogconv.blockReduceSum(y, 3).shape == (2, 8, 8)
        # Up/Down(1) behave as NoSampling (CogFunctionAPI.scala:1140-1144)
        z = cogconv.projectFrame(xf, odd, cogconv.BorderZero, cogconv.DownsampleOutputConvolution(1))
        assert z.shape == (6, 8, 8)


def test_epilogue_checks_are_host_side():
    """cogconv_plan_set_epilogue checks ops and operand types against the output
    type (ResultType.scala:29-58) without a device."""
    import cogconv._lib as L
    desc = L.make_desc(L.PROJECT_FRAME_BLOCK_REDUCE_SUM, L.BORDER_ZERO, L.CROSS_CORRELATION, L.NO_SAMPLING, 1, 2,
                       (2, 16, 16, 8), (2, 3, 3, 16))
    plan = L.Plan(desc)  # output: 2 images x 4 planes of 16x16 = 8 planes
    ok = [(2, 16, 16, 8, 1), (0, 1, 1, 8, 1), (2, 16, 16, 1, 0), (0, 1, 1, 1, 0)]
    for t in ok:
        plan.set_epilogue(L.make_epilogue([(L.EPI_ADD_FIELD, 0.0, 0)], [t]))
    for t in [(2, 16, 16, 3, 1), (2, 8, 16, 8, 1), (0, 1, 1, 3, 1)]:
        with pytest.raises(L.CogconvError) as e:
            plan.set_epilogue(L.make_epilogue([(L.EPI_ADD_FIELD, 0.0, 0)], [t]))
        assert e.value.status == L.ERR_ILLEGAL_OP
    with pytest.raises(L.CogconvError) as e:
        plan.set_epilogue(L.make_epilogue([(99, 0.0, 0)]))
    assert e.value.status == L.ERR_INVALID_ARG
    with pytest.raises(L.CogconvError) as e:
        plan.set_epilogue(L.make_epilogue([(L.EPI_MUL_FIELD, 0.0, 1)], [(2, 16, 16, 8, 1)]))
    assert e.value.status == L.ERR_INVALID_ARG
    plan.set_epilogue(None)
    plan.close()


def test_graph_file_written_and_parse_errors(tmp_path):
    """ComputeGraph.write produces the documented text header + data layout, and
    cogconv_function_create rejects malformed files before touching a device."""
    rng = np.random.default_rng(9)
    cg = cogconv.ComputeGraph()
    with cg:
        x = cogconv.VectorField(rng.uniform(-1, 1, (4, 10, 10)).astype(np.float32))
        w = cogconv.VectorField(rng.uniform(-1, 1, (8, 3, 3)).astype(np.float32))
        y = cogconv.max(cogconv.blockReduceSum(cogconv.projectFrame(x, w, cogconv.BorderZero, batchSize=2), 2), 0.0)
        cogconv.probe(y)
    path = cg.write(str(tmp_path / "g.cgraph"), inputs=[x], name="tiny")
    raw = open(path, "rb").read()
    head, data = raw.split(b"end\n", 1)
    lines = head.decode().splitlines()
    assert lines[0] == "COGCONV-GRAPH 1 0" and lines[1] == "name tiny"
    assert [ln.split()[7] for ln in lines if ln.startswith("field")] == ["input", "const", "computed"]
    step = [ln for ln in lines if ln.startswith("step")]
    assert len(step) == 1 and step[0].split()[1] == "conv"
    assert int(step[0].split()[5]) == L.PROJECT_FRAME_BLOCK_REDUCE_SUM
    assert [ln.split()[1] for ln in lines if ln.startswith("epi")] == [str(L.EPI_MAX_CONST)]
    assert np.array_equal(np.frombuffer(data, "<f4"), cg._fields[1].data.ravel())
    for bad in (b"NOT-A-GRAPH\n", b"COGCONV-GRAPH 2 0\nend\n",
                b"COGCONV-GRAPH 1 0\nfield 0 2 4 4 1 0 input a\nstep reduce 1 0 2\nend\n",
                b"COGCONV-GRAPH 1 0\nfield 0 2 4 4 2 1 input a\nfield 1 2 4 4 1 0 computed b\n"
                b"step elementwise 1 0\nend\n",
                b"COGCONV-GRAPH 1 0\nfield 0 2 4 4 1 0 const a\nend\n"):
        p = tmp_path / "bad.cgraph"
        p.write_bytes(bad)
        h = ctypes.c_void_p()
        st = L.lib().cogconv_function_create(str(p).encode(), 0, ctypes.byref(h))
        assert st == L.ERR_INVALID_ARG, (bad, st, L.last_error())
