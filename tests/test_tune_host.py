"""Host-side checks of the plan candidates and the autotuning cache (no GPU).

- every descriptor has at least one candidate configuration; candidate 0 is the
  plan an untuned COGCONV_ALGO_AUTO descriptor gets; candidates are distinct and
  all describe the same result FieldType;
- the algorithm hints pick the first candidate of their kernel family;
- the cache file (the counterpart of ProfilerCache) round-trips and rejects
  malformed samples.
"""
import os

import pytest

import cogconv._lib as L
import oracle as O
from gpuhelp import desc_pair

DESCS = [
    # (mode, border, orient, sampling, step, batch, in0 shape, in1 shape)
    (O.PF_BRS, O.VALID, O.XCORR, O.NONE, 1, 16, (16 * 16, 64, 64), (32 * 16, 3, 3)),
    (O.PF_BRS, O.CYCLIC, O.XCORR, O.DOWN, 2, 16, (16 * 16, 64, 64), (32 * 16, 3, 3)),
    (O.BPF_BRS, O.FULL, O.CONV, O.NONE, 1, 16, (16 * 32, 62, 62), (32 * 16, 3, 3)),
    (O.FA_BRS, O.VALID, O.XCORR, O.NONE, 1, 16, (16 * 16, 64, 64), (16 * 32, 62, 62)),
    (O.PF_BRS, O.ZERO, O.XCORR, O.NONE, 1, 1, (3, 32, 32), (24, 5, 5)),
    (O.PBP, O.CLAMP, O.CONV, O.NONE, 1, 1, (1, 256, 256), (1, 11, 11)),
    (O.FA, O.ZERO, O.XCORR, O.NONE, 1, 2, (2 * 3, 10, 12), (2 * 2, 5, 5)),
    (O.FA_BRS, O.VALID, O.CONV, O.UP, 2, 2, (2 * 3, 21, 23), (2 * 2, 5, 6)),
]


def cdesc(case, **kw):
    return desc_pair(*case, **kw)[1]


@pytest.mark.parametrize("case", DESCS)
def test_candidates(case):
    d = cdesc(case)
    n = L.variant_count(d)
    assert n >= 1
    default = L.Plan(d)
    assert default.variant == 0
    names = []
    for i in range(n):
        p = L.Plan(d, i)
        assert p.variant == i
        assert p.out_type == default.out_type
        assert p.flops == default.flops
        names.append(p.description)
        p.close()
    assert names[0] == default.description
    assert len(set(names)) == n
    with pytest.raises(L.CogconvError) as e:
        L.Plan(d, n)
    assert e.value.status == L.ERR_INVALID_ARG


def test_precision_candidates():
    d = cdesc((O.PF_BRS, O.ZERO, O.XCORR, O.NONE, 1, 32, (32 * 64, 128, 128), (64 * 64, 3, 3)), precision=L.BF16)
    n = L.variant_count(d)
    assert n >= 1
    for i in range(n):
        p = L.Plan(d, i)
        assert p.description.startswith("igemm_bf16")
        p.close()


@pytest.mark.parametrize("algo,prefix", [(L.ALGO_DIRECT, "direct"), (L.ALGO_IMPLICIT_GEMM, "fa_igemm"),
                                         (L.ALGO_SMALL_OUTPUT, "small_output")])
def test_algo_hint_picks_first_of_family(algo, prefix):
    case = DESCS[3]
    d = cdesc(case, algo=algo)
    p = L.Plan(d)
    assert p.description.startswith(prefix)
    auto = cdesc(case)
    firsts = [i for i in range(L.variant_count(auto)) if L.Plan(auto, i).description.startswith(prefix)]
    assert p.variant == firsts[0]


def test_tune_cache_file_round_trip(tmp_path):
    L.tune_cache_clear()
    src = tmp_path / "in.txt"
    lines = ["gfx950/AMD_Instinct_MI355X\tm6 b3 o0 s0 st1 B16 a2,64,64,256,1 f2,62,62,512,1 p0\t3\t12.500\t1\t5\t"
             "1700000000000\tfa_igemm_f32 x",
             "gfx950/AMD_Instinct_MI355X\tm2 b3 o0 s0 st1 B16 a2,64,64,256,1 f2,3,3,512,1 p0\t0\t8.250\t1\t5\t"
             "1700000000001\tigemm_f32 y"]
    src.write_text("# cogconv tune cache 1\n" + "\n".join(lines) + "\n")
    L.tune_cache_load(str(src))
    out = tmp_path / "out.txt"
    L.tune_cache_save(str(out))
    got = [ln for ln in out.read_text().splitlines() if not ln.startswith("#")]
    assert sorted(got) == sorted(lines)
    L.tune_cache_clear()
    L.tune_cache_save(str(out))
    assert [ln for ln in out.read_text().splitlines() if not ln.startswith("#")] == []


def test_tune_cache_rejects_malformed(tmp_path):
    L.tune_cache_clear()
    bad = tmp_path / "bad.txt"
    bad.write_text("dev\tkey\tnot-a-number\t1.0\t1\t5\t0\tname\n")
    with pytest.raises(L.CogconvError) as e:
        L.tune_cache_load(str(bad))
    assert e.value.status == L.ERR_INVALID_ARG
    L.tune_cache_load(os.path.join(str(tmp_path), "missing.txt"))  # a missing file is not an error
    L.tune_cache_clear()
