"""Autotuner on the GPU: every candidate configuration of a plan gives the
oracle's result, and cogconv_autotune picks, caches and applies one.

Marked `gpu`; each test prints its case before touching the device.
"""
import sys
import zlib

import numpy as np
import pytest

import cogconv._lib as L
import oracle as O
from gpuhelp import assert_parity, desc_pair, oracle_ref, run_gpu

pytestmark = pytest.mark.gpu


def announce(*what):
    print("CASE", *what, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def device():
    L.init(0)
    L.tune_cache_clear()
    yield
    L.tune_cache_clear()
    L.synchronize()


CASES = [
    # (mode, border, orient, sampling, step, batch, in0 shape, in1 shape)
    (O.PF_BRS, O.VALID, O.XCORR, O.NONE, 1, 2, (2 * 5, 20, 24), (3 * 5, 3, 3)),
    (O.PF_BRS, O.CYCLIC, O.XCORR, O.DOWN, 2, 2, (2 * 6, 20, 24), (4 * 6, 5, 5)),
    (O.BPF_BRS, O.FULL, O.CONV, O.NONE, 1, 2, (2 * 3, 14, 18), (3 * 5, 3, 3)),
    (O.FA_BRS, O.VALID, O.XCORR, O.NONE, 1, 3, (3 * 4, 20, 22), (3 * 5, 18, 20)),
    (O.FA_BRS, O.VALID, O.CONV, O.UP, 2, 2, (2 * 3, 21, 23), (2 * 2, 9, 10)),
    # FilterAdjoint with 2, 3 and 4 row blocks (K 32, 40, 64): the 4- and
    # 8-wave kernels, one- and two-pass dy rows (68 tap columns), down-sampling.
    (O.FA_BRS, O.VALID, O.XCORR, O.NONE, 1, 2, (2 * 16, 34, 36), (2 * 32, 32, 34)),
    (O.FA_BRS, O.VALID, O.CONV, O.DOWN, 2, 2, (2 * 6, 34, 38), (2 * 40, 15, 17)),
    (O.FA_BRS, O.VALID, O.XCORR, O.NONE, 1, 1, (1 * 3, 90, 70), (1 * 64, 85, 66)),
    (O.PBP, O.CLAMP, O.CONV, O.NONE, 1, 1, (2, 40, 44), (2, 7, 7)),
]


@pytest.mark.parametrize("case", CASES)
def test_every_candidate_matches_oracle(case):
    rng = np.random.default_rng(zlib.crc32(repr(case).encode()))
    mode, border, orient, sampling, step, batch, s0, s1 = case
    a = rng.uniform(-1, 1, s0).astype(np.float32)
    k = rng.uniform(-1, 1, s1).astype(np.float32)
    od, cd = desc_pair(mode, border, orient, sampling, step, batch, s0, s1)
    ref, mag = oracle_ref(od, a, k)
    for v in range(L.variant_count(cd)):
        p = L.Plan(cd, v)
        announce("candidate", v, p.description)
        p.close()
        got = run_gpu(cd, a, k, variant=v)
        assert_parity(got, ref, mag, what="candidate %d" % v)


def test_autotune_picks_caches_and_applies(tmp_path):
    """c2 FilterAdjoint+BRS, BorderValid: batch 16, x 64x64x16, dy 62x62x32."""
    od, cd = desc_pair(O.FA_BRS, O.VALID, O.XCORR, O.NONE, 1, 16, (16 * 16, 64, 64), (16 * 32, 62, 62))
    announce("autotune c2 FA")
    r = L.autotune(cd, warmup_steps=1, run_steps=3, flush_bytes=64 << 20)
    assert 0 <= r["variant"] < r["variants"] and not r["from_cache"] and r["avg_us"] > 0
    assert L.Plan(cd).variant == r["variant"]
    assert L.Plan(cd).description == r["description"]
    again = L.autotune(cd, warmup_steps=1, run_steps=3, flush_bytes=64 << 20)
    assert again["from_cache"] and again["variant"] == r["variant"]
    more = L.autotune(cd, warmup_steps=1, run_steps=4, flush_bytes=64 << 20)  # more steps than cached: re-timed
    assert not more["from_cache"]
    path = str(tmp_path / "tune.txt")
    L.tune_cache_save(path)
    L.tune_cache_clear()
    assert L.Plan(cd).variant == 0
    L.tune_cache_load(path)
    assert L.Plan(cd).variant == more["variant"]
    L.tune_cache_clear()
