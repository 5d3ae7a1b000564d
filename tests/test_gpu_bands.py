"""Banded passes of the FP32 pair-once self field.

The pair-once kernel keeps one FP32 partial field per particle and block.  When those
would take more than 13 GiB, the block pairs (I, J) run in bands of columns J, one pass
each, and each pass adds its partials (in partner order, FP64) to the result.  Every
item computes the same FP32 partials either way, so banded and single-pass results
differ only by the grouping of FP64 sums; both are checked against the FP64 oracle
at the direct-sum tolerance (median 1e-6, p99 1e-4).
"""
import numpy as np
import pytest

from conftest import rel_err
from gravhopper_amd import _native, workloads

pytestmark = pytest.mark.gpu

FP32, FP64 = _native.GH_FP32, _native.GH_FP64
EPS = 1e-5


def banded(ctx, pos, mass, eps, band, ipt=1):
    ctx.set_tuning(_native.TUNE_DIRECT_SELF, ipt)
    ctx.set_tuning(_native.TUNE_SYM_BAND, band)
    try:
        return ctx.direct(pos, mass, eps, FP32)
    finally:
        ctx.set_tuning(_native.TUNE_SYM_BAND, 0)
        ctx.set_tuning(_native.TUNE_DIRECT_SELF, 0)


def two_spheres(n, seed):
    a = workloads.plummer(n // 2, b=1e-3, mass=1e6, seed=seed)
    b = workloads.plummer(n - n // 2, b=1e-3, mass=1e6, seed=seed + 1)
    pos = np.vstack([a["pos"] + [50.0, 0, 0], b["pos"] - [50.0, 0, 0]])
    return pos, np.concatenate([a["mass"], b["mass"]])


@pytest.mark.parametrize("case", ["plummer", "two_spheres"])
@pytest.mark.parametrize("band", [1, 7, 40])
def test_bands_match_single_pass(ctx, oracle, case, band):
    # 20000 particles in 79 blocks of 256: bands of 1, 7 (the last one partial) and 40
    if case == "plummer":
        p = workloads.plummer(20000, b=1e-3, mass=1e6, seed=61)
        pos, mass = p["pos"], p["mass"]
    else:
        pos, mass = two_spheres(20000, 62)  # blocks straddle the spheres: the DS pass too
    one = banded(ctx, pos, mass, EPS, 0)
    got = banded(ctx, pos, mass, EPS, band)
    scale = np.abs(one).max()
    np.testing.assert_allclose(got, one, rtol=1e-11, atol=1e-13 * scale)
    err = rel_err(got, oracle.direct(pos, mass, EPS))
    print("%s band %d: median %.2e p99 %.2e" % (case, band, np.median(err),
                                                np.percentile(err, 99)), flush=True)
    assert np.median(err) <= 1e-6 and np.percentile(err, 99) <= 1e-4


def test_bands_repeat_bitwise(ctx):
    p = workloads.plummer(9000, b=1e-3, mass=1e6, seed=63)
    a = banded(ctx, p["pos"], p["mass"], EPS, 5)
    b = banded(ctx, p["pos"], p["mass"], EPS, 5)
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
