"""Parity of the GPU direct sum (libgravhop.so through the C ABI) with the reference.

FP64 mode must be bit-identical to the reference's direct_summation and
direct_summation_position (golden vectors written by the reference's own C code, and
the oracle restatement at other sizes).  FP32 mode is checked against the FP64 result
with a stated tolerance.
"""
import numpy as np
import pytest

from conftest import assert_bitwise, rel_err
from gravhopper_amd import _native, workloads

pytestmark = pytest.mark.gpu

FP64, FP32 = _native.GH_FP64, _native.GH_FP32


def test_fp64_direct_bitwise_golden(ctx, golden):
    for name, c in golden.items():
        print("case", name, flush=True)
        got = ctx.direct(c["pos"], c["mass"], float(c["eps"]), FP64)
        assert_bitwise(got, c["direct"], name)
        got = ctx.direct(c["pos"], c["mass"], float(c["eps"]), FP64, force_pos=c["force_pos"])
        assert_bitwise(got, c["direct_position"], name + " position")


@pytest.mark.parametrize("n", [1, 2, 3, 255, 256, 257, 1000, 2048])
def test_fp64_direct_bitwise_oracle(ctx, oracle, n):
    rng = np.random.default_rng(100 + n)
    pos = rng.standard_cauchy(size=(n, 3))
    mass = rng.uniform(0.1, 3.0, size=n)
    for eps in (0.0, 1e-3, 0.7):
        if n == 1 or eps > 0 or n > 1:
            got = ctx.direct(pos, mass, eps, FP64)
            assert_bitwise(got, oracle.direct(pos, mass, eps), "n=%d eps=%g" % (n, eps))
    fpos = np.concatenate([pos[: min(n, 7)], rng.normal(size=(5, 3))])
    for eps in (0.0, 0.01):
        got = ctx.direct(pos, mass, eps, FP64, force_pos=fpos)
        assert_bitwise(got, oracle.direct_position(pos, mass, fpos, eps), "pos n=%d" % n)


def test_fp64_direct_plummer_n2000(ctx, oracle):
    ic = workloads.plummer(2000, b=1e-3, mass=1e6, seed=20260001)
    got = ctx.direct(ic["pos"], ic["mass"], 5e-5, FP64)
    assert_bitwise(got, oracle.direct(ic["pos"], ic["mass"], 5e-5), "plummer 2000")


def test_empty_and_single(ctx):
    out = ctx.direct(np.zeros((0, 3)), np.zeros(0), 0.1, FP64)
    assert out.shape == (0, 3)
    out = ctx.direct(np.ones((1, 3)), np.ones(1), 0.0, FP64)
    assert np.all(out == 0.0)
    out = ctx.direct(np.ones((1, 3)), np.ones(1), 0.0, FP32)
    assert np.all(out == 0.0)
    out = ctx.direct(np.zeros((0, 3)), np.zeros(0), 0.1, FP32, force_pos=np.ones((4, 3)))
    assert np.all(out == 0.0)


# FP32 tolerance: pair forces in FP32 (v_rsq_f32 is accurate to 1 ulp), FP32 sums of
# 256-source tiles, compensated across tiles.  Median relative error <= 1e-6 and p99
# <= 1e-4 against the FP64 oracle on Plummer spheres.
FP32_MEDIAN = 1e-6
FP32_P99 = 1e-4


@pytest.mark.parametrize("n", [1000, 4096, 16384])
def test_fp32_direct_accuracy(ctx, oracle, n):
    ic = workloads.plummer(n, b=1e-3, mass=1e6, seed=n)
    ref = oracle.direct(ic["pos"], ic["mass"], 1e-5)
    got = ctx.direct(ic["pos"], ic["mass"], 1e-5, FP32)
    e = rel_err(got, ref)
    print("n=%d median %.3g p99 %.3g max %.3g" % (n, np.median(e), np.percentile(e, 99), e.max()))
    assert np.median(e) <= FP32_MEDIAN
    assert np.percentile(e, 99) <= FP32_P99


def test_fp32_position_and_zero_softening(ctx, oracle):
    rng = np.random.default_rng(5)
    pos = rng.normal(size=(3000, 3))
    mass = rng.uniform(size=3000)
    fpos = np.concatenate([pos[:50], rng.normal(size=(200, 3))])
    for eps in (0.0, 0.05):
        ref = oracle.direct_position(pos, mass, fpos, eps)
        got = ctx.direct(pos, mass, eps, FP32, force_pos=fpos)
        assert np.all(np.isfinite(got))
        e = rel_err(got, ref)
        assert np.median(e) < 1e-6, eps
    # self-exclusion with eps = 0 (the self pair is guarded to zero)
    got = ctx.direct(pos, mass, 0.0, FP32)
    e = rel_err(got, oracle.direct(pos, mass, 0.0))
    assert np.all(np.isfinite(got)) and np.median(e) < 1e-6


def test_fp32_split_sources_consistent(ctx, oracle):
    # few targets, many sources: the kernel splits sources over workgroups
    rng = np.random.default_rng(6)
    pos = rng.normal(size=(50000, 3))
    mass = rng.uniform(size=50000)
    fpos = rng.normal(size=(37, 3))
    ref = oracle.direct_position(pos, mass, fpos, 0.01)
    got = ctx.direct(pos, mass, 0.01, FP32, force_pos=fpos)
    assert np.median(rel_err(got, ref)) < 1e-6


def test_energy_matches_oracle(ctx, oracle):
    ic = workloads.plummer(3000, seed=9)
    ke, pe, p = ctx.energy(ic["pos"], ic["vel"], ic["mass"], 5e-5)
    ke0, pe0, p0 = oracle.energy(ic["pos"], ic["vel"], ic["mass"], 5e-5)
    assert abs(ke - ke0) <= 1e-12 * abs(ke0)
    assert abs(pe - pe0) <= 1e-12 * abs(pe0)
    assert np.allclose(p, p0, atol=1e-9 * np.abs(ic["mass"][:, None] * ic["vel"]).sum())


@pytest.mark.parametrize("ipt", [-1, 1, 2, 4])
@pytest.mark.parametrize("n,eps", [(5000, 1e-5), (20001, 1e-5), (3000, 0.0)])
def test_fp32_self_kernels(ctx, oracle, ipt, n, eps):
    """Every FP32 self-field kernel (one-sided, and pair-once with 1/2/4 particles per
    lane) against the FP64 oracle, including padding (n not a block multiple) and
    eps = 0 (zero-separation guard)."""
    ic = workloads.plummer(n, b=1e-3, mass=1e6, seed=n + 7)
    ref = oracle.direct(ic["pos"], ic["mass"], eps)
    ctx.set_tuning(_native.TUNE_DIRECT_SELF, ipt)
    try:
        got = ctx.direct(ic["pos"], ic["mass"], eps, FP32)
    finally:
        ctx.set_tuning(_native.TUNE_DIRECT_SELF, 0)
    assert np.all(np.isfinite(got))
    e = rel_err(got, ref)
    print("ipt=%d n=%d eps=%g median %.3g p99 %.3g" % (ipt, n, eps, np.median(e),
                                                      np.percentile(e, 99)))
    assert np.median(e) <= FP32_MEDIAN
    assert np.percentile(e, 99) <= FP32_P99
    # momentum: the pair-once kernels give exactly opposite pair forces
    if ipt > 0:
        tot = (ic["mass"][:, None] * got).sum(0)
        assert np.all(np.abs(tot) <= 1e-5 * (ic["mass"][:, None] * np.abs(got)).sum(0))


def test_fp32_self_kernels_agree_unequal_masses(ctx):
    rng = np.random.default_rng(55)
    pos = rng.normal(size=(17000, 3))
    mass = rng.lognormal(size=17000)
    out = {}
    for ipt in (-1, 1, 4):
        ctx.set_tuning(_native.TUNE_DIRECT_SELF, ipt)
        out[ipt] = ctx.direct(pos, mass, 0.01, FP32)
    ctx.set_tuning(_native.TUNE_DIRECT_SELF, 0)
    for ipt in (1, 4):
        assert np.median(rel_err(out[ipt], out[-1])) < 1e-6


@pytest.mark.parametrize("tpw", [1, 2, 4, 8, 16, 32, 64, -1, -2, -4, -8, -16])
def test_fp64_direct_targets_per_wave_bitwise(ctx, oracle, tpw):
    # every spread of a target's terms over lanes (positive: the spread kernel, negative:
    # a lane per component sum) adds them in the reference's order
    rng = np.random.default_rng(300 + tpw)
    n = 777
    pos = rng.standard_cauchy(size=(n, 3))
    mass = rng.uniform(0.1, 3.0, size=n)
    fpos = np.concatenate([pos[:9], rng.normal(size=(40, 3))])
    try:
        ctx.set_tuning(_native.TUNE_DIRECT_F64_TPW, tpw)
        for eps in (0.0, 1e-3):
            assert_bitwise(ctx.direct(pos, mass, eps, FP64), oracle.direct(pos, mass, eps),
                           "self eps=%g" % eps)
            assert_bitwise(ctx.direct(pos, mass, eps, FP64, force_pos=fpos),
                           oracle.direct_position(pos, mass, fpos, eps), "position eps=%g" % eps)
            # fewer sources than one chunk, and fewer targets than one wavefront's
            assert_bitwise(ctx.direct(pos[:5], mass[:5], eps, FP64),
                           oracle.direct(pos[:5], mass[:5], eps), "5 particles eps=%g" % eps)
            assert_bitwise(ctx.direct(pos[:70], mass[:70], eps, FP64, force_pos=fpos[:3]),
                           oracle.direct_position(pos[:70], mass[:70], fpos[:3], eps),
                           "3 targets eps=%g" % eps)
    finally:
        ctx.set_tuning(_native.TUNE_DIRECT_F64_TPW, 0)


@pytest.mark.parametrize("n", [1024, 1025, 4096, 4097])
def test_fp64_direct_automatic_shapes_bitwise(ctx, oracle, n):
    # both sides of the automatic shape steps (1, 2 and 4 targets per wavefront of the
    # per-sum-lane kernel) give the reference's bits
    rng = np.random.default_rng(400 + n)
    pos = rng.standard_cauchy(size=(n, 3))
    mass = rng.uniform(0.5, 2.0, size=n)
    assert_bitwise(ctx.direct(pos, mass, 1e-3, FP64), oracle.direct(pos, mass, 1e-3),
                   "self n=%d" % n)
    fpos = rng.normal(size=(n, 3))
    assert_bitwise(ctx.direct(pos, mass, 1e-3, FP64, force_pos=fpos),
                   oracle.direct_position(pos, mass, fpos, 1e-3), "position n=%d" % n)


@pytest.mark.parametrize("n", [2, 3, 255, 2047, 2049, 16384, 65536, 65537])
def test_fp32_sort_forms_agree(ctx, n):
    # the spatial order of up to 32 radix chunks (65536 points) takes one kernel per
    # pass; the classic histogram / scan / scatter passes must give the same order, so
    # the same FP32 forces bit for bit; both within the FP32 tolerance of the FP64 sum
    ic = workloads.plummer(n, b=1e-3, mass=1e6, seed=77 + n)
    pos = ic["pos"] + np.array([3.0, -8.0, 40.0])  # away from the origin
    got = ctx.direct(pos, ic["mass"], 1e-5, FP32)
    ctx.set_tuning(_native.TUNE_SORT_DIRECT, -1)
    try:
        classic = ctx.direct(pos, ic["mass"], 1e-5, FP32)
    finally:
        ctx.set_tuning(_native.TUNE_SORT_DIRECT, 0)
    assert_bitwise(got, classic, "n=%d" % n)
    if n >= 16:
        ref = ctx.direct(pos, ic["mass"], 1e-5, FP64)
        e = rel_err(got, ref)
        assert np.median(e) <= FP32_MEDIAN and np.percentile(e, 99) <= FP32_P99
