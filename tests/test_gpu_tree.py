"""Parity of the device-built Barnes-Hut octree with the reference tree.

Topology (node levels, branches, centres, sizes and per-leaf particle) must be
bit-identical to the reference's insertion-built octree (oracle restatement, itself
bit-identical to gravhopper/_jbgrav.c's tree forces).  The FP64 walk must accept the
same nodes as the reference walk (exact interaction counts per target) and give the
same forces up to summation order; the FP32 walk is checked with a stated tolerance.
"""
import numpy as np
import pytest

from conftest import assert_bitwise, rel_err
from gravhopper_amd import Simulation, _native, workloads

pytestmark = pytest.mark.gpu

FP64, FP32 = _native.GH_FP64, _native.GH_FP32

# FP32 walk tolerance against the FP64 reference tree: FP32 monopoles and opening test
# (a node on the opening boundary may go either way), FP32 partial sums added into FP64
# every 64 node visits.
FP32_TREE_MEDIAN = 1e-5


def compare_topology(got, want, name):
    assert len(got) == len(want), "%s: %d vs %d nodes" % (name, len(got), len(want))
    for f in ("level", "branch", "leaf"):
        assert np.array_equal(got[f], want[f]), "%s: %s differs" % (name, f)
    assert_bitwise(got["center"], want["center"], name + " centre")
    assert_bitwise(got["size"], want["size"], name + " size")
    # moments: double-double prefix sums vs the reference's insertion-order sums
    np.testing.assert_allclose(got["mass"], want["mass"], rtol=1e-13)
    scale = np.abs(want["center"]).max() + want["size"][0]
    np.testing.assert_allclose(got["com"], want["com"], rtol=1e-12, atol=1e-13 * scale)


def tree_cases():
    rng = np.random.default_rng(2026)
    cases = []
    cases.append(("one", np.array([[0.5, -0.25, 2.0]]), np.ones(1), 0.1))
    cases.append(("two", np.array([[0.0, 0.0, 0.0], [1.0, 1.0, 1.0]]), np.array([1.0, 2.0]), 0.0))
    cases.append(("cauchy_3000", rng.standard_cauchy(size=(3000, 3)), rng.uniform(1, 2, 3000), 1e-3))
    p = workloads.plummer(8192, seed=77)
    cases.append(("plummer_8192", p["pos"], p["mass"], 1e-5))
    g = workloads.galaxy(n_halo=8192, n_disk=8192, seed=78)
    cases.append(("galaxy_16384", g["pos"], g["mass"], 0.05))
    # coordinates exactly on dyadic cell walls of the root box
    w = np.unique(rng.integers(-16, 17, size=(3000, 3)).astype(float) / 16.0, axis=0)
    cases.append(("walls", w, np.ones(len(w)), 0.0))
    # a very close pair needing more than 21 (and 42) levels of octant digits
    deep = rng.normal(size=(500, 3))
    deep[1] = deep[0] + np.array([1e-9, -2e-9, 3e-10])
    deep[2] = deep[0] + np.array([1e-14, 0, 0])
    cases.append(("deep", deep, rng.uniform(1, 2, 500), 1e-3))
    return cases


@pytest.mark.parametrize("case", tree_cases(), ids=lambda c: c[0])
def test_topology_bitwise(ctx, oracle, case):
    name, pos, mass, eps = case
    print("case", name, len(pos), flush=True)
    compare_topology(ctx.tree_dump(pos, mass, eps), oracle.tree_dump(pos, mass, eps), name)


@pytest.mark.parametrize("n", [4096, 4097, 65536, 70000])
def test_topology_bitwise_search_levels(ctx, oracle, n):
    # node ends come from a search over group minima whose number of levels grows with
    # n (3 up to 4096 particles, 4 up to 65536, 5 beyond): sizes on either side of the
    # steps, clustered so that large nodes reach the upper levels
    rng = np.random.default_rng(n)
    pos = np.concatenate([rng.normal(size=(n // 2, 3)) * 1e-3,
                          rng.standard_cauchy(size=(n - n // 2, 3))])
    mass = rng.uniform(1, 2, n)
    print("n", n, flush=True)
    compare_topology(ctx.tree_dump(pos, mass, 1e-3), oracle.tree_dump(pos, mass, 1e-3),
                     "n=%d" % n)


@pytest.mark.parametrize("case", tree_cases(), ids=lambda c: c[0])
def test_fp64_walk_matches_reference(ctx, oracle, case):
    name, pos, mass, eps = case
    print("case", name, len(pos), flush=True)
    for theta in (0.0, 0.7):
        want, wcnt = oracle.tree(pos, mass, eps, theta, count=True)
        got, gcnt = ctx.tree_interactions(pos, mass, eps, theta)
        assert np.array_equal(gcnt, wcnt), "%s theta=%g: accepted-node counts" % (name, theta)
        # same accepted nodes, different summation order and moment rounding
        scale = np.abs(want).max() if len(want) else 1.0
        np.testing.assert_allclose(got, want, rtol=1e-10, atol=1e-12 * scale)


def test_fp64_tree_golden(ctx, golden):
    for name, c in golden.items():
        for theta in (0.3, 0.7):
            want = c["tree_%g" % theta]
            got = ctx.tree(c["pos"], c["mass"], float(c["eps"]), theta, FP64)
            scale = np.abs(want).max()
            np.testing.assert_allclose(got, want, rtol=1e-10, atol=1e-12 * scale, err_msg=name)
            want = c["tree_position_%g" % theta]
            got = ctx.tree(c["pos"], c["mass"], float(c["eps"]), theta, FP64,
                           force_pos=c["force_pos"])
            np.testing.assert_allclose(got, want, rtol=1e-10, atol=1e-12 * np.abs(want).max(),
                                       err_msg=name + " position")


@pytest.mark.parametrize("n_targets", [400, 9000])
def test_position_walk_counts(ctx, oracle, n_targets):
    # 9000 targets are walked in the library's own Morton order of the targets (from
    # 4096 on); results must come back in the caller's order.  Some targets lie
    # outside the particles' root box.
    rng = np.random.default_rng(91)
    pos = rng.normal(size=(5000, 3))
    mass = rng.uniform(size=5000)
    fpos = np.concatenate([pos[:100], rng.normal(size=(n_targets - 100, 3)) * 3])
    want, wcnt = oracle.tree(pos, mass, 0.01, 0.7, force_pos=fpos, count=True)
    got, gcnt = ctx.tree_interactions(pos, mass, 0.01, 0.7, force_pos=fpos)
    assert np.array_equal(gcnt, wcnt)
    np.testing.assert_allclose(got, want, rtol=1e-10, atol=1e-12 * np.abs(want).max())
    got = ctx.tree(pos, mass, 0.01, 0.7, FP64, force_pos=fpos)
    np.testing.assert_allclose(got, want, rtol=1e-10, atol=1e-12 * np.abs(want).max())
    got32 = ctx.tree(pos, mass, 0.01, 0.7, FP32, force_pos=fpos)
    assert np.median(rel_err(got32, want)) < FP32_TREE_MEDIAN


def test_fp32_walk_accuracy(ctx, oracle):
    g = workloads.galaxy(n_halo=16384, n_disk=16384, seed=79)
    want = oracle.tree(g["pos"], g["mass"], 0.05, 0.7)
    got = ctx.tree(g["pos"], g["mass"], 0.05, 0.7, FP32)
    e = rel_err(got, want)
    print("fp32 tree vs fp64 tree: median %.3g p99 %.3g" % (np.median(e), np.percentile(e, 99)))
    assert np.median(e) < FP32_TREE_MEDIAN
    # and the approximation error of the reference algorithm itself, for the record
    direct = oracle.direct(g["pos"], g["mass"], 0.05)
    e2 = rel_err(got, direct)
    print("tree (theta=0.7) vs direct: median %.3g p99 %.3g" % (np.median(e2),
                                                               np.percentile(e2, 99)))
    assert np.median(e2) < 1e-2


def test_tree_edge_cases(ctx):
    with pytest.raises(_native.GravhopError, match="share a position"):
        ctx.tree(np.array([[0.0, 0, 0], [1, 1, 1], [0, 0, 0]]), np.ones(3), 0.1, 0.7)
    out = ctx.tree(np.zeros((0, 3)), np.zeros(0), 0.1, 0.7, FP64, force_pos=np.ones((3, 3)))
    assert np.all(out == 0)
    out = ctx.tree(np.ones((1, 3)), np.ones(1), 0.0, 0.7, FP64)
    assert np.all(out == 0)


def test_tree_deterministic(ctx):
    g = workloads.galaxy(n_halo=30000, n_disk=30000, seed=80)
    a = ctx.tree(g["pos"], g["mass"], 0.05, 0.7, FP32)
    b = ctx.tree(g["pos"], g["mass"], 0.05, 0.7, FP32)
    assert_bitwise(a, b)


def test_simulation_tree_device_vs_host(ctx):
    ic = workloads.plummer(2000, seed=81)
    a = Simulation(dt=0.005, eps=5e-5, algorithm='tree')
    a.add_IC(ic)
    a.run(4)
    b = Simulation(dt=0.005, eps=5e-5, algorithm='tree', device_loop=False)
    b.add_IC(ic)
    b.add_external_force(lambda pos, args: np.zeros_like(pos))
    b.run(4)
    assert_bitwise(a.positions, b.positions)
    assert_bitwise(a.velocities, b.velocities)


def test_walks_per_wave_bitwise(ctx):
    # one or two independent walks per wavefront, static or queue-driven scheduling:
    # every target accepts the same nodes in the same order, so the forces are
    # bit-identical
    g = workloads.galaxy(n_halo=20000, n_disk=20000, seed=82)
    out = []
    try:
        for sched in (-1, 1):
            for wpw in (1, 2):
                ctx.set_tuning(_native.TUNE_WALK_SCHEDULE, sched)
                ctx.set_tuning(_native.TUNE_WALKS_PER_WAVE, wpw)
                out.append(ctx.tree(g["pos"], g["mass"], 0.05, 0.7, FP32))
                # a second call reuses the queue counters, which must start from zero
                out.append(ctx.tree(g["pos"], g["mass"], 0.05, 0.7, FP32))
    finally:
        ctx.set_tuning(_native.TUNE_WALKS_PER_WAVE, 0)
        ctx.set_tuning(_native.TUNE_WALK_SCHEDULE, 0)
    for o in out[1:]:
        assert_bitwise(out[0], o)


def test_fp64_walk_independent_of_target_grouping(ctx):
    # The FP64 walk sums each target's accepted nodes in tree order, whatever the other
    # targets of its wavefront: a target's force is bit-identical among 300 targets
    # (5 per walk), among 20000 (sorted, 5 per walk) or as a particle of the self walk
    # (Morton order, 64 per walk).
    g = workloads.galaxy(n_halo=150000, n_disk=150000, seed=86)
    pos, mass = g["pos"], g["mass"]
    rng = np.random.default_rng(87)
    pick = rng.choice(len(mass), 20000, replace=False)
    fpos = pos[pick]
    self_acc = ctx.tree(pos, mass, 0.05, 0.7, FP64)
    all_pos = ctx.tree(pos, mass, 0.05, 0.7, FP64, force_pos=fpos)
    few_pos = ctx.tree(pos, mass, 0.05, 0.7, FP64, force_pos=fpos[:300])
    assert_bitwise(all_pos[:300], few_pos, "300 vs 20000 targets")
    assert_bitwise(all_pos, self_acc[pick], "position targets vs self walk")


def test_fp32_walk_few_targets(ctx, oracle):
    # few position targets are spread over many short walks (down to one target each)
    rng = np.random.default_rng(88)
    pos = rng.normal(size=(20000, 3))
    mass = rng.uniform(size=20000)
    for nt in (1, 7, 300):
        fpos = rng.normal(size=(nt, 3)) * 2
        want = oracle.tree(pos, mass, 0.01, 0.7, force_pos=fpos)
        got = ctx.tree(pos, mass, 0.01, 0.7, FP32, force_pos=fpos)
        assert np.median(rel_err(got, want)) < FP32_TREE_MEDIAN, nt


def test_fp32_walk_zero_softening(ctx, oracle):
    # eps = 0: the walk's guarded form, where a zero separation (a particle's own leaf,
    # or a target on a particle) gives zero force, as in the reference's tree walk
    rng = np.random.default_rng(89)
    pos = rng.normal(size=(6000, 3))
    mass = rng.uniform(1, 2, 6000)
    want = oracle.tree(pos, mass, 0.0, 0.7)
    got = ctx.tree(pos, mass, 0.0, 0.7, FP32)
    assert np.all(np.isfinite(got))
    assert np.median(rel_err(got, want)) < FP32_TREE_MEDIAN
    fpos = np.concatenate([pos[:50], rng.normal(size=(250, 3))])
    want = oracle.tree(pos, mass, 0.0, 0.7, force_pos=fpos)
    got = ctx.tree(pos, mass, 0.0, 0.7, FP32, force_pos=fpos)
    assert np.all(np.isfinite(got))
    assert np.median(rel_err(got, want)) < FP32_TREE_MEDIAN
