"""Parity of the device-built Barnes-Hut octree with the reference tree.

Topology (node levels, branches, centres, sizes and per-leaf particle) must be
bit-identical to the reference's insertion-built octree (oracle restatement, itself
bit-identical to gravhopper/_jbgrav.c's tree forces).  The FP64 walk must accept the
same nodes as the reference walk (exact interaction counts per target) and give the
same forces up to summation order.

The FP32 walk opens nodes by a group test that accepts a node only where the
reference's test accepts it for every target of the group, so each target's
interactions refine the reference's.  It is checked three ways: (1) its FP32 arithmetic
against the same interaction lists evaluated in FP64 (GH_TUNE_WALK_F64EVAL), at the
direct sum's FP32 tolerance; (2) its accuracy against the FP64 direct sum, which must be
no worse than the reference tree's; (3) at theta = 0 every list is every particle, so
the walk must equal the direct sum.
"""
import numpy as np
import pytest

from conftest import assert_bitwise, rel_err
from gravhopper_amd import Simulation, _native, workloads

pytestmark = pytest.mark.gpu

FP64, FP32 = _native.GH_FP64, _native.GH_FP32

# FP32 walk arithmetic against the same lists in FP64: FP32 monopoles, FP32 sums of at
# most 64 / (64 / G) list entries per lane added into FP64 (the direct sum's tolerance)
FP32_LIST_MEDIAN, FP32_LIST_P99 = 1e-6, 1e-4


def fp32_walk(ctx, pos, mass, eps, theta, fpos=None):
    """FP32 group walk, and the same interaction lists evaluated in FP64."""
    got = ctx.tree(pos, mass, eps, theta, FP32, force_pos=fpos)
    ctx.set_tuning(_native.TUNE_WALK_F64EVAL, 1)
    try:
        ref = ctx.tree(pos, mass, eps, theta, FP32, force_pos=fpos)
    finally:
        ctx.set_tuning(_native.TUNE_WALK_F64EVAL, 0)
    return got, ref


def check_fp32_walk(ctx, oracle, pos, mass, eps, theta, fpos=None, what=""):
    got, lists64 = fp32_walk(ctx, pos, mass, eps, theta, fpos)
    e = rel_err(got, lists64)
    print("%s: fp32 vs same lists in fp64: median %.2e p99 %.2e"
          % (what, np.median(e), np.percentile(e, 99)), flush=True)
    assert np.median(e) <= FP32_LIST_MEDIAN and np.percentile(e, 99) <= FP32_LIST_P99, what
    # accuracy against the FP64 direct sum: no worse than the reference tree's
    if fpos is None:
        direct = oracle.direct(pos, mass, eps)
        reftree = oracle.tree(pos, mass, eps, theta)
    else:
        direct = oracle.direct_position(pos, mass, fpos, eps)
        reftree = oracle.tree(pos, mass, eps, theta, force_pos=fpos)
    eg, er = rel_err(lists64, direct), rel_err(reftree, direct)
    print("%s: vs direct: group walk median %.2e p99 %.2e, reference tree %.2e / %.2e"
          % (what, np.median(eg), np.percentile(eg, 99), np.median(er),
             np.percentile(er, 99)), flush=True)
    assert np.median(eg) <= np.median(er) * 1.001 + 1e-15, what
    assert np.percentile(eg, 99) <= np.percentile(er, 99) * 1.05 + 1e-15, what
    return got


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


def clustered(n_field, sizes, seed):
    """Field particles plus tight clusters (spread 1e-6 of a root box of about 10), each
    inside one cell of octree level 12, so their keys agree over levels 0-12."""
    rng = np.random.default_rng(seed)
    parts = [rng.normal(size=(n_field, 3)) * 2.0]
    for k in sizes:
        parts.append(rng.normal(size=(1, 3)) + rng.normal(size=(k, 3)) * 1e-6)
    pos = np.concatenate(parts)
    return pos, rng.uniform(1, 2, len(pos))


@pytest.mark.parametrize("mode", [1, -1, 0])
def test_topology_bitwise_key_runs(ctx, oracle, mode):
    # runs of 2..64 keys equal over levels 0-12 are ordered by k_sort_runs after the radix
    # passes over those levels (mode 1); -1 sorts over all levels; each gives the same tree
    pos, mass = clustered(20000, [2, 3, 5, 17, 40, 63, 64], seed=11)
    ctx.set_tuning(_native.TUNE_TREE_SORT, mode)
    try:
        print("mode", mode, len(pos), flush=True)
        compare_topology(ctx.tree_dump(pos, mass, 1e-3), oracle.tree_dump(pos, mass, 1e-3),
                         "runs, mode %d" % mode)
    finally:
        ctx.set_tuning(_native.TUNE_TREE_SORT, 0)


@pytest.mark.parametrize("n", [2, 3, 5])
def test_topology_bitwise_key_runs_small(ctx, oracle, n):
    # n particles 1e-7 apart in a root box of about 10 (set by one far particle): their
    # keys agree over levels 0-12, one run of n
    rng = np.random.default_rng(n)
    pos = np.concatenate([rng.normal(size=(1, 3)) + rng.normal(size=(n, 3)) * 1e-7,
                          [[10.0, -3.0, 2.0]]])
    mass = rng.uniform(1, 2, n + 1)
    ctx.set_tuning(_native.TUNE_TREE_SORT, 1)
    try:
        compare_topology(ctx.tree_dump(pos, mass, 0.0), oracle.tree_dump(pos, mass, 0.0),
                         "one run of %d" % n)
    finally:
        ctx.set_tuning(_native.TUNE_TREE_SORT, 0)


@pytest.mark.parametrize("mode", [1, 0])
def test_topology_bitwise_long_key_run(ctx, oracle, mode):
    # a run of 65 or 300 keys is longer than k_sort_runs takes: the build sorts again over
    # all levels; in the automatic mode the next builds start that way (same results)
    ctx.set_tuning(_native.TUNE_TREE_SORT, mode)
    try:
        for sizes in ([65], [300, 4], [3]):
            pos, mass = clustered(5000, sizes, seed=12)
            print("mode", mode, "clusters", sizes, flush=True)
            compare_topology(ctx.tree_dump(pos, mass, 1e-3),
                             oracle.tree_dump(pos, mass, 1e-3), "clusters %s" % sizes)
            want, wcnt = oracle.tree(pos, mass, 1e-3, 0.7, count=True)
            got, gcnt = ctx.tree_interactions(pos, mass, 1e-3, 0.7)
            assert np.array_equal(gcnt, wcnt), "clusters %s: accepted-node counts" % sizes
            # a cluster member's accepted nodes lie about 1e-6 away, so the reference's
            # centres of mass (running sums, rounding about 1e-15 absolute) move its
            # force by about 1e-8 relative; the device's come from double-double sums
            np.testing.assert_allclose(got, want, rtol=1e-6, atol=1e-10 * np.abs(want).max())
    finally:
        ctx.set_tuning(_native.TUNE_TREE_SORT, 0)


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
    check_fp32_walk(ctx, oracle, pos, mass, 0.01, 0.7, fpos=fpos, what="%d targets" % n_targets)


@pytest.mark.parametrize("theta", [0.3, 0.7, 1.0])
def test_fp32_walk_accuracy(ctx, oracle, theta):
    g = workloads.galaxy(n_halo=16384, n_disk=16384, seed=79)
    check_fp32_walk(ctx, oracle, g["pos"], g["mass"], 0.05, theta,
                    what="galaxy 32768 theta %g" % theta)


def test_fp32_walk_theta0_is_direct_sum(ctx, oracle):
    # theta = 0 opens every internal node: every list is every particle exactly once
    g = workloads.galaxy(n_halo=10000, n_disk=10000, seed=92)
    got = ctx.tree(g["pos"], g["mass"], 0.05, 0.0, FP32)
    e = rel_err(got, oracle.direct(g["pos"], g["mass"], 0.05))
    assert np.median(e) < 1e-6 and np.percentile(e, 99) < 1e-4


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


@pytest.mark.parametrize("group", [1, 2, 4, 8, 16])
def test_fp32_walk_group_sizes(ctx, oracle, group):
    # any group size: FP32 arithmetic against the same lists in FP64, accuracy no worse
    # than the reference tree, and repeated calls bit-identical
    g = workloads.galaxy(n_halo=12000, n_disk=12000, seed=82)
    ctx.set_tuning(_native.TUNE_WALK_GROUP, group)
    try:
        a = check_fp32_walk(ctx, oracle, g["pos"], g["mass"], 0.05, 0.7,
                            what="group %d" % group)
        b = ctx.tree(g["pos"], g["mass"], 0.05, 0.7, FP32)
    finally:
        ctx.set_tuning(_native.TUNE_WALK_GROUP, 0)
    assert_bitwise(a, b)


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


@pytest.mark.parametrize("group", [0, 1, 4, 16, 64])
def test_fp32_walk_prefix_same_lists(ctx, group):
    # the walks of a block of groups start from the nodes k_walk_prefix decided for all of
    # them (accepted into the list, or opened) and the nodes it deferred; every target's
    # list must hold as many nodes as from the root (the same nodes, in another order),
    # and the forces differ only by FP32 summation order
    rng = np.random.default_rng(404)
    g = workloads.galaxy(n_halo=40000, n_disk=40000, seed=405)
    fpos = np.concatenate([g["pos"][:5000], rng.normal(size=(9000, 3)) * 300.0])
    fpos[7000, 1] = np.nan  # its block of groups starts from the root
    cases = [("galaxy", g["pos"], g["mass"], 0.05, None),
             ("cauchy", rng.standard_cauchy(size=(30000, 3)), rng.uniform(1, 2, 30000), 1e-3, None),
             ("positions", g["pos"], g["mass"], 0.05, fpos),
             ("tiny", rng.normal(size=(50, 3)), rng.uniform(1, 2, 50), 1e-2,
              rng.normal(size=(7, 3)))]
    ctx.set_tuning(_native.TUNE_WALK_GROUP, group)
    try:
        for name, pos, mass, eps, fpos in cases:
            for theta in (0.0, 0.4, 0.7, 1.2):
                what = "%s G=%d theta=%g" % (name, group, theta)
                print(what, flush=True)
                got, gcnt = ctx.tree_interactions(pos, mass, eps, theta, force_pos=fpos,
                                                  precision=FP32)
                ctx.set_tuning(_native.TUNE_WALK_PREFIX, -1)
                try:
                    want, wcnt = ctx.tree_interactions(pos, mass, eps, theta, force_pos=fpos,
                                                       precision=FP32)
                finally:
                    ctx.set_tuning(_native.TUNE_WALK_PREFIX, 0)
                assert np.array_equal(gcnt, wcnt), what + ": list lengths"
                fin = np.isfinite(want).all(axis=1)
                assert np.array_equal(fin, np.isfinite(got).all(axis=1)), what
                e = rel_err(got[fin], want[fin])
                assert np.median(e) <= FP32_LIST_MEDIAN and np.percentile(e, 99) <= FP32_LIST_P99, what
    finally:
        ctx.set_tuning(_native.TUNE_WALK_GROUP, 0)


def test_fp32_walk_few_targets(ctx, oracle):
    # few position targets are spread over many short walks (down to one target each)
    rng = np.random.default_rng(88)
    pos = rng.normal(size=(20000, 3))
    mass = rng.uniform(size=20000)
    for nt in (1, 7, 300):
        fpos = rng.normal(size=(nt, 3)) * 2
        check_fp32_walk(ctx, oracle, pos, mass, 0.01, 0.7, fpos=fpos, what="%d targets" % nt)


def test_fp32_walk_zero_softening(ctx, oracle):
    # eps = 0: the walk's guarded form, where a zero separation (a particle's own leaf,
    # or a target on a particle) gives zero force, as in the reference's tree walk
    rng = np.random.default_rng(89)
    pos = rng.normal(size=(6000, 3))
    mass = rng.uniform(1, 2, 6000)
    got = check_fp32_walk(ctx, oracle, pos, mass, 0.0, 0.7, what="eps 0")
    assert np.all(np.isfinite(got))
    fpos = np.concatenate([pos[:50], rng.normal(size=(250, 3))])
    got = check_fp32_walk(ctx, oracle, pos, mass, 0.0, 0.7, fpos=fpos, what="eps 0 targets")
    assert np.all(np.isfinite(got))


@pytest.mark.parametrize("group", [0, 2, 16])
def test_fp32_walk_list_lengths(ctx, group):
    # every target's list refines the reference's accepted nodes, so it is never shorter;
    # and the group test must not open far more than the reference (a loose group, whose
    # sorted targets jump across the system, is walked target by target)
    g = workloads.galaxy(n_halo=32768, n_disk=32768, seed=3)
    _, cref = ctx.tree_interactions(g["pos"], g["mass"], 0.05, 0.7)
    ctx.set_tuning(_native.TUNE_WALK_GROUP, group)
    try:
        _, c32 = ctx.tree_interactions(g["pos"], g["mass"], 0.05, 0.7, precision=FP32)
    finally:
        ctx.set_tuning(_native.TUNE_WALK_GROUP, 0)
    print("group %d: list mean %.0f max %d; reference mean %.0f max %d"
          % (group, c32.mean(), c32.max(), cref.mean(), cref.max()), flush=True)
    assert np.all(c32 >= cref)
    assert c32.mean() < 2.5 * cref.mean()
    assert c32.max() < 8 * cref.max()
