"""Pins the CPU restatement (oracle/) before it is trusted as the checker.

Each oracle function is checked against (1) the reference's own C engine compiled
from its source into oracle/_ref (where that was built), (2) the golden vectors the
reference wrote (tests/golden/reference_vectors.npz) and (3) closed forms and
invariants chosen so that a dropped term, a wrong sign or a transposed index fails.
"""
import numpy as np
import pytest

from conftest import assert_bitwise


# --------------------------------------------------------------------------- golden
def test_direct_matches_reference_vectors(oracle, golden):
    for name, c in golden.items():
        assert_bitwise(oracle.direct(c["pos"], c["mass"], float(c["eps"])), c["direct"],
                       name + " direct")


def test_direct_position_matches_reference_vectors(oracle, golden):
    for name, c in golden.items():
        out = oracle.direct_position(c["pos"], c["mass"], c["force_pos"], float(c["eps"]))
        assert_bitwise(out, c["direct_position"], name + " direct_position")


@pytest.mark.parametrize("theta", [0.0, 0.3, 0.7, 1e30])
def test_tree_matches_reference_vectors(oracle, golden, theta):
    for name, c in golden.items():
        out = oracle.tree(c["pos"], c["mass"], float(c["eps"]), theta)
        assert_bitwise(out, c["tree_%g" % theta], "%s tree %g" % (name, theta))
        out = oracle.tree(c["pos"], c["mass"], float(c["eps"]), theta, force_pos=c["force_pos"])
        assert_bitwise(out, c["tree_position_%g" % theta], "%s tree_position %g" % (name, theta))


# ---------------------------------------------------------------- live reference build
def test_against_reference_build(oracle):
    ref = oracle.ref_module()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference source absent on this machine)")
    rng = np.random.default_rng(7)
    for n in (1, 2, 5, 333):
        pos = rng.standard_cauchy(size=(n, 3))  # heavy tails: deep trees
        mass = rng.uniform(0.1, 5.0, size=n)
        fpos = rng.normal(size=(4, 3))
        for eps in (0.0, 1e-3):
            if n > 1 or eps > 0:
                assert_bitwise(oracle.direct(pos, mass, eps), ref.direct_summation(pos, mass, eps))
            assert_bitwise(oracle.direct_position(pos, mass, fpos, eps),
                           ref.direct_summation_position(pos, mass, fpos, eps))
            for th in (0.0, 0.5, 0.7, 2.0):
                assert_bitwise(oracle.tree(pos, mass, eps, th), ref.tree_force(pos, mass, eps, th))
                assert_bitwise(oracle.tree(pos, mass, eps, th, force_pos=fpos),
                               ref.tree_force_position(pos, mass, fpos, eps, th))


# ------------------------------------------------------------------- closed forms
def test_two_body_closed_form(oracle):
    pos = np.array([[0.0, 0.0, 0.0], [2.0, 0.0, 0.0]])
    mass = np.array([3.0, 5.0])
    eps = 0.5
    a = oracle.direct(pos, mass, eps)
    s3 = (4.0 + eps * eps) ** 1.5
    np.testing.assert_allclose(a[0], [5.0 * 2.0 / s3, 0, 0], rtol=1e-15)
    np.testing.assert_allclose(a[1], [-3.0 * 2.0 / s3, 0, 0], rtol=1e-15)


def test_equilateral_three_body(oracle):
    # three unit masses on a circle of radius 1: each feels |a| = sqrt(3)/s^2 towards
    # the centre with side s = sqrt(3) (eps = 0)
    ang = np.array([0.0, 2 * np.pi / 3, 4 * np.pi / 3])
    pos = np.stack([np.cos(ang), np.sin(ang), np.zeros(3)], axis=1)
    a = oracle.direct(pos, np.ones(3), 0.0)
    mag = np.linalg.norm(a, axis=1)
    np.testing.assert_allclose(mag, np.sqrt(3) / 3.0, rtol=1e-14)
    np.testing.assert_allclose(a / mag[:, None], -pos, atol=1e-14)


def test_direct_invariants(oracle):
    rng = np.random.default_rng(3)
    n = 300
    pos = rng.normal(size=(n, 3))
    mass = rng.uniform(0.5, 2.0, size=n)
    eps = 0.05
    a = oracle.direct(pos, mass, eps)
    # Newton's third law: total momentum change is zero to rounding
    tot = (mass[:, None] * a).sum(axis=0)
    assert np.all(np.abs(tot) < 1e-12 * np.abs(mass[:, None] * a).sum())
    # translation invariance (to rounding of the shifted differences)
    a2 = oracle.direct(pos + np.array([1e-3, -2e-3, 5e-4]), mass, eps)
    np.testing.assert_allclose(a2, a, rtol=1e-9, atol=1e-9 * np.abs(a).max())
    # linear in the masses: doubling all masses doubles the field exactly
    assert_bitwise(oracle.direct(pos, 2 * mass, eps), 2 * a)
    # scaling: a(l x, l eps) = a(x, eps) / l^2 exactly for l a power of two
    assert_bitwise(oracle.direct(4.0 * pos, mass, 4.0 * eps), a / 16.0)
    # brute force in numpy, vectorised, FP64 (different summation order)
    d = pos[None, :, :] - pos[:, None, :]
    r2 = (d ** 2).sum(-1) + eps * eps
    np.fill_diagonal(r2, np.inf)
    ab = (mass[None, :, None] * d / r2[..., None] ** 1.5).sum(1)
    np.testing.assert_allclose(a, ab, rtol=1e-11, atol=1e-12 * np.abs(ab).max())


def test_direct_position_self_is_direct(oracle):
    rng = np.random.default_rng(4)
    pos = rng.normal(size=(100, 3))
    mass = rng.uniform(size=100)
    a = oracle.direct(pos, mass, 0.1)
    b = oracle.direct_position(pos, mass, pos, 0.1)
    # the self term adds exactly +0.0, so the two agree bit for bit (up to -0.0)
    np.testing.assert_array_equal(a, b)


# ----------------------------------------------------------------------- tree
def _walk_from_dump(dump, pos, mass, x, eps, theta):
    """Recursive walk over a pre-order dump (independent of the C walk)."""
    n = len(dump)
    # reconstruct children lists from levels
    children = [[] for _ in range(n)]
    stack = []
    for k in range(n):
        lv = dump["level"][k]
        while stack and dump["level"][stack[-1]] >= lv:
            stack.pop()
        if stack:
            children[stack[-1]].append(k)
        stack.append(k)

    def acc(k):
        c = dump["center"][k]
        nd = np.sqrt(((c[0] - x[0]) * (c[0] - x[0]) + (c[1] - x[1]) * (c[1] - x[1]))
                     + (c[2] - x[2]) * (c[2] - x[2]))
        with np.errstate(divide="ignore", invalid="ignore"):
            accept = dump["leaf"][k] >= 0 or (dump["size"][k] / nd) < theta
        if accept:
            com = pos[dump["leaf"][k]] if dump["leaf"][k] >= 0 else dump["com"][k]
            dp = com - x
            d2 = (dp[0] * dp[0] + dp[1] * dp[1]) + dp[2] * dp[2]
            d = d2 + eps * eps
            inv = 0.0 if d == 0.0 else 1.0 / d / np.sqrt(d)
            return dp * dump["mass"][k] * inv
        f = np.zeros(3)
        for ch in children[k]:
            f = f + acc(ch)
        return f

    return acc(0)


def _paths(pos, eps, depth):
    """Octant digits per level from the FP64 centre recurrence of _jbgrav.c:406,456,
    restated independently in numpy."""
    mn, mx = pos.min(axis=0), pos.max(axis=0)
    size = mx[0] - mn[0] + eps
    for i in (1, 2):
        if (mx[i] - mn[i]) > size:
            size = mx[i] - mn[i] + eps
    c = np.tile(0.5 * (mn + mx), (len(pos), 1))
    h = 0.5 * size
    digits = np.zeros((len(pos), depth), dtype=np.int64)
    centers = np.zeros((len(pos), depth + 1, 3))
    centers[:, 0] = c
    for lv in range(depth):
        gt = pos > c
        digits[:, lv] = gt[:, 0] * 1 + gt[:, 1] * 2 + gt[:, 2] * 4
        c = c + np.where(gt, 1.0, -1.0) * 0.5 * h
        centers[:, lv + 1] = c
        h = 0.5 * h
    return digits, centers, size


def test_tree_dump_structure(oracle):
    rng = np.random.default_rng(11)
    n = 300
    pos = rng.standard_cauchy(size=(n, 3))
    mass = rng.uniform(0.5, 2.0, size=n)
    eps = 1e-3
    dump = oracle.tree_dump(pos, mass, eps)
    leaves = dump[dump["leaf"] >= 0]
    # one particle per leaf, every particle in exactly one leaf
    assert sorted(leaves["leaf"].tolist()) == list(range(n))
    # node size halves per level exactly
    assert np.all(dump["size"] == dump["size"][0] * 0.5 ** dump["level"])
    # leaf level = 1 + longest shared octant prefix with any other particle, and the
    # leaf centre is the recurrence centre at that level
    depth = int(dump["level"].max()) + 1
    digits, centers, size = _paths(pos, eps, depth)
    assert dump["size"][0] == size
    for rec in leaves:
        p = rec["leaf"]
        same = (digits == digits[p])
        lcp = np.argmin(np.concatenate([same, np.zeros((n, 1), bool)], axis=1), axis=1)
        lcp[p] = -1
        assert rec["level"] == (lcp.max() + 1 if n > 1 else 0)
        assert np.array_equal(rec["center"], centers[p, rec["level"]])
    # internal node mass equals the sum of its leaves' masses (to rounding)
    assert abs(dump["mass"][0] - mass.sum()) < 1e-12 * mass.sum()


def test_tree_dump_walk_reproduces_reference_forces(oracle):
    ref = oracle.ref_module()
    rng = np.random.default_rng(12)
    n = 150
    pos = rng.normal(size=(n, 3)) * np.exp(rng.normal(size=(n, 1)))
    mass = rng.uniform(0.5, 2.0, size=n)
    eps, theta = 0.01, 0.7
    dump = oracle.tree_dump(pos, mass, eps)
    want = ref.tree_force(pos, mass, eps, theta) if ref is not None else oracle.tree(
        pos, mass, eps, theta)
    got = np.array([_walk_from_dump(dump, pos, mass, pos[i], eps, theta) for i in range(n)])
    assert_bitwise(got, want, "walk over dumped tree")


def test_tree_topology_independent_of_order(oracle):
    rng = np.random.default_rng(13)
    n = 400
    pos = rng.normal(size=(n, 3)) ** 3
    mass = np.ones(n)
    d1 = oracle.tree_dump(pos, mass, 0.01)
    perm = rng.permutation(n)
    d2 = oracle.tree_dump(pos[perm], mass[perm], 0.01)
    assert len(d1) == len(d2)
    for f in ("level", "branch", "center", "size"):
        assert np.array_equal(d1[f], d2[f]), f
    leaf2 = np.where(d2["leaf"] >= 0, perm[np.maximum(d2["leaf"], 0)], -1)
    assert np.array_equal(d1["leaf"], leaf2)


def test_tree_limits(oracle):
    rng = np.random.default_rng(14)
    n = 200
    pos = rng.normal(size=(n, 3))
    mass = rng.uniform(size=n)
    eps = 0.02
    # theta = 0 opens every internal node: the direct sum in another order
    a0 = oracle.tree(pos, mass, eps, 0.0)
    ad = oracle.direct(pos, mass, eps)
    np.testing.assert_allclose(a0, ad, rtol=1e-11, atol=1e-13 * np.abs(ad).max())
    # theta -> infinity: the root monopole
    big = oracle.tree(pos, mass, eps, 1e300, force_pos=np.array([[30.0, -40.0, 5.0]]))
    com = (mass[:, None] * pos).sum(0) / mass.sum()
    dp = com - np.array([30.0, -40.0, 5.0])
    d = (dp ** 2).sum() + eps * eps
    np.testing.assert_allclose(big[0], dp * mass.sum() / d ** 1.5, rtol=1e-13)
    # a single particle feels nothing from itself
    assert np.all(oracle.tree(pos[:1], mass[:1], eps, 0.7) == 0.0)


def test_tree_coincident_is_an_error(oracle):
    pos = np.array([[0.0, 0.0, 0.0], [1.0, 1.0, 1.0], [0.0, 0.0, 0.0]])
    with pytest.raises(RuntimeError, match="coincident"):
        oracle.tree(pos, np.ones(3), 0.01, 0.7)


# ------------------------------------------------------------------------ DKD
def test_dkd_one_step_matches_formula(oracle):
    rng = np.random.default_rng(21)
    n = 64
    pos = rng.normal(size=(n, 3)) * 1e-3
    vel = rng.normal(size=(n, 3))
    mass = rng.uniform(size=n) * 1e4
    dt, eps = 0.01, 1e-4
    x1, v1 = oracle.dkd_direct(pos, vel, mass, dt, eps, 1)
    xh = pos + 0.5 * vel * dt * oracle.DRIFT_SCALE
    a = oracle.direct(xh, mass, eps) * oracle.ACCEL_SCALE + 0.0
    v = vel + a * dt
    x = xh + 0.5 * v * dt * oracle.DRIFT_SCALE
    assert_bitwise(v1, v)
    assert_bitwise(x1, x)


def test_dkd_free_particle(oracle):
    pos = np.array([[1.0, 2.0, 3.0]])
    vel = np.array([[10.0, -20.0, 5.0]])
    x, v = oracle.dkd_direct(pos, vel, np.ones(1), 0.5, 0.1, 8)
    assert_bitwise(v, vel)
    # 16 rounded half-drifts: a few ulp of accumulated rounding
    np.testing.assert_allclose(x, pos + vel * 4.0 * oracle.DRIFT_SCALE, rtol=1e-14)


def test_dkd_time_reversible_and_energy_bounded(oracle):
    # circular two-body orbit: energy error stays small and bounded, and running the
    # map backwards returns to the start
    G = oracle.ACCEL_SCALE / oracle.DRIFT_SCALE  # kpc (km/s)^2 / Msun
    m = np.array([1e10, 1e10])
    r = 1.0
    vc = np.sqrt(G * m[0] / (4 * r))  # each body at distance r from the COM
    pos = np.array([[r, 0, 0], [-r, 0, 0]], dtype=float)
    vel = np.array([[0, vc, 0], [0, -vc, 0]], dtype=float)
    period = 2 * np.pi * r / vc / oracle.DRIFT_SCALE  # Myr
    dt = period / 200
    x, v = oracle.dkd_direct(pos, vel, m, dt, 0.0, 400)
    ke0, pe0, _ = oracle.energy(pos, vel, m, 0.0)
    ke1, pe1, _ = oracle.energy(x, v, m, 0.0)
    e0 = ke0 + G * pe0
    e1 = ke1 + G * pe1
    assert abs(e1 - e0) < 2e-3 * abs(e0)
    xb, vb = oracle.dkd_direct(x, -v, m, dt, 0.0, 400)
    np.testing.assert_allclose(xb, pos, atol=1e-9)
    np.testing.assert_allclose(-vb, vel, atol=1e-6 * vc)


def test_energy_two_body_closed_form(oracle):
    pos = np.array([[0.0, 0.0, 0.0], [3.0, 4.0, 0.0]])
    vel = np.array([[1.0, 0.0, 0.0], [0.0, 2.0, 0.0]])
    m = np.array([2.0, 3.0])
    ke, pe, p = oracle.energy(pos, vel, m, 0.0)
    assert ke == 0.5 * 2 * 1 + 0.5 * 3 * 4
    assert pe == -2.0 * 3.0 / 5.0
    np.testing.assert_array_equal(p, [2.0, 6.0, 0.0])
    _, pe_s, _ = oracle.energy(pos, vel, m, 1.0)
    assert abs(pe_s + 6.0 / np.sqrt(26.0)) < 1e-15
