"""CPU check of the adjoint algebra used by the GPU kernels (gh_adjoint.hip).

The reverse step and the split of the tidal-tensor product into the accumulators A, B,
C, L (i side) and A', B', C', L' (j side) are restated in FP64 numpy and compared with
the direct definition T(d) = I s^-3 - 3 d d^T s^-5 and with central finite differences
of the leapfrog itself."""
import numpy as np

ACCEL, DRIFT = 4.398600412921223e-09, 1.022712165045695e-03


def force(x, m, eps):
    d = x[None, :, :] - x[:, None, :]          # d[k, j] = x_j - x_k
    r2 = (d * d).sum(-1) + eps * eps
    w = r2 ** -1.5
    np.fill_diagonal(w, 0.0)
    return (m[None, :, None] * w[:, :, None] * d).sum(1)


def run(x, v, m, dt, eps, n, keep=False):
    c = 0.5 * dt * DRIFT
    xs = []
    for _ in range(n):
        xh = x + c * v
        xs.append(xh)
        v = v + dt * ACCEL * force(xh, m, eps)
        x = xh + c * v
    return (x, v, xs) if keep else (x, v)


def vjp_split(x, m, eps, mu):
    """The kernel's accumulator form, both sides of every unordered pair."""
    n = len(m)
    g = np.zeros_like(x)
    lm = np.zeros(n)
    for k in range(n):
        for j in range(k + 1, n):
            d = x[j] - x[k]
            s2 = d @ d + eps * eps
            ri3 = s2 ** -1.5
            ri2 = 1.0 / s2
            pj, pk = d @ mu[j], d @ mu[k]
            dw = m[k] * pj - m[j] * pk
            cu = ri3 * (-3.0 * ri2) * dw
            # i side: g_k += m_k (ri3 mu_j) - (ri3 m_j) mu_k + cu d ; lm_k -= ri3 p_j
            g[k] += m[k] * ri3 * mu[j] - ri3 * m[j] * mu[k] + cu * d
            lm[k] -= ri3 * pj
            # j side: g_j += -mu_j (ri3 m_k) + m_j (ri3 mu_k) - cu d ; lm_j += ri3 p_k
            g[j] += -mu[j] * ri3 * m[k] + m[j] * ri3 * mu[k] - cu * d
            lm[j] += ri3 * pk
    return g, lm


def vjp_tensor(x, m, eps, mu):
    n = len(m)
    g = np.zeros_like(x)
    lm = np.zeros(n)
    for k in range(n):
        for j in range(n):
            if j == k:
                continue
            d = x[j] - x[k]
            s2 = d @ d + eps * eps
            T = np.eye(3) * s2 ** -1.5 - 3.0 * np.outer(d, d) * s2 ** -2.5
            g[k] += T @ (m[k] * mu[j] - m[j] * mu[k])
            lm[k] += mu[j] @ (-d) * s2 ** -1.5
    return g, lm


def backward(x0, v0, m, dt, eps, n, lx, lv):
    c = 0.5 * dt * DRIFT
    _, _, xs = run(x0, v0, m, dt, eps, n, keep=True)
    lm = np.zeros(len(m))
    for s in reversed(range(n)):
        mv = lv + c * lx
        g, gm = vjp_split(xs[s], m, eps, dt * ACCEL * mv)
        lxh = lx + g
        lx, lv, lm = lxh, mv + c * lxh, lm + gm
    return lx, lv, lm


def test_split_equals_tidal_tensor():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(7, 3))
    m = rng.uniform(1, 3, 7)
    mu = rng.normal(size=(7, 3))
    a, am = vjp_split(x, m, 0.1, mu)
    b, bm = vjp_tensor(x, m, 0.1, mu)
    np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(am, bm, rtol=1e-12, atol=1e-14)


def test_vjp_is_the_force_jacobian():
    rng = np.random.default_rng(2)
    x = rng.normal(size=(6, 3))
    m = rng.uniform(1, 3, 6)
    mu = rng.normal(size=(6, 3))
    g, gm = vjp_split(x, m, 0.2, mu)
    h = 1e-6
    for k in range(6):
        for a in range(3):
            e = np.zeros_like(x)
            e[k, a] = h
            fd = ((force(x + e, m, 0.2) - force(x - e, m, 0.2)) * mu).sum() / (2 * h)
            assert abs(fd - g[k, a]) < 1e-7 * max(1.0, abs(fd))
        em = np.zeros(6)
        em[k] = h
        fd = ((force(x, m + em, 0.2) - force(x, m - em, 0.2)) * mu).sum() / (2 * h)
        assert abs(fd - gm[k]) < 1e-7 * max(1.0, abs(fd))


def test_reverse_sweep_matches_finite_differences():
    rng = np.random.default_rng(3)
    n_p, steps, dt, eps = 6, 5, 0.005, 5e-5
    x0 = rng.normal(size=(n_p, 3)) * 1e-3
    v0 = rng.normal(size=(n_p, 3))
    m = rng.uniform(1e5, 2e5, n_p)
    tgt = rng.normal(size=(n_p, 3)) * 1e-3

    def L(x0, v0, m):
        x, _ = run(x0, v0, m, dt, eps, steps)
        return np.mean(((x - tgt) ** 2).sum(1))

    x, _ = run(x0, v0, m, dt, eps, steps)
    gx, gv, gm = backward(x0, v0, m, dt, eps, steps, 2 * (x - tgt) / n_p, np.zeros_like(x))
    dx, dv, dm = (1e-9 * rng.normal(size=(n_p, 3)), 1e-6 * rng.normal(size=(n_p, 3)),
                  1e-1 * rng.normal(size=n_p))
    fd = (L(x0 + dx, v0 + dv, m + dm) - L(x0 - dx, v0 - dv, m - dm)) / 2
    ad = (gx * dx).sum() + (gv * dv).sum() + (gm * dm).sum()
    assert abs(fd - ad) <= 1e-6 * abs(fd)
