"""Seeded synthetic initial conditions for the benchmark and test workloads.

All outputs are in the engine's internal units: positions kpc, velocities km/s,
masses Msun.  The Plummer and Hernquist samplers follow the formulas of the reference's
IC generators (gravhopper/gravhopper.py:1449-1493 and :1544-1607) without astropy;
the exponential disk and the ensemble are generators of this project.
"""
import numpy as np

from .units import G_KPC_KMS2_MSUN as G


def _isotropic(rng, n):
    cost = rng.uniform(-1.0, 1.0, size=n)
    phi = rng.uniform(0.0, 2.0 * np.pi, size=n)
    sint = np.sqrt(1.0 - cost ** 2)
    return np.stack([sint * np.cos(phi), sint * np.sin(phi), cost], axis=1)


def _recentre(pos, vel, mass=None):
    """Mean position and mean velocity to zero (force_centers, gravhopper.py:1740-1785)."""
    pos = pos - pos.mean(axis=0)
    vel = vel - vel.mean(axis=0)
    return pos, vel


def plummer(n, b=1e-3, mass=1e6, seed=0, velocities="rejection"):
    """Isotropic Plummer sphere.

    Radii invert the cumulative mass profile, r = b / sqrt(xi^(-2/3) - 1)
    (gravhopper.py:1456-1457).  Speeds v = q sqrt(2 G M / b) (1 + r^2/b^2)^(-1/4) with
    q drawn from q^2 (1 - q^2)^(7/2) (Aarseth, Henon & Wielen 1974), either by rejection
    sampling ("rejection") or by inverting the reference's 101-point tabulated CDF
    ("tabulated", gravhopper.py:1469-1476).
    """
    rng = np.random.default_rng(seed)
    xi = rng.uniform(0.0, 1.0, size=n)
    r = b / np.sqrt(xi ** (-2.0 / 3.0) - 1.0)
    pos = r[:, None] * _isotropic(rng, n)
    if velocities == "tabulated":
        qax = np.arange(0, 1.01, 0.01)
        prob = qax ** 2 * (1.0 - qax ** 2) ** 3.5
        cum = np.cumsum(prob)
        cum /= cum[-1]
        q = np.interp(rng.uniform(0.0, 1.0, size=n), cum, qax)
    elif velocities == "rejection":
        gmax = (2.0 / 9.0) * (7.0 / 9.0) ** 3.5  # maximum of q^2 (1-q^2)^3.5, at q^2 = 2/9
        q = np.empty(n)
        filled = 0
        while filled < n:
            m = max(2 * (n - filled), 1024)
            qq = rng.uniform(0.0, 1.0, size=m)
            yy = rng.uniform(0.0, gmax, size=m)
            ok = qq[yy < qq ** 2 * (1.0 - qq ** 2) ** 3.5]
            take = min(len(ok), n - filled)
            q[filled:filled + take] = ok[:take]
            filled += take
    else:
        raise ValueError("velocities must be 'rejection' or 'tabulated'")
    speed = q * np.sqrt(2.0 * G * mass / b) * (1.0 + (r / b) ** 2) ** (-0.25)
    vel = speed[:, None] * _isotropic(rng, n)
    pos, vel = _recentre(pos, vel)
    return {"pos": pos, "vel": vel, "mass": np.full(n, mass / n)}


def hernquist(n, a=10.0, mass=1e12, cutoff=10.0, seed=0):
    """Isotropic Hernquist (1990) sphere truncated at cutoff*a.

    Radii r/a = 1/(xi^(-1/2) - 1), xi uniform below cutoff^2/(1+cutoff)^2; binding
    energies drawn from f(E) (Hernquist 1990 eq. 17) below -Phi(r) by inverting its
    tabulated cumulative integral (gravhopper.py:1550-1590).
    """
    from scipy import integrate

    rng = np.random.default_rng(seed)
    xi_cut = cutoff ** 2 / (1.0 + cutoff) ** 2
    xi = rng.uniform(0.0, xi_cut, size=n)
    roa = 1.0 / (1.0 / np.sqrt(xi) - 1.0)
    pos = (roa * a)[:, None] * _isotropic(rng, n)

    def fE(E):
        return (np.sqrt(E) * (1 - 2.0 * E) * (8.0 * E * E - 8 * E - 3.0) / ((1.0 - E) ** 2)
                + 3.0 * np.arcsin(np.sqrt(E)) / ((1.0 - E) ** 2.5)) / (8.0 * np.sqrt(2) * np.pi ** 3)

    psi = 1.0 / (1.0 + roa)  # -Phi in units of G M / a
    top = max(0.998, float(psi.max()))
    eax = np.append(np.arange(0.0, 0.998, 0.002), top)
    cum = np.array([integrate.quad(fE, 0.0, e)[0] for e in eax])
    cum /= cum.max()
    cmax = np.interp(psi, eax, cum)
    e = np.interp(rng.uniform(0.0, cmax), cum, eax)
    speed = np.sqrt(np.maximum(2.0 * (psi - e), 0.0) * G * mass / a)
    vel = speed[:, None] * _isotropic(rng, n)
    pos, vel = _recentre(pos, vel)
    return {"pos": pos, "vel": vel, "mass": np.full(n, mass / n)}


def _hernquist_menc(r, a, mass):
    return mass * r ** 2 / (r + a) ** 2


def exp_disk(n, rd=3.0, z0=0.3, mass=5e10, seed=0, halo_a=10.0, halo_mass=1e12,
             sigma_frac=0.1):
    """Exponential disk, Sigma ~ exp(-R/Rd), rho ~ sech^2(z/z0).

    Circular speed from the enclosed mass of disk (spherical approximation) plus a
    Hernquist halo, v_c^2 = G (M_disk(<R) + M_halo(<R)) / R, with an isotropic Gaussian
    dispersion of sigma_frac * v_c per component.
    """
    rng = np.random.default_rng(seed)
    # cumulative disk mass fraction 1 - (1 + x) e^-x, inverted on a fine grid
    x = np.linspace(0.0, 40.0, 400001)
    cdf = 1.0 - (1.0 + x) * np.exp(-x)
    R = rd * np.interp(rng.uniform(0.0, cdf[-1], size=n), cdf, x)
    phi = rng.uniform(0.0, 2.0 * np.pi, size=n)
    z = z0 * np.arctanh(rng.uniform(-1.0, 1.0, size=n) * (1.0 - 1e-12))
    pos = np.stack([R * np.cos(phi), R * np.sin(phi), z], axis=1)
    mdisk = mass * (1.0 - (1.0 + R / rd) * np.exp(-R / rd))
    menc = mdisk + _hernquist_menc(R, halo_a, halo_mass)
    vc = np.sqrt(G * menc / np.maximum(R, 1e-12))
    vel = np.stack([-vc * np.sin(phi), vc * np.cos(phi), np.zeros(n)], axis=1)
    vel += rng.normal(0.0, 1.0, size=(n, 3)) * (sigma_frac * vc)[:, None]
    return {"pos": pos, "vel": vel, "mass": np.full(n, mass / n)}


def galaxy(n_halo=524288, n_disk=524288, seed=3):
    """Hernquist halo (a = 10 kpc, 1e12 Msun, cut at 100 kpc) plus exponential disk
    (Rd = 3 kpc, z0 = 0.3 kpc, 5e10 Msun)."""
    h = hernquist(n_halo, a=10.0, mass=1e12, cutoff=10.0, seed=seed)
    d = exp_disk(n_disk, rd=3.0, z0=0.3, mass=5e10, seed=seed + 1000)
    return {k: np.concatenate([h[k], d[k]]) for k in ("pos", "vel", "mass")}


def plummer_crossing_time(b, mass):
    """t_cr = sqrt(b^3 / (G M)) in Myr for b in kpc, mass in Msun."""
    g_kpc3_msun_myr2 = G * 1.022712165045695e-03 ** 2
    return np.sqrt(b ** 3 / (g_kpc3_msun_myr2 * mass))


def ensemble(n_sims=128, n_lo=1024, n_hi=65536, seed=4):
    """Independent Plummer spheres: N, b and M log-uniform in [n_lo, n_hi],
    [0.5, 2] pc and [1e5, 1e7] Msun; eps = 0.02 b; dt = 0.005 t_cr."""
    rng = np.random.default_rng(seed)
    ns = np.exp(rng.uniform(np.log(n_lo), np.log(n_hi), size=n_sims)).astype(np.int64)
    bs = np.exp(rng.uniform(np.log(0.5e-3), np.log(2e-3), size=n_sims))
    ms = np.exp(rng.uniform(np.log(1e5), np.log(1e7), size=n_sims))
    sims = []
    for k in range(n_sims):
        ic = plummer(int(ns[k]), b=bs[k], mass=ms[k], seed=seed * 100003 + k)
        ic["eps"] = 0.02 * bs[k]
        ic["dt"] = 0.005 * plummer_crossing_time(bs[k], ms[k])
        ic["b"] = bs[k]
        sims.append(ic)
    return sims
