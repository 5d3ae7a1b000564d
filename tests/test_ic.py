"""IC generators with the reference's interface (gravhopper.py:1326-1785), CPU only.

Their exact samples follow numpy's default_rng draw order of the reference, but the
reference itself cannot be imported here (astropy is absent), so sample-level parity is
unpinned; the tests pin the distributions by their closed-form properties."""
import numpy as np
import pytest

from gravhopper_amd import IC, ICException, force_centers
from gravhopper_amd.units import G_KPC_KMS2_MSUN as G


def test_missing_arguments():
    with pytest.raises(ICException):
        IC.Plummer(N=10, b=1.0)
    with pytest.raises(ICException):
        IC.Hernquist(N=10, totmass=1.0)
    with pytest.raises(ICException):
        IC.TSIS(maxrad=1.0, totmass=1.0)


def test_plummer_profile_and_virial():
    ic = IC.Plummer(N=20000, b=1e-3, totmass=1e6, seed=1)
    r = np.linalg.norm(ic['pos'], axis=1)
    # half-mass radius of a Plummer sphere: b / sqrt(2^(2/3) - 1) = 1.305 b
    assert abs(np.median(r) / 1e-3 - 1.305) < 0.03
    # kinetic energy of the isotropic Plummer DF: 3 pi G M^2 / (64 b)
    ke = 0.5 * (ic['mass'][:, None] * ic['vel'] ** 2).sum()
    assert abs(ke / (3 * np.pi * G * 1e12 / (64 * 1e-3)) - 1) < 0.03
    assert np.abs(ic['pos'].mean(0)).max() < 1e-15 and np.abs(ic['vel'].mean(0)).max() < 1e-12
    # reproducible with the seed
    assert np.array_equal(IC.Plummer(N=100, b=1.0, totmass=1.0, seed=3)['pos'],
                          IC.Plummer(N=100, b=1.0, totmass=1.0, seed=3)['pos'])


def test_hernquist_profile():
    ic = IC.Hernquist(N=20000, a=10.0, totmass=1e12, seed=2)
    r = np.linalg.norm(ic['pos'], axis=1)
    assert r.max() <= 100.0 * (1 + 1e-12) + 1.0
    # mass within r: M r^2/(r+a)^2 normalised to the cutoff mass fraction
    frac = (r < 10.0).mean()
    assert abs(frac - 0.25 / (100.0 / 121.0)) < 0.02
    v = np.linalg.norm(ic['vel'], axis=1)
    assert np.all(np.isfinite(v))


def test_tsis_dispersion():
    ic = IC.TSIS(N=30000, maxrad=100.0, totmass=1e11, seed=3)
    sigma = np.sqrt(1e11 * G / 200.0)
    assert abs(ic['vel'].std() / sigma - 1) < 0.02
    assert np.linalg.norm(ic['pos'], axis=1).max() <= 100.0 + 1.0


def test_expdisk_mass_and_rotation():
    ic = IC.expdisk(sigma0=2e8, Rd=2.0, z0=0.5, sigmaR_Rd=10.0,
                    external_rotcurve=lambda r: 200.0 + 0 * r, N=20000, seed=4)
    assert abs(ic['mass'].sum() - np.pi * 4.0 * 2e8) < 1e-6 * ic['mass'].sum()
    R = np.hypot(ic['pos'][:, 0], ic['pos'][:, 1])
    vphi = (ic['pos'][:, 0] * ic['vel'][:, 1] - ic['pos'][:, 1] * ic['vel'][:, 0]) / R
    outer = R > 6.0
    assert abs(np.median(vphi[outer]) - 210.0) < 25.0   # halo 200 km/s plus the disk


def test_force_centers():
    pos = np.ones((4, 3))
    vel = np.zeros((4, 3))
    p, v = force_centers(pos, vel, center_pos=[1.0, 2.0, 3.0], center_vel=[0, 0, 5.0])
    assert np.allclose(p.mean(0), [1, 2, 3]) and np.allclose(v.mean(0), [0, 0, 5])
    p, v = force_centers(pos, vel, force_origin=False)
    assert np.array_equal(p, pos)
