"""Initial-condition generators with the reference's interface, astropy optional.

Restates `IC.Plummer`, `IC.Hernquist`, `IC.TSIS`, `IC.expdisk` and `force_centers` of
gravhopper/gravhopper.py:1326-1785: same arguments, same formulas and the same sequence
of draws from numpy's `default_rng(seed)`.  Without astropy, inputs and outputs are plain
numbers in the engine's units (kpc, km/s, Msun, Msun/kpc^2); with astropy, Quantities
are converted on the way in and returned on the way out.  galpy/pynbody-based
generators (`from_galpy_df`, `from_pyn_snap`) are not provided.
"""
import numpy as np

from .simulation import ICException
from .units import G_KPC_KMS2_MSUN as G, HAVE_ASTROPY, is_quantity, u, value_in


def _val(x, unit_name):
    return None if x is None else np.asarray(value_in(x, unit_name), dtype=np.float64)


def _out(pos, vel, mass, like):
    if HAVE_ASTROPY and is_quantity(like):  # pragma: no cover
        return {'pos': pos * u.kpc, 'vel': vel * u.km / u.s, 'mass': mass * u.Msun}
    return {'pos': pos, 'vel': vel, 'mass': mass}


def force_centers(positions, velocities, center_pos=None, center_vel=None, force_origin=True):
    """Shift to the requested centre of mass position and mean velocity
    (gravhopper.py:1740-1785)."""
    newpos = np.array(positions, dtype=np.float64)
    newvel = np.array(velocities, dtype=np.float64)
    if force_origin:
        if center_pos is None:
            center_pos = np.zeros(3)
        if center_vel is None:
            center_vel = np.zeros(3)
    if center_pos is not None:
        newpos += _val(center_pos, 'kpc') - np.mean(newpos, axis=0)
    if center_vel is not None:
        newvel += _val(center_vel, 'km/s') - np.mean(newvel, axis=0)
    return newpos, newvel


def _isotropic(rng, N):
    costheta = rng.uniform(-1.0, 1.0, size=N)
    phi = rng.uniform(0.0, 2.0 * np.pi, size=N)
    sintheta = np.sqrt(1.0 - costheta ** 2)
    return sintheta * np.cos(phi), sintheta * np.sin(phi), costheta


class IC(object):
    """Static initial-condition generators (gravhopper.py:1167-1734)."""

    @staticmethod
    def TSIS(N=None, maxrad=None, totmass=None, center_pos=None, center_vel=None,
             force_origin=True, seed=None):
        """Truncated singular isothermal sphere (gravhopper.py:1326-1400)."""
        if (N is None) or (maxrad is None) or (totmass is None):
            raise ICException("TSIS requires N, maxrad, and totmass.")
        rng = np.random.default_rng(seed)
        R = float(value_in(maxrad, 'kpc'))
        M = float(value_in(totmass, 'Msun'))
        sigma = np.sqrt(M * G / (2 * R))
        radius = rng.uniform(0.0, R, size=N)
        sx, sy, sz = _isotropic(rng, N)
        pos = np.stack([radius * sx, radius * sy, radius * sz], axis=1)
        vel = np.stack([rng.normal(0.0, sigma, size=N) for _ in range(3)], axis=1)
        pos, vel = force_centers(pos, vel, center_pos, center_vel, force_origin)
        return _out(pos, vel, np.ones(N) * (M / N), totmass)

    @staticmethod
    def Plummer(N=None, b=None, totmass=None, center_pos=None, center_vel=None,
                force_origin=True, seed=None):
        """Isotropic Plummer sphere (gravhopper.py:1403-1493): radii from the inverse
        cumulative mass profile, speeds q sqrt(2GM/b) (1 + r^2/b^2)^(-1/4) with q from
        the inverse of the tabulated cumulative q^2 (1-q^2)^(7/2)."""
        if (N is None) or (b is None) or (totmass is None):
            raise ICException("Plummer requires N, b, and totmass.")
        from scipy.interpolate import interp1d
        rng = np.random.default_rng(seed)
        bb = float(value_in(b, 'kpc'))
        M = float(value_in(totmass, 'Msun'))
        rad_xi = rng.uniform(0.0, 1.0, size=N)
        radius = bb / np.sqrt(rad_xi ** (-2. / 3) - 1)
        sx, sy, sz = _isotropic(rng, N)
        qax = np.arange(0, 1.01, 0.01)
        q_prob = qax ** 2 * (1. - qax ** 2) ** 3.5
        q_cumprob = np.cumsum(q_prob)
        q_cumprob /= q_cumprob[-1]
        q = interp1d(q_cumprob, qax)(rng.uniform(0.0, 1.0, size=N))
        velocity = q * np.sqrt(2. * G * M / bb) * (1. + (radius / bb) ** 2) ** (-0.25)
        vx, vy, vz = _isotropic(rng, N)
        pos = np.stack([radius * sx, radius * sy, radius * sz], axis=1)
        vel = np.stack([velocity * vx, velocity * vy, velocity * vz], axis=1)
        pos, vel = force_centers(pos, vel, center_pos, center_vel, force_origin)
        return _out(pos, vel, np.ones(N) * (M / N), totmass)

    @staticmethod
    def Hernquist(N=None, a=None, totmass=None, cutoff=10., center_pos=None, center_vel=None,
                  force_origin=True, seed=None):
        """Isotropic Hernquist (1990) sphere truncated at cutoff*a (gravhopper.py:1496-1607).
        The binding-energy table is extended to the most bound particle when needed."""
        if (N is None) or (a is None) or (totmass is None):
            raise ICException("Hernquist requires N, a, and totmass.")
        from scipy import integrate
        from scipy.interpolate import interp1d
        rng = np.random.default_rng(seed)
        aa = float(value_in(a, 'kpc'))
        M = float(value_in(totmass, 'Msun'))
        xi_cutoff = cutoff ** 2 / ((1. + cutoff) ** 2)
        rad_xi = rng.uniform(0.0, xi_cutoff, size=N)
        r_over_a = 1. / (1. / np.sqrt(rad_xi) - 1.)
        sx, sy, sz = _isotropic(rng, N)
        radius = r_over_a * aa

        def fE(E):
            return (np.sqrt(E) * (1 - 2. * E) * (8. * E * E - 8 * E - 3.) / ((1. - E) ** 2) +
                    3. * np.arcsin(np.sqrt(E)) / ((1. - E) ** (5. / 2))) / (8. * np.sqrt(2) * np.pi ** 3)

        Eax = np.arange(0.0, 1.0, 0.002)
        cum = [integrate.quad(fE, 0.0, Etop)[0] for Etop in Eax]
        potential = -1. / (1. + r_over_a)
        most_bound = np.max(-potential)
        if most_bound > Eax.max():
            Eax = np.append(Eax, most_bound)
            cum.append(integrate.quad(fE, 0.0, most_bound)[0])
        cum = np.array(cum) / np.max(cum)
        Einterp = interp1d(cum, Eax)
        inverse_Einterp = interp1d(Eax, cum)
        max_possible_xi = inverse_Einterp(-potential)
        E_xi = rng.uniform(0.0, max_possible_xi, size=N)
        bindingE = Einterp(E_xi)
        velocity = np.sqrt(2. * (-(bindingE + potential)) * (G * M / aa))
        vx, vy, vz = _isotropic(rng, N)
        pos = np.stack([radius * sx, radius * sy, radius * sz], axis=1)
        vel = np.stack([velocity * vx, velocity * vy, velocity * vz], axis=1)
        pos, vel = force_centers(pos, vel, center_pos, center_vel, force_origin)
        return _out(pos, vel, np.ones(N) * (M / N), totmass)

    @staticmethod
    def expdisk(sigma0=None, Rd=None, z0=None, sigmaR_Rd=None, external_rotcurve=None, N=None,
                center_pos=None, center_vel=None, force_origin=True, seed=None):
        """Exponential disk with a sech^2 vertical profile in approximate equilibrium
        (gravhopper.py:1610-1734).  sigma0 in Msun/kpc^2 (or a Quantity), Rd and z0 in
        kpc, sigmaR_Rd in km/s; external_rotcurve(R_kpc) returns km/s."""
        from scipy import special
        from scipy.differentiate import derivative
        from scipy.interpolate import interp1d
        rng = np.random.default_rng(seed)
        s0 = float(value_in(sigma0, 'Msun/kpc2')) if is_quantity(sigma0) else float(sigma0)
        Rd_kpc = float(value_in(Rd, 'kpc'))
        z0_kpc = float(value_in(z0, 'kpc'))
        sRd = float(value_in(sigmaR_Rd, 'km/s'))
        totmass = np.pi * Rd_kpc ** 2 * s0
        Rax = np.arange(0.001 * Rd_kpc, 10 * Rd_kpc, 0.01 * Rd_kpc)
        R_cumprob = Rd_kpc ** 2 - Rd_kpc * np.exp(-Rax / Rd_kpc) * (Rax + Rd_kpc)
        R_cumprob /= R_cumprob[-1]
        R = interp1d(R_cumprob, Rax)(rng.uniform(0.0, 1.0, size=N))
        phi = rng.uniform(0.0, 2.0 * np.pi, size=N)
        x = R * np.cos(phi)
        y = R * np.sin(phi)
        z = 2 * z0_kpc * np.arctanh(rng.uniform(0, 1.0, size=N))
        z *= (2 * (rng.uniform(0, 1, size=N) < 0.5)) - 1

        def om2(rad):
            y_R = rad / (2. * Rd_kpc)
            omega2 = np.pi * G * s0 / Rd_kpc * (special.iv(0, y_R) * special.kv(0, y_R) -
                                                special.iv(1, y_R) * special.kv(1, y_R))
            if external_rotcurve is not None:
                omega2 = omega2 + (np.asarray(value_in(external_rotcurve(rad), 'km/s')) / rad) ** 2
            return omega2

        Omega2 = om2(R)
        kappa2 = 4. * Omega2 + R * derivative(om2, R, initial_step=1e-3).df
        sigma_R = sRd * np.exp(0.25 * (1 - R / Rd_kpc))
        sigma2_phi = sigma_R ** 2 * 4 * Omega2 / kappa2
        sigma2_z = np.pi * G * z0_kpc * s0 * 0.5 * np.exp(-R / Rd_kpc)
        vphi_mean = R * np.sqrt(Omega2)
        vphi = np.sqrt(sigma2_phi) * rng.normal(size=N) + vphi_mean
        vR = sigma_R * rng.normal(size=N)
        vx = -vphi * np.sin(phi) + vR * np.cos(phi)
        vy = vphi * np.cos(phi) + vR * np.sin(phi)
        vz = np.sqrt(sigma2_z) * rng.normal(size=N)
        pos, vel = force_centers(np.stack([x, y, z], axis=1), np.stack([vx, vy, vz], axis=1),
                                 center_pos, center_vel, force_origin)
        return _out(pos, vel, np.ones(N) * (totmass / N), Rd)
