"""Drop-in for `gravhopper.jbgrav`: the four public force functions.

Same names and arguments as gravhopper/jbgrav.py:14,52,91,136.  Each takes a
snapshot dict with 'pos' (N,3) and 'mass' (N,), converts to kpc and Msun, calls the
GPU engine in code units and returns the acceleration in km/s/Myr, multiplying by
G*Msun/kpc^2 = 4.398600412921223e-9 km/s/Myr (jbgrav.py:38-48).

astropy is optional: with it, Quantities go in and a Quantity comes out, as in the
reference; without it, plain arrays are taken as kpc / Msun and a plain array in
km/s/Myr is returned.  `precision` ("fp64" default, or "fp32") is an addition.
"""
from . import _jbgrav
from .units import ACCEL_SCALE, value_in, with_unit

__all__ = ['direct_summation', 'direct_summation_position', 'tree_force',
           'tree_force_position']


def _finish(force, like):
    return with_unit(force * ACCEL_SCALE, "km/s/Myr", like=like)


def direct_summation(snap, eps, precision=None):
    """Acceleration on every particle of `snap` from every other particle, by direct
    summation (jbgrav.py:14-48)."""
    positions = snap['pos']
    masses = snap['mass']
    posarray = value_in(positions, 'kpc')
    massarray = value_in(masses, 'Msun')
    eps_in_units = value_in(eps, 'kpc')
    forcearray = _jbgrav.direct_summation(posarray, massarray, eps_in_units,
                                          precision=precision)
    return _finish(forcearray, positions)


def tree_force(snap, eps, theta=0.7, precision=None):
    """Acceleration on every particle of `snap` from a Barnes-Hut octree with opening
    angle theta (jbgrav.py:52-88)."""
    positions = snap['pos']
    masses = snap['mass']
    posarray = value_in(positions, 'kpc')
    massarray = value_in(masses, 'Msun')
    eps_in_units = value_in(eps, 'kpc')
    forcearray = _jbgrav.tree_force(posarray, massarray, eps_in_units, theta,
                                    precision=precision)
    return _finish(forcearray, positions)


def direct_summation_position(snap, force_pos, eps, precision=None):
    """Acceleration at force_pos from every particle of `snap`, by direct summation
    (jbgrav.py:91-133).  A target that coincides with a particle gets no force from
    it when eps is 0."""
    positions = snap['pos']
    masses = snap['mass']
    posarray = value_in(positions, 'kpc')
    massarray = value_in(masses, 'Msun')
    eps_in_units = value_in(eps, 'kpc')
    forceposarray = value_in(force_pos, 'kpc')
    forcearray = _jbgrav.direct_summation_position(posarray, massarray, forceposarray,
                                                   eps_in_units, precision=precision)
    return _finish(forcearray, force_pos)


def tree_force_position(snap, force_pos, eps, theta=0.7, precision=None):
    """Acceleration at force_pos from a Barnes-Hut octree of `snap`
    (jbgrav.py:136-177)."""
    positions = snap['pos']
    masses = snap['mass']
    posarray = value_in(positions, 'kpc')
    massarray = value_in(masses, 'Msun')
    forceposarray = value_in(force_pos, 'kpc')
    eps_in_units = value_in(eps, 'kpc')
    forcearray = _jbgrav.tree_force_position(posarray, massarray, forceposarray,
                                             eps_in_units, theta, precision=precision)
    return _finish(forcearray, force_pos)
