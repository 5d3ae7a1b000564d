"""ctypes front end of the CPU restatement in jbgrav_oracle.c (test infrastructure only).

Imported by tests/, the smoke check and bench.py's cpu_baseline leg as the checker.
The product package never imports this module.

Units are the reference's code units: positions in kpc, masses in Msun, G = 1
(gravhopper/_jbgrav.c works in these, jbgrav.py:34-48 converts around it).
"""
import ctypes
import glob
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

# astropy-derived scales used by the reference's unit layer (jbgrav.py:40-48,
# gravhopper.py:409,416); derived from CODATA-2018 G, IAU-2015 GM_sun, kpc, Julian Myr
ACCEL_SCALE = 4.398600412921223e-09   # G*Msun/kpc^2 -> km/s/Myr
DRIFT_SCALE = 1.022712165045695e-03   # km/s * Myr -> kpc

ORC_ERRORS = {
    -1: "out of memory",
    -2: "coincident particles",
    -3: "tree depth exhausted",
    -4: "bad argument",
}

NODE_DTYPE = np.dtype([
    ("level", np.int32), ("branch", np.int32), ("leaf", np.int64),
    ("center", np.float64, (3,)), ("size", np.float64), ("mass", np.float64),
    ("com", np.float64, (3,)),
])

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle/liboracle.so is not built (run `make -C oracle`)")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        I64 = ctypes.c_int64
        D = ctypes.c_double
        I = ctypes.c_int
        L.orc_direct.argtypes = [P, P, I64, D, P, I]
        L.orc_direct_position.argtypes = [P, P, I64, P, I64, D, P, I]
        L.orc_tree.argtypes = [P, P, I64, P, I64, D, D, P, P, I]
        L.orc_tree_dump.argtypes = [P, P, I64, D, P, I64]
        L.orc_tree_dump.restype = I64
        L.orc_dkd_direct.argtypes = [P, P, P, I64, D, D, I64, D, D, P, P, I]
        L.orc_energy.argtypes = [P, P, P, I64, D, P, P, P, I]
        for f in (L.orc_direct, L.orc_direct_position, L.orc_tree, L.orc_dkd_direct,
                  L.orc_energy):
            f.restype = I
        _lib = L
    return _lib


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        a = a.reshape(shape)
    return a


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _check(rc):
    if rc < 0:
        raise RuntimeError("oracle: " + ORC_ERRORS.get(int(rc), str(rc)))
    return rc


def default_threads():
    return max(1, min(16, os.cpu_count() or 1))


def direct(pos, mass, eps, nthreads=None):
    """Restatement of _jbgrav.direct_summation (_jbgrav.c:140-193)."""
    pos = _f64(pos)
    mass = _f64(mass)
    out = np.empty_like(pos)
    _check(lib().orc_direct(_ptr(pos), _ptr(mass), len(mass), float(eps), _ptr(out),
                            nthreads or default_threads()))
    return out


def direct_position(pos, mass, force_pos, eps, nthreads=None):
    """Restatement of _jbgrav.direct_summation_position (_jbgrav.c:299-353)."""
    pos = _f64(pos)
    mass = _f64(mass)
    fpos = _f64(force_pos)
    out = np.empty_like(fpos)
    _check(lib().orc_direct_position(_ptr(pos), _ptr(mass), len(mass), _ptr(fpos),
                                     len(fpos), float(eps), _ptr(out),
                                     nthreads or default_threads()))
    return out


def tree(pos, mass, eps, theta=0.7, force_pos=None, count=False, nthreads=None):
    """Restatement of _jbgrav.tree_force / tree_force_position (_jbgrav.c:737-806).

    With count=True also returns the number of accepted nodes per target."""
    pos = _f64(pos)
    mass = _f64(mass)
    fpos = pos if force_pos is None else _f64(force_pos)
    out = np.empty_like(fpos)
    cnt = np.empty(len(fpos), dtype=np.int64) if count else None
    _check(lib().orc_tree(_ptr(pos), _ptr(mass), len(mass), _ptr(fpos), len(fpos),
                          float(eps), float(theta), _ptr(out), _ptr(cnt),
                          nthreads or default_threads()))
    return (out, cnt) if count else out


def tree_dump(pos, mass, eps):
    """Canonical pre-order dump of the reference octree (children in branch order)."""
    pos = _f64(pos)
    mass = _f64(mass)
    n = _check(lib().orc_tree_dump(_ptr(pos), _ptr(mass), len(mass), float(eps), None, 0))
    out = np.zeros(n, dtype=NODE_DTYPE)
    _check(lib().orc_tree_dump(_ptr(pos), _ptr(mass), len(mass), float(eps), _ptr(out), n))
    return out


def dkd_direct(pos, vel, mass, dt, eps, nsteps, history=False, nthreads=None):
    """DKD leapfrog of Simulation.perform_timestep (gravhopper.py:405-416) with the
    direct-sum force, in code units (kpc, km/s, Msun, Myr)."""
    pos = _f64(pos).copy()
    vel = _f64(vel).copy()
    mass = _f64(mass)
    n = len(mass)
    hp = np.empty((nsteps, n, 3)) if history else None
    hv = np.empty((nsteps, n, 3)) if history else None
    _check(lib().orc_dkd_direct(_ptr(pos), _ptr(vel), _ptr(mass), n, float(dt), float(eps),
                                int(nsteps), DRIFT_SCALE, ACCEL_SCALE, _ptr(hp), _ptr(hv),
                                nthreads or default_threads()))
    if history:
        return pos, vel, hp, hv
    return pos, vel


def energy(pos, vel, mass, eps, nthreads=None):
    """(kinetic [Msun (km/s)^2], potential [Msun^2/kpc, G=1], momentum [Msun km/s])."""
    pos = _f64(pos)
    vel = _f64(vel)
    mass = _f64(mass)
    ke = ctypes.c_double()
    pe = ctypes.c_double()
    mom = np.zeros(3)
    _check(lib().orc_energy(_ptr(pos), _ptr(vel), _ptr(mass), len(mass), float(eps),
                            ctypes.byref(ke), ctypes.byref(pe), _ptr(mom),
                            nthreads or default_threads()))
    return ke.value, pe.value, mom


def ref_module():
    """The reference's own _jbgrav extension built from its source into oracle/_ref,
    or None where it was not built."""
    import importlib.util
    cands = glob.glob(os.path.join(HERE, "_ref", "_jbgrav*.so"))
    if not cands:
        return None
    spec = importlib.util.spec_from_file_location("_jbgrav", cands[0])
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod
