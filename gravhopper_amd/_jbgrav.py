"""Drop-in for the reference's C extension module `gravhopper._jbgrav`.

Same four functions, same positional arguments, same code units (kpc, Msun, G = 1),
same validation order and RuntimeError messages (gravhopper/_jbgrav.c:25-31, 68-135,
200-294, 564-630, 634-727).  The arithmetic runs in libgravhop.so on the GPU.

The extra keyword `precision` selects the pair arithmetic: "fp64" (default) keeps the
reference's IEEE FP64 operation order -- bit-identical results for the direct sum --
and "fp32" selects FP32 pair forces with compensated accumulation.
"""
import numpy as np

from . import _native

_default_precision = "fp64"


def set_default_precision(precision):
    """Sets the precision used when a call does not pass one ("fp64" or "fp32")."""
    global _default_precision
    _prec(precision)
    _default_precision = precision


def get_default_precision():
    return _default_precision


def _prec(precision):
    p = _default_precision if precision is None else precision
    try:
        return _native.PRECISIONS[p]
    except KeyError:
        raise ValueError("precision must be 'fp64' or 'fp32'") from None


def _as_float(x, name):
    try:
        return float(x)
    except (TypeError, ValueError):
        raise TypeError("%s must be a real number" % name) from None


def _particles(pos, mass, prefix):
    """PyArray_FROM_OTF(NPY_DOUBLE, NPY_ARRAY_IN_ARRAY) plus the shape checks of
    _jbgrav.c:89-108 (prefix "") and :224-246 (prefix "Particle ")."""
    posarray = np.ascontiguousarray(pos, dtype=np.float64)
    massarray = np.ascontiguousarray(mass, dtype=np.float64)
    if posarray.ndim != 2:
        raise RuntimeError("%s does not have 2 dimensions." %
                           ("Particle position array" if prefix else "Position array"))
    if posarray.shape[1] != 3:
        raise RuntimeError("%s is not Nx3." %
                           ("Particle position array" if prefix else "Position array"))
    n = posarray.shape[0]
    if massarray.ndim == 0 or massarray.shape[0] != n:
        raise RuntimeError("Mass array and %s contain different numbers of particles." %
                           ("particle position array" if prefix else "position array"))
    # the C code reads element j of the first axis (PyArray_GETPTR1)
    if massarray.ndim > 1:
        massarray = np.ascontiguousarray(massarray.reshape(n, -1)[:, 0])
    return posarray, massarray


def _targets(force_pos):
    """Force-position checks of _jbgrav.c:248-263 / :682-697."""
    fp = np.ascontiguousarray(force_pos, dtype=np.float64)
    if fp.ndim != 2:
        raise RuntimeError("Force position array does not have 2 dimensions.")
    if fp.shape[1] != 3:
        raise RuntimeError("Force position array is not Nx3.")
    return fp


def _ctx():
    return _native.default_context()


def direct_summation(pos, mass, eps, precision=None):
    """Acceleration on every particle from every other particle (code units)."""
    eps = _as_float(eps, "eps")
    prec = _prec(precision)
    p, m = _particles(pos, mass, "")
    try:
        return _ctx().direct(p, m, eps, prec)
    except _native.GravhopError as e:
        raise RuntimeError("Error in direct summation C code. (%s)" % e) from e


def direct_summation_position(pos, mass, force_pos, eps, precision=None):
    """Acceleration at force_pos from every particle (code units)."""
    eps = _as_float(eps, "eps")
    prec = _prec(precision)
    p, m = _particles(pos, mass, "Particle ")
    fp = _targets(force_pos)
    try:
        return _ctx().direct(p, m, eps, prec, force_pos=fp)
    except _native.GravhopError as e:
        raise RuntimeError("Error in direct summation position C code. (%s)" % e) from e


def tree_force(pos, mass, eps, theta, precision=None):
    """Barnes-Hut acceleration on every particle (code units)."""
    eps = _as_float(eps, "eps")
    theta = _as_float(theta, "theta")
    prec = _prec(precision)
    p, m = _particles(pos, mass, "")
    try:
        return _ctx().tree(p, m, eps, theta, prec)
    except _native.GravhopError as e:
        raise RuntimeError("Error in tree C code. (%s)" % e) from e


def tree_force_position(pos, mass, force_pos, eps, theta, precision=None):
    """Barnes-Hut acceleration at force_pos (code units)."""
    eps = _as_float(eps, "eps")
    theta = _as_float(theta, "theta")
    prec = _prec(precision)
    p, m = _particles(pos, mass, "Particle ")
    fp = _targets(force_pos)
    try:
        return _ctx().tree(p, m, eps, theta, prec, force_pos=fp)
    except _native.GravhopError as e:
        raise RuntimeError("Error in tree C code. (%s)" % e) from e
