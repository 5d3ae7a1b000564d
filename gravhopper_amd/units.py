"""Units of the force engine, with astropy optional.

Internally (as in the reference, gravhopper.py:150-154) lengths are kpc, velocities
km/s, masses Msun and times Myr.  The C engine works with G = 1 in kpc and Msun;
jbgrav.py:38-48 converts its output with G*Msun/kpc^2 -> km/s/Myr.

The two scale factors below are what astropy computes from CODATA-2018 G, IAU-2015
GM_sun, the IAU kpc and the Julian year:
  ACCEL_SCALE = G * Msun / kpc^2 in km/s/Myr
  DRIFT_SCALE = (km/s) * Myr in kpc
When astropy is importable, Quantities are accepted and returned; otherwise plain
numbers are taken to be in the internal units above.
"""
import numpy as np

ACCEL_SCALE = 4.398600412921223e-09
DRIFT_SCALE = 1.022712165045695e-03
# G in kpc (km/s)^2 / Msun: ACCEL_SCALE * kpc/Myr in km/s, i.e. ACCEL_SCALE / DRIFT_SCALE
G_KPC_KMS2_MSUN = ACCEL_SCALE / DRIFT_SCALE

try:  # pragma: no cover - astropy is absent in the build image
    from astropy import units as u  # noqa: F401
    HAVE_ASTROPY = True
except ImportError:
    u = None
    HAVE_ASTROPY = False


def is_quantity(x):
    return HAVE_ASTROPY and hasattr(x, "unit") and hasattr(x, "to")


def value_in(x, unit_name):
    """x as a float or ndarray in `unit_name` ('kpc', 'Msun', 'Myr', 'km/s',
    'km/s/Myr'); plain numbers are taken to be in that unit already."""
    if is_quantity(x):
        return x.to(getattr(u, unit_name) if "/" not in unit_name else u.Unit(unit_name)).value
    return x


def with_unit(value, unit_name, like=None):
    """Returns a Quantity in `unit_name` when astropy is present and `like` (when
    given) is a Quantity; otherwise the plain value."""
    if HAVE_ASTROPY and (like is None or is_quantity(like)):
        return value * u.Unit(unit_name)
    return value


def check_dimension(x, unit_name, message):
    """Raises ValueError(message) unless x converts to unit_name (Quantity) or is a
    plain real number/array (taken in internal units)."""
    if is_quantity(x):
        try:
            x.to(u.Unit(unit_name))
        except u.UnitConversionError:
            raise ValueError(message) from None
        return
    try:
        np.asarray(x, dtype=np.float64)
    except (TypeError, ValueError):
        raise ValueError(message) from None
