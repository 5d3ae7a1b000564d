"""gravhopper_amd: an MI355X (gfx950) gravity engine behind GravHopper's force API.

Mirrors the namespace of the reference package (gravhopper/__init__.py:1-9):
`Simulation`, `grav` (the jbgrav force functions) and the exceptions.  All force
arithmetic runs in libgravhop.so (hand-written HIP) through ctypes.
"""
__version__ = '1.2.0+mi355x.1'

from .simulation import (Simulation, GravHopperException, UninitializedSimulationException,
                         ICException, UnknownAlgorithmException, ExternalPackageException)
from . import jbgrav as grav
from . import _jbgrav
from .ic import IC, force_centers

__all__ = ['Simulation', 'IC', 'force_centers', 'grav', '_jbgrav', 'GravHopperException',
           'UninitializedSimulationException', 'ICException', 'UnknownAlgorithmException',
           'ExternalPackageException', '__version__']
