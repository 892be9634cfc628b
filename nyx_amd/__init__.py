"""nyx_amd -- MI355X-native particle-mesh gravity step for Nyx.

The product is ``libnyxpm.so`` (hand-written gfx950 HIP kernels behind the C ABI
in ``include/nyxpm.h``).  This package is its host-side interface:

* ``PMContext``                    one call per C-ABI entry point
* ``DarkMatterParticleContainer``  Nyx's particle-container interface
* ``Gravity``                      Nyx's single-level periodic Gravity interface
* ``ics``                          seeded synthetic initial conditions
"""
from ._lib import NyxPMError, load, LIB_PATH, NYXPM_FP64, NYXPM_MIXED  # noqa: F401
from .pm import (BoxGeom, PMContext, DarkMatterParticleContainer, Gravity, advance_particles_only,  # noqa: F401
                 integrate_comoving_a)
from . import ics  # noqa: F401
