"""pyphasefield_amd: MI355X-native engine hot path with pyphasefield's API.

``import pyphasefield_amd as ppf`` gives ``ppf.Simulation``, ``ppf.Field``
and the utilities; ``import pyphasefield_amd.Engines as engines`` gives the
engine classes (Diffusion, CahnAllen, CahnHilliard, Warren1995, NCGPU_new),
as ``pyphasefield/__init__.py:1-3`` and ``Engines/__init__.py`` do for the
reference.  Every engine step runs in libppf_hip.so (hand-written HIP for
gfx950); there is no CPU fallback.
"""

from .field import Field  # noqa: F401
from .simulation import Simulation  # noqa: F401
from .utils import *  # noqa: F401,F403
from . import thermo  # noqa: F401

__version__ = "0.1.0"
