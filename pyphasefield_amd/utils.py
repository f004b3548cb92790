"""Host utilities exported at package level (ppf_utils.py of the reference):
the colormaps, ``make_seed``, the ThermoModel container that stands in for
TDBContainer on this engine path, and the thermal-history helpers CSVtoXDMF
and XDMFLoader (``ppf_utils.py:63-123``, implemented in ``thermal.py``)."""

import numpy as np

from .thermo import ThermoModel as TDBContainer  # noqa: F401
from .thermal import CSVtoXDMF, XDMFLoader  # noqa: F401

__all__ = ["COLORMAP_OTHER", "COLORMAP_PHASE", "COLORMAP_PHASE_INV", "make_seed", "TDBContainer",
           "CSVtoXDMF", "XDMFLoader",
           "successfully_imported_pycalphad", "successfully_imported_numba"]

_OTHER = [(0, 0, 1), (0, 1, 1), (0, 1, 0), (1, 1, 0), (1, 0, 0)]
_PHASE = [(1, 0, 0), (1, 1, 0), (0, 0, 1)]
_PHASE_INV = [(0, 0, 1), (1, 1, 0), (1, 0, 0)]

try:  # colormaps as in ppf_utils.py:8-13 when matplotlib is present
    from matplotlib.colors import LinearSegmentedColormap
    COLORMAP_OTHER = LinearSegmentedColormap.from_list("rgb", _OTHER)
    COLORMAP_PHASE = LinearSegmentedColormap.from_list("rgb", _PHASE)
    COLORMAP_PHASE_INV = LinearSegmentedColormap.from_list("rgb", _PHASE_INV)
except ImportError:  # pragma: no cover
    COLORMAP_OTHER, COLORMAP_PHASE, COLORMAP_PHASE_INV = "jet", "jet", "jet_r"


def successfully_imported_pycalphad():
    """This engine path never needs pycalphad (ppf_utils.py:15-27)."""
    return False


def successfully_imported_numba():
    """The GPU path is libppf_hip.so, not numba (ppf_utils.py:29-44)."""
    return False


def make_seed(phi, q1, q4, x, y, angle, seed_radius):
    """ppf_utils.py:46-61: phi = 1 on a disc of radius r, orientation
    (cos(angle/2), sin(angle/2)) on the disc of radius r+5, wrapped."""
    shape = phi.shape
    qrad = seed_radius + 5
    x_size = shape[1]
    y_size = shape[0]
    for i in range(int(y - seed_radius), int(y + seed_radius)):
        for j in range(int(x - seed_radius), int(x + seed_radius)):
            if (i - y) * (i - y) + (j - x) * (j - x) < seed_radius ** 2:
                phi[i % y_size][j % x_size] = 1
    for i in range(int(y - qrad), int(y + qrad)):
        for j in range(int(x - qrad), int(x + qrad)):
            if (i - y) * (i - y) + (j - x) * (j - x) < qrad ** 2:
                q1[i % y_size][j % x_size] = np.cos(0.5 * angle)
                q4[i % y_size][j % x_size] = np.sin(0.5 * angle)
    return phi, q1, q4
