"""Boundary-condition halo rules of the reference GPU path (oracle).

Follows ``ppf_gpu_utils.py:11-96`` (field kernels) and ``:197-253`` (the
"single" temperature kernels), dispatched as in ``ppf_gpu_utils.py:430-514``:
axis x (the LAST array axis) first, then y, then z, each over the whole
padded extent of the other axes, so corner cells are the y-rule applied to
values the x-rule already wrote.  The CPU path of the reference only ever
touches array axis 0 (``simulation.py:561-566``); the GPU semantics are the
ones restated here.

The per-axis list order is therefore [x, y, z] = reversed array axes, exactly
as ``bc[0]`` selects the x-kernel in ``ppf_gpu_utils.py:452``.
"""

import numpy as np

PERIODIC = "PERIODIC"
NEUMANN = "NEUMANN"
DIRICHLET = "DIRCHLET"  # reference spelling (simulation.py:587)


def normalize_bc(bc, ndim):
    """Per-axis list [x, y, z][:ndim] of canonical names (ppf_gpu_utils.py:431-436)."""
    if isinstance(bc, str) or bc is None:
        bc = [bc or PERIODIC] * ndim
    out = []
    for b in list(bc)[:ndim]:
        b = str(b).upper()
        if b == "DIRICHLET":
            b = DIRICHLET
        if b not in (PERIODIC, NEUMANN, DIRICHLET):
            raise ValueError("unknown boundary condition %r" % b)
        out.append(b)
    if len(out) != ndim:
        raise ValueError("need one boundary condition per axis")
    return out


def _sl(nd, ax, idx):
    s = [slice(None)] * nd
    s[ax] = idx
    return tuple(s)


def apply_bc(fields, bc, bcarray=None, dx=1.0):
    """Fill the 1-cell halos of ``fields`` [F, *padded] in place.

    PERIODIC:  halo <- opposite interior plane   (ppf_gpu_utils.py:38-56)
    NEUMANN:   halo <- neighbour -/+ dx*bcarray  (ppf_gpu_utils.py:78-96)
    DIRCHLET:  halo <- bcarray                   (ppf_gpu_utils.py:58-76)
    """
    nd = fields.ndim
    ndim = nd - 1
    bc = normalize_bc(bc, ndim)
    if bcarray is None:
        bcarray = np.zeros_like(fields)
    for a in range(ndim):
        ax = nd - 1 - a
        lo, lo1 = _sl(nd, ax, 0), _sl(nd, ax, 1)
        hi, hi1 = _sl(nd, ax, -1), _sl(nd, ax, -2)
        t = bc[a]
        if t == PERIODIC:
            fields[lo] = fields[hi1]
            fields[hi] = fields[lo1]
        elif t == NEUMANN:
            fields[lo] = fields[lo1] - dx * bcarray[lo]
            fields[hi] = fields[hi1] + dx * bcarray[hi]
        else:
            fields[lo] = bcarray[lo]
            fields[hi] = bcarray[hi]
    return fields


def apply_bc_single(T, bc):
    """Temperature halo (ppf_gpu_utils.py:197-253, dispatch :439-475):
    periodic wrap for PERIODIC, zero-gradient copy for NEUMANN and DIRCHLET."""
    nd = T.ndim
    bc = normalize_bc(bc, nd)
    for a in range(nd):
        ax = nd - 1 - a
        lo, lo1 = _sl(nd, ax, 0), _sl(nd, ax, 1)
        hi, hi1 = _sl(nd, ax, -1), _sl(nd, ax, -2)
        if bc[a] == PERIODIC:
            T[lo] = T[hi1]
            T[hi] = T[lo1]
        else:
            T[lo] = T[lo1]
            T[hi] = T[hi1]
    return T


def pad(interior):
    """Field storage: +1 halo cell per axis, zero filled (field.py:6-20)."""
    out = np.zeros([s + 2 for s in interior.shape], dtype=np.float64)
    out[tuple(slice(1, -1) for _ in interior.shape)] = interior
    return out


def interior(a, lead=0):
    """View of the interior cells (field.py:69-70); ``lead`` leading axes kept."""
    return a[(slice(None),) * lead + tuple(slice(1, -1) for _ in a.shape[lead:])]
