"""Drive the oracle the way Simulation.simulate drives an engine
(simulation.py:262-266): counter++, engine step, T update, boundary rules."""

import numpy as np

from oracle import bc as obc
from oracle import engines as oe


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def run_point(kind, interior, steps, bc, dx, dt, bcarray=None, **p):
    """kind: 'diffusion' | 'allencahn' | 'cahnhilliard'; returns padded array."""
    a = obc.pad(interior)
    stack = a[None]
    obc.apply_bc(stack, bc, bcarray, dx)
    a = stack[0]
    for _ in range(steps):
        if kind == "diffusion":
            a = oe.diffusion_step(a, p["D"], dt, dx)
        elif kind == "allencahn":
            a = oe.allencahn_step(a, p["M"], p["W"], p["epsilon"], dt, dx)
        else:
            a = oe.cahnhilliard_step(a, p["M"], p["epsilon"], dt, dx, bc)
        stack = a[None].copy()
        obc.apply_bc(stack, bc, bcarray, dx)
        a = stack[0]
    return a


def run_warren(phi, c, steps, bc, dx, params, seed=0, step0=0, bcarray=None):
    f = np.stack([obc.pad(phi), obc.pad(c)])
    obc.apply_bc(f, bc, bcarray, dx)
    for s in range(step0 + 1, step0 + steps + 1):
        p2, c2 = oe.warren_step(f[0], f[1], params, params["dt"], dx, seed=seed, step=s)
        f = np.stack([p2, c2])
        obc.apply_bc(f, bc, bcarray, dx)
    return f


def run_ncomp(fields_int, steps, bc, T_of_step, params, c_params, thermo_l, thermo_s, seed=0, step0=0,
              bcarray=None):
    """fields_int: [N+2, ny, nx] interiors; T_of_step(s) -> padded T used by step s."""
    dx = params[0]
    f = np.stack([obc.pad(x) for x in fields_int])
    obc.apply_bc(f, bc, bcarray, dx)
    for s in range(step0 + 1, step0 + steps + 1):
        T = T_of_step(s)
        tr = oe.ncomp_helper(f, T, params, c_params, thermo_l, thermo_s, seed=seed, step=s)
        f = oe.ncomp_main(f, T, tr, params, c_params, seed=seed, step=s)
        obc.apply_bc(f, bc, bcarray, dx)
    return f


def padded_T(T_int, bc):
    T = obc.pad(T_int)
    obc.apply_bc_single(T, bc)
    return T


def interface_band(phi_a, phi_b, lo=0.01, hi=0.99, width=1):
    """Cells that are not bulk liquid (phi < lo) or bulk solid (phi > hi) in
    both fields, dilated by ``width`` cells (Chebyshev)."""
    bulk = ((phi_a < lo) & (phi_b < lo)) | ((phi_a > hi) & (phi_b > hi))
    band = ~bulk
    out = band.copy()
    for dy in range(-width, width + 1):
        for dx in range(-width, width + 1):
            out |= np.roll(np.roll(band, dy, 0), dx, 1)
    return out


def composition_errors(c_got, c_want, phi_got, phi_want):
    """(full-domain L-inf, bulk L-inf outside the interface band, relative L2)
    over the composition fields (lists of interior arrays)."""
    band = interface_band(phi_got, phi_want)
    full = max(float(np.abs(a - b).max()) for a, b in zip(c_got, c_want))
    bulk = max(float(np.abs(a - b)[~band].max()) for a, b in zip(c_got, c_want))
    l2 = max(float(np.linalg.norm(a - b) / np.linalg.norm(b)) for a, b in zip(c_got, c_want))
    return full, bulk, l2


def run_aniso(fields_int, steps, bc, p, dx, seed=0, step0=0, bcarray=None):
    """AnisoDorr oracle: fields_int [4, ny, nx] interiors (phi, q1, q4, c)."""
    from oracle import aniso as oa
    f = np.stack([obc.pad(x) for x in fields_int])
    obc.apply_bc(f, bc, bcarray, dx)
    for s in range(step0 + 1, step0 + steps + 1):
        tr = oa.helper(f, p, dx)
        f = oa.step(f, tr, p, dx, seed=seed, step=s)
        obc.apply_bc(f, bc, bcarray, dx)
    return f
