"""One explicit-Euler step of each engine, FP64 numpy restatement (oracle).

All arrays are halo-padded (+1 cell per axis, ``field.py:6-20``), axes
ordered (z, y, x) with x the last axis; on entry their halos hold the values
the boundary rules of ``oracle.bc`` wrote.  Every step function returns new
padded arrays whose INTERIOR is the updated state; the caller re-applies the
boundary rules afterwards, as ``Simulation.simulate`` does
(``simulation.py:262-266``).

Defect resolutions (reference defect ids D1-D8 of the design notes):
  * Diffusion reads D from user_data and multiplies by dt (D1, D2).
  * Allen-Cahn / Cahn-Hilliard read user_data (D3); Cahn-Hilliard multiplies
    by dt and evaluates mu on the +1 ring from a consistent 2-cell halo:
    periodic axes wrap, non-periodic axes give mu a zero-flux ring
    (mu_halo = mu of the adjacent interior cell), which conserves mass.
  * Warren1995 reads its parameters from user_data (D4); noise from the
    counter RNG instead of np.random.
  * N-component: noise from the counter RNG (D8).
"""

import numpy as np

from . import bc as _bc
from . import rng as _rng
from . import thermo as _thermo


# ---------------------------------------------------------------------------
# shared helpers
# ---------------------------------------------------------------------------

def laplacian(a, dx):
    """Field.laplacian (field.py:60-67): np.roll over the padded array."""
    idx2 = (1.0 / dx) ** 2
    r = (np.roll(a, -1, 0) + np.roll(a, 1, 0) - 2 * a) * idx2
    for ax in range(1, a.ndim):
        r += (np.roll(a, -1, ax) + np.roll(a, 1, ax) - 2 * a) * idx2
    return r


def _set_interior(dst, src):
    sl = tuple(slice(1, -1) for _ in dst.shape)
    dst[sl] = src[sl]


# ---------------------------------------------------------------------------
# Diffusion  (Engines/Diffusion.py:198-208)
# ---------------------------------------------------------------------------

def diffusion_step(c, D, dt, dx):
    dc = dt * (D * laplacian(c, dx))
    out = c.copy()
    _set_interior(out, c + dc)
    return out


# ---------------------------------------------------------------------------
# Allen-Cahn  (Engines/CahnAllen.py:128-135)
# ---------------------------------------------------------------------------

def allencahn_step(phi, M, W, epsilon, dt, dx):
    deltaphi = dt * M * (epsilon ** 2 * laplacian(phi, dx)
                         - 16. * W * (4. * phi ** 3 - 6. * phi ** 2 + 2. * phi))
    out = phi.copy()
    _set_interior(out, phi + deltaphi)
    return out


# ---------------------------------------------------------------------------
# Cahn-Hilliard  (Engines/CahnHilliard.py:134-146, D3 resolution)
# ---------------------------------------------------------------------------

def cahnhilliard_mu(phi, epsilon, dx, bc):
    """mu = -eps^2 lap(phi) + (4phi^3 - 6phi^2 + 2phi) on the padded extent.

    phi must have BC-filled 1-cell halos.  Returns mu with the same shape whose
    ring (non-corner) cells are valid: periodic axes from a wrapped second
    halo ring, non-periodic axes copied from the adjacent interior cell.
    """
    nd = phi.ndim
    bcs = _bc.normalize_bc(bc, nd)
    ext = np.pad(phi, 1, mode="edge")
    for a in range(nd):
        ax = nd - 1 - a
        if bcs[a] == _bc.PERIODIC:
            n = phi.shape[ax] - 2
            s0 = [slice(None)] * nd
            s1 = [slice(None)] * nd
            s0[ax] = 0
            s1[ax] = n          # ext index n  <-> interior cell n-1 (1-based)
            ext[tuple(s0)] = ext[tuple(s1)]
            s0[ax] = n + 3
            s1[ax] = 3          # ext index 3  <-> interior cell 2
            ext[tuple(s0)] = ext[tuple(s1)]
    mu_ext = -epsilon ** 2 * laplacian(ext, dx) + (4 * ext ** 3 - 6 * ext ** 2 + 2 * ext)
    mu = mu_ext[tuple(slice(1, -1) for _ in range(nd))].copy()
    for a in range(nd):
        ax = nd - 1 - a
        if bcs[a] != _bc.PERIODIC:
            lo, lo1 = _bc._sl(nd, ax, 0), _bc._sl(nd, ax, 1)
            hi, hi1 = _bc._sl(nd, ax, -1), _bc._sl(nd, ax, -2)
            mu[lo] = mu[lo1]
            mu[hi] = mu[hi1]
    return mu


def cahnhilliard_step(phi, M, epsilon, dt, dx, bc):
    mu = cahnhilliard_mu(phi, epsilon, dx, bc)
    deltaphi = 0
    for ax in range(phi.ndim):
        deltaphi = deltaphi + (np.roll(mu, 1, ax) + np.roll(mu, -1, ax) - 2 * mu)
    deltaphi = deltaphi / (dx ** 2)
    out = phi.copy()
    _set_interior(out, phi + dt * M * deltaphi)
    return out


def cahnhilliard_free_energy(phi, epsilon, dx):
    """F = sum over interior of [c^2(1-c)^2 + eps^2/2 |grad c|^2] dx^d, with
    forward differences (uses the BC-filled halo on the high side)."""
    nd = phi.ndim
    sl = tuple(slice(1, -1) for _ in range(nd))
    c = phi[sl]
    f = c * c * (1 - c) * (1 - c)
    g2 = np.zeros_like(c)
    for ax in range(nd):
        s = [slice(1, -1)] * nd
        s[ax] = slice(2, None)
        g = (phi[tuple(s)] - c) / dx
        g2 += g * g
    return float(np.sum(f + 0.5 * epsilon ** 2 * g2) * dx ** nd)


# ---------------------------------------------------------------------------
# Warren1995  (Engines/Warren1995.py:17-92; parameters :138-169, :197-203)
# ---------------------------------------------------------------------------

def _p(phi):
    return phi * phi * phi * (10 - 15 * phi + 6 * phi * phi)


def _g(phi):
    return (phi * phi * (1 - phi) * (1 - phi))


def _gprime(phi):
    return (4 * phi * phi * phi - 6 * phi * phi + 2 * phi)


_p_vec = np.vectorize(_p)
_g_vec = np.vectorize(_g)
_gprime_vec = np.vectorize(_gprime)


def gradx(phi, dx):
    return (np.roll(phi, 1, 0) - np.roll(phi, -1, 0)) / (2 * dx)


def grady(phi, dx):
    return (np.roll(phi, 1, 1) - np.roll(phi, -1, 1)) / (2 * dx)


def gradxx(phi, dx):
    return (np.roll(phi, 1, 0) + np.roll(phi, -1, 0) - 2 * phi) / (dx * dx)


def gradyy(phi, dx):
    return (np.roll(phi, 1, 1) + np.roll(phi, -1, 1) - 2 * phi) / (dx * dx)


def warren_derived(ud, dx):
    """Engines/Warren1995.py:197-203."""
    d = dx / 0.94
    out = dict(ud)
    out["dt"] = dx ** 2 / 5. / ud["D_L"]
    out["d"] = d
    out["ebar"] = np.sqrt(6 * np.sqrt(2) * ud["s_A"] * d / ud["T_mA"])
    out["W_A"] = 3 * ud["s_A"] / (np.sqrt(2) * ud["T_mA"] * d)
    out["W_B"] = 3 * ud["s_B"] / (np.sqrt(2) * ud["T_mB"] * d)
    out["M_A"] = (ud["T_mA"] ** 2) * ud["B_A"] / (6 * np.sqrt(2) * ud["L_A"] * d)
    out["M_B"] = (ud["T_mB"] ** 2) * ud["B_B"] / (6 * np.sqrt(2) * ud["L_B"] * d)
    return out


def warren_step(phi, c, p, dt, dx, seed=0, step=0, vectorize=False):
    """Engines/Warren1995.py:50-92.  ``p`` holds the derived parameters.

    ``step`` is the 1-based time-step counter of the step being taken (the
    reference increments the counter before the engine runs,
    ``simulation.py:263-264``); it keys the noise draw.
    ``vectorize=True`` evaluates the interpolants with np.vectorize exactly as
    the reference does (``Warren1995.py:26-28``), for the CPU baseline.
    """
    deltaphi, deltac = warren_rates(phi, c, p, dx, seed, step, vectorize)
    phi_o = phi.copy()
    c_o = c.copy()
    _set_interior(phi_o, phi + deltaphi * dt)
    _set_interior(c_o, c + deltac * dt)
    return phi_o, c_o


def warren_rates(phi, c, p, dx, seed=0, step=0, vectorize=False):
    """(dphi/dt, dc/dt) of Engines/Warren1995.py:56-86 on the padded arrays."""
    if vectorize:
        g, pp, gprime = _g_vec(phi), _p_vec(phi), _gprime_vec(phi)
    else:
        g, pp, gprime = _g(phi), _p(phi), _gprime(phi)
    T = p["T"]
    H_A = p["W_A"] * gprime + 30 * p["L_A"] * (1 / T - 1 / p["T_mA"]) * g
    H_B = p["W_B"] * gprime + 30 * p["L_B"] * (1 / T - 1 / p["T_mB"]) * g
    phixx = gradxx(phi, dx)
    phiyy = gradyy(phi, dx)
    lphi = phixx + phiyy
    phix = gradx(phi, dx)
    phiy = grady(phi, dx)
    phixy = grady(phix, dx)

    D_C = p["D_S"] + pp * (p["D_L"] - p["D_S"])
    temp = D_C * p["v_m"] * c * (1 - c) * (H_B - H_A) / p["R"]
    deltac = (D_C * (gradxx(c, dx) + gradyy(c, dx))
              + (gradx(D_C, dx) * gradx(c, dx) + grady(D_C, dx) * grady(c, dx))
              + temp * lphi + (gradx(temp, dx) * phix + grady(temp, dx) * phiy))

    theta = np.arctan2(phiy, phix)
    eta = 1 + p["y_e"] * np.cos(4 * theta)
    etap = -4 * p["y_e"] * np.sin(4 * theta)
    etapp = -16 * (eta - 1)
    c2 = np.cos(2 * theta)
    s2 = np.sin(2 * theta)
    M_phi = (1 - c) * p["M_A"] + c * p["M_B"]
    ebar2 = p["ebar"] ** 2
    deltaphi = M_phi * ((ebar2 * eta * eta * lphi - (1 - c) * H_A - c * H_B)
                        + ebar2 * eta * etap * (s2 * (phiyy - phixx) + 2 * c2 * phixy)
                        + 0.5 * ebar2 * (etap * etap + eta * etapp)
                        * (-2 * s2 * phixy + lphi + c2 * (phiyy - phixx)))
    if p["alpha"] != 0:
        cells = _rng.padded_cells(phi.shape)
        u = _rng.uniform_f64(seed, step, _rng.S_WARREN_PHI, phi.size, cells)
        randArray = 2 * u - 1
        deltaphi = deltaphi + M_phi * p["alpha"] * randArray * (16 * g) * ((1 - c) * H_A + c * H_B)
    return deltaphi, deltac


# ---------------------------------------------------------------------------
# N-component  (Engines/NCGPU_new.py:28-313)
# ---------------------------------------------------------------------------

PARAM_NAMES = ["dx", "d", "v_m", "M_qmax", "H", "y_e", "beta", "D_L", "D_S", "dt",
               "ebar", "eqbar", "noise_phi", "noise_c", "noise_q"]  # NCGPU_new.py:497-511


def nc_g(phi):
    """Double-well g(phi) of NCGPU_new.py:298 (numba integer powers: products)."""
    return (phi ** 2) * (1 - phi) ** 2


def nc_h(phi):
    """Interpolant h(phi) of NCGPU_new.py:299."""
    return (phi * phi * phi) * (6. * phi ** 2 - 15. * phi + 10.)


def ncomp_helper(fields, T, params, c_params, thermo_l, thermo_s, seed=0, step=0):
    """NComponent_helper_kernel (NCGPU_new.py:256-313) over every padded cell.

    fields: [N+2, *padded] = phi, q1, q4, c_1..c_{N-1};  T: padded.
    thermo_l / thermo_s: (g0, L, R) of the liquid (ufunc 1) and solid (ufunc 0)
    phases (NCGPU_new.py:353-354).
    Returns transfer [2N, *padded] = G_L, G_S, M_c x (N-1), dFdc x (N-1):
      G_phase and its h = 1e-7 forward differences along e_l - e_N
      (:295-305, thermo.fd_derivatives),
      M_c,l = v_m c_l (D_L + h(phi) (D_S - D_L)) / (8.314 T)        (:308),
      dF/dc_l = (dGL_l + h(phi) (dGS_l - dGL_l)) / v_m
                + (W_l - W_N) g(phi) T + noise_c,l                  (:311),
      noise_c,l = noise_amp_c sqrt(2 * 8.314 T / v_m) xi,  xi ~ N(0, 1) (:309).
    """
    pr = dict(zip(PARAM_NAMES, params))
    v_m, D_L, D_S, noise_amp_c = pr["v_m"], pr["D_L"], pr["D_S"], pr["noise_c"]
    W = c_params[4]
    F = fields.shape[0]
    ncomp = F - 2
    phi = fields[0]
    cs = [fields[l] for l in range(3, F)]
    c_N = np.ones_like(phi)
    for cl in cs:
        c_N = c_N - cl
    arr = [c.copy() for c in cs] + [c_N]
    G_L, dGL = _thermo.fd_derivatives(arr, T, *thermo_l)
    G_S, dGS = _thermo.fd_derivatives(arr, T, *thermo_s)
    g = nc_g(phi)
    h = nc_h(phi)
    transfer = np.zeros((2 * ncomp,) + phi.shape)
    transfer[0] = G_L
    transfer[1] = G_S
    cells = _rng.padded_cells(phi.shape)
    for l in range(ncomp - 1):
        M_c = v_m * cs[l] * (D_L + h * (D_S - D_L)) / (8.314 * T)
        if noise_amp_c != 0:
            xi = _rng.normal_f64(seed, step, _rng.S_NC_C0 + l, phi.size, cells)
            noise_c = noise_amp_c * np.sqrt(2. * 8.314 * T / v_m) * xi
        else:
            noise_c = 0.
        dFdc = (dGL[l] + h * (dGS[l] - dGL[l])) / v_m + (W[l] - W[ncomp - 1]) * g * T + noise_c
        transfer[2 + l] = M_c
        transfer[2 + l + ncomp - 1] = dFdc
    return transfer


def _divagradb(a, axp, axm, ayp, aym, b, bxp, bxm, byp, bym, idx):
    """NCGPU_new.py:28-30."""
    return 0.5 * idx * idx * ((axp + a) * (bxp - b) - (a + axm) * (b - bxm)
                              + (ayp + a) * (byp - b) - (a + aym) * (b - bym))


def _f_ori_term(D_q, D_q_xp, D_q_xm, D_q_yp, D_q_ym, mgq_xp, mgq_xm, mgq_yp, mgq_ym,
                q, q_xp, q_xm, q_yp, q_ym, idx):
    """NCGPU_new.py:33-35."""
    return 0.5 * idx * idx * ((D_q + D_q_xp) * (q_xp - q) / mgq_xp
                              - (D_q + D_q_xm) * (q - q_xm) / mgq_xm
                              + (D_q + D_q_yp) * (q_yp - q) / mgq_yp
                              - (D_q + D_q_ym) * (q - q_ym) / mgq_ym)


class _Nb:
    """Five/nine-point neighbour views of a 2D padded array (interior cells)."""

    def __init__(self, a):
        self.c = a[1:-1, 1:-1]
        self.xp = a[1:-1, 2:]
        self.xm = a[1:-1, :-2]
        self.yp = a[2:, 1:-1]
        self.ym = a[:-2, 1:-1]
        self.pp = a[2:, 2:]      # i+1, j+1
        self.pm = a[2:, :-2]     # i+1, j-1
        self.mp = a[:-2, 2:]     # i-1, j+1
        self.mm = a[:-2, :-2]    # i-1, j-1


def ncomp_anisotropy_terms(phi, q1, q4, T, dx, y_e, ebar2):
    """d_term_dx + d_term_dy of NCGPU_new.py:89-160 on the interior (2D)."""
    P, Q1, Q4, TT = _Nb(phi), _Nb(q1), _Nb(q4), _Nb(T)
    idx = 1. / dx
    dphidx = 0.5 * (P.xp - P.xm) * idx
    dphidx2 = dphidx ** 2
    dphidx3 = dphidx2 * dphidx
    dphidy = 0.5 * (P.yp - P.ym) * idx
    dphidy2 = dphidy ** 2
    dphidy3 = dphidy2 * dphidy
    dTdx = 0.5 * idx * (TT.xp - TT.xm)
    dTdy = 0.5 * idx * (TT.yp - TT.ym)
    d2phidx2 = (P.xp + P.xm - 2. * P.c) * idx * idx
    d2phidy2 = (P.yp + P.ym - 2. * P.c) * idx * idx
    d2phidxy = 0.25 * (P.pp - P.pm - P.mp + P.mm) * idx * idx
    mag_grad_phi2 = dphidx ** 2 + dphidy ** 2
    mag_grad_phi2 = np.where(mag_grad_phi2 < 1e-6, 1e-6, mag_grad_phi2)
    mag_grad_phi4 = mag_grad_phi2 ** 2
    mag_grad_phi8 = mag_grad_phi4 ** 2
    dq1dx = 0.5 * idx * (Q1.xp - Q1.xm)
    dq1dy = 0.5 * idx * (Q1.yp - Q1.ym)
    dq4dx = 0.5 * idx * (Q4.xp - Q4.xm)
    dq4dy = 0.5 * idx * (Q4.yp - Q4.ym)
    q1c, q4c, Tc = Q1.c, Q4.c, TT.c
    q2q2 = (q1c ** 2 - q4c ** 2)
    qq2 = 2. * q1c * q4c
    psix = q2q2 * dphidx - qq2 * dphidy
    psiy = qq2 * dphidx + q2q2 * dphidy
    psix2 = psix ** 2
    psix3 = psix2 * psix
    psix4 = psix2 ** 2
    psiy2 = psiy ** 2
    psiy3 = psiy2 * psiy
    psiy4 = psiy2 ** 2
    dq2q2dx = (2. * q1c * dq1dx - 2. * q4c * dq4dx)
    dqq2dx = (2. * q4c * dq1dx + 2. * q1c * dq4dx)
    dq2q2dy = (2 * q1c * dq1dy - 2. * q4c * dq4dy)
    dqq2dy = (2. * q4c * dq1dy + 2. * q1c * dq4dy)
    dpsixdx = dq2q2dx * dphidx + q2q2 * d2phidx2 - dqq2dx * dphidy - qq2 * d2phidxy
    dpsixdy = dq2q2dy * dphidx + q2q2 * d2phidxy - dqq2dy * dphidy - qq2 * d2phidy2
    dpsiydx = dq2q2dx * dphidy + q2q2 * d2phidxy + dqq2dx * dphidx + qq2 * d2phidx2
    dpsiydy = dq2q2dy * dphidy + q2q2 * d2phidy2 + dqq2dy * dphidx + qq2 * d2phidxy

    d_term_dx = dTdx * ((2. * psix3 * q2q2 + 2. * psiy3 * qq2) / mag_grad_phi2
                        - dphidx * (psix4 + psiy4) / mag_grad_phi4)
    d_term_dx = d_term_dx + Tc * (6. * psix2 * dpsixdx * q2q2 + 2. * psix3 * dq2q2dx
                                  + 6. * psiy2 * dpsiydx * qq2 + 2. * psiy3 * dqq2dx) / mag_grad_phi2
    d_term_dx = d_term_dx - Tc * (2. * psix3 * q2q2 + 2. * psiy3 * qq2) \
        * (2. * dphidx * d2phidx2 + 2. * dphidy * d2phidxy) / mag_grad_phi4
    d_term_dx = d_term_dx - Tc * (d2phidx2 * (psix4 + psiy4)
                                  + dphidx * (4. * psix3 * dpsixdx + 4. * psiy3 * dpsiydx)) / mag_grad_phi4
    d_term_dx = d_term_dx + Tc * (dphidx * (psix4 + psiy4)
                                  * (4. * dphidx3 * d2phidx2 + 4. * dphidx * d2phidx2 * dphidy2
                                     + 4. * dphidy * d2phidxy * dphidx2 + 4. * dphidy3 * d2phidxy)) / mag_grad_phi8
    d_term_dx = d_term_dx * (4. * y_e * ebar2)

    d_term_dy = dTdy * ((-2. * psix3 * qq2 + 2. * psiy3 * q2q2) / mag_grad_phi2
                        - dphidy * (psix4 + psiy4) / mag_grad_phi4)
    d_term_dy = d_term_dy + Tc * (-6. * psix2 * dpsixdy * qq2 - 2. * psix3 * dqq2dy
                                  + 6. * psiy2 * dpsiydy * q2q2 + 2. * psiy3 * dq2q2dy) / mag_grad_phi2
    d_term_dy = d_term_dy - Tc * (-2. * psix3 * qq2 + 2. * psiy3 * q2q2) \
        * (2. * dphidx * d2phidxy + 2. * dphidy * d2phidy2) / mag_grad_phi4
    d_term_dy = d_term_dy - Tc * (d2phidy2 * (psix4 + psiy4)
                                  + dphidy * (4. * psix3 * dpsixdy + 4. * psiy3 * dpsiydy)) / mag_grad_phi4
    d_term_dy = d_term_dy + Tc * (dphidy * (psix4 + psiy4)
                                  * (4. * dphidx3 * d2phidxy + 4. * dphidx * d2phidxy * dphidy2
                                     + 4. * dphidy * d2phidy2 * dphidx2 + 4. * dphidy3 * d2phidy2)) / mag_grad_phi8
    d_term_dy = d_term_dy * (4. * y_e * ebar2)
    return d_term_dx, d_term_dy


def ncomp_main(fields, T, transfer, params, c_params, seed=0, step=0, clamp=(0.0001, 0.9999), ebar2=None,
               eqbar2=None):
    """NComponent_kernel (NCGPU_new.py:37-250) over the interior cells (2D).

    Returns fields_out (padded; interior updated, halos copied from input).
    clamp: the bounds of phi (NCGPU_new.py:234-237).
    """
    dt = dict(zip(PARAM_NAMES, params))["dt"]
    dphidt, dcdt, dq1dt, dq4dt = ncomp_rates(fields, T, transfer, params, c_params, seed, step, ebar2, eqbar2)
    F = fields.shape[0]
    out = fields.copy()
    p = fields[0][1:-1, 1:-1]
    for l in range(3, F):
        out[l][1:-1, 1:-1] = dcdt[l - 3] * dt + fields[l][1:-1, 1:-1]
    phi_new = p + dt * dphidt
    phi_new = np.where(phi_new < clamp[0], clamp[0], phi_new)
    phi_new = np.where(phi_new > clamp[1], clamp[1], phi_new)
    q1c, q4c = fields[1][1:-1, 1:-1], fields[2][1:-1, 1:-1]
    q1n = q1c + dt * dq1dt
    q4n = q4c + dt * dq4dt
    renorm = np.sqrt(q1n ** 2 + q4n ** 2)
    out[0][1:-1, 1:-1] = phi_new
    out[1][1:-1, 1:-1] = q1n / renorm
    out[2][1:-1, 1:-1] = q4n / renorm
    return out


# ---- the legacy engine (class NCGPU, Engines/NCGPU.py) -----------------------
# Its params vector has the first ten entries of PARAM_NAMES (NCGPU.py:
# 486-497); c_params is the same.  Kernels: NCGPU.py:37-242 (main) and
# :248-303 (helper).

LEGACY_PARAM_NAMES = PARAM_NAMES[:10]


def _legacy_full_params(params10):
    """The NCGPU_new-layout vector the shared restatement reads: no phi or
    orientation noise (NCGPU.py:185, 221-224), composition noise amplitude 1
    (:296); ebar / eqbar are not read (the legacy values enter as ebar2,
    eqbar2)."""
    return np.concatenate([np.asarray(params10, dtype=np.float64)[:10], [0., 0., 0., 1., 0.]])


def ncomp_helper_legacy(fields, T, params10, c_params, thermo_l, thermo_s, seed=0, step=0):
    """NComponent_helper_kernel of the legacy engine (NCGPU.py:248-303): as
    ncomp_helper, but one forward difference, along e_1 - e_N (c_1 + h,
    c_N - h, :287-290), serves every component (:303), and one noise draw
    per cell, sqrt(2 * 8.314 T / v_m) xi (:296), is added to every
    component."""
    pr = dict(zip(LEGACY_PARAM_NAMES, params10))
    v_m, D_L, D_S = pr["v_m"], pr["D_L"], pr["D_S"]
    W = c_params[4]
    F = fields.shape[0]
    ncomp = F - 2
    phi = fields[0]
    cs = [fields[l] for l in range(3, F)]
    c_N = np.ones_like(phi)
    for cl in cs:
        c_N = c_N - cl
    arr = [c.copy() for c in cs] + [c_N]
    G_L, dGL = _thermo.fd_derivatives(arr, T, *thermo_l)
    G_S, dGS = _thermo.fd_derivatives(arr, T, *thermo_s)
    g = nc_g(phi)
    h = nc_h(phi)
    cells = _rng.padded_cells(phi.shape)
    xi = _rng.normal_f64(seed, step, _rng.S_NC_C0, phi.size, cells)
    noise_c = np.sqrt(2. * 8.314 * T / v_m) * xi
    transfer = np.zeros((2 * ncomp,) + phi.shape)
    transfer[0] = G_L
    transfer[1] = G_S
    for l in range(ncomp - 1):
        transfer[2 + l] = v_m * cs[l] * (D_L + h * (D_S - D_L)) / (8.314 * T)
        transfer[2 + l + ncomp - 1] = ((dGL[0] + h * (dGS[0] - dGL[0])) / v_m + (W[l] - W[ncomp - 1]) * g * T
                                       + noise_c)
    return transfer


def ncomp_main_legacy(fields, T, transfer, params10, c_params, seed=0, step=0):
    """NComponent_kernel of the legacy engine (NCGPU.py:37-242): the
    NCGPU_new update with ebar^2 = 6 sqrt(2) S_1 d / T_M,1 and
    eqbar^2 = ebar^2 / 4 (:74-75), no phi noise (:185), no orientation noise
    (:221-224) and phi clamped to [1e-6, 0.999999] (:229-232)."""
    pr = dict(zip(LEGACY_PARAM_NAMES, params10))
    S, T_M = c_params[2], c_params[1]
    ebar2 = 6. * np.sqrt(2.) * S[1] * pr["d"] / T_M[1]
    return ncomp_main(fields, T, transfer, _legacy_full_params(params10), c_params, seed, step,
                      clamp=(0.000001, 0.999999), ebar2=ebar2, eqbar2=0.25 * ebar2)


def ncomp_rates(fields, T, transfer, params, c_params, seed=0, step=0, ebar2=None, eqbar2=None):
    """(dphi/dt, [dc_l/dt], dq1/dt, dq4/dt) on the interior (NCGPU_new.py:84-232).
    ebar2 / eqbar2 override ebar^2 / eqbar^2 (the legacy engine's values)."""
    pr = dict(zip(PARAM_NAMES, params))
    dx, v_m, M_qmax, H, y_e, beta = pr["dx"], pr["v_m"], pr["M_qmax"], pr["H"], pr["y_e"], pr["beta"]
    dt, ebar, eqbar = pr["dt"], pr["ebar"], pr["eqbar"]
    noise_amp_phi, noise_amp_q = pr["noise_phi"], pr["noise_q"]
    W = c_params[4]
    Mo = c_params[5]
    F = fields.shape[0]
    phi, q1, q4 = fields[0], fields[1], fields[2]
    P, Q1, Q4, TT = _Nb(phi), _Nb(q1), _Nb(q4), _Nb(T)
    ebar2 = ebar * ebar if ebar2 is None else ebar2
    eqbar2 = eqbar * eqbar if eqbar2 is None else eqbar2
    idx = 1. / dx

    p = P.c
    g = (p ** 2) * (1 - p) ** 2
    h = (p * p * p) * (6. * p ** 2 - 15. * p + 10.)  # numba int power: products
    hprime = 30. * g
    gprime = 4. * (p * p * p) - 6. * p ** 2 + 2 * p
    dphidx = 0.5 * (P.xp - P.xm) * idx
    dphidy = 0.5 * (P.yp - P.ym) * idx
    dTdx = 0.5 * idx * (TT.xp - TT.xm)
    dTdy = 0.5 * idx * (TT.yp - TT.ym)
    d2phidx2 = (P.xp + P.xm - 2. * p) * idx * idx
    d2phidy2 = (P.yp + P.ym - 2. * p) * idx * idx
    lphi = d2phidx2 + d2phidy2

    mgq_xp = idx * np.sqrt((Q1.xp - Q1.c) ** 2 + (Q4.xp - Q4.c) ** 2)
    mgq_xm = idx * np.sqrt((Q1.xm - Q1.c) ** 2 + (Q4.xm - Q4.c) ** 2)
    mgq_yp = idx * np.sqrt((Q1.yp - Q1.c) ** 2 + (Q4.yp - Q4.c) ** 2)
    mgq_ym = idx * np.sqrt((Q1.ym - Q1.c) ** 2 + (Q4.ym - Q4.c) ** 2)
    mag_grad_q = 0.5 * (mgq_xp + mgq_xm + mgq_yp + mgq_ym)
    mgq_xp = np.where(mgq_xp < beta, beta, mgq_xp)
    mgq_xm = np.where(mgq_xm < beta, beta, mgq_xm)
    mgq_yp = np.where(mgq_yp < beta, beta, mgq_yp)
    mgq_ym = np.where(mgq_ym < beta, beta, mgq_ym)

    d_term_dx, d_term_dy = ncomp_anisotropy_terms(phi, q1, q4, T, dx, y_e, ebar2)

    cW = np.zeros_like(p)
    c_N = np.ones_like(p)
    M_phi = np.zeros_like(p)
    for l in range(3, F):
        c_i = fields[l][1:-1, 1:-1]
        cW = cW + c_i * W[l - 3]
        c_N = c_N - c_i
        M_phi = M_phi + c_i * Mo[l - 3]
    cW = cW + c_N * W[F - 3]
    M_phi = M_phi + c_N * Mo[F - 3]

    M_q = M_qmax + (1e-6 - M_qmax) * h
    Tc = TT.c
    G_L = transfer[0][1:-1, 1:-1]
    G_S = transfer[1][1:-1, 1:-1]
    dphidt = ebar2 * (1. - 3. * y_e) * (Tc * lphi + dTdx * dphidx + dTdy * dphidy)
    dphidt = dphidt + d_term_dx + d_term_dy
    dphidt = dphidt - hprime * (G_S - G_L) / v_m
    dphidt = dphidt - gprime * Tc * cW
    dphidt = dphidt - 4. * H * Tc * p * mag_grad_q
    dphidt = dphidt * M_phi

    cells = _rng.padded_cells(phi.shape)[1:-1, 1:-1]
    if noise_amp_phi != 0:
        xi = _rng.normal_f64(seed, step, _rng.S_NC_PHI, phi.size, cells)
        dphidt = dphidt + np.sqrt(2. * 8.314 * Tc * M_phi / v_m) * xi * noise_amp_phi

    dcdt = []
    for l in range(3, F):
        M_c = _Nb(transfer[l - 1])
        dF = _Nb(transfer[l - 1 + F - 3])
        acc = _divagradb(M_c.c, M_c.xp, M_c.xm, M_c.yp, M_c.ym,
                         dF.c, dF.xp, dF.xm, dF.yp, dF.ym, idx)
        for m in range(3, F):
            cj = _Nb(fields[m])
            dFj = _Nb(transfer[m - 1 + F - 3])
            acc = acc - _divagradb(M_c.c * cj.c, M_c.xp * cj.xp, M_c.xm * cj.xm,
                                   M_c.yp * cj.yp, M_c.ym * cj.ym,
                                   dFj.c, dFj.xp, dFj.xm, dFj.yp, dFj.ym, idx)
        dcdt.append(acc)

    D_q = 2. * H * Tc * (p ** 2)
    D_q_xp = 2. * H * TT.xp * (P.xp ** 2)
    D_q_xm = 2. * H * TT.xm * (P.xm ** 2)
    D_q_yp = 2. * H * TT.yp * (P.yp ** 2)
    D_q_ym = 2. * H * TT.ym * (P.ym ** 2)
    f_ori_1 = _f_ori_term(D_q, D_q_xp, D_q_xm, D_q_yp, D_q_ym, mgq_xp, mgq_xm, mgq_yp, mgq_ym,
                          Q1.c, Q1.xp, Q1.xm, Q1.yp, Q1.ym, idx)
    f_ori_4 = _f_ori_term(D_q, D_q_xp, D_q_xm, D_q_yp, D_q_ym, mgq_xp, mgq_xm, mgq_yp, mgq_ym,
                          Q4.c, Q4.xp, Q4.xm, Q4.yp, Q4.ym, idx)
    lq1 = (Q1.xp + Q1.xm + Q1.yp + Q1.ym - 4 * Q1.c) * idx * idx
    lq4 = (Q4.xp + Q4.xm + Q4.yp + Q4.ym - 4 * Q4.c) * idx * idx
    q_noise_coeff = 0.0000000001
    if noise_amp_q != 0:
        xi1 = _rng.normal_f64(seed, step, _rng.S_NC_Q1, phi.size, cells)
        xi4 = _rng.normal_f64(seed, step, _rng.S_NC_Q4, phi.size, cells)
        noise_q1 = noise_amp_q * np.sqrt(q_noise_coeff * 8.314 * Tc / v_m) * xi1
        noise_q4 = noise_amp_q * np.sqrt(q_noise_coeff * 8.314 * Tc / v_m) * xi4
    else:
        noise_q1 = noise_q4 = 0.
    q1c, q4c = Q1.c, Q4.c
    t1 = f_ori_1 + lq1 * eqbar2 + noise_q1
    t4 = f_ori_4 + lq4 * eqbar2 + noise_q4
    dq1dt = M_q * ((1 - q1c ** 2) * t1 - q1c * q4c * t4)
    dq4dt = M_q * ((1 - q4c ** 2) * t4 - q1c * q4c * t1)
    return dphidt, dcdt, dq1dt, dq4dt


# ---------------------------------------------------------------------------
# Simulation-level loop  (simulation.py:245-271 with the D5 resolution)
# ---------------------------------------------------------------------------

def temperature_linear(shape, T0, dTdx, dTdy, dx, dTdt, dt, step, dTdz=0.0):
    """Interior temperature of a LINEAR_GRADIENT run after ``step`` steps
    (simulation.py:134-156 with counter*dTdt*dt; ppf_gpu_utils.py:332-355),
    for 2D shapes (ny, nx) and 3D shapes (nz, ny, nx).

    T = (((T0 + i*(dx*dTdx)) + j*(dx*dTdy)) + k*(dx*dTdz)) + step*(dTdt*dt),
    i along x, j along y, k along z.
    """
    T = np.zeros(shape) + T0
    nd = len(shape)
    if dTdx:
        T = T + np.arange(shape[-1], dtype=np.float64) * (dx * dTdx)
    if nd > 1 and dTdy:
        T = T + (np.arange(shape[-2], dtype=np.float64) * (dx * dTdy))[:, None]
    if nd > 2 and dTdz:
        T = T + (np.arange(shape[-3], dtype=np.float64) * (dx * dTdz))[:, None, None]
    if dTdt:
        T = T + step * (dTdt * dt)
    return T


def temperature_history(times, frames, time, form="kernel"):
    """Interior temperature of an XDMF_FILE run at ``time``.

    Window (simulation.py:169-174, 319-326; ppf_gpu_utils.py:534-543):
    starting from frames (0, 1) it moves one frame at a time while time > the
    later frame's time.  Interpolation between T_a (t_a) and T_b (t_b):

    * form="kernel" -- update_thermal_file_*_kernel (ppf_gpu_utils.py:357-394),
      the temperature a GPU run steps with:
      r1 = (time - t_a)/(t_b - t_a), r0 = 1 - r1, T = T_a*r0 + T_b*r1;
    * form="host" -- simulation.py:175,329 (CPU runs and the initial field):
      T = T_a*(t_b - time)/(t_b - t_a) + T_b*(time - t_a)/(t_b - t_a).

    The two agree to rounding (about one ulp of T)."""
    a, b = 0, 1
    while time > times[b]:
        if b + 1 >= len(times):
            raise ValueError("thermal history ends before t=%g" % time)
        a, b = b, b + 1
    ta, tb = times[a], times[b]
    if form == "kernel":
        r1 = (time - ta) / (tb - ta)
        r0 = 1 - r1
        return frames[a] * r0 + frames[b] * r1
    if form == "host":
        return frames[a] * (tb - time) / (tb - ta) + frames[b] * (time - ta) / (tb - ta)
    raise ValueError("form must be 'kernel' or 'host'")
