"""AnisoDorrGPU engine, FP64 numpy restatement (oracle).

Test infrastructure only: the product path never imports this module.

Follows Engines/AnisoDorrGPU.py: AnisoDorr_helper_kernel (:262-375) and
AnisoDorr_kernel (:54-245), with the device helpers of :25-52.  Binary
(Ni-Cu) anisotropic polycrystalline solidification with fields
(phi, q1, q4, c) on a 2D halo-padded grid.  The reference hard-codes every
material constant inside both kernels; here they come from ``params`` (the
reference values are the defaults of the engine).  Resolutions of defects:

  * noise: the counter RNG (uniform in [0, 1), stream ``S_ANISO_PHI``) in
    place of xoroshiro128p states indexed by the thread id;
  * the vertex terms at padded row 0 / column 0 read index -1 in the
    reference (numba wrap-around) and are never used; they are set to zero;
  * ``simulation_loop`` (:592-619) never swaps the field buffers, so the
    reference class does not advance; the step here returns the new state
    (as ``engine_AnisoDorrGPU``, :394-405, does with its swap).

Operation order follows the reference expressions (Python evaluates left
to right), which fixes FP64 parity.
"""

import numpy as np

from . import rng as _rng

S_ANISO_PHI = 8

PARAM_NAMES = ["T", "M_qmax", "H", "T_mA", "T_mB", "L_A", "L_B", "D_L", "D_S", "v_m", "R", "y_e", "beta",
               "alpha", "ebar", "eqbar", "W_A", "W_B", "M_A", "M_B", "dt"]


def derived_params(dx, d_ratio=1.0 / 0.94, T=1574., M_qmax=80000000., H=1e-11, T_mA=1728., T_mB=1358.,
                   L_A=2350., L_B=1728., s_A=0.000037, s_B=0.000029, D_L=1e-5, D_S=1e-9, B_A=0.33, B_B=0.39,
                   v_m=7.42, R=8.314, y_e=0.12, beta=1.5, alpha=0.3):
    """The constants of AnisoDorrGPU.py:63-105 (d = dx*d_ratio, dt = dx^2/5/D_L/8)."""
    d = dx * d_ratio
    dt = dx * dx / 5. / D_L / 8.
    ebar = np.sqrt(6. * np.sqrt(2.) * s_A * d / T_mA)
    return dict(T=T, M_qmax=M_qmax, H=H, T_mA=T_mA, T_mB=T_mB, L_A=L_A, L_B=L_B, D_L=D_L, D_S=D_S, v_m=v_m, R=R,
                y_e=y_e, beta=beta, alpha=alpha, ebar=ebar, eqbar=0.5 * ebar,
                W_A=3. * s_A / (np.sqrt(2.) * T_mA * d), W_B=3. * s_B / (np.sqrt(2.) * T_mB * d),
                M_A=T_mA * T_mA * B_A / (6. * np.sqrt(2.) * L_A * d),
                M_B=T_mB * T_mB * B_B / (6. * np.sqrt(2.) * L_B * d), dt=dt)


def _g(p):
    return p * p * (1 - p) * (1 - p)


def _gprime(p):
    return 4 * p * p * p - 6 * p * p + 2 * p


def _h(p):
    return p * p * p * (10 - 15 * p + 6 * p * p)


def helper(fields, p, dx):
    """transfer[0..5] = pf_comp_xmm, pf_comp_ymm, t1_temp, t4_temp, temp, D_C
    for every padded cell (vertex terms only where i, j >= 1)."""
    phi, q1, q4, c = fields
    idx = 1. / dx
    T = p["T"]
    tr = np.zeros((6,) + phi.shape)
    g = _g(phi)
    gprime = _gprime(phi)
    m = 1 - _h(phi)
    H_A = p["W_A"] * gprime - 30 * p["L_A"] * (1 / T - 1 / p["T_mA"]) * g
    H_B = p["W_B"] * gprime - 30 * p["L_B"] * (1 / T - 1 / p["T_mB"]) * g
    D_C = p["D_S"] + m * (p["D_L"] - p["D_S"])
    tr[5] = D_C
    tr[4] = D_C * p["v_m"] * c * (1 - c) * (H_B - H_A) / p["R"]
    # vertex (i-1/2, j-1/2): cells (i,j), (i,j-1), (i-1,j), (i-1,j-1)
    P, Pl, Pu, Pul = phi[1:, 1:], phi[1:, :-1], phi[:-1, 1:], phi[:-1, :-1]
    gphi_x = 0.5 * (P - Pl + Pu - Pul) * idx
    gphi_y = 0.5 * (P - Pu + Pl - Pul) * idx
    q1m = 0.25 * (q1[1:, 1:] + q1[:-1, 1:] + q1[1:, :-1] + q1[:-1, :-1])
    q4m = 0.25 * (q4[1:, 1:] + q4[:-1, 1:] + q4[1:, :-1] + q4[:-1, :-1])
    a2_b2 = q1m * q1m - q4m * q4m
    ab2 = 2. * q1m * q4m
    gpsi_x = a2_b2 * gphi_x - ab2 * gphi_y
    gpsi_y = a2_b2 * gphi_y + ab2 * gphi_x
    psi_xxy = gpsi_x * gpsi_x * gpsi_y
    psi_xyy = gpsi_x * gpsi_y * gpsi_y
    mg2 = gpsi_x * gpsi_x + gpsi_y * gpsi_y
    mg2 = np.maximum(mg2, 0.000000001)
    psix3 = gpsi_x * gpsi_x * gpsi_x
    psiy3 = gpsi_y * gpsi_y * gpsi_y
    y_e, ebar = p["y_e"], p["ebar"]
    tr[0][1:, 1:] = 4 * y_e * ((2 * a2_b2 * psix3 + 2 * ab2 * psiy3) / mg2
                               - gphi_x * (psix3 * gphi_x + psiy3 * gphi_y) / (mg2 * mg2))
    tr[1][1:, 1:] = 4 * y_e * ((2 * a2_b2 * psiy3 - 2 * ab2 * psix3) / mg2
                               - gphi_y * (psix3 * gphi_x + psiy3 * gphi_y) / (mg2 * mg2))
    q1px, q1py = q1m * gphi_x, q1m * gphi_y
    q4px, q4py = q4m * gphi_x, q4m * gphi_y
    k = 16 * ebar * ebar * y_e / mg2
    tr[2][1:, 1:] = k * (psi_xyy * (q1px - q4py) + psi_xxy * (q1py + q4px))
    tr[3][1:, 1:] = k * (psi_xyy * (-q4px - q1py) + psi_xxy * (-q4py + q1px))
    return tr


def _grad2(a, idx):
    return (a[1:-1, 2:] + a[1:-1, :-2] + a[2:, 1:-1] + a[:-2, 1:-1] - 4 * a[1:-1, 1:-1]) * (idx * idx)


def _divagradb(a, b, idx):
    """AnisoDorrGPU.py:25-27 on the interior."""
    ac, axp, axm, ayp, aym = a[1:-1, 1:-1], a[1:-1, 2:], a[1:-1, :-2], a[2:, 1:-1], a[:-2, 1:-1]
    bc, bxp, bxm, byp, bym = b[1:-1, 1:-1], b[1:-1, 2:], b[1:-1, :-2], b[2:, 1:-1], b[:-2, 1:-1]
    return 0.5 * idx * idx * ((axp + ac) * (bxp - bc) - (ac + axm) * (bc - bxm)
                              + (ayp + ac) * (byp - bc) - (ac + aym) * (bc - bym))


def _f_ori(Dq, Dxp, Dxm, Dyp, Dym, mxp, mxm, myp, mym, q, idx):
    """AnisoDorrGPU.py:30-31 on the interior."""
    qc, qxp, qxm, qyp, qym = q[1:-1, 1:-1], q[1:-1, 2:], q[1:-1, :-2], q[2:, 1:-1], q[:-2, 1:-1]
    return 0.5 * idx * idx * ((Dq + Dxp) * (qxp - qc) / mxp - (Dq + Dxm) * (qc - qxm) / mxm
                              + (Dq + Dyp) * (qyp - qc) / myp - (Dq + Dym) * (qc - qym) / mym)


def step(fields, tr, p, dx, seed=0, step=0):
    """AnisoDorr_kernel: new padded fields (interior updated, halos copied)."""
    phi, q1, q4, c = fields
    idx = 1. / dx
    T, H, y_e, ebar, eqbar = p["T"], p["H"], p["y_e"], p["ebar"], p["eqbar"]
    P = phi[1:-1, 1:-1]
    Q1, Q4, C = q1[1:-1, 1:-1], q4[1:-1, 1:-1], c[1:-1, 1:-1]
    g = _g(P)
    h = _h(P)
    m = 1 - h
    M_q = 1e-6 + (p["M_qmax"] - 1e-6) * m
    lq1 = _grad2(q1, idx)
    lq4 = _grad2(q4, idx)
    pp = 2 * P
    gprime = _gprime(P)
    H_A = p["W_A"] * gprime - 30 * p["L_A"] * (1 / T - 1 / p["T_mA"]) * g
    H_B = p["W_B"] * gprime - 30 * p["L_B"] * (1 / T - 1 / p["T_mB"]) * g

    def dm(a, di, dj):  # a[i][j] - a[i+di][j+dj] on the interior
        ny, nx = a.shape[0] - 2, a.shape[1] - 2
        return a[1:-1, 1:-1] - a[1 + di:1 + di + ny, 1 + dj:1 + dj + nx]

    gq1 = {k: dm(q1, *o) * idx for k, o in (("xm", (0, -1)), ("ym", (-1, 0)), ("xp", (0, 1)), ("yp", (1, 0)))}
    gq4 = {k: dm(q4, *o) * idx for k, o in (("xm", (0, -1)), ("ym", (-1, 0)), ("xp", (0, 1)), ("yp", (1, 0)))}
    gqs = {k: gq1[k] * gq1[k] + gq4[k] * gq4[k] for k in gq1}
    rgqs_0 = np.sqrt(0.5 * (gqs["xm"] + gqs["ym"] + gqs["xp"] + gqs["yp"]))
    beta = p["beta"]
    rg = {k: np.sqrt(np.maximum(v, beta)) for k, v in gqs.items()}

    deltac = _divagradb(tr[4], phi, idx)
    deltac = deltac + _divagradb(tr[5], c, idx)

    lphi = _grad2(phi, idx)
    xmm, ymm = tr[0], tr[1]
    pf_x = 0.5 * idx * (xmm[1:-1, 2:] + xmm[2:, 2:] - xmm[1:-1, 1:-1] - xmm[2:, 1:-1])
    pf_y = 0.5 * idx * (ymm[2:, 1:-1] + ymm[2:, 2:] - ymm[1:-1, 1:-1] - ymm[1:-1, 2:])
    M_phi = (1 - C) * p["M_A"] + C * p["M_B"]
    deltaphi = M_phi * (ebar * ebar * ((1 - 3 * y_e) * lphi + pf_x + pf_y) - (1 - C) * H_A - C * H_B
                        - 2 * H * T * pp * rgqs_0)
    cells = _rng.padded_cells(phi.shape)[1:-1, 1:-1]
    rand = _rng.uniform_f64(seed, step, S_ANISO_PHI, phi.size, cells)
    deltaphi = deltaphi + M_phi * p["alpha"] * rand * (16 * g) * ((1 - C) * H_A + C * H_B)

    base = 2 * H * T
    Dq = base * (P * P)
    Dxp = base * phi[1:-1, 2:] * phi[1:-1, 2:]
    Dxm = base * phi[1:-1, :-2] * phi[1:-1, :-2]
    Dyp = base * phi[2:, 1:-1] * phi[2:, 1:-1]
    Dym = base * phi[:-2, 1:-1] * phi[:-2, 1:-1]
    gaq1 = _f_ori(Dq, Dxp, Dxm, Dyp, Dym, rg["xp"], rg["xm"], rg["yp"], rg["ym"], q1, idx)
    gaq4 = _f_ori(Dq, Dxp, Dxm, Dyp, Dym, rg["xp"], rg["xm"], rg["yp"], rg["ym"], q4, idx)
    t1t, t4t = tr[2], tr[3]
    cc1 = 0.25 * (t1t[1:-1, 1:-1] + t1t[2:, 1:-1] + t1t[1:-1, 2:] + t1t[2:, 2:])
    cc4 = 0.25 * (t4t[1:-1, 1:-1] + t4t[2:, 1:-1] + t4t[1:-1, 2:] + t4t[2:, 2:])
    t1 = eqbar * eqbar * lq1 + gaq1 + cc1
    t4 = eqbar * eqbar * lq4 + gaq4 + cc4
    lmbda = Q1 * t1 + Q4 * t4
    dq1 = M_q * (t1 - Q1 * lmbda)
    dq4 = M_q * (t4 - Q4 * lmbda)

    dt = p["dt"]
    out = fields.copy()
    out[0][1:-1, 1:-1] = P + deltaphi * dt
    q1n = Q1 + dq1 * dt
    q4n = Q4 + dq4 * dt
    out[3][1:-1, 1:-1] = C + deltac * dt
    renorm = np.sqrt(q1n * q1n + q4n * q4n)
    out[1][1:-1, 1:-1] = q1n / renorm
    out[2][1:-1, 1:-1] = q4n / renorm
    return out
