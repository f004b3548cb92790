"""Pin the oracle's engine steps with closed forms and continuum limits.

Each check is chosen so that a plausible transcription mistake (a dropped
term, a wrong sign, a swapped axis, a wrong coefficient) fails it:
  * Diffusion: exact discrete Fourier eigenvalue, mass conservation.
  * Allen-Cahn: double-well fixed points, exact linear growth factors about
    them, O(dx^2) stationarity of the closed-form tanh kink.
  * Cahn-Hilliard: exact mass conservation (periodic and zero-flux), discrete
    dispersion relation, free-energy decrease.
  * Warren1995 / N-component: the discrete rates converge at O(dx^2) to the
    continuum operators built by symbolic differentiation of divergence forms
    (Warren-Boettinger anisotropy; the quaternion-rotated 4-fold anisotropy as
    the divergence of the gradient-energy derivative; div(a grad b) fluxes).
"""

import numpy as np
import pytest
import sympy as sp

from oracle import bc as obc
from oracle import engines as oe
from oracle import thermo as oth


def _mode(shape, ks):
    grids = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    out = np.ones(shape)
    for g, k, n in zip(grids, ks, shape):
        out = out * np.cos(2 * np.pi * k * g / n + 0.3)
    return out


def _lam(shape, ks, dx):
    """discrete -laplacian eigenvalue of the mode"""
    return sum(4.0 / dx ** 2 * np.sin(np.pi * k / n) ** 2 for k, n in zip(ks, shape))


def _step_point(kind, interior, bc="PERIODIC", dx=1.0, dt=0.1, **p):
    a = obc.pad(interior)
    obc.apply_bc(a[None], bc, None, dx)
    if kind == "diffusion":
        return oe.diffusion_step(a, p["D"], dt, dx)[tuple(slice(1, -1) for _ in interior.shape)]
    if kind == "allencahn":
        return oe.allencahn_step(a, p["M"], p["W"], p["epsilon"], dt, dx)[tuple(slice(1, -1) for _ in interior.shape)]
    return oe.cahnhilliard_step(a, p["M"], p["epsilon"], dt, dx, bc)[tuple(slice(1, -1) for _ in interior.shape)]


# ------------------------------------------------------------- Diffusion --
@pytest.mark.parametrize("shape,ks", [((32, 24), (3, 5)), ((40,), (7,)), ((8, 10, 12), (1, 2, 3))])
def test_diffusion_fourier_eigenvalue(shape, ks):
    dx, dt, D = 0.5, 0.01, 1.3
    u = _mode(shape, ks)
    got = _step_point("diffusion", u, "PERIODIC", dx, dt, D=D)
    lam = 1 - dt * D * _lam(shape, ks, dx)
    np.testing.assert_allclose(got, lam * u, atol=1e-13)


@pytest.mark.parametrize("bc", ["PERIODIC", "NEUMANN"])
def test_diffusion_mass_conservation(bc):
    c = np.random.default_rng(0).random((20, 30))
    out = _step_point("diffusion", c, bc, 1.0, 0.2, D=1.0)
    assert abs(out.sum() - c.sum()) < 1e-11


# ------------------------------------------------------------- Allen-Cahn --
def test_allencahn_fixed_points_and_linear_growth():
    shape, ks = (24, 20), (2, 3)
    M, W, eps, dt = 0.7, 0.3, 1.2, 0.01
    for v in (0.0, 0.5, 1.0):
        out = _step_point("allencahn", np.full(shape, v), M=M, W=W, epsilon=eps, dt=dt)
        np.testing.assert_allclose(out, v, atol=1e-15)
    # linearisation: f'(phi) = 4phi^3-6phi^2+2phi has slope 2 at 0 and 1, -1 at 1/2
    lam = _lam(shape, ks, 1.0)
    for base, slope in ((0.0, 2.0), (0.5, -1.0), (1.0, 2.0)):
        d = 1e-7
        u = _mode(shape, ks)
        out = _step_point("allencahn", base + d * u, M=M, W=W, epsilon=eps, dt=dt)
        factor = 1 + dt * M * (-eps ** 2 * lam - 16 * W * slope)
        np.testing.assert_allclose((out - base) / d, factor * u, atol=1e-6)


def test_allencahn_tanh_kink_is_stationary_to_second_order():
    """eps^2 phi'' = 16 W f'(phi) with f = phi^2(1-phi)^2 has the kink
    phi = (1 + tanh(sqrt(8W) x / eps))/2; the discrete rate is O(dx^2)."""
    W, eps, M, dt = 0.5, 4.0, 1.0, 1e-3
    errs = []
    for n in (128, 256, 512):
        L = 48.0
        dx = L / n
        x = (np.arange(n) + 0.5) * dx - L / 2
        phi = 0.5 * (1 + np.tanh(np.sqrt(8 * W) * x / eps))
        out = _step_point("allencahn", np.tile(phi, (3, 1)), "NEUMANN", dx, dt, M=M, W=W, epsilon=eps)
        rate = (out[1] - phi) / (dt * M)
        errs.append(np.max(np.abs(rate[5:-5])))
    assert errs[0] / errs[1] > 3.5 and errs[1] / errs[2] > 3.5
    assert errs[-1] < 3e-3  # terms themselves are O(1)


# ---------------------------------------------------------- Cahn-Hilliard --
@pytest.mark.parametrize("bc,shape", [("PERIODIC", (30, 26)), ("NEUMANN", (30, 26)),
                                      (["PERIODIC", "NEUMANN"], (22, 18)), ("PERIODIC", (8, 9, 10)),
                                      ("NEUMANN", (8, 9, 10))])
def test_cahnhilliard_mass_conservation(bc, shape):
    c = 0.5 + 0.1 * np.random.default_rng(1).standard_normal(shape)
    out = _step_point("cahnhilliard", c, bc, 1.0, 0.01, M=1.0, epsilon=1.0)
    assert abs(out.sum() - c.sum()) < 1e-11 * c.size
    assert np.max(np.abs(out - c)) > 1e-5  # it did evolve


def test_cahnhilliard_dispersion_relation():
    shape, ks = (32, 32), (3, 2)
    dt, M, eps = 0.01, 1.0, 1.0
    d = 1e-8
    u = _mode(shape, ks)
    out = _step_point("cahnhilliard", 0.5 + d * u, M=M, epsilon=eps, dt=dt)
    lam = _lam(shape, ks, 1.0)
    fpp = -1.0  # f''(1/2) for f = c^2(1-c)^2
    factor = 1 - dt * M * lam * (fpp + eps ** 2 * lam)
    np.testing.assert_allclose((out - 0.5) / d, factor * u, atol=1e-5)


def test_cahnhilliard_free_energy_decreases():
    c = 0.5 + 0.02 * (2 * np.random.default_rng(2).random((32, 32)) - 1)
    a = obc.pad(c)
    obc.apply_bc(a[None], "PERIODIC")
    F = [oe.cahnhilliard_free_energy(a, 1.0, 1.0)]
    for _ in range(300):
        a = oe.cahnhilliard_step(a, 1.0, 1.0, 0.01, 1.0, "PERIODIC")
        obc.apply_bc(a[None], "PERIODIC")
        F.append(oe.cahnhilliard_free_energy(a, 1.0, 1.0))
    assert all(F[i + 1] <= F[i] + 1e-12 for i in range(len(F) - 1))
    assert F[-1] < F[0]


# -------------------------------------------------- continuum-limit helpers --
X, Y = sp.symbols("x y", real=True)  # x = array axis 1 (columns), y = axis 0 (rows)


def _grid(n, L=1.0):
    h = L / n
    coords = (np.arange(-1, n + 1) + 0.5) * h  # padded cell centres
    yy, xx = np.meshgrid(coords, coords, indexing="ij")
    return h, xx, yy


def _ev(expr, xx, yy):
    f = sp.lambdify((X, Y), expr, "numpy")
    return np.broadcast_to(np.asarray(f(xx, yy), dtype=np.float64), xx.shape).copy()


def _order(errs):
    return [errs[i] / errs[i + 1] for i in range(len(errs) - 1)]


# --------------------------------------------------------------- Warren --
def test_warren_rates_converge_to_divergence_form():
    """dc/dt -> div(D_C grad c + temp grad phi); dphi/dt ->
    M_phi[div(eb^2 eta^2 grad phi) - d_r(eb^2 eta eta' d_k phi) + d_k(eb^2 eta eta' d_r phi)
    - (1-c)H_A - cH_B] with r = array axis 0, k = axis 1, theta = atan2(d_k phi, d_r phi)."""
    p = dict(T=1574., T_mA=1728., T_mB=1358., L_A=2.35, L_B=1.728, D_L=1.0, D_S=0.1, v_m=1.0, R=8.314,
             y_e=0.04, alpha=0.0, ebar=0.05, W_A=0.7, W_B=0.5, M_A=1.0, M_B=2.0)
    phi_e = 0.5 + 0.3 * sp.sin(2 * sp.pi * X + 0.4) * sp.cos(2 * sp.pi * Y - 0.2) + 0.1 * sp.sin(2 * sp.pi * Y)
    c_e = 0.4 + 0.1 * sp.cos(2 * sp.pi * (X + 2 * Y)) + 0.05 * sp.sin(2 * sp.pi * X)
    dr = lambda e: sp.diff(e, Y)  # axis 0
    dk = lambda e: sp.diff(e, X)  # axis 1
    pp = phi_e ** 3 * (10 - 15 * phi_e + 6 * phi_e ** 2)
    g = phi_e ** 2 * (1 - phi_e) ** 2
    gp = 4 * phi_e ** 3 - 6 * phi_e ** 2 + 2 * phi_e
    H_A = p["W_A"] * gp + 30 * p["L_A"] * (1 / p["T"] - 1 / p["T_mA"]) * g
    H_B = p["W_B"] * gp + 30 * p["L_B"] * (1 / p["T"] - 1 / p["T_mB"]) * g
    D_C = p["D_S"] + pp * (p["D_L"] - p["D_S"])
    temp = D_C * p["v_m"] * c_e * (1 - c_e) * (H_B - H_A) / p["R"]
    dc_exact = dr(D_C * dr(c_e) + temp * dr(phi_e)) + dk(D_C * dk(c_e) + temp * dk(phi_e))
    th = sp.atan2(dk(phi_e), dr(phi_e))
    eta = 1 + p["y_e"] * sp.cos(4 * th)
    etap = -4 * p["y_e"] * sp.sin(4 * th)
    eb2 = p["ebar"] ** 2
    M_phi = (1 - c_e) * p["M_A"] + c_e * p["M_B"]
    dphi_exact = M_phi * (dr(eb2 * eta ** 2 * dr(phi_e)) + dk(eb2 * eta ** 2 * dk(phi_e))
                          - dr(eb2 * eta * etap * dk(phi_e)) + dk(eb2 * eta * etap * dr(phi_e))
                          - (1 - c_e) * H_A - c_e * H_B)
    ec, ep = [], []
    for n in (32, 64, 128):
        h, xx, yy = _grid(n)
        phi = _ev(phi_e, xx, yy)
        c = _ev(c_e, xx, yy)
        dphi, dc = oe.warren_rates(phi, c, p, h)
        sl = (slice(1, -1), slice(1, -1))
        ec.append(np.max(np.abs(dc[sl] - _ev(dc_exact, xx, yy)[sl])))
        ep.append(np.max(np.abs(dphi[sl] - _ev(dphi_exact, xx, yy)[sl])))
    assert min(_order(ec)) > 3.5, ec
    assert min(_order(ep)) > 3.5, ep


def test_warren_noise_is_localised_uniform():
    """alpha term = M_phi alpha r 16 g ((1-c)H_A + cH_B), r in [-1, 1)."""
    p = dict(T=1574., T_mA=1728., T_mB=1358., L_A=2.35, L_B=1.728, D_L=1.0, D_S=0.1, v_m=1.0, R=8.314,
             y_e=0.04, alpha=0.3, ebar=0.05, W_A=0.7, W_B=0.5, M_A=1.0, M_B=2.0)
    rng = np.random.default_rng(3)
    phi = rng.random((12, 14))
    c = 0.4 + 0.1 * rng.random((12, 14))
    a, _ = oe.warren_rates(phi, c, p, 0.1, seed=4, step=9)
    b, _ = oe.warren_rates(phi, c, dict(p, alpha=0.0), 0.1, seed=4, step=9)
    g = phi ** 2 * (1 - phi) ** 2
    gp = 4 * phi ** 3 - 6 * phi ** 2 + 2 * phi
    H_A = p["W_A"] * gp + 30 * p["L_A"] * (1 / p["T"] - 1 / p["T_mA"]) * g
    H_B = p["W_B"] * gp + 30 * p["L_B"] * (1 / p["T"] - 1 / p["T_mB"]) * g
    M_phi = (1 - c) * p["M_A"] + c * p["M_B"]
    r = (a - b) / (M_phi * p["alpha"] * 16 * g * ((1 - c) * H_A + c * H_B))
    assert np.all(r[1:-1, 1:-1] >= -1 - 1e-9) and np.all(r[1:-1, 1:-1] < 1 + 1e-9)
    assert r[1:-1, 1:-1].std() > 0.3


# ---------------------------------------------------------- N-component --
def _nc_setup(n, beta, y_e=0.12):
    q_ang = 0.4 * sp.sin(2 * sp.pi * X) * sp.cos(2 * sp.pi * Y) + 0.2
    q1_e, q4_e = sp.cos(q_ang), sp.sin(q_ang)
    phi_e = 0.5 + 0.25 * sp.sin(2 * sp.pi * (X + 0.3 * Y) + 0.1)
    T_e = 2.0 + 0.3 * sp.cos(2 * sp.pi * X) + 0.2 * sp.sin(2 * sp.pi * Y)
    c1_e = 0.3 + 0.05 * sp.sin(2 * sp.pi * Y)
    c2_e = 0.3 + 0.05 * sp.cos(2 * sp.pi * (X - Y))
    dG_e = 0.2 * sp.sin(2 * sp.pi * X)
    M1_e = 1.0 + 0.2 * sp.cos(2 * sp.pi * Y)
    M2_e = 0.8 + 0.1 * sp.sin(2 * sp.pi * X)
    mu1_e = 0.5 * sp.sin(2 * sp.pi * (X + Y))
    mu2_e = 0.3 * sp.cos(2 * sp.pi * X) * sp.sin(2 * sp.pi * Y)
    params = dict(dx=None, d=1.0, v_m=1.3, M_qmax=0.7, H=0.05, y_e=y_e, beta=beta, D_L=1.0, D_S=0.1,
                  dt=1e-6, ebar=0.4, eqbar=0.04, noise_phi=0.0, noise_c=0.0, noise_q=0.0)
    W = [0.3, 0.5, 0.2]
    Mo = [1.0, 1.5, 0.5]
    return locals()


def test_ncomp_rates_converge_to_continuum_operators():
    exprs = None
    errs = {"phi": [], "c1": [], "c2": [], "q1": [], "q4": []}
    for n in (32, 64, 128):
        s = _nc_setup(n, beta=1e3)  # beta above every face gradient: f_ori = div(D_q grad q)/beta
        h, xx, yy = _grid(n)
        ev = lambda e: _ev(e, xx, yy)
        fields = np.stack([ev(s["phi_e"]), ev(s["q1_e"]), ev(s["q4_e"]), ev(s["c1_e"]), ev(s["c2_e"])])
        T = ev(s["T_e"])
        transfer = np.stack([np.zeros_like(T), ev(s["dG_e"]), ev(s["M1_e"]), ev(s["M2_e"]), ev(s["mu1_e"]),
                             ev(s["mu2_e"])])
        transfer[0] = 0.0  # G_L
        # G_S - G_L = dG
        pr = dict(s["params"], dx=h)
        params = [pr[k] for k in oe.PARAM_NAMES]
        c_params = np.array([[1, 1, 1], [1, 1, 1], [1, 1, 1], [1, 1, 1], s["W"], s["Mo"]], dtype=float)
        dphi, dc, dq1, dq4 = oe.ncomp_rates(fields, T, transfer, params, c_params)
        if exprs is None:
            exprs = _nc_exact(s)
        sl = (slice(1, -1), slice(1, -1))
        errs["phi"].append(np.max(np.abs(dphi - ev(exprs["phi"])[sl])))
        errs["c1"].append(np.max(np.abs(dc[0] - ev(exprs["c1"])[sl])))
        errs["c2"].append(np.max(np.abs(dc[1] - ev(exprs["c2"])[sl])))
        errs["q1"].append(np.max(np.abs(dq1 - ev(exprs["q1"])[sl])))
        errs["q4"].append(np.max(np.abs(dq4 - ev(exprs["q4"])[sl])))
    for k, e in errs.items():
        assert min(_order(e)) > 3.3, (k, e)


def _nc_exact(s):
    pr = s["params"]
    dxf = lambda e: sp.diff(e, X)
    dyf = lambda e: sp.diff(e, Y)
    phi, q1, q4, T = s["phi_e"], s["q1_e"], s["q4_e"], s["T_e"]
    c1, c2 = s["c1_e"], s["c2_e"]
    eb2 = pr["ebar"] ** 2
    y_e = pr["y_e"]
    a = q1 ** 2 - q4 ** 2
    b = 2 * q1 * q4
    px, py = sp.symbols("px py", real=True)
    psix = a * px - b * py
    psiy = b * px + a * py
    A = (psix ** 4 + psiy ** 4) / (px ** 2 + py ** 2)
    Ax = sp.diff(A, px).subs({px: dxf(phi), py: dyf(phi)})
    Ay = sp.diff(A, py).subs({px: dxf(phi), py: dyf(phi)})
    # gradient-energy part: div(T eb^2 [(1-3y_e) grad phi + 2 y_e dA/dgrad phi])
    grad_part = (dxf(T * eb2 * ((1 - 3 * y_e) * dxf(phi) + 2 * y_e * Ax))
                 + dyf(T * eb2 * ((1 - 3 * y_e) * dyf(phi) + 2 * y_e * Ay)))
    g = phi ** 2 * (1 - phi) ** 2
    h = phi ** 3 * (6 * phi ** 2 - 15 * phi + 10)
    gp = 4 * phi ** 3 - 6 * phi ** 2 + 2 * phi
    cN = 1 - c1 - c2
    W, Mo = s["W"], s["Mo"]
    cW = c1 * W[0] + c2 * W[1] + cN * W[2]
    M_phi = c1 * Mo[0] + c2 * Mo[1] + cN * Mo[2]
    mgq = sp.sqrt(dxf(q1) ** 2 + dxf(q4) ** 2) + sp.sqrt(dyf(q1) ** 2 + dyf(q4) ** 2)
    dphi = M_phi * (grad_part - 30 * g * s["dG_e"] / pr["v_m"] - gp * T * cW - 4 * pr["H"] * T * phi * mgq)
    M = [s["M1_e"], s["M2_e"]]
    mu = [s["mu1_e"], s["mu2_e"]]
    cs = [c1, c2]
    divagradb = lambda aa, bb: dxf(aa * dxf(bb)) + dyf(aa * dyf(bb))
    dc = [divagradb(M[l], mu[l]) - sum(divagradb(M[l] * cs[m], mu[m]) for m in range(2)) for l in range(2)]
    D_q = 2 * pr["H"] * T * phi ** 2
    fori = [divagradb(D_q, q) / pr["beta"] for q in (q1, q4)]
    t1 = fori[0] + pr["eqbar"] ** 2 * (dxf(dxf(q1)) + dyf(dyf(q1)))
    t4 = fori[1] + pr["eqbar"] ** 2 * (dxf(dxf(q4)) + dyf(dyf(q4)))
    M_q = pr["M_qmax"] + (1e-6 - pr["M_qmax"]) * h
    dq1 = M_q * ((1 - q1 ** 2) * t1 - q1 * q4 * t4)
    dq4 = M_q * ((1 - q4 ** 2) * t4 - q1 * q4 * t1)
    return {"phi": dphi, "c1": dc[0], "c2": dc[1], "q1": dq1, "q4": dq4}


def test_ncomp_orientation_flux_with_active_clip():
    """With beta below the face gradients the orientation flux is
    sum_a d_a(D_q d_a q / |d_a q_vec|): check it converges too."""
    errs = []
    for n in (32, 64, 128):
        s = _nc_setup(n, beta=1e-6, y_e=0.0)
        h, xx, yy = _grid(n)
        ev = lambda e: _ev(e, xx, yy)
        q1, q4 = ev(s["q1_e"]), ev(s["q4_e"])
        phi, T = ev(s["phi_e"]), ev(s["T_e"])
        P, Q1, Q4, TT = oe._Nb(phi), oe._Nb(q1), oe._Nb(q4), oe._Nb(T)
        idx = 1 / h
        mg = {}
        for f in ("xp", "xm", "yp", "ym"):
            m = idx * np.sqrt((getattr(Q1, f) - Q1.c) ** 2 + (getattr(Q4, f) - Q4.c) ** 2)
            mg[f] = np.where(m < 1e-6, 1e-6, m)  # the kernel's clip at beta (NCGPU_new.py:118-125)
        Dq = {f: 2 * 0.05 * getattr(TT, f) * getattr(P, f) ** 2 for f in ("c", "xp", "xm", "yp", "ym")}
        got = oe._f_ori_term(Dq["c"], Dq["xp"], Dq["xm"], Dq["yp"], Dq["ym"], mg["xp"], mg["xm"], mg["yp"],
                             mg["ym"], Q1.c, Q1.xp, Q1.xm, Q1.yp, Q1.ym, idx)
        D_q = 2 * 0.05 * s["T_e"] * s["phi_e"] ** 2
        dxf = lambda e: sp.diff(e, X)
        dyf = lambda e: sp.diff(e, Y)
        nx_ = sp.sqrt(dxf(s["q1_e"]) ** 2 + dxf(s["q4_e"]) ** 2)
        ny_ = sp.sqrt(dyf(s["q1_e"]) ** 2 + dyf(s["q4_e"]) ** 2)
        exact = dxf(D_q * dxf(s["q1_e"]) / nx_) + dyf(D_q * dyf(s["q1_e"]) / ny_)
        e = np.abs(got - ev(exact)[1:-1, 1:-1])
        errs.append(np.nanmedian(e))  # |d_a q| vanishes on lines: use the median
    assert min(_order(errs)) > 2.5, errs


# ------------------------------------------------------------------ thermo --
def _model(n=3, seed=0):
    rng = np.random.default_rng(seed)
    g0 = np.zeros((n, 5))
    g0[:, 0] = rng.normal(0, 1e4, n)
    g0[:, 1] = rng.normal(0, 10, n)
    g0[:, 2] = rng.normal(0, 1, n)
    g0[:, 3] = rng.normal(0, 1e-3, n)
    g0[:, 4] = rng.normal(0, 1e4, n)
    L = np.zeros((n * (n - 1) // 2, 3, 2))
    L[:, :, 0] = rng.normal(0, 1e4, L.shape[:2])
    L[:, :, 1] = rng.normal(0, 1, L.shape[:2])
    return g0, L, 8.314


def test_thermo_closed_form_binary():
    g0, L, R = _model(2)
    c, T = 0.3, 1500.0
    gi = [g0[i, 0] + g0[i, 1] * T + g0[i, 2] * T * np.log(T) + g0[i, 3] * T * T + g0[i, 4] / T for i in range(2)]
    Lk = L[0, :, 0] + L[0, :, 1] * T
    d = c - (1 - c)
    want = (c * gi[0] + (1 - c) * gi[1] + c * (1 - c) * (Lk[0] + Lk[1] * d + Lk[2] * d * d)
            + R * T * (c * np.log(c) + (1 - c) * np.log(1 - c)))
    got = oth.gibbs([np.array(c), np.array(1 - c)], np.array(T), g0, L, R)
    assert abs(got - want) < 1e-9 * abs(want)


def test_thermo_rk_symmetry_under_component_swap():
    """L^k_ij = (-1)^k L^k_ji: swapping components 0 and 1 with odd-order
    terms negated leaves G unchanged."""
    g0, L, R = _model(3, 1)
    rng = np.random.default_rng(2)
    c = rng.dirichlet([2, 2, 2], 50).T
    T = rng.uniform(800, 2500, 50)
    G = oth.gibbs(list(c), T, g0, L, R)
    g0s = g0[[1, 0, 2]]
    Ls = L.copy()  # pairs: (0,1), (0,2), (1,2) -> (1,0)=(0,1) flipped, (1,2)->(0,2), (0,2)->(1,2)
    Ls[0] = L[0] * np.array([1, -1, 1])[:, None]
    Ls[1] = L[2]
    Ls[2] = L[1]
    Gs = oth.gibbs([c[1], c[0], c[2]], T, g0s, Ls, R)
    np.testing.assert_allclose(Gs, G, rtol=1e-12)


def test_thermo_derivatives_analytic_vs_differences():
    g0, L, R = _model(3, 3)
    rng = np.random.default_rng(4)
    c = rng.dirichlet([3, 3, 3], 200).T
    T = rng.uniform(900, 2500, 200)
    an = oth.analytic_derivatives(list(c), T, g0, L, R)
    # central difference along e_l - e_N (independent of the forward scheme)
    h = 1e-6
    for l in range(2):
        cp = [x.copy() for x in c]
        cm = [x.copy() for x in c]
        cp[l] += h
        cp[2] -= h
        cm[l] -= h
        cm[2] += h
        cd = (oth.gibbs(cp, T, g0, L, R) - oth.gibbs(cm, T, g0, L, R)) / (2 * h)
        np.testing.assert_allclose(an[l], cd, rtol=1e-6, atol=1e-3)
    G, fd = oth.fd_derivatives(list(c), T, g0, L, R)
    for l in range(2):
        # forward difference with h = 1e-7: truncation (h/2) G'' ~ 1e-7 * RT/c
        bound = 1e-7 * 8.314 * T / np.minimum(np.minimum(c[l], c[2]), 1.0) + 1e-3
        assert np.all(np.abs(fd[l] - an[l]) < bound)
