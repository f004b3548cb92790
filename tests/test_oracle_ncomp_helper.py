"""Pins of the oracle's restatement of NComponent_helper_kernel
(Engines/NCGPU_new.py:256-313, oracle.engines.ncomp_helper) by what the
mathematics fixes, independently of the expressions it evaluates:

* uniform states at phi = 0, 1/2, 1, where the interpolants take exact values
  (h = 0, 1/2, 1; g = 0, 1/16, 0): the mobility reduces to the liquid, mean
  and solid diffusivity, the chemical potential to the liquid, mean and solid
  derivative, and the (W_l - W_N) g T term vanishes in the bulk and equals
  (W_l - W_N) T / 16 at phi = 1/2 -- with W distinct per component, a wrong
  component index, a swapped g / h or D_L / D_S changes these numbers;
* the chemical potentials are forward differences along e_l - e_N with
  h = 1e-7: they agree with the exact directional derivative (thermo's
  analytic form, itself pinned by central differences) to within the
  truncation bound h/2 |G''| plus rounding, with the sign of a forward
  difference; a perturbation along another direction, or a backward
  difference, fails;
* the noise of the chemical potential has zero mean, unit variance after
  division by sqrt(2 R T / v_m), independent streams per component.

The last test plants plausible slips in the oracle (on the CPU) and checks
that these pins catch each of them.
"""

import numpy as np
import pytest

from oracle import engines as oe
from oracle import thermo as ot

R_GAS = 8.314
W = np.array([1.3e3, 2.9e3, 0.4e3])          # distinct per component
D_L, D_S, V_M = 1.0e-5, 1.0e-9, 7.42
T0 = 1500.0
C = (0.31, 0.27)                              # c_1, c_2 (c_3 = 0.42)


def _model():
    from pyphasefield_amd import thermo
    model, _ = thermo.synthetic_rk_model(seed=0)
    return (model.liquid.g0, model.liquid.L, model.liquid.R), (model.solid.g0, model.solid.L, model.solid.R)


def _params(noise_c=0.0):
    p = dict(dx=4.6e-6, d=5e-6, v_m=V_M, M_qmax=8e7, H=1e-11, y_e=0.12, beta=1.5, D_L=D_L, D_S=D_S,
             dt=1e-7, ebar=1e-3, eqbar=5e-4, noise_phi=0.0, noise_c=noise_c, noise_q=0.0)
    params = np.array([p[k] for k in oe.PARAM_NAMES])
    c_params = np.zeros((6, 3))
    c_params[4] = W
    return params, c_params


def _uniform(phi, shape=(4, 5), c=C):
    f = np.zeros((5,) + shape)
    f[0] = phi
    f[1] = 1.0
    f[3] = c[0]
    f[4] = c[1]
    return f


def _exact_dG(thermo_phase, c):
    cs = [np.array([c[0]]), np.array([c[1]]), np.array([1.0 - c[0] - c[1]])]
    return [float(d[0]) for d in ot.analytic_derivatives(cs, np.array([T0]), *thermo_phase)]


def _second_derivative(thermo_phase, c, l, step=1e-4):
    """d2G along e_l - e_3 by a central second difference (a bound only)."""
    def G(t):
        cc = [c[0], c[1], 1.0 - c[0] - c[1]]
        cc[l] += t
        cc[2] -= t
        return float(ot.gibbs([np.array([x]) for x in cc], np.array([T0]), *thermo_phase)[0])
    return (G(step) - 2 * G(0.0) + G(-step)) / step ** 2


def check_uniform_states(helper=None):
    """Raise AssertionError if the helper's uniform-state values leave their
    closed forms."""
    helper = helper or oe.ncomp_helper
    liq, sol = _model()
    params, c_params = _params()
    dGL, dGS = _exact_dG(liq, C), _exact_dG(sol, C)
    for phi, h, g in ((0.0, 0.0, 0.0), (0.5, 0.5, 1.0 / 16.0), (1.0, 1.0, 0.0)):
        f = _uniform(phi)
        T = np.full(f.shape[1:], T0)
        tr = helper(f, T, params, c_params, liq, sol)
        for l in range(2):
            M_c = tr[2 + l]
            want_M = V_M * C[l] * (D_L + h * (D_S - D_L)) / (R_GAS * T0)
            np.testing.assert_allclose(M_c, want_M, rtol=1e-13)
            mu = tr[4 + l]
            want_mu = (dGL[l] + h * (dGS[l] - dGL[l])) / V_M + (W[l] - W[2]) * g * T0
            # forward-difference truncation and rounding, in units of mu
            bound = (0.5 * ot.FD_H * max(abs(_second_derivative(liq, C, l)), abs(_second_derivative(sol, C, l)))
                     + 8 * 2.3e-16 * 1e5 / ot.FD_H) / V_M
            assert np.all(np.abs(mu - want_mu) <= bound + 1e-12 * abs(want_mu)), (phi, l)
    return True


def test_uniform_states_closed_forms():
    assert check_uniform_states()


def test_interpolants_exact_values():
    np.testing.assert_array_equal(oe.nc_h(np.array([0.0, 0.5, 1.0])), [0.0, 0.5, 1.0])
    np.testing.assert_array_equal(oe.nc_g(np.array([0.0, 0.5, 1.0])), [0.0, 1.0 / 16.0, 0.0])


def check_forward_difference(helper=None):
    """Liquid state (h = g = 0): mu_l v_m is the h = 1e-7 forward difference of
    G_L along e_l - e_N -- within the truncation bound of the exact
    derivative, on the side of it that a forward difference takes."""
    helper = helper or oe.ncomp_helper
    liq, sol = _model()
    params, c_params = _params()
    f = _uniform(0.0, shape=(1, 1))
    tr = helper(f, np.full((1, 1), T0), params, c_params, liq, sol)
    exact = _exact_dG(liq, C)
    for l in range(2):
        fd = float(tr[4 + l][0, 0]) * V_M
        d2 = _second_derivative(liq, C, l)
        trunc = 0.5 * ot.FD_H * d2
        rounding = 8 * 2.3e-16 * 1e5 / ot.FD_H
        assert abs((fd - exact[l]) - trunc) <= 0.5 * abs(trunc) + rounding, (l, fd, exact[l], trunc)
    return True


def test_chemical_potential_is_forward_difference_along_el_minus_eN():
    assert check_forward_difference()


def test_noise_moments():
    liq, sol = _model()
    n = 160
    f = _uniform(0.3, shape=(n, n))
    T = np.full(f.shape[1:], T0)
    p0, cp = _params(0.0)
    p1, _ = _params(1.0)
    clean = oe.ncomp_helper(f, T, p0, cp, liq, sol, seed=3, step=7)
    noisy = oe.ncomp_helper(f, T, p1, cp, liq, sol, seed=3, step=7)
    other = oe.ncomp_helper(f, T, p1, cp, liq, sol, seed=3, step=8)
    scale = np.sqrt(2.0 * R_GAS * T0 / V_M)
    m = f[0].size
    xi = [((noisy[4 + l] - clean[4 + l]) / scale).ravel() for l in range(2)]
    for x in xi:
        assert abs(x.mean()) < 4.0 / np.sqrt(m)
        assert abs(x.var() - 1.0) < 5.0 * np.sqrt(2.0 / m)
    assert abs(np.corrcoef(xi[0], xi[1])[0, 1]) < 4.0 / np.sqrt(m)
    x8 = ((other[4] - clean[4]) / scale).ravel()
    assert abs(np.corrcoef(xi[0], x8)[0, 1]) < 4.0 / np.sqrt(m)
    # the mobilities and G carry no noise
    np.testing.assert_array_equal(noisy[:4], clean[:4])


def _swap_interpolants(monkeypatch):
    g, h = oe.nc_g, oe.nc_h
    monkeypatch.setattr(oe, "nc_g", h)
    monkeypatch.setattr(oe, "nc_h", g)


def _fd_along_el(monkeypatch):
    # perturbs c_l only (forgets the compensating -h on c_N)
    def fd(cs, T, g0, L, R, G=None):
        arr = [np.array(c, dtype=np.float64, copy=True) for c in cs]
        G = ot.gibbs(arr, T, g0, L, R) if G is None else G
        out = []
        for l in range(len(cs) - 1):
            arr[l] = arr[l] + ot.FD_H
            out.append(ot.FD_INV * (ot.gibbs(arr, T, g0, L, R) - G))
            arr[l] = arr[l] - ot.FD_H
        return G, out
    monkeypatch.setattr(ot, "fd_derivatives", fd)


def _fd_backward(monkeypatch):
    orig = ot.fd_derivatives

    def fd(cs, T, g0, L, R, G=None):
        G0, _ = orig(cs, T, g0, L, R, G)
        arr = [np.array(c, dtype=np.float64, copy=True) for c in cs]
        arr[-1] = arr[-1] + ot.FD_H
        out = []
        for l in range(len(cs) - 1):
            arr[l] = arr[l] - ot.FD_H
            out.append(ot.FD_INV * (G0 - ot.gibbs(arr, T, g0, L, R)))
            arr[l] = arr[l] + ot.FD_H
        return G0, out
    monkeypatch.setattr(ot, "fd_derivatives", fd)


def _w_index(monkeypatch):
    # (W_l - W_0) in place of (W_l - W_N)
    orig = oe.ncomp_helper

    def helper(fields, T, params, c_params, *a, **k):
        cp = np.array(c_params, copy=True)
        cp[4] = cp[4][[0, 1, 0]]
        return orig(fields, T, params, cp, *a, **k)
    return helper


def _swap_diffusivities(monkeypatch):
    orig = oe.ncomp_helper

    def helper(fields, T, params, c_params, *a, **k):
        p = np.array(params, copy=True)
        i, j = oe.PARAM_NAMES.index("D_L"), oe.PARAM_NAMES.index("D_S")
        p[i], p[j] = p[j], p[i]
        return orig(fields, T, p, c_params, *a, **k)
    return helper


@pytest.mark.parametrize("slip", ["swap_g_h", "fd_along_el", "fd_backward", "w_index", "swap_D"])
def test_pins_catch_planted_slips(monkeypatch, slip):
    helper = None
    if slip == "swap_g_h":
        _swap_interpolants(monkeypatch)
    elif slip == "fd_along_el":
        _fd_along_el(monkeypatch)
    elif slip == "fd_backward":
        _fd_backward(monkeypatch)
    elif slip == "w_index":
        helper = _w_index(monkeypatch)
    else:
        helper = _swap_diffusivities(monkeypatch)
    caught = False
    for check in (check_uniform_states, check_forward_difference):
        try:
            check(helper)
        except AssertionError:
            caught = True
    assert caught, slip
