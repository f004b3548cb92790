"""AnisoDorr oracle (Engines/AnisoDorrGPU.py) pinned without a GPU: closed
forms of uniform states, conservation of c, unit quaternions, and the
mirror / quarter-turn covariance of the staggered anisotropy terms (a slip
in a vertex index or a sign of the x or y terms breaks one of them)."""

import numpy as np
import pytest

from oracle import aniso as oa
from oracle import bc as obc

DX = 4.6e-6


def _params(alpha=0.0):
    p = oa.derived_params(DX)
    p["alpha"] = alpha
    return p


def _one_step(f_int, p, bc="PERIODIC", seed=0, step=1):
    F = np.stack([obc.pad(a) for a in f_int])
    obc.apply_bc(F, bc, None, DX)
    out = oa.step(F, oa.helper(F, p, DX), p, DX, seed=seed, step=step)
    return [o[1:-1, 1:-1] for o in out]


def _state(n=24, seed=1):
    y, x = np.mgrid[0:n, 0:n].astype(float)
    rng = np.random.default_rng(seed)
    phi = 0.5 * (1 + np.tanh((6.0 - np.sqrt((x - 11.3) ** 2 + 1.3 * (y - 12.6) ** 2)) / 1.5)) \
        + 0.01 * rng.random((n, n))
    th = 0.3 + 0.2 * np.sin(x / 4.0) + 0.1 * rng.random((n, n))
    c = 0.4 + 0.05 * rng.random((n, n))
    return [phi, np.cos(th / 2), np.sin(th / 2), c]


def test_uniform_states_follow_the_bulk_driving_force():
    """No gradients: q and c stay, phi moves by dt*M_phi*(-(1-c)H_A - c H_B)
    with H = W g'(phi) - 30 L (1/T - 1/T_m) g(phi); liquid (0) and solid (1)
    are fixed points."""
    p = _params()
    n = 8
    for phi0, c0 in ((0.3, 0.41), (0.0, 0.5), (1.0, 0.2)):
        f = [np.full((n, n), phi0), np.full((n, n), np.cos(0.2)), np.full((n, n), np.sin(0.2)),
             np.full((n, n), c0)]
        out = _one_step(f, p)
        g = phi0 ** 2 * (1 - phi0) ** 2
        gp = 2 * phi0 * (1 - phi0) * (1 - 2 * phi0)
        H_A = p["W_A"] * gp - 30 * p["L_A"] * (1 / p["T"] - 1 / p["T_mA"]) * g
        H_B = p["W_B"] * gp - 30 * p["L_B"] * (1 / p["T"] - 1 / p["T_mB"]) * g
        M_phi = (1 - c0) * p["M_A"] + c0 * p["M_B"]
        want = phi0 + p["dt"] * M_phi * (-(1 - c0) * H_A - c0 * H_B)
        assert np.allclose(out[0], want, rtol=1e-12, atol=1e-15)
        assert np.allclose(out[1], f[1], rtol=0, atol=1e-15) and np.allclose(out[2], f[2], rtol=0, atol=1e-15)
        assert np.array_equal(out[3], f[3])


def test_noise_scales_with_the_interface_function():
    """The noise term alpha*rand*16g*((1-c)H_A + c H_B)*M_phi vanishes in the
    bulk and is bounded by its amplitude in between."""
    p0, p1 = _params(0.0), _params(0.3)
    f = _state()
    a, b = _one_step(f, p0, seed=5), _one_step(f, p1, seed=5)
    phi, c = f[0], f[3]
    g = phi ** 2 * (1 - phi) ** 2
    gp = 2 * phi * (1 - phi) * (1 - 2 * phi)
    H_A = p0["W_A"] * gp - 30 * p0["L_A"] * (1 / p0["T"] - 1 / p0["T_mA"]) * g
    H_B = p0["W_B"] * gp - 30 * p0["L_B"] * (1 / p0["T"] - 1 / p0["T_mB"]) * g
    M_phi = (1 - c) * p0["M_A"] + c * p0["M_B"]
    bound = p0["dt"] * M_phi * 0.3 * 16 * g * np.abs((1 - c) * H_A + c * H_B)
    diff = np.abs(b[0] - a[0])
    assert np.all(diff <= bound * (1 + 1e-9) + 1e-18)
    assert diff.max() > 0.2 * bound.max()
    for i in (1, 2, 3):
        assert np.array_equal(a[i], b[i])


def test_concentration_is_conserved_and_quaternions_stay_unit():
    f = _state()
    out = _one_step(f, _params(0.3), seed=2)
    assert abs(out[3].sum() - f[3].sum()) <= 1e-13 * f[3].sum()
    assert np.abs(out[1] ** 2 + out[2] ** 2 - 1).max() < 1e-14
    assert np.abs(out[0] - f[0]).max() > 1e-4  # the state does evolve


@pytest.mark.parametrize("axis", [0, 1])
def test_mirror_covariance(axis):
    """Reflecting the grid flips the orientation angle (q4 -> -q4)."""
    p = _params()
    f = _state()
    base = _one_step(f, p)
    fl = lambda a: np.flip(a, axis)
    m = _one_step([fl(f[0]), fl(f[1]), -fl(f[2]), fl(f[3])], p)
    for got, want, s in zip(m, base, (1, 1, -1, 1)):
        assert np.allclose(s * fl(got), want, rtol=0, atol=1e-15 * np.abs(want).max() + 1e-300)


@pytest.mark.parametrize("k", [1, 2, 3])
def test_quarter_turn_covariance(k):
    """Rotating the grid by k quarter turns (orientation unchanged: the
    psi = R(theta) grad phi construction commutes with rotations)."""
    p = _params()
    f = _state()
    base = _one_step(f, p)
    r = _one_step([np.rot90(a, k) for a in f], p)
    for got, want in zip(r, base):
        assert np.allclose(np.rot90(got, -k), want, rtol=0, atol=4e-16 * np.abs(want).max())
