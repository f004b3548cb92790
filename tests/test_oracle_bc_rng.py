"""Pin the oracle's boundary rules and counter RNG.

Boundary rules are checked cell by cell against the rules of
ppf_gpu_utils.py:11-96 (x kernel, then y kernel over every row including the
halo rows) written as explicit loops over a tiny grid, independently of the
slicing the oracle uses.
"""

import numpy as np
import pytest

from oracle import bc as obc
from oracle import rng as orng


def reference_rules_2d(f, bcx, bcy, bca, dx):
    """Explicit loops: x kernel over all rows, then y kernel over all columns."""
    f = f.copy()
    F, H, W = f.shape
    for k in range(F):
        for i in range(H):  # x kernel (ppf_gpu_utils.py:39-46, 59-66, 79-86)
            if bcx == "PERIODIC":
                f[k, i, 0] = f[k, i, W - 2]
                f[k, i, W - 1] = f[k, i, 1]
            elif bcx == "NEUMANN":
                f[k, i, 0] = f[k, i, 1] - dx * bca[k, i, 0]
                f[k, i, W - 1] = f[k, i, W - 2] + dx * bca[k, i, W - 1]
            else:
                f[k, i, 0] = bca[k, i, 0]
                f[k, i, W - 1] = bca[k, i, W - 1]
        for j in range(W):  # y kernel (ppf_gpu_utils.py:49-56, 69-76, 89-96)
            if bcy == "PERIODIC":
                f[k, 0, j] = f[k, H - 2, j]
                f[k, H - 1, j] = f[k, 1, j]
            elif bcy == "NEUMANN":
                f[k, 0, j] = f[k, 1, j] - dx * bca[k, 0, j]
                f[k, H - 1, j] = f[k, H - 2, j] + dx * bca[k, H - 1, j]
            else:
                f[k, 0, j] = bca[k, 0, j]
                f[k, H - 1, j] = bca[k, H - 1, j]
    return f


KINDS = ["PERIODIC", "NEUMANN", "DIRCHLET"]


@pytest.mark.parametrize("bcx", KINDS)
@pytest.mark.parametrize("bcy", KINDS)
def test_apply_bc_2d_equals_kernel_loops(bcx, bcy):
    rng = np.random.default_rng(0)
    f = rng.random((2, 6, 7))
    bca = rng.random((2, 6, 7))
    want = reference_rules_2d(f, bcx, bcy, bca, 0.3)
    got = obc.apply_bc(f.copy(), [bcx, bcy], bca, 0.3)
    np.testing.assert_array_equal(got, want)


def test_apply_bc_invariants():
    rng = np.random.default_rng(1)
    f = rng.random((1, 8, 9))
    p = obc.apply_bc(f.copy(), "PERIODIC")
    # periodic: halo equals the wrapped interior, corners included
    np.testing.assert_array_equal(p[0, 0, 1:-1], p[0, -2, 1:-1])
    np.testing.assert_array_equal(p[0, 1:-1, -1], p[0, 1:-1, 1])
    assert p[0, 0, 0] == p[0, -2, -2]
    n = obc.apply_bc(f.copy(), "NEUMANN")
    # zero-valued Neumann: zero one-sided gradient on every side
    np.testing.assert_array_equal(n[0, 0, :], n[0, 1, :])
    np.testing.assert_array_equal(n[0, :, -1], n[0, :, -2])


def test_apply_bc_single_temperature_rule():
    T = np.arange(30.0).reshape(5, 6)
    Tp = obc.apply_bc_single(T.copy(), "DIRCHLET")
    # Dirichlet fields give T a zero-gradient halo (ppf_gpu_utils.py:442-445)
    np.testing.assert_array_equal(Tp[0], Tp[1])
    np.testing.assert_array_equal(Tp[:, 0], Tp[:, 1])
    Tq = obc.apply_bc_single(T.copy(), "PERIODIC")
    np.testing.assert_array_equal(Tq[:, 0], Tq[:, -2])


def test_bc_names_and_order():
    assert obc.normalize_bc("neumann", 2) == ["NEUMANN", "NEUMANN"]
    assert obc.normalize_bc(["periodic", "dirichlet"], 2) == ["PERIODIC", "DIRCHLET"]
    with pytest.raises(ValueError):
        obc.normalize_bc("SOMETHING", 1)


# ------------------------------------------------------------------ RNG --
def test_rng_known_answers():
    """splitmix64 finalizer of (seed + counter*golden): first outputs for
    seed 0 are the published SplitMix64 sequence (seed 0)."""
    h = orng.hash64(0, np.arange(1, 4, dtype=np.uint64))
    assert [int(x) for x in h] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_rng_uniform_and_normal_moments():
    cells = np.arange(200000, dtype=np.uint64)
    u = orng.uniform_f64(7, 3, 0, 200000, cells)
    assert 0.0 <= u.min() and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 3e-3 and abs(u.var() - 1 / 12) < 2e-3
    z = orng.normal_f64(7, 3, 1, 200000, cells)
    assert abs(z.mean()) < 1e-2 and abs(z.var() - 1) < 2e-2
    u32 = orng.uniform_f32(7, 3, 0, 200000, cells)
    assert np.max(np.abs(u32 - u)) < 2 ** -23


def test_rng_streams_distinct():
    cells = np.arange(1000, dtype=np.uint64)
    a = orng.uniform_f64(1, 1, 0, 1000, cells)
    b = orng.uniform_f64(1, 1, 1, 1000, cells)
    c = orng.uniform_f64(1, 2, 0, 1000, cells)
    assert len(np.unique(np.concatenate([a, b, c]))) == 3000
