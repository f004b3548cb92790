"""Host logic of the batched ensemble: seeded instance draws, stability
limits (checked by growth of the unstable mode in the FP64 oracle), arena
layout, tile work-lists and XCD segments."""

import numpy as np
import pytest

from oracle import bc as obc
from oracle import engines as oe
from pyphasefield_amd.ensemble import (InstanceSpec, XCDS, ensemble_key, layout, random_specs, stable_dt,
                                       tile_list)


def test_random_specs_distribution_and_limits():
    a = random_specs(256, seed=0)
    b = random_specs(256, seed=0)
    assert a == b
    sides = np.array([[s.nx, s.ny] for s in a])
    assert sides.min() >= 64 and sides.max() < 1024
    # log-uniform on [64, 1024): median near the geometric mean 256
    assert 180 < np.median(sides) < 360
    kinds = [s.kind for s in a]
    assert 90 < kinds.count("AC") < 166
    bcs = {s.bc for s in a}
    assert len(bcs) == 4
    for s in a:
        lim = stable_dt(s.kind, s.M, s.eps, s.W, s.dx)
        assert 0.45 * lim - 1e-15 <= s.dt <= 0.9 * lim + 1e-15
        assert 0.5 <= s.M <= 2 and 0.5 <= s.eps ** 2 <= 2


def _checkerboard_growth(kind, M, eps, W, dt, steps=60):
    """Amplitude ratio of a 1e-6 checkerboard around the double-well state
    phi = 0 (f'' = 2, the stiffest linear mode) after `steps` oracle steps."""
    n = 16
    yy, xx = np.mgrid[0:n, 0:n]
    f = 1e-6 * (-1.0) ** (xx + yy)
    a = obc.pad(f)
    obc.apply_bc(a[None], "PERIODIC", None, 1.0)
    for _ in range(steps):
        if kind == "AC":
            a = oe.allencahn_step(a, M, W, eps, dt, 1.0)
        else:
            a = oe.cahnhilliard_step(a, M, eps, dt, 1.0, "PERIODIC")
        obc.apply_bc(a[None], "PERIODIC", None, 1.0)
    return np.abs(a[1:-1, 1:-1]).max() / 1e-6


@pytest.mark.parametrize("kind,M,eps,W", [("AC", 1.3, 1.2, 0.8), ("CH", 0.7, 0.9, 1.0), ("CH", 2.0, 1.4, 1.0)])
def test_stable_dt_is_the_explicit_limit(kind, M, eps, W):
    lim = stable_dt(kind, M, eps, W)
    assert _checkerboard_growth(kind, M, eps, W, 0.97 * lim) <= 1.0
    assert _checkerboard_growth(kind, M, eps, W, 1.05 * lim) > 10.0


def test_layout_alignment_and_disjoint():
    specs = random_specs(40, seed=2, side=(3, 300))
    pitches, offsets, total = layout(specs)
    end = -1
    for s, p, o in zip(specs, pitches, offsets):
        assert p % 64 == 0 and p >= s.nx + 2
        assert (o + 1) % 64 == 0          # padded column 1 on a 256-byte boundary
        assert o > end
        end = o + p * (s.ny + 2) - 1
    assert total > end


@pytest.mark.parametrize("shape", [(64, 8, 64, 8), (128, 16, 32, 16), (256, 4, 128, 4), (62, 32, 60, 24)])
def test_tile_list_covers_every_cell_once(shape):
    specs = [InstanceSpec("AC", 70, 40), InstanceSpec("CH", 33, 65), InstanceSpec("AC", 1, 17),
             InstanceSpec("CH", 200, 3), InstanceSpec("AC", 257, 31)]
    tiles, seg = tile_list(specs, shape)
    cover = [np.zeros((s.ny, s.nx), int) for s in specs]
    for i, x0, y0, z in tiles:
        s = specs[i]
        tx, ty = (shape[0], shape[1]) if s.kind == "AC" else (shape[2], shape[3])
        assert (x0 - 1) % tx == 0 and (y0 - 1) % ty == 0 and z == 0
        cover[i][y0 - 1:y0 - 1 + ty, x0 - 1:x0 - 1 + tx] += 1
    assert all((c == 1).all() for c in cover)
    assert len(seg) == XCDS + 1 and seg[0] == 0 and seg[-1] == len(tiles)
    assert (np.diff(seg) >= 0).all()
    # instance order, row-major inside an instance
    keys = [(i, y0, x0) for i, x0, y0, _ in tiles]
    assert keys == sorted(keys)


def test_configurations_name_their_kernel():
    """Block tiles (LDS) first, then wave tiles whose output widths are the
    64 lanes less the halo (1 Allen-Cahn, 2 Cahn-Hilliard columns a side)."""
    from pyphasefield_amd import _lib
    from pyphasefield_amd.ensemble import WAVE_TILE_WIDTHS, Ensemble
    n = _lib.lib().ppf_ensemble_num_configs()
    kinds = [Ensemble.kernel(c) for c in range(n)]
    assert kinds == sorted(kinds) and "k_ensemble" in kinds and "k_ensemble_w" in kinds
    assert WAVE_TILE_WIDTHS == (64 - 2, 64 - 4)
    for c in range(n):
        shape = Ensemble.shape(c)
        if kinds[c] == "k_ensemble_w":
            assert shape[::2] == WAVE_TILE_WIDTHS and shape[1] > 0 and shape[3] > 0
        else:
            assert shape[0] % 32 == 0 and shape[2] % 32 == 0


def test_segments_balance_cost():
    specs = random_specs(256, seed=0)
    tiles, seg, cost = tile_list(specs, (64, 16, 64, 16), return_cost=True)
    parts = np.array([cost[seg[k]:seg[k + 1]].sum() for k in range(XCDS)])
    # each split point is off by at most one tile
    assert parts.max() - parts.min() <= 2 * cost.max()
    assert parts.max() / parts.min() < 1.01


def test_ensemble_key_depends_on_shapes_only():
    a = random_specs(8, seed=1)
    b = [InstanceSpec(s.kind, s.nx, s.ny, s.bc, s.M * 2, s.eps, s.dt) for s in a]
    assert ensemble_key(a, "dev") == ensemble_key(b, "dev")
    b[0] = InstanceSpec(a[0].kind, a[0].nx + 1, a[0].ny)
    assert ensemble_key(a, "dev") != ensemble_key(b, "dev")
