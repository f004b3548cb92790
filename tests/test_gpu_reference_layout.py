"""The C ABI on the reference's own memory layout.

The reference hands its kernels C-contiguous FP64 stacks ``[F, ny+2, nx+2]``
built by ``np.array`` and copied to the device as they are
(ppf_gpu_utils.py:396-420): no row pitch, no alignment.  A binding on the
reference side passes such buffers straight through, so here the entry points
are called through ctypes on ``torch.empty((F, ny+2, nx+2), float64)`` with an
odd nx (every row starts at an unaligned address) and compared with the FP64
oracle at the 1e-10 bar.
"""

import ctypes

import numpy as np
import pytest

from oracle import bc as obc
from tests.oracle_runner import rel_l2, run_point, run_warren
from tests.test_gpu_parity import _nc_oracle, _nc_sim, _warren_sim

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _contiguous(host):
    import torch
    t = torch.empty(host.shape, dtype=torch.float64, device="cuda")
    assert t.is_contiguous()
    t.copy_(torch.from_numpy(np.ascontiguousarray(host)))
    return t


def _grid(t, bc, dx):
    from pyphasefield_amd import _lib
    g = _lib.make_grid(t, t.dim() - 1, bc, dx)
    assert g.sy == t.shape[-1] and g.sf == t.shape[-1] * t.shape[-2]
    return g


@pytest.mark.parametrize("bc", ["PERIODIC", "NEUMANN", ["DIRCHLET", "PERIODIC"]])
def test_apply_bc_and_diffusion_contiguous(gpu, bc):
    from pyphasefield_amd import _lib
    L = _lib.lib()
    rng = np.random.default_rng(12)
    ny, nx = 41, 65
    c0 = rng.random((ny, nx))
    bca = rng.random((1, ny + 2, nx + 2))
    host = obc.pad(c0)[None].copy()
    a = _contiguous(host)
    b = _contiguous(np.zeros_like(host))
    bcd = _contiguous(bca)
    g = _grid(a, bc, 1.0)
    s = _lib.stream_handle()
    _lib.check(L.ppf_apply_bc(g, _lib.ptr(a), _lib.ptr(bcd), s), "apply_bc")
    want0 = host.copy()
    obc.apply_bc(want0, bc, bca, 1.0)
    np.testing.assert_array_equal(a.cpu().numpy(), want0)
    steps = 30
    for _ in range(steps):
        _lib.check(L.ppf_diffusion_step(g, _lib.ptr(a), _lib.ptr(b), _lib.ptr(bcd), 0.9, 0.2, s), "diffusion_step")
        a, b = b, a
    _lib.check(L.ppf_apply_bc(g, _lib.ptr(a), _lib.ptr(bcd), s), "apply_bc")
    want = run_point("diffusion", c0, steps, bc, 1.0, 0.2, bcarray=bca, D=0.9)
    assert rel_l2(a.cpu().numpy()[0], want) <= TOL


def test_warren_step_contiguous(gpu):
    from pyphasefield_amd import _lib
    from pyphasefield_amd.Engines.Warren1995 import warren_derived
    E = _engines()
    ny, nx = 50, 65
    sim = _warren_sim(E, 8, "FP64", 0.3, seed=5)  # parameters only
    sim.just_before_simulating()
    rng = np.random.default_rng(3)
    phi0 = np.ones((ny, nx))
    phi0[20:30, 25:40] = 0.0
    c0 = 0.40 + 0.01 * rng.random((ny, nx))
    host = np.stack([obc.pad(phi0), obc.pad(c0)])
    a = _contiguous(host)
    b = _contiguous(np.zeros_like(host))
    g = _grid(a, "NEUMANN", sim.dx)
    L = _lib.lib()
    s = _lib.stream_handle()
    steps = 40
    for step in range(1, steps + 1):
        _lib.check(L.ppf_warren_step(g, _lib.ptr(a), _lib.ptr(b), None, sim._params, ctypes.c_uint64(5), step, s),
                   "warren_step")
        a, b = b, a
    _lib.check(L.ppf_apply_bc(g, _lib.ptr(a), None, s), "apply_bc")
    want = run_warren(phi0, c0, steps, "NEUMANN", sim.dx, warren_derived(sim.user_data, sim.dx), seed=5)
    got = a.cpu().numpy()
    assert rel_l2(got[0], want[0]) <= TOL and rel_l2(got[1], want[1]) <= TOL


@pytest.mark.parametrize("fused", [False, True])
def test_ncomp_contiguous(gpu, fused):
    """ppf_ncomp_thermo + ppf_ncomp_step (and the one-launch step) on
    contiguous FP64 stacks, transfer stack included, noise on."""
    from pyphasefield_amd import _lib
    E = _engines()
    sim, model = _nc_sim(E, 37, 61, "FP64", 1.0)
    sim.just_before_simulating()
    sim._begun_simulation = True
    fields0, run = _nc_oracle(sim, model, 20, "NEUMANN")
    host = np.stack([obc.pad(f) for f in fields0])
    a = _contiguous(host)
    b = _contiguous(np.zeros_like(host))
    ncomp = host.shape[0] - 2
    tr = _contiguous(np.zeros((2 * ncomp - 1,) + host.shape[1:]))
    g = _grid(a, "NEUMANN", sim.dx)
    tg = _grid(tr, "NEUMANN", sim.dx)
    L = _lib.lib()
    s = _lib.stream_handle()
    t = sim.temperature_descriptor()
    for step in range(1, 21):
        if fused:
            _lib.check(L.ppf_ncomp_fused_step(g, _lib.ptr(a), _lib.ptr(b), None, t, sim._ncparams, sim._thermo,
                                              ctypes.c_uint64(sim.seed), step, s), "ncomp_fused_step")
        else:
            _lib.check(L.ppf_ncomp_thermo(g, _lib.ptr(a), None, t, tg, _lib.ptr(tr), sim._ncparams, sim._thermo,
                                          ctypes.c_uint64(sim.seed), step, s), "ncomp_thermo")
            _lib.check(L.ppf_ncomp_step(g, _lib.ptr(a), _lib.ptr(b), None, t, tg, _lib.ptr(tr), sim._ncparams,
                                        ctypes.c_uint64(sim.seed), step, s), "ncomp_step")
        a, b = b, a
    _lib.check(L.ppf_apply_bc(g, _lib.ptr(a), None, s), "apply_bc")
    want = run()
    got = a.cpu().numpy()
    for i in range(host.shape[0]):
        assert rel_l2(got[i], want[i]) <= TOL, i


def _engines():
    import pyphasefield_amd.Engines as engines
    return engines
