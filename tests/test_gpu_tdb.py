"""The N-component engine on the reference's own TDB databases (read by
pyphasefield_amd.tdb, evaluated on the device from ppf_tdb_tables) against
the FP64 oracle evaluating the same tables."""

import os

import numpy as np
import pytest

from oracle import engines as oe
from tests.oracle_runner import padded_T, rel_l2, run_ncomp

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "tdb")
FP64_TOL = 1e-10


@pytest.mark.parametrize("name,comps", [("Ni-Cu-Al_Ideal.tdb", None), ("mc_simplified.tdb", ["AL", "CR", "NI"])])
def test_thermo_eval_with_tdb_tables(gpu, name, comps):
    """Device tables against the oracle's evaluation of the same tables;
    mc_simplified adds binary orders up to 4 and a ternary term."""
    import torch
    from oracle import thermo as oth
    from pyphasefield_amd import _lib, thermo
    model = thermo.load_tdb_model(os.path.join(GOLD, name), components=comps)
    rng = np.random.default_rng(1)
    n = 2000
    T = rng.uniform(400, 2900, n)   # every temperature range of the Al, Cu, Ni functions
    c = 0.01 + 0.97 * rng.dirichlet([2, 2, 2], n)[:, :2]
    Td = torch.from_numpy(T).to(gpu)
    cd = torch.from_numpy(np.ascontiguousarray(c)).to(gpu)
    out = torch.empty(n * 6, dtype=torch.float64, device=gpu)
    th = model.to_ctypes()
    assert th.tables
    _lib.check(_lib.lib().ppf_thermo_eval(th, n, _lib.ptr(Td), _lib.ptr(cd), _lib.ptr(out),
                                          _lib.stream_handle()), "thermo_eval")
    o = out.cpu().numpy().reshape(n, 6)
    cs = [c[:, 0], c[:, 1], 1 - c[:, 0] - c[:, 1]]
    for k, which in ((0, "liquid"), (3, "solid")):
        g0, L, R = model.oracle_phase(which)
        G = oth.gibbs(cs, T, g0, L, R)
        d = oth.analytic_derivatives(cs, T, g0, L, R)
        assert np.max(np.abs(o[:, k] - G) / np.abs(G).max()) <= 1e-14
        for l in range(2):
            assert np.max(np.abs(o[:, k + 1 + l] - d[l])) <= 1e-12 * np.abs(d[l]).max()


_COMPONENTS = {"mc_simplified.tdb": ["AL", "CR", "NI"]}


def _tdb_sim(name, precision, ny=30, nx=44, T0=1300.0):
    import pyphasefield_amd.Engines as E
    path = os.path.join(GOLD, name)
    sim = E.NCGPU_new(dimensions=[ny, nx], seed=11, tdb_path=path, tdb_components=_COMPONENTS.get(name))
    sim.set_framework("GPU_SERIAL")
    sim.set_dx(0.0000046)
    sim.set_temperature_type("LINEAR_GRADIENT")
    sim.set_temperature_initial_T(T0)
    sim.set_temperature_dTdx(1.0 / 0.0000046)
    sim.set_temperature_dTdy(0.0)
    sim.set_temperature_dTdt(0.0)
    sim.set_boundary_conditions("NEUMANN")
    sim.set_precision(precision)
    ncomp = 3 if ("Al" in name or name in _COMPONENTS) else 2
    ica = [0.3, 0.3][:ncomp - 1]
    sim.set_user_data({"d_ratio": 1.1, "sim_type": "seeds", "number_of_seeds": 3, "seed_radius": 4,
                       "initial_concentration_array": ica, "noise_phi": 0.0, "noise_c": 0.0, "noise_q": 0.0})
    sim.initialize_fields_and_imported_data()
    return sim


@pytest.mark.parametrize("name", ["Ni-Cu-Al_Ideal.tdb", "Ni-Nb_Simplified.tdb", "mc_simplified.tdb"])
def test_ncomp_fp64_on_reference_tdb(gpu, name):
    sim = _tdb_sim(name, "FP64")
    sim.just_before_simulating()
    sim._begun_simulation = True
    u = sim.user_data
    model = sim._tdb
    f0 = np.stack([f.get_cells().copy() for f in sim.fields])
    ny, nx = sim.dimensions

    def T_of(s):
        return padded_T(oe.temperature_linear((ny, nx), sim._initial_T, sim._dTdx, 0.0, sim.dx, 0.0, sim.dt,
                                              s - 1), "NEUMANN")
    steps = 20
    sim.simulate(steps)
    want = run_ncomp(f0, steps, "NEUMANN", T_of, u["params"], u["c_params"], model.oracle_phase("liquid"),
                     model.oracle_phase("solid"), seed=sim.seed)
    for i in range(len(sim.fields)):
        assert rel_l2(sim.fields[i].data, want[i]) <= FP64_TOL, sim.fields[i].name


def test_ncomp_fp32_on_reference_tdb(gpu):
    sim = _tdb_sim("Ni-Cu-Al_Ideal.tdb", "FP32", 48, 48)
    sim.simulate(200)
    for f in sim.fields:
        assert np.isfinite(f.get_cells()).all()
    for i in (3, 4):
        c = sim.fields[i].get_cells()
        assert 0.0 < c.min() and c.max() < 1.0
