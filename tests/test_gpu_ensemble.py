"""Batched ensemble on the GPU: equal bit for bit to the single-instance FP32
engines for every tile configuration, close to the FP64 oracle, HIP-graph
replays equal to eager steps, autotuner and Simulation interoperation."""

import numpy as np
import pytest

from tests.oracle_runner import rel_l2, run_point

pytestmark = pytest.mark.gpu

# mixed kinds, boundaries and awkward sides (1-cell and prime extents,
# partial tiles in both directions)
SPECS = [("AC", 70, 40, ("PERIODIC", "PERIODIC")), ("CH", 33, 65, ("NEUMANN", "PERIODIC")),
         ("AC", 129, 7, ("NEUMANN", "NEUMANN")), ("CH", 64, 64, ("PERIODIC", "PERIODIC")),
         ("CH", 5, 130, ("PERIODIC", "NEUMANN")), ("AC", 1, 17, ("PERIODIC", "NEUMANN")),
         ("CH", 200, 3, ("NEUMANN", "NEUMANN")), ("AC", 257, 31, ("PERIODIC", "NEUMANN")),
         ("CH", 97, 41, ("NEUMANN", "NEUMANN"))]


def _specs():
    from pyphasefield_amd.ensemble import InstanceSpec, stable_dt
    rng = np.random.default_rng(5)
    out = []
    for kind, nx, ny, bc in SPECS:
        M = float(rng.uniform(0.5, 2.0))
        eps = float(np.sqrt(rng.uniform(0.5, 2.0)))
        W = float(rng.uniform(0.5, 1.5)) if kind == "AC" else 1.0
        dt = 0.8 * stable_dt(kind, M, eps, W)
        out.append(InstanceSpec(kind, nx, ny, bc, M, eps, dt, W))
    return out


def _initial(specs, seed=3):
    rng = np.random.default_rng(seed)
    return [(0.5 + 0.05 * (2 * rng.random((s.ny, s.nx)) - 1)).astype(np.float32) for s in specs]


def _single_engine(spec, field, steps):
    import pyphasefield_amd.Engines as engines
    cls = engines.CahnAllen if spec.kind == "AC" else engines.CahnHilliard
    sim = cls(dimensions=[spec.ny, spec.nx], seed=0)
    ud = {"M": spec.M, "epsilon": spec.eps}
    if spec.kind == "AC":
        ud["W"] = spec.W
    sim.set_user_data(ud)
    sim.set_framework("GPU_SERIAL")
    sim.set_dx(spec.dx)
    sim.set_dt(spec.dt)
    sim.set_boundary_conditions(list(spec.bc))
    sim.set_precision("FP32")
    sim.initialize_fields_and_imported_data()
    sim.fields[0].data[1:-1, 1:-1] = field
    sim.simulate(steps)
    return sim.fields[0].get_cells()


def test_every_config_equals_single_instance_engines(gpu):
    from pyphasefield_amd import _lib
    from pyphasefield_amd.ensemble import Ensemble
    specs = _specs()
    init = _initial(specs)
    steps = 12
    want = [_single_engine(s, f, steps) for s, f in zip(specs, init)]
    for config in range(_lib.lib().ppf_ensemble_num_configs()):
        # fewer blocks than XCD segments (1, 3), exactly one per segment (8), more
        for blocks in (1, 3, 8, 13, 64):
            ens = Ensemble(specs, init)
            ens.step(steps, config=config, blocks=blocks)
            for i, got in enumerate(ens.fields()):
                assert np.array_equal(got, want[i]), (config, blocks, i, specs[i])


def test_ensemble_matches_fp64_oracle(gpu):
    from pyphasefield_amd.ensemble import Ensemble
    specs = _specs()
    init = _initial(specs, seed=4)
    steps = 40
    ens = Ensemble(specs, init)
    ens.step(steps)
    for s, f0, got in zip(specs, init, ens.fields()):
        if s.kind == "AC":
            want = run_point("allencahn", f0.astype(np.float64), steps, list(s.bc), s.dx, s.dt, M=s.M, W=s.W,
                             epsilon=s.eps)
        else:
            want = run_point("cahnhilliard", f0.astype(np.float64), steps, list(s.bc), s.dx, s.dt, M=s.M,
                             epsilon=s.eps)
        # FP32 rounding (6e-8 per step) grown by the phase-separation modes the
        # instances amplify over 40 steps at 0.8 of the stability limit
        assert rel_l2(got, want[1:-1, 1:-1]) <= 3e-5, s
        if s.kind == "CH":  # conservation (zero flux or periodic on every side)
            assert abs(float(got.astype(np.float64).sum()) - float(f0.astype(np.float64).sum())) <= 1e-4 * f0.size


def test_graph_replays_equal_eager_steps(gpu):
    import torch
    from pyphasefield_amd.ensemble import Ensemble
    specs = _specs()
    init = _initial(specs, seed=6)
    eager = Ensemble(specs, init)
    graph = Ensemble(specs, init)
    for rep in range(4):
        eager.step(10)
        graph.simulate(10, graph_steps=10)
        # unrelated eager work and a sync between replays
        torch.ones(1000, device="cuda").sum().item()
        torch.cuda.synchronize()
        for a, b in zip(eager.fields(), graph.fields()):
            assert np.array_equal(a, b), rep
    # an odd number of steps: graph part + eager remainder
    eager.step(25)
    graph.simulate(25, graph_steps=10)
    for a, b in zip(eager.fields(), graph.fields()):
        assert np.array_equal(a, b)


def test_autotune_restores_state_and_caches(gpu, tmp_path, monkeypatch):
    from pyphasefield_amd.ensemble import Ensemble, random_specs, random_initial, ensemble_key
    import json
    monkeypatch.setenv("PPF_AUTOTUNE_CACHE", str(tmp_path / "tune.json"))
    specs = random_specs(12, seed=3, side=(16, 200))
    init = random_initial(specs, seed=3)
    ens = Ensemble(specs, init)
    ens.step(3)
    before = ens.fields()
    best = ens.autotune(steps=4, blocks_per_cu=(1, 2), configs=[0, 2, 5])
    assert best["config"] in (0, 2, 5) and best["blocks"] in (ens.cus, 2 * ens.cus)
    for a, b in zip(before, ens.fields()):
        assert np.array_equal(a, b)
    db = json.load(open(tmp_path / "tune.json"))
    assert db[ensemble_key(specs, ens.device_name)]["config"] == best["config"]
    ens2 = Ensemble(specs, init)
    assert ens2.autotune(steps=4) == db[ensemble_key(specs, ens.device_name)]


def test_from_simulations_round_trip(gpu):
    import pyphasefield_amd.Engines as engines
    from pyphasefield_amd.ensemble import Ensemble
    sims, ref = [], []
    for k, (cls, dims, bc) in enumerate([(engines.CahnAllen, [30, 50], "PERIODIC"),
                                         (engines.CahnHilliard, [45, 20], ["NEUMANN", "PERIODIC"])]):
        for lst in (sims, ref):
            s = cls(dimensions=dims, seed=k)
            s.set_user_data({"M": 0.7, "epsilon": 1.1})
            s.set_framework("GPU_SERIAL")
            s.set_dx(1.0)
            s.set_dt(0.01)
            s.set_boundary_conditions(bc)
            s.set_precision("FP32")
            s.initialize_fields_and_imported_data()
            lst.append(s)
    ens = Ensemble.from_simulations(sims)
    ens.simulate(30, graph_steps=10)
    ens.write_back()
    for s, r in zip(sims, ref):
        r.simulate(30)
        assert np.array_equal(s.fields[0].get_cells(), r.fields[0].get_cells())
