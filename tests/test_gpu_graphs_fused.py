"""HIP-graph replay of the engines whose steps read the step counter (Warren1995,
N-component, AnisoDorr), and the one-launch N-component step.

Graph replays take the step from a device counter that the captured chunk
advances itself; every replayed chunk must equal the same steps launched
eagerly with the host counter, bit for bit, with noise on, a temperature ramp
(dTdt != 0) and unrelated eager work and synchronisations between the calls.
"""

import numpy as np
import pytest

from tests.oracle_runner import rel_l2
from tests.test_gpu_parity import _nc_oracle, _nc_sim, _warren_sim

pytestmark = pytest.mark.gpu

CALLS = (45, 40, 7, 80)  # two partial chunks, an eager-only call, several replays


def _engines():
    import pyphasefield_amd.Engines as engines
    return engines


def _compare_runs(sims):
    import torch
    for n in CALLS:
        for sim in sims:
            sim.simulate(n)
        torch.ones(4096, device="cuda").sum().item()
        torch.cuda.synchronize()
        assert sims[0].time_step_counter == sims[1].time_step_counter
        for fa, fb in zip(sims[0].fields, sims[1].fields):
            assert np.array_equal(fa.get_cells(), fb.get_cells()), (n, fa.name)


@pytest.mark.parametrize("precision", ["FP64", "FP32"])
def test_warren_graph_replay_equals_eager(gpu, precision):
    E = _engines()
    sims = [_warren_sim(E, 70, precision, 0.3, seed=9) for _ in range(2)]
    sims[1]._GRAPH_CHUNK = 0
    _compare_runs(sims)
    assert sims[0]._graph_cache is not None


@pytest.mark.parametrize("precision,mode", [("FP64", "two_pass"), ("FP64", "fused"), ("FP32", "fused"),
                                            ("FP32", "two_pass")])
def test_ncomp_graph_replay_equals_eager(gpu, precision, mode):
    E = _engines()
    sims = []
    for graph in (True, False):
        sim, _ = _nc_sim(E, 40, 70, precision, 1.0)
        sim.user_data["ncomp_kernel"] = mode
        if not graph:
            sim._GRAPH_CHUNK = 0
        sims.append(sim)
    _compare_runs(sims)
    assert sims[0]._kernel_mode() == mode


def test_aniso_graph_replay_equals_eager(gpu):
    E = _engines()
    sims = []
    for graph in (True, False):
        sim = E.AnisoDorrGPU(dimensions=[40, 70], seed=3)
        sim.set_user_data({"sim_type": "seeds", "number_of_seeds": 3})
        sim.set_framework("GPU_SERIAL")
        sim.set_boundary_conditions("NEUMANN")
        sim.set_precision("FP64")
        sim.initialize_fields_and_imported_data()
        if not graph:
            sim._GRAPH_CHUNK = 0
        sims.append(sim)
    _compare_runs(sims)


def test_warren_fp32_wave_tiles_graph_replay_equals_eager(gpu):
    """512 x 512 FP32 has 1152 tiles of 62 x 4: the register-tile kernel.
    With noise, graph replays equal eager launches bit for bit."""
    E = _engines()
    sims = [_warren_sim(E, 512, "FP32", 0.3, seed=9) for _ in range(2)]
    sims[1]._GRAPH_CHUNK = 0
    _compare_runs(sims)


def test_aniso_fp32_wave_tiles_graph_replay_equals_eager(gpu):
    """1030 x 200 FP32 AnisoDorr runs the register-tile kernel; its noise
    reads the device step counter inside graph replays."""
    E = _engines()
    sims = []
    for graph in (True, False):
        sim = E.AnisoDorrGPU(dimensions=[1030, 200], seed=3)
        sim.set_user_data({"sim_type": "seeds", "number_of_seeds": 12})
        sim.set_framework("GPU_SERIAL")
        sim.set_boundary_conditions("NEUMANN")
        sim.set_precision("FP32")
        sim.initialize_fields_and_imported_data()
        if not graph:
            sim._GRAPH_CHUNK = 0
        sims.append(sim)
    _compare_runs(sims)


@pytest.mark.parametrize("noise,bc", [(1.0, "NEUMANN"), (0.0, "PERIODIC"), (1.0, ["PERIODIC", "DIRCHLET"])])
def test_ncomp_fused_equals_two_pass_fp64(gpu, noise, bc):
    """The fused step evaluates the helper's quantities with the same device
    function on the same inputs: FP64 results equal the two-kernel path bit
    for bit (partial tiles in x and y, boundary tiles, noise)."""
    E = _engines()
    sims = []
    for mode in ("fused", "two_pass"):
        sim, _ = _nc_sim(E, 45, 133, "FP64", noise, bc=bc)
        sim.user_data["ncomp_kernel"] = mode
        sims.append(sim)
    for sim in sims:
        sim.simulate(25)
    for fa, fb in zip(sims[0].fields, sims[1].fields):
        assert np.array_equal(fa.data, fb.data), fa.name


def test_ncomp_fused_fp64_oracle_parity(gpu):
    E = _engines()
    sim, model = _nc_sim(E, 40, 52, "FP64", 1.0)
    sim.user_data["ncomp_kernel"] = "fused"
    sim.just_before_simulating()
    sim._begun_simulation = True
    _, run = _nc_oracle(sim, model, 30, "NEUMANN")
    sim.simulate(30)
    want = run()
    for i in range(len(sim.fields)):
        assert rel_l2(sim.fields[i].data, want[i]) <= 1e-10, sim.fields[i].name


def test_ncomp_fused_fp32_close_to_two_pass(gpu):
    """FP32: the fused and two-kernel paths run the same arithmetic but may be
    contracted differently by the compiler; after 20 steps without noise the
    compositions agree to 1e-5 and the orientation stays a unit quaternion."""
    E = _engines()
    sims = []
    for mode in ("fused", "two_pass"):
        sim, _ = _nc_sim(E, 48, 96, "FP32", 0.0)
        sim.user_data["ncomp_kernel"] = mode
        sims.append(sim)
    for sim in sims:
        sim.simulate(20)
    for f in (3, 4):
        a, b = sims[0].fields[f].get_cells(), sims[1].fields[f].get_cells()
        assert np.max(np.abs(a - b)) <= 1e-5
    q1, q4 = sims[0].fields[1].get_cells(), sims[0].fields[2].get_cells()
    assert np.max(np.abs(q1 * q1 + q4 * q4 - 1.0)) <= 1e-5


def test_ncomp_fp32_wave_step_close_to_fused(gpu):
    """FP32 grids of 1024 or more 62 x 4 tiles run the step from the transfer
    stack in the register-tile kernel (lane shifts for x, the lane's rows for
    y; 1030 x 200: 4 strips, the last partial, x 258 tile rows).  It calls
    the same cell update as the LDS tile of the fused step, so one step agrees
    to a few FP32 ulp (contractions differ) and ten steps to 1e-5 in the
    compositions; graph replays equal eager launches bit for bit with noise."""
    import torch
    E = _engines()
    sims = []
    for mode in ("fused", "two_pass"):
        sim, _ = _nc_sim(E, 1030, 200, "FP32", 0.0, seeds=12)
        sim.user_data["ncomp_kernel"] = mode
        sims.append(sim)
    for steps, tol in ((1, 1e-6), (10, 1e-5)):
        for sim in sims:
            sim.simulate(steps - sim.time_step_counter)
        for f in range(5):
            a, b = sims[0].fields[f].get_cells(), sims[1].fields[f].get_cells()
            assert np.all(np.isfinite(b)), (steps, f)
            assert np.max(np.abs(a - b)) <= tol * max(1.0, float(np.abs(a).max())), (steps, f, float(np.max(np.abs(a - b))))
    q1, q4 = sims[1].fields[1].get_cells(), sims[1].fields[2].get_cells()
    assert np.max(np.abs(q1 * q1 + q4 * q4 - 1.0)) <= 1e-5
    # graph replay of the wave step with noise equals eager launches
    runs = []
    for graph in (True, False):
        sim, _ = _nc_sim(E, 1030, 200, "FP32", 1.0, seeds=12)
        sim.user_data["ncomp_kernel"] = "two_pass"
        if not graph:
            sim._GRAPH_CHUNK = 0
        runs.append(sim)
    for n in (25, 7):
        for sim in runs:
            sim.simulate(n)
        torch.cuda.synchronize()
        for fa, fb in zip(runs[0].fields, runs[1].fields):
            assert np.array_equal(fa.get_cells(), fb.get_cells()), (n, fa.name)
