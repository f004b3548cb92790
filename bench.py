"""Benchmark of the engine hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config NAME] [--no-cpu]

Default workload (``--config warren``): Warren1995 binary solidification,
2D 2048x2048, FP32 (phi, c), circular solid seed of radius 8 at the centre,
c = 0.40 in the solid and 0.41 in the liquid, isothermal 1574 K, 4-fold
anisotropy 0.04, interface noise amplitude 0.01, Neumann boundaries.  A step
is one explicit-Euler time step of the whole grid; inputs are resident in HBM
before the timed region; the steps replay a captured HIP graph.  Prints ONE
JSON line with the metric (cell-updates/s), the roofline of the dominant
kernel, the CPU baseline (the FP64 numpy restatement of the reference engine
timed on this host, with its CPU model and thread counts), the achievable
HBM bandwidth of a 1 GiB device copy measured in the same run, and
``ncomp_surrogate_same_run``: the N-component 2048^2 workload with the
surrogate thermodynamics against the plain FP64 and FP32 engines, timed in
this process (``--no-surrogate`` skips it).

Other configs (one JSON line each): ``diffusion`` (2D 256^2 FP64),
``ch3d`` (Cahn-Hilliard 3D 512^3 FP32), ``ncomp`` (N-component 2048^2, add
``--variant fp64|fp32|surrogate``), ``ensemble`` (256 batched AC/CH
instances), ``copy`` (device copy of 1 GiB, the HBM calibration kernel for
rocprofv3 counters).
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic.json")


def _traffic(kernel_key):
    """HBM bytes per launch from the committed rocprofv3 counter summary."""
    try:
        with open(TRAFFIC_FILE) as f:
            return json.load(f).get(kernel_key)
    except (OSError, ValueError):
        return None


def _events():
    import torch
    s = torch.cuda.current_stream()
    return torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True), s


def _timed(fn, steps):
    """Run fn(steps) between a barrier/sync pair; return seconds (HIP events on
    the stream the kernels are launched on)."""
    import torch
    torch.cuda.synchronize()
    e0, e1, s = _events()
    e0.record(s)
    fn(steps)
    e1.record(s)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / 1e3


def _line(metric, value, unit, steps, warmup, t, dtype, config, roofline, cpu, higher=True, scaling="weak"):
    return {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": steps, "warmup": warmup,
            "ms_per_step": t / steps * 1e3, "higher_is_better": higher, "scaling": scaling, "vs_baseline": None,
            "dtype": dtype, "data": "synthetic", "config": config, "roofline": roofline, "cpu_baseline": cpu}


def _roofline(bytes_per_launch, seconds_per_launch, kernel_key):
    achieved = bytes_per_launch / seconds_per_launch / 1e9
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": _traffic(kernel_key), "kernel": kernel_key,
            "bytes_per_launch_algorithmic": bytes_per_launch}


# ------------------------------------------------------------------ Warren --
def warren_sim(n, precision="FP32", alpha=0.01):
    import pyphasefield_amd.Engines as engines
    sim = engines.Warren1995(dimensions=[n, n], seed=1234)
    sim.set_user_data({"alpha": alpha, "y_e": 0.04, "T": 1574., "seed_shape": "circle", "seed_radius": 8,
                       "c_solid": 0.40, "c_liquid": 0.41})
    sim.set_framework("GPU_SERIAL")
    sim.set_dx(4.6e-8)
    sim.set_boundary_conditions("NEUMANN")
    sim.set_precision(precision)
    sim.initialize_fields_and_imported_data()
    return sim


def cpu_warren(n, k=3):
    """FP64 numpy restatement of engine_Warren1995 as the reference writes it
    (np.roll, np.vectorize interpolants), 1 warm-up + k timed steps."""
    from oracle import bc as obc
    from oracle import engines as oe
    from pyphasefield_amd.Engines.Warren1995 import _DEFAULTS, warren_derived
    p = warren_derived(dict(_DEFAULTS, alpha=0.01), 4.6e-8)
    yy, xx = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    solid = (yy - n / 2) ** 2 + (xx - n / 2) ** 2 < 64
    f = np.stack([obc.pad(np.where(solid, 0.0, 1.0)), obc.pad(np.where(solid, 0.40, 0.41))])
    obc.apply_bc(f, "NEUMANN", None, 4.6e-8)
    t0 = 0.0
    for s in range(k + 1):
        if s == 1:
            t0 = time.perf_counter()
        a, b = oe.warren_step(f[0], f[1], p, p["dt"], 4.6e-8, seed=1, step=s + 1, vectorize=True)
        f = np.stack([a, b])
        obc.apply_bc(f, "NEUMANN", None, 4.6e-8)
    dt = time.perf_counter() - t0
    return {"value": n * n * k / dt, "unit": "cell-updates/s", "cores": 1, "kind": "port",
            "sample": "%d timed steps (+1 warm-up) of the %dx%d FP64 numpy restatement with np.vectorize "
                      "interpolants (Warren1995.py:26-28), single-threaded numpy" % (k, n, n)}


def _warm(sim, warmup):
    """Untimed steps; engines that replay captured HIP graphs also capture
    theirs here, so the timed region holds replays only."""
    sim.simulate(warmup)
    chunk = sim._graph_chunk()
    if chunk:
        sim.simulate(chunk)


def host_info():
    """CPU model, logical CPUs and OpenMP threads of this host (the CPU
    baseline's context)."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count()
    return {"cpu_model": model, "nproc": usable, "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def copy_peak(steps=20):
    """Achievable HBM bandwidth on this GPU: a 1 GiB device-to-device copy
    (read + write bytes), timed like the engines."""
    import torch
    nbytes = 1 << 30
    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1.0)
    for _ in range(3):
        b.copy_(a)
    t = _timed(lambda k: [b.copy_(a) for _ in range(k)], steps)
    del a, b
    torch.cuda.empty_cache()
    return 2 * nbytes * steps / t / 1e9


# per network 3-64-64-3: (3*64 + 64*64 + 64*3) multiply-adds; two networks per cell
MLP_FLOP_PER_CELL = 2 * 2 * (3 * 64 + 64 * 64 + 64 * 3)
FP16_DENSE_PEAK_TFLOPS = 2500.0  # dense FP16 / BF16 MFMA peak of the MI355X


def ncomp_same_run(steps=100, warmup=20, rounds=3):
    """The N-component workload (2048^2, T0 = 1440 K + 1 K/cell) with the
    surrogate thermodynamics, timed in this process next to the plain FP64
    engine (forward differences: the reference's method and precision) and
    the plain FP32 engine (analytic derivatives).  The networks are trained
    here on 1e6 samples of the plain evaluation (time reported).  The three
    engines are built and warmed first, then timed in `rounds` interleaved
    rounds of `steps` steps; each engine's median round is reported, so all
    three are measured under the same GPU conditions."""
    import torch
    out = {"workload": "N-component 2D 2048x2048, 3 components, 16 seeds r=6, T = %g K + 1 K/cell" % NC_T0}
    t0 = time.perf_counter()
    sur = ncomp_sim(2048, "FP32", "analytic", thermo_eval="surrogate")
    sur.simulate(1)  # uploads, trains the networks, captures nothing yet
    torch.cuda.synchronize()
    out["training_seconds"] = time.perf_counter() - t0
    sims = {"surrogate": sur,
            "plain_fp64_forward_difference": ncomp_sim(2048, "FP64", "forward_difference"),
            "plain_fp32_analytic": ncomp_sim(2048, "FP32", "analytic")}
    for sim in sims.values():
        _warm(sim, warmup)
    times = {k: [] for k in sims}
    for _ in range(rounds):
        for k, sim in sims.items():
            times[k].append(_timed(sim.simulate, steps) / steps)
    med = {k: sorted(v)[len(v) // 2] for k, v in times.items()}
    t_sur = med["surrogate"]
    # cells of the last step handed to the FP64 fallback (outside the trained box)
    padded = 2050 * 2050
    marks = sur._surrogate.packed()[1][-1]
    fallback = int(marks[int(sur.time_step_counter) & 1].item())
    k = 20
    t_mlp = _timed(lambda m: [sur.thermo_pass() for _ in range(m)], k) / k
    tflops = MLP_FLOP_PER_CELL * padded / t_mlp / 1e12
    out["surrogate"] = {"ms_per_step": t_sur * 1e3, "rounds_ms": [t * 1e3 for t in times["surrogate"]],
                        "network_pass_ms": t_mlp * 1e3, "network_tflops": tflops,
                        "network_frac_of_fp16_dense_peak": tflops / FP16_DENSE_PEAK_TFLOPS,
                        "fallback_cell_fraction": fallback / padded}
    for name in ("plain_fp64_forward_difference", "plain_fp32_analytic"):
        out[name] = {"ms_per_step": med[name] * 1e3, "rounds_ms": [t * 1e3 for t in times[name]],
                     "surrogate_speedup": med[name] / t_sur}
    out["steps"] = steps
    out["rounds"] = rounds
    del sims, sur
    torch.cuda.empty_cache()
    return out


def bench_warren(args):
    n = 2048
    sim = warren_sim(n, "FP32")
    if args.graph_chunk is not None:
        sim._GRAPH_CHUNK = args.graph_chunk
    _warm(sim, args.warmup)
    t = _timed(sim.simulate, args.steps)
    cells = n * n
    value = cells * args.steps / t
    roof = _roofline(16.0 * cells, t / args.steps, "k_warren_w<float>")
    cpu = None
    if not args.no_cpu:
        cpu = cpu_warren(n)
        cpu.update(host_info())
    line = _line("cell-updates/s", value, "cell-updates/s", args.steps, args.warmup, t, "f32",
                 {"workload": "Warren1995 2D 2048x2048 FP32, Neumann, seed r=8, noise 0.01",
                  "engine": "Warren1995", "grid": [n, n], "bytes_per_cell_update": 16}, roof, cpu)
    del sim
    line["hbm_copy_peak_GBps"] = copy_peak()
    roof["frac_of_copy_peak"] = roof["achieved"] / line["hbm_copy_peak_GBps"]
    if not args.no_surrogate:
        line["ncomp_surrogate_same_run"] = ncomp_same_run()
    return line


# --------------------------------------------------------------- Diffusion --
def cpu_diffusion(n, k):
    from oracle import bc as obc
    from oracle import engines as oe
    a = obc.pad(np.random.default_rng(0).random((n, n)))
    obc.apply_bc(a[None], "PERIODIC")
    t0 = time.perf_counter()
    for _ in range(k):
        a = oe.diffusion_step(a, 1.0, 0.1, 1.0)
        obc.apply_bc(a[None], "PERIODIC")
    dt = time.perf_counter() - t0
    return {"value": n * n * k / dt, "unit": "cell-updates/s", "cores": 1, "kind": "port",
            "sample": "%d steps of the %dx%d FP64 numpy restatement (np.roll laplacian)" % (k, n, n)}


def _gpu_multistep_max(sim):
    from pyphasefield_amd import _lib
    mx = _lib.lib().ppf_multistep_max(sim._grid, _lib.ptr(sim._boundary_conditions_gpu_device))
    chunk = sim._graph_chunk()
    if mx < 2 or not chunk:
        return 0
    launches = -(-chunk // mx)
    launches += launches % 2
    return chunk // launches if chunk % launches == 0 else 0


def bench_diffusion(args):
    import pyphasefield_amd.Engines as engines
    n = 256
    sim = engines.Diffusion(dimensions=[n, n])
    sim.set_user_data({"D": 1.0})
    sim.set_dx(1.0)
    sim.set_dt(0.1)
    sim.set_boundary_conditions("PERIODIC")
    sim.set_precision("FP64")
    sim.initialize_fields_and_imported_data()
    sim.fields[0].data[1:-1, 1:-1] = np.random.default_rng(0).random((n, n))
    args.warmup = max(args.warmup, sim._GRAPH_CHUNK)  # one graph capture inside the warm-up
    _warm(sim, args.warmup)
    t = _timed(sim.simulate, args.steps)
    cells = n * n
    # each launch of the multi-step kernel reads and writes the field once
    # for up to 8 steps (48-step graph chunks: 6 launches of 8 steps)
    per_launch = _gpu_multistep_max(sim)
    if per_launch >= 2:
        roof = _roofline(16.0 * cells, t / args.steps * per_launch, "k_point2d_multi<double,0,16,8>")
        roof["steps_per_launch"] = per_launch
    else:
        roof = _roofline(16.0 * cells, t / args.steps, "k_point2d<double,0,4>")
    cpu = None if args.no_cpu else cpu_diffusion(n, 1000)
    return _line("cell-updates/s", cells * args.steps / t, "cell-updates/s", args.steps, args.warmup, t, "f64",
                 {"workload": "Diffusion 2D 256x256 FP64 periodic, D=1 dx=1 dt=0.1", "engine": "Diffusion",
                  "grid": [n, n], "bytes_per_cell_update": 16}, roof, cpu)


# ------------------------------------------------------------ CH 3D 512^3 --
def cpu_ch3d(n=256, k=2):
    """FP64 numpy restatement of the Cahn-Hilliard step (mu with its periodic
    2-cell ring, then the Laplacian of mu), n^3 sample, 1 warm-up + k steps."""
    from oracle import bc as obc
    from oracle import engines as oe
    rng = np.random.default_rng(7)
    a = obc.pad(0.5 + 0.01 * (2 * rng.random((n, n, n)) - 1))
    obc.apply_bc(a[None], "PERIODIC")
    t0 = 0.0
    for s in range(k + 1):
        if s == 1:
            t0 = time.perf_counter()
        a = oe.cahnhilliard_step(a, 1.0, 1.0, 0.01, 1.0, "PERIODIC")
        obc.apply_bc(a[None], "PERIODIC")
    dt = time.perf_counter() - t0
    return {"value": n ** 3 * k / dt, "unit": "cell-updates/s", "cores": 1, "kind": "port",
            "sample": "%d timed steps (+1 warm-up) of the FP64 numpy Cahn-Hilliard restatement at %d^3 "
                      "(the benchmark grid is 512^3), single-threaded numpy" % (k, n)}


def bench_ch3d(args):
    import pyphasefield_amd.Engines as engines
    n = args.size or 512
    sim = engines.CahnHilliard(dimensions=[n, n, n], seed=7)
    sim.set_user_data({"M": 1.0, "epsilon": 1.0, "c0": 0.5, "noise": 0.01})
    sim.set_dx(1.0)
    sim.set_dt(0.01)
    sim.set_boundary_conditions("PERIODIC")
    sim.set_precision("FP32")
    sim.initialize_fields_and_imported_data()
    # the warm-up covers at least one graph chunk, so the captured step graph
    # is recorded there and replayed in the timed region
    args.warmup = max(args.warmup, sim._GRAPH_CHUNK)
    _warm(sim, args.warmup)
    t = _timed(sim.simulate, args.steps)
    cells = n ** 3
    from pyphasefield_amd import _lib
    per_launch = _lib.lib().ppf_cahnhilliard_multistep_max(sim._grid) if sim._graph_chunk() else 1
    if per_launch == 2 and args.steps % sim._graph_chunk() == 0:
        # two steps per launch: one read and one write of the field per launch
        roof = _roofline(8.0 * cells, t / args.steps * 2, "k_ch3w2<13>")
        roof["steps_per_launch"] = 2
        roof["frac_per_step_equivalent"] = 8.0 * cells / (t / args.steps) / 1e9 / HBM_PEAK_GBS
    else:
        roof = _roofline(8.0 * cells, t / args.steps, "k_ch3w<float,16>")
    mass = sim.field_sums()[0]
    line = _line("cell-updates/s", cells * args.steps / t, "cell-updates/s", args.steps, args.warmup, t, "f32",
                 {"workload": "Cahn-Hilliard 3D %d^3 FP32 periodic, M=1 kappa=1 dt=0.01" % n,
                  "engine": "CahnHilliard", "grid": [n, n, n], "bytes_per_cell_update": 8}, roof,
                 None if args.no_cpu else cpu_ch3d())
    line["fp64_mass"] = mass
    line["fp64_free_energy"] = sim.free_energy()
    return line


# ---------------------------------------------------------- N-component --
NC_DX = 0.0000046
# T0 about 31 K below the liquidus of the initial liquid (c = 0.30, 0.30, 0.40
# melts at 1471 K in the synthetic model): solid grows at the cold edge and
# melts back up the gradient; far colder starts drive |dG| to tens of times
# the double-well barrier, where the reference scheme breaks interfaces into
# checkerboards
NC_T0 = 1440.0


def ncomp_sim(n, precision, dg=None, seeds=16, noise=0.0, thermo_eval=None, surrogate=None, T0=None):
    """N-component 2048^2: AL-CU-NI synthetic Redlich-Kister model (seed 0),
    16 random-orientation seeds of radius 6, T = 1440 K + 1 K/cell along x,
    Neumann, engine parameters of the Ni-Cu test database."""
    import pyphasefield_amd.Engines as engines
    from pyphasefield_amd import thermo
    model, _ = thermo.synthetic_rk_model(seed=0)
    sim = engines.NCGPU_new(dimensions=[n, n], seed=2024)
    sim.set_framework("GPU_SERIAL")
    sim.set_dx(NC_DX)
    sim.set_temperature_type("LINEAR_GRADIENT")
    sim.set_temperature_initial_T(NC_T0 if T0 is None else T0)
    sim.set_temperature_dTdx(1.0 / NC_DX)
    sim.set_temperature_dTdy(0.0)
    sim.set_temperature_dTdt(0.0)
    sim.set_tdb_container(model)
    sim.set_boundary_conditions("NEUMANN")
    sim.set_precision(precision)
    ud = {"d_ratio": 1.1, "sim_type": "seeds", "number_of_seeds": seeds, "seed_radius": 6,
          "initial_concentration_array": [0.30, 0.30], "noise_phi": noise, "noise_c": noise, "noise_q": noise}
    if dg:
        ud["dg_mode"] = dg
    if thermo_eval:
        ud["thermo_eval"] = thermo_eval
    if surrogate is not None:
        ud["surrogate"] = surrogate
    sim.set_user_data(ud)
    sim.initialize_fields_and_imported_data()
    return sim


def cpu_ncomp(n=512, k=3):
    """FP64 numpy restatement of NComponent_helper_kernel + NComponent_kernel
    on an n x n sample of the same workload, 1 warm-up + k steps."""
    from oracle import bc as obc
    from oracle import engines as oe
    sim = ncomp_sim(n, "FP64", "forward_difference")
    sim._pack_params()
    u = sim.user_data
    model = sim._tdb
    liq = (model.liquid.g0, model.liquid.L, model.liquid.R)
    sol = (model.solid.g0, model.solid.L, model.solid.R)
    f = np.stack([obc.pad(fl.get_cells()) for fl in sim.fields])
    obc.apply_bc(f, "NEUMANN", None, sim.dx)
    T = obc.pad(oe.temperature_linear((n, n), NC_T0, 1.0 / NC_DX, 0.0, sim.dx, 0.0, sim.dt, 0))
    obc.apply_bc_single(T, "NEUMANN")
    t0 = 0.0
    for s in range(k + 1):
        if s == 1:
            t0 = time.perf_counter()
        tr = oe.ncomp_helper(f, T, u["params"], u["c_params"], liq, sol, seed=sim.seed, step=s + 1)
        f = oe.ncomp_main(f, T, tr, u["params"], u["c_params"], seed=sim.seed, step=s + 1)
        obc.apply_bc(f, "NEUMANN", None, sim.dx)
    dt = time.perf_counter() - t0
    return {"value": n * n * k / dt, "unit": "cell-updates/s", "cores": 1, "kind": "port",
            "sample": "%d timed steps (+1 warm-up) of the FP64 numpy restatement of NComponent_helper_kernel + "
                      "NComponent_kernel (forward-difference thermodynamics) on a %dx%d sample of the same "
                      "workload (the benchmark grid is 2048x2048), single-threaded numpy" % (k, n, n)}


def bench_ncomp(args):
    import torch
    n = args.size or 2048
    variant = args.variant
    precision = "FP64" if variant == "fp64" else "FP32"
    dg = "forward_difference" if variant == "fp64" else "analytic"
    sim = ncomp_sim(n, precision, dg, thermo_eval="surrogate" if variant == "surrogate" else None,
                    surrogate=args.surrogate if variant == "surrogate" and args.surrogate and
                    os.path.exists(args.surrogate) else None)
    if args.nc_kernel:
        sim.user_data["ncomp_kernel"] = args.nc_kernel
    if variant == "surrogate":
        sim._setup_surrogate()
        sim._surrogate.split2 = "2" in args.mlp_split
        sim._surrogate.split3 = "3" in args.mlp_split
        sim._surrogate._packed = None
    _warm(sim, args.warmup)
    if variant == "surrogate" and args.surrogate and not os.path.exists(args.surrogate):
        sim.user_data["surrogate"].save(args.surrogate)
    t = _timed(sim.simulate, args.steps)
    cells = n * n
    # per-kernel averages from separate timed loops on the same stream
    k = max(10, args.steps // 10)
    t_thermo = 0.0 if sim._kernel_mode() == "fused" else _timed(lambda m: [sim.thermo_pass() for _ in range(m)], k) / k
    t_total = t / args.steps
    esize = 8 if precision == "FP64" else 4
    nf = len(sim.fields)
    bytes_cell = 2 * nf * esize
    kernels = sim.kernel_names()
    roof = _roofline(bytes_cell * cells, t_total, "+".join(kernels))
    roof["seconds_per_step_thermo"] = t_thermo
    roof["seconds_per_step_stencil"] = t_total - t_thermo
    line = _line("cell-updates/s", cells * args.steps / t, "cell-updates/s", args.steps, args.warmup, t,
                 "f64" if precision == "FP64" else "f32",
                 {"workload": "N-component 2D %dx%d, 3 components (AL, CU, NI), %s" % (n, n, variant),
                  "engine": "NCGPU_new", "grid": [n, n], "fields": nf, "bytes_per_cell_update": bytes_cell,
                  "thermo": sim.user_data.get("thermo_eval", "plain") + "/" + dg}, roof,
                 None if args.no_cpu else cpu_ncomp())
    if variant == "surrogate":
        # the plain GPU engines on the same workload, timed in this run
        del sim
        torch.cuda.empty_cache()
        ks = max(20, args.steps // 5)
        same = {}
        for name, prec, mode in (("plain_fp64_forward_difference", "FP64", "forward_difference"),
                                 ("plain_fp32_analytic", "FP32", "analytic")):
            other = ncomp_sim(n, prec, mode)
            _warm(other, max(2, args.warmup // 2))
            tt = _timed(other.simulate, ks) / ks
            same[name] = {"ms_per_step": tt * 1e3, "surrogate_speedup": tt / t_total, "steps": ks}
            del other
            torch.cuda.empty_cache()
        line["config"]["same_run_plain_engines"] = same
    return line


# -------------------------------------------------------------- copy ref --
def bench_copy(args):
    import torch
    nbytes = 1 << 30
    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1.0)

    def run(k):
        for _ in range(k):
            b.copy_(a)
    run(args.warmup)
    t = _timed(run, args.steps)
    gbs = 2 * nbytes * args.steps / t / 1e9
    return {"metric": "device copy GB/s (read + write)", "value": gbs, "unit": "GB/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": t / args.steps * 1e3,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic", "config": {"workload": "copy 1 GiB fp32 -> fp32"},
            "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": gbs / HBM_PEAK_GBS, "traffic": 2 * nbytes}, "cpu_baseline": None}


# ------------------------------------------------------------- AnisoDorr --
def bench_aniso(args):
    """AnisoDorrGPU 2D, 16 seeds, the reference's constants; FP32 (default) or
    FP64 (--variant fp64).  4 fields read and written per cell-update."""
    import pyphasefield_amd.Engines as engines
    n = args.size or 2048
    precision = "FP64" if args.variant == "fp64" else "FP32"
    sim = engines.AnisoDorrGPU(dimensions=[n, n], seed=5)
    sim.set_boundary_conditions("PERIODIC")
    sim.set_precision(precision)
    sim.set_user_data({"sim_type": "seeds", "number_of_seeds": 16, "initial_concentration_array": [0.40831]})
    sim.initialize_fields_and_imported_data()
    _warm(sim, args.warmup)
    t = _timed(sim.simulate, args.steps)
    cells = n * n
    esize = 8 if precision == "FP64" else 4
    # FP32 grids of 1024 or more 62 x 4 wave tiles run the register-tile kernel
    wave = esize == 4 and -(-n // 62) * -(-n // 4) >= 1024
    kernel = "k_aniso_w<float>" if wave else "k_aniso_fused<%s>" % ("double" if esize == 8 else "float")
    roof = _roofline(2 * 4 * esize * cells, t / args.steps, kernel)
    return _line("cell-updates/s", cells * args.steps / t, "cell-updates/s", args.steps, args.warmup, t,
                 "f64" if precision == "FP64" else "f32",
                 {"workload": "AnisoDorrGPU 2D %dx%d, 16 seeds, %s" % (n, n, precision), "engine": "AnisoDorrGPU",
                  "grid": [n, n], "bytes_per_cell_update": 2 * 4 * esize}, roof, None)


# ---------------------------------------------------- implicit diffusion --
def bench_implicit(args):
    """Diffusion 2D periodic, Crank-Nicolson: --variant adi (two cyclic line
    solves per step, x then y) or cg (full system by conjugate gradients);
    FP64, a = D dt / dx^2 = 4 (eight times the explicit stability limit)."""
    import pyphasefield_amd.Engines as engines
    n = args.size or 2048
    adi = args.variant != "cg"
    sim = engines.Diffusion(dimensions=[n, n], seed=3)
    ud = {"D": 1.0, "solver": "crank-nicolson", "adi": adi}
    if os.environ.get("PPF_SCAN_CHUNK"):
        ud["scan_chunk"] = int(os.environ["PPF_SCAN_CHUNK"])
    sim.set_user_data(ud)
    sim.set_dx(1.0)
    sim.set_dt(4.0)
    sim.set_boundary_conditions("PERIODIC")
    sim.set_precision("FP64")
    sim.initialize_fields_and_imported_data()
    sim.fields[0].data[1:-1, 1:-1] = np.random.default_rng(1).random((n, n))
    _warm(sim, args.warmup)
    t = _timed(sim.simulate, args.steps)
    cells = n * n
    # algorithmic bytes: ADI streams each cell 3 times per line solve (forward,
    # back, correction: 6 x 8 B) for 2 solves; CG is reported per step
    per = 2 * 48.0 if adi else 16.0
    roof = _roofline(per * cells, t / args.steps, "k_line_solve<double>" if adi else "k_cg")
    cfg = {"workload": "Diffusion 2D %dx%d FP64 periodic, Crank-Nicolson %s, a = 4" % (n, n, "ADI" if adi else "CG"),
           "engine": "Diffusion", "grid": [n, n], "bytes_per_cell_update": per}
    if not adi:
        cfg["cg_iterations_last_step"] = getattr(sim, "_cg_iterations", None)
    return _line("cell-updates/s", cells * args.steps / t, "cell-updates/s", args.steps, args.warmup, t, "f64", cfg,
                 roof, None)


# ---------------------------------------------------------------- ensemble --
def bench_ensemble(args):
    """256-instance ensemble: autotune, then time args.steps graph-replayed steps."""
    from pyphasefield_amd.ensemble import Ensemble, random_initial, random_specs
    n = args.size or 256
    specs = random_specs(n, seed=0)
    ens = Ensemble(specs, random_initial(specs, seed=0))
    if getattr(args, "ens_config", None) is not None:
        ens.config = int(args.ens_config)
        ens.blocks = int(args.ens_blocks or 8 * ens.cus)
        best = {"config": ens.config, "blocks": ens.blocks, "shape": list(ens.shape(ens.config))}
    else:
        best = ens.autotune(steps=20, cache=False)
    gs = getattr(args, "graph_steps", None)
    if gs is None:
        gs = 50 if args.steps % 50 == 0 else 2
    ens.simulate(args.warmup - args.warmup % 2 if args.warmup > 1 else 0, graph_steps=gs)
    t = _timed(lambda k: ens.simulate(k, graph_steps=gs), args.steps)
    roof = _roofline(8.0 * ens.cells, t / args.steps, Ensemble.kernel(best["config"]))
    cpu = None if args.no_cpu else cpu_ensemble(specs)
    nac = sum(s.kind == "AC" for s in specs)
    return _line("cell-updates/s", ens.cells * args.steps / t, "cell-updates/s", args.steps, args.warmup, t, "f32",
                 {"workload": "batched ensemble: %d 2D instances (%d Allen-Cahn, %d Cahn-Hilliard), sides "
                              "log-uniform in [64, 1024] per axis, mixed periodic/Neumann, FP32" % (n, nac, n - nac),
                  "instances": n, "cells": ens.cells, "bytes_per_cell_update": 8,
                  "tile_config": best["config"], "tile_shape": best["shape"], "blocks": best["blocks"],
                  "graph_steps": gs}, roof, cpu)


def cpu_ensemble(specs, k=2, every=8):
    """FP64 numpy oracle over every `every`-th instance, k steps each (one
    warm-up) -- a bounded sample of the ensemble's work."""
    from oracle import bc as obc
    from oracle import engines as oe
    from pyphasefield_amd.ensemble import random_initial
    fields = random_initial(specs, seed=0)[::every]
    specs = specs[::every]
    padded = []
    for s, f in zip(specs, fields):
        p = obc.pad(f.astype(np.float64))
        obc.apply_bc(p[None], list(s.bc), None, s.dx)
        padded.append(p)

    def one(i, p):
        s = specs[i]
        if s.kind == "AC":
            q = oe.allencahn_step(p, s.M, s.W, s.eps, s.dt, s.dx)
        else:
            q = oe.cahnhilliard_step(p, s.M, s.eps, s.dt, s.dx, list(s.bc))
        obc.apply_bc(q[None], list(s.bc), None, s.dx)
        return q

    padded = [one(i, p) for i, p in enumerate(padded)]
    t0 = time.perf_counter()
    for _ in range(k):
        padded = [one(i, p) for i, p in enumerate(padded)]
    dt = time.perf_counter() - t0
    cells = sum(s.nx * s.ny for s in specs)
    return {"value": cells * k / dt, "unit": "cell-updates/s", "cores": 1, "kind": "port",
            "sample": "%d steps (+1 warm-up) of %d instances (every %dth of the ensemble, %d cells) with the "
                      "FP64 numpy restatements of engine_CahnAllenExplicit / engine_CahnHilliardExplicit, "
                      "single-threaded numpy" % (k, len(specs), every, cells)}


CONFIGS = {"warren": bench_warren, "diffusion": bench_diffusion, "ch3d": bench_ch3d, "copy": bench_copy,
           "ncomp": bench_ncomp, "aniso": bench_aniso, "implicit": bench_implicit, "ensemble": bench_ensemble}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--config", default="warren", choices=sorted(CONFIGS))
    ap.add_argument("--variant", default="fp32")
    ap.add_argument("--size", type=int, default=None)
    ap.add_argument("--nc-kernel", default=None, choices=["fused", "two_pass"],
                    help="ncomp: one-launch step or the helper + step pair")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--graph-chunk", type=int, default=None,
                    help="warren: steps per captured HIP graph (0: eager launches)")
    ap.add_argument("--no-surrogate", action="store_true",
                    help="warren: skip the same-run surrogate / plain N-component comparison")
    ap.add_argument("--ens-config", type=int, default=None,
                    help="ensemble: tile configuration (skips the autotuner)")
    ap.add_argument("--ens-blocks", type=int, default=None, help="ensemble: persistent grid size")
    ap.add_argument("--graph-steps", type=int, default=None,
                    help="ensemble: steps per captured HIP graph (0: eager launches)")
    ap.add_argument("--mlp-split", default="none",
                    help="ncomp surrogate: layers (2, 3) whose inputs enter as BF16 hi + lo pairs")
    ap.add_argument("--surrogate", default=None,
                    help="ncomp surrogate: load the trained networks from this .npz, or train and save there")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this engine path has no multi-GPU mode "
              "(the domain would shard by slabs with a 1-2 cell halo exchange; not built)", file=sys.stderr)
        sys.exit(2)
    defaults = {"warren": (2000, 100), "diffusion": (4800, 96), "ch3d": (60, 20), "copy": (50, 5),
                "ncomp": (500, 20), "ensemble": (5000, 100), "aniso": (500, 20),
                "implicit": (50, 5)}
    st, wu = defaults[args.config]
    args.steps = args.steps if args.steps is not None else st
    args.warmup = args.warmup if args.warmup is not None else wu
    line = CONFIGS[args.config](args)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
