"""Batched ensembles of independent 2D Allen-Cahn / Cahn-Hilliard instances.

The reference runs one ``Simulation`` per instance, each stepping with
``engine_CahnAllenExplicit`` (``Engines/CahnAllen.py:128-135``) or
``engine_CahnHilliardExplicit`` (``Engines/CahnHilliard.py:134-146``).  Here
all instances share two device arenas and advance together: one persistent
``ppf_ensemble_step`` launch per step walks a tile work-list over every
instance, and a fixed number of steps is captured once in a HIP graph and
replayed.  Tile shapes and the persistent grid size come from an autotuner
whose choices are cached in a JSON file keyed by device and ensemble.

Instances use FP32, dx given per instance, and periodic or zero-gradient
Neumann boundaries per axis.  Parameters of an instance: Allen-Cahn
(M, W, eps, dt) with dphi/dt = M (eps^2 lap phi - 16 W (4phi^3 - 6phi^2 + 2phi));
Cahn-Hilliard (M, eps, dt) with dc/dt = M lap(-eps^2 lap c + 4c^3 - 6c^2 + 2c).
"""

import ctypes
import hashlib
import json
import os
import time
from dataclasses import dataclass

import numpy as np

from . import _lib

KINDS = {"AC": 0, "CH": 1}
BCS = {"PERIODIC": 0, "NEUMANN": 1}
XCDS = 8
# relative cost of one Cahn-Hilliard cell-update (mu ring + second Laplacian)
CH_WEIGHT = 2.0
# output tile widths (Allen-Cahn, Cahn-Hilliard) of the wave-tile configurations
WAVE_TILE_WIDTHS = (62, 60)


class EnsInstance(ctypes.Structure):
    """include/ppf_hip.h: ppf_ens_instance"""
    _fields_ = [("kind", ctypes.c_int32), ("bc", ctypes.c_int32 * 2), ("pad_", ctypes.c_int32),
                ("n", ctypes.c_int64 * 2), ("pitch", ctypes.c_int64), ("offset", ctypes.c_int64),
                ("M", ctypes.c_double), ("W", ctypes.c_double), ("eps", ctypes.c_double),
                ("dt", ctypes.c_double), ("dx", ctypes.c_double)]


@dataclass
class InstanceSpec:
    kind: str          # "AC" or "CH"
    nx: int
    ny: int
    bc: tuple = ("PERIODIC", "PERIODIC")   # (x, y)
    M: float = 1.0
    eps: float = 1.0
    dt: float = 0.01
    W: float = 1.0
    dx: float = 1.0

    def validate(self):
        if self.kind not in KINDS:
            raise ValueError("instance kind must be AC or CH, not %r" % (self.kind,))
        if self.nx < 1 or self.ny < 1:
            raise ValueError("instance sides must be positive")
        for b in self.bc:
            if b not in BCS:
                raise ValueError("ensemble boundaries are PERIODIC or NEUMANN (zero gradient), not %r" % (b,))
        if not (self.dx > 0 and self.dt > 0 and self.M > 0):
            raise ValueError("dx, dt and M must be positive")


def stable_dt(kind, M, eps, W=1.0, dx=1.0):
    """Explicit-Euler limit of the linearised 2D step (largest |eigenvalue|
    of the 5-point operator, double-well curvature f'' in [-1, 2]):
    Allen-Cahn 2 / (M (8 eps^2/dx^2 + 32 W)), Cahn-Hilliard
    2 / (M (8/dx^2) (8 eps^2/dx^2 + 2))."""
    if kind == "AC":
        return 2.0 / (M * (8.0 * eps * eps / dx ** 2 + 32.0 * W))
    return 2.0 / (M * (8.0 / dx ** 2) * (8.0 * eps * eps / dx ** 2 + 2.0))


def random_specs(n=256, seed=0, side=(64, 1024), M=(0.5, 2.0), kappa=(0.5, 2.0), dt_frac=(0.45, 0.9)):
    """Seeded ensemble: sides floor(lo (hi/lo)^U) drawn per axis (log-uniform),
    kind AC or CH and the boundary of each axis drawn uniformly, M and
    kappa = eps^2 uniform, dt a uniform fraction of the instance's stability
    limit."""
    rng = np.random.default_rng(seed)
    lo, hi = side
    specs = []
    for _ in range(n):
        nx, ny = (int(np.floor(lo * (hi / lo) ** u)) for u in rng.random(2))
        kind = "AC" if rng.random() < 0.5 else "CH"
        bc = tuple("PERIODIC" if rng.random() < 0.5 else "NEUMANN" for _ in range(2))
        m = float(rng.uniform(*M))
        eps = float(np.sqrt(rng.uniform(*kappa)))
        dt = float(rng.uniform(*dt_frac)) * stable_dt(kind, m, eps)
        specs.append(InstanceSpec(kind, nx, ny, bc, m, eps, dt))
    return specs


def random_initial(specs, seed=0):
    """c = 0.5 + U[-0.01, 0.01] (Cahn-Hilliard), phi = 0.5 + U[-0.05, 0.05]
    (Allen-Cahn), interiors only."""
    rng = np.random.default_rng(seed + 1)
    out = []
    for s in specs:
        amp = 0.01 if s.kind == "CH" else 0.05
        out.append((0.5 + rng.uniform(-amp, amp, (s.ny, s.nx))).astype(np.float32))
    return out


def layout(specs):
    """Arena layout: per instance (pitch, offset) with rows rounded up to 64
    floats and padded column 1 on a 256-byte boundary; returns the arena
    length in floats."""
    pitches, offsets = [], []
    cursor = 0
    for s in specs:
        pitch = -(-(s.nx + 2) // 64) * 64
        off = -(-(cursor + 1) // 64) * 64 - 1
        pitches.append(pitch)
        offsets.append(off)
        cursor = off + pitch * (s.ny + 2)
    return pitches, offsets, cursor + 64


def tile_list(specs, shape, ch_weight=CH_WEIGHT, return_cost=False):
    """Work-list [n_tiles, 4] int32 (instance, x0, y0, 0) for one tile
    configuration shape = (AC TX, TY, CH TX, TY), ordered by instance and
    row, and the XCD segments: 9 tile indices splitting the list into 8
    contiguous runs of about equal cost (cells x kind weight + a per-tile
    share for the halo)."""
    rows, cost = [], []
    for i, s in enumerate(specs):
        tx, ty = (shape[0], shape[1]) if s.kind == "AC" else (shape[2], shape[3])
        halo = 1 if s.kind == "AC" else 2
        w = 1.0 if s.kind == "AC" else ch_weight
        for y0 in range(1, s.ny + 1, ty):
            hy = min(ty, s.ny - y0 + 1)
            for x0 in range(1, s.nx + 1, tx):
                hx = min(tx, s.nx - x0 + 1)
                rows.append((i, x0, y0, 0))
                cost.append(w * hx * hy + (tx + 2 * halo) * (ty + 2 * halo) * 0.25)
    tiles = np.asarray(rows, np.int32).reshape(-1, 4)
    c = np.cumsum(cost)
    seg = np.zeros(XCDS + 1, np.int32)
    for k in range(1, XCDS):
        seg[k] = int(np.searchsorted(c, c[-1] * k / XCDS))
    seg[XCDS] = len(tiles)
    seg = np.maximum.accumulate(seg)
    if return_cost:
        return tiles, seg, np.asarray(cost)
    return tiles, seg


def autotune_cache_path():
    return os.environ.get("PPF_AUTOTUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "pyphasefield_amd", "autotune.json"))


def ensemble_key(specs, device_name):
    h = hashlib.sha1()
    for s in specs:
        h.update(("%s:%d:%d;" % (s.kind, s.nx, s.ny)).encode())
    return "ensemble|f32|%s|%d|%s" % (device_name, len(specs), h.hexdigest()[:16])


class Ensemble:
    """Device state of a batched ensemble and its stepping."""

    def __init__(self, specs, initial=None, device="cuda"):
        import torch
        for s in specs:
            s.validate()
        if not specs:
            raise ValueError("empty ensemble")
        self.specs = list(specs)
        self.device = torch.device(device)
        self.pitches, self.offsets, total = layout(self.specs)
        self.total = total
        desc = (EnsInstance * len(self.specs))()
        for i, s in enumerate(self.specs):
            d = desc[i]
            d.kind = KINDS[s.kind]
            d.bc[0], d.bc[1] = BCS[s.bc[0]], BCS[s.bc[1]]
            d.n[0], d.n[1] = s.nx, s.ny
            d.pitch, d.offset = self.pitches[i], self.offsets[i]
            d.M, d.W, d.eps, d.dt, d.dx = s.M, s.W, s.eps, s.dt, s.dx
        raw = np.frombuffer(bytes(desc), np.uint8).copy()
        self._desc = torch.from_numpy(raw).to(self.device)
        self._arena = [torch.zeros(total, dtype=torch.float32, device=self.device) for _ in range(2)]
        self._cur = 0
        self._tiles = {}
        self.cells = int(sum(s.nx * s.ny for s in self.specs))
        self.steps_done = 0
        self.config = 0
        props = torch.cuda.get_device_properties(self.device)
        self.cus = int(props.multi_processor_count)
        self.device_name = str(props.name)
        self.blocks = 4 * self.cus
        self._graph = None
        if initial is not None:
            self.set_fields(initial)

    # ------------------------------------------------------------ fields --
    def set_fields(self, fields):
        """Interior arrays [ny, nx] of every instance (host)."""
        import torch
        if len(fields) != len(self.specs):
            raise ValueError("one field per instance")
        host = np.zeros(self.total, np.float32)
        for i, (s, f) in enumerate(zip(self.specs, fields)):
            f = np.asarray(f, np.float32)
            if f.shape != (s.ny, s.nx):
                raise ValueError("instance %d: field shape %s, expected %s" % (i, f.shape, (s.ny, s.nx)))
            v = host[self.offsets[i]:self.offsets[i] + self.pitches[i] * (s.ny + 2)].reshape(s.ny + 2,
                                                                                              self.pitches[i])
            v[1:-1, 1:s.nx + 1] = f
        self._arena[self._cur].copy_(torch.from_numpy(host))
        self._graph = None

    def fields(self):
        """Interior arrays of every instance (host copies)."""
        host = self._arena[self._cur].cpu().numpy()
        out = []
        for i, s in enumerate(self.specs):
            v = host[self.offsets[i]:self.offsets[i] + self.pitches[i] * (s.ny + 2)].reshape(s.ny + 2,
                                                                                              self.pitches[i])
            out.append(v[1:-1, 1:s.nx + 1].copy())
        return out

    # ---------------------------------------------------------- stepping --
    def _worklist(self, config):
        import torch
        if config not in self._tiles:
            shape = (ctypes.c_int32 * 4)()
            _lib.check(_lib.lib().ppf_ensemble_config(config, shape), "ensemble_config")
            tiles, seg = tile_list(self.specs, tuple(shape))
            self._tiles[config] = (torch.from_numpy(tiles.ravel().copy()).to(self.device),
                                   (ctypes.c_int32 * (XCDS + 1))(*[int(v) for v in seg]), len(tiles))
        return self._tiles[config]

    def _launch(self, src, dst, config, blocks):
        tiles, seg, n = self._worklist(config)
        _lib.check(_lib.lib().ppf_ensemble_step(_lib.ptr(self._desc), len(self.specs), _lib.ptr(tiles), n, seg,
                                                config, blocks, _lib.ptr(src), _lib.ptr(dst),
                                                _lib.stream_handle()), "ensemble_step")

    def step(self, n=1, config=None, blocks=None):
        """n eager steps (one launch each)."""
        c = self.config if config is None else config
        b = self.blocks if blocks is None else blocks
        for _ in range(n):
            self._launch(self._arena[self._cur], self._arena[1 - self._cur], c, b)
            self._cur = 1 - self._cur
        self.steps_done += n

    def _capture(self, k):
        """HIP graph of k (even) steps from arena 0 back to arena 0."""
        import torch
        self._worklist(self.config)
        g = torch.cuda.CUDAGraph()
        torch.cuda.synchronize()
        with torch.cuda.graph(g):
            for i in range(k):
                self._launch(self._arena[i % 2], self._arena[1 - i % 2], self.config, self.blocks)
        return g

    def simulate(self, n, graph_steps=50):
        """n steps: whole graph replays of graph_steps (even) steps, the rest eager."""
        if graph_steps % 2:
            raise ValueError("graph_steps must be even")
        reps = n // graph_steps if graph_steps > 0 else 0
        if reps:
            if self._cur != 0:
                self.step(1)
                n -= 1
                reps = n // graph_steps
            if reps and (self._graph is None or self._graph[1] != (graph_steps, self.config, self.blocks)):
                self._graph = (self._capture(graph_steps), (graph_steps, self.config, self.blocks))
            for _ in range(reps):
                self._graph[0].replay()
            self.steps_done += reps * graph_steps
            n -= reps * graph_steps
        if n:
            self.step(n)

    # ---------------------------------------------------------- autotune --
    def autotune(self, steps=20, blocks_per_cu=(4, 8, 16, 32), configs=None, cache=True, verbose=False):
        """Time every tile configuration x persistent grid size on this
        ensemble (graph of `steps` steps, device state restored afterwards);
        keep the fastest.  Results are cached per device and ensemble."""
        import torch
        key = ensemble_key(self.specs, self.device_name)
        path = autotune_cache_path()
        if cache and os.path.exists(path):
            try:
                with open(path) as f:
                    hit = json.load(f).get(key)
            except (OSError, ValueError):
                hit = None
            if hit:
                self.config, self.blocks = int(hit["config"]), int(hit["blocks"])
                return hit
        saved = self._arena[self._cur].clone()
        cur, done = self._cur, self.steps_done
        nconf = _lib.lib().ppf_ensemble_num_configs()
        results = []
        for c in (range(nconf) if configs is None else configs):
            for bpc in blocks_per_cu:
                self.config, self.blocks = c, bpc * self.cus
                self._graph = None
                self._arena[0].copy_(saved)
                self._cur = 0
                g = self._capture(steps)
                g.replay()
                torch.cuda.synchronize()
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                g.replay()
                e1.record()
                torch.cuda.synchronize()
                ms = e0.elapsed_time(e1) / steps
                results.append((ms, c, self.blocks))
                if verbose:
                    print("config %d blocks %d: %.4f ms/step" % (c, self.blocks, ms), flush=True)
                del g
        ms, c, b = min(results)
        self.config, self.blocks = c, b
        self._graph = None
        self._arena[cur].copy_(saved)
        self._cur, self.steps_done = cur, done
        best = {"config": c, "blocks": b, "ms_per_step": ms, "tried": len(results),
                "shape": list(self.shape(c)), "timestamp": time.time()}
        if cache:
            os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
            try:
                with open(path) as f:
                    db = json.load(f)
            except (OSError, ValueError):
                db = {}
            db[key] = best
            tmp = path + ".tmp"
            with open(tmp, "w") as f:
                json.dump(db, f, indent=1, sort_keys=True)
            os.replace(tmp, path)
        return best

    @staticmethod
    def shape(config):
        shape = (ctypes.c_int32 * 4)()
        _lib.check(_lib.lib().ppf_ensemble_config(config, shape), "ensemble_config")
        return tuple(shape)

    @staticmethod
    def kernel(config):
        """The kernel a configuration runs: wave tiles in registers
        (k_ensemble_w, output widths 62 / 60 = 64 lanes less the halo) or
        block tiles through LDS (k_ensemble)."""
        return "k_ensemble_w" if Ensemble.shape(config)[::2] == WAVE_TILE_WIDTHS else "k_ensemble"

    # ------------------------------------------- Simulation interoperation --
    @classmethod
    def from_simulations(cls, sims):
        """Batch CahnAllen / CahnHilliard Simulation objects (2D, FP32,
        periodic or zero-gradient Neumann): their parameters and current
        fields become the ensemble's instances."""
        specs, fields = [], []
        for sim in sims:
            name = type(sim).__name__
            if name not in ("CahnAllen", "CahnHilliard"):
                raise ValueError("ensembles batch CahnAllen and CahnHilliard engines, not %s" % name)
            if len(sim.dimensions) != 2:
                raise ValueError("ensemble instances are 2D")
            if getattr(sim, "_precision", "FP32") != "FP32":
                raise ValueError("ensemble instances run in FP32")
            bc = sim._boundary_conditions_type
            bc = [bc, bc] if isinstance(bc, str) else list(bc)
            u = sim.user_data
            ny, nx = sim.dimensions
            kind = "AC" if name == "CahnAllen" else "CH"
            specs.append(InstanceSpec(kind, nx, ny, tuple(bc[:2]), float(u["M"]), float(u["epsilon"]),
                                      float(sim.dt), float(u.get("W", 1.0)), float(sim.dx)))
            fields.append(sim.fields[0].get_cells())
        ens = cls(specs, fields)
        ens.sims = list(sims)
        return ens

    def write_back(self):
        """Copy the ensemble's fields into the Simulation objects it was built from."""
        for sim, f in zip(self.sims, self.fields()):
            sim.fields[0].data[1:-1, 1:-1] = f
            sim.time_step_counter = getattr(sim, "time_step_counter", 0) + self.steps_done
