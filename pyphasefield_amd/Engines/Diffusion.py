"""Diffusion engine (Engines/Diffusion.py:635-733 of the reference) on the GPU.

user_data["solver"]:
* "explicit": c <- c + dt*D*lap(c) (Diffusion.py:198-208, with the D1
  resolution: D from user_data).  Replaces DiffusionGPU (DiffusionGPU.py:18-117),
  which omitted dt and could not run.
* "implicit" / "crank-nicolson", with user_data["adi"] and ["gmres"]
  (Diffusion.py:210-633): periodic line solves and conjugate gradients, see
  ``_implicit``."""

import numpy as np

from ..simulation import Simulation
from . import _gpu, _implicit


class Diffusion(Simulation):
    # explicit steps read no step counter: Simulation.simulate replays them
    # from a HIP graph; 48 steps = 6 launches of 8 fused steps where the grid
    # allows several steps per launch
    _GRAPH_CHUNK = 48
    _DEFAULT_USER_DATA = {"D": 0.1, "solver": "explicit", "gmres": False, "adi": False}

    def init_fields(self):
        """Square (cube) of c = 1 in the middle half (Diffusion.py:641-659)."""
        dim = self.dimensions
        c = np.zeros(dim)
        sl = tuple(slice(n // 4, 3 * n // 4) for n in dim)
        c[sl] = 1
        self.add_field(c, "c")

    def just_before_simulating(self):
        if self.dx is None:
            self.dx = 1.0
        super().just_before_simulating()

    def _graph_key(self):
        return (float(self.user_data["D"]), float(self.dt))

    def _graph_chunk(self):
        if self.user_data.get("solver", "explicit") != "explicit":
            return 0  # conjugate gradients read their state back between iterations
        return super()._graph_chunk()

    def _capture_chunk(self, chunk):
        # explicit steps of 2D periodic / zero-gradient grids fuse into
        # launches of up to 8 steps each (ppf_diffusion_multistep)
        if not _gpu.multistep(self, "ppf_diffusion_multistep", chunk, float(self.user_data["D"]), float(self.dt),
                              nactive=1):
            super()._capture_chunk(chunk)

    def simulation_loop(self):
        if self.user_data.get("solver", "explicit") != "explicit":
            _implicit.step(self)
            return
        _gpu.stencil_step(self, "ppf_diffusion_step", float(self.user_data["D"]), float(self.dt), nactive=1)
