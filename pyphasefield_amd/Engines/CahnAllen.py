"""Allen-Cahn engine (class CahnAllen, Engines/CahnAllen.py:285-362 of the
reference) on the GPU.

user_data["solver"] = "explicit":
phi <- phi + dt*M*(eps^2 lap(phi) - 16 W (4phi^3 - 6phi^2 + 2phi))
(CahnAllen.py:128-135, D3 resolution: parameters from user_data).
"implicit" / "imex" / "crank-nicolson" (with "gmres" and "adi"): the
semi-implicit schemes of CahnAllen.py:137-280 on periodic boxes, see
``_implicit.allencahn_step``."""

import numpy as np

from ..simulation import Simulation
from . import _gpu, _implicit


class CahnAllen(Simulation):
    # explicit steps read no step counter: Simulation.simulate replays them
    # from a HIP graph; 48 steps = 6 launches of 8 fused steps where the grid
    # allows several steps per launch
    _GRAPH_CHUNK = 48
    _DEFAULT_USER_DATA = {"M": 0.1, "epsilon": 1., "W": 1., "solver": "explicit", "gmres": False, "adi": False}

    def init_fields(self):
        """phi = 0.4995 + 0.001 U[0,1) (CahnAllen.py:307-312), seeded."""
        rng = np.random.default_rng(self.seed)
        phi = 0.001 * rng.random(self.dimensions) + 0.4995
        self.add_field(phi, "phi")

    def just_before_simulating(self):
        if self.dx is None:
            self.dx = 1.0
        super().just_before_simulating()

    def _graph_key(self):
        u = self.user_data
        return (float(u["M"]), float(u["W"]), float(u["epsilon"]), float(self.dt))

    def _graph_chunk(self):
        if self.user_data.get("solver", "explicit") != "explicit":
            return 0
        return super()._graph_chunk()

    def _capture_chunk(self, chunk):
        # explicit steps of 2D periodic / zero-gradient grids fuse into
        # launches of up to 8 steps each (ppf_allencahn_multistep)
        u = self.user_data
        if not _gpu.multistep(self, "ppf_allencahn_multistep", chunk, float(u["M"]), float(u["W"]),
                              float(u["epsilon"]), float(self.dt), nactive=1):
            super()._capture_chunk(chunk)

    def simulation_loop(self):
        if self.user_data.get("solver", "explicit") != "explicit":
            _implicit.allencahn_step(self)
            return
        u = self.user_data
        _gpu.stencil_step(self, "ppf_allencahn_step", float(u["M"]), float(u["W"]), float(u["epsilon"]),
                          float(self.dt), nactive=1)
