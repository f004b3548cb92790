"""Cahn-Hilliard engine (Engines/CahnHilliard.py:347-421 of the reference)
on the GPU.

user_data["solver"] = "explicit":
mu = -eps^2 lap(phi) + (4phi^3 - 6phi^2 + 2phi);  phi <- phi + dt*M*lap(mu)
(CahnHilliard.py:134-146 with the D3 resolution: dt applied, mu on the +1
ring from a consistent 2-cell halo -- periodic wrap, or zero flux on
non-periodic sides, which conserves mass).
"implicit" / "imex" / "crank-nicolson" (with "gmres" and "adi"): the
semi-implicit schemes of CahnHilliard.py:148-341 on periodic boxes, see
``_implicit.cahnhilliard_step``."""

import numpy as np

from ..simulation import Simulation
from . import _gpu, _implicit


class CahnHilliard(Simulation):
    # explicit steps read no step counter: Simulation.simulate replays them from a HIP graph
    _GRAPH_CHUNK = 20
    _DEFAULT_USER_DATA = {"M": 0.1, "epsilon": 1., "W": 1., "solver": "explicit", "gmres": False, "adi": False}

    def init_fields(self):
        """phi = 0.4995 + 0.001 U[0,1) (CahnHilliard.py:365-370), seeded;
        user_data "c0" and "noise" override the mean and the half-width."""
        rng = np.random.default_rng(self.seed)
        c0 = self.user_data.get("c0", 0.5)
        amp = self.user_data.get("noise", 0.0005)
        phi = c0 + amp * (2 * rng.random(self.dimensions) - 1)
        self.add_field(phi, "phi")

    def just_before_simulating(self):
        if self.dx is None:
            self.dx = 1.0
        super().just_before_simulating()

    def _graph_key(self):
        u = self.user_data
        return (float(u["M"]), float(u["epsilon"]), float(self.dt))

    def _graph_chunk(self):
        if self.user_data.get("solver", "explicit") != "explicit":
            return 0
        return super()._graph_chunk()

    def _capture_chunk(self, chunk):
        # FP32 3D periodic grids: two explicit steps per launch
        # (ppf_cahnhilliard_multistep, one pass of the register march)
        from .. import _lib
        u = self.user_data
        mx = _lib.lib().ppf_cahnhilliard_multistep_max(self._grid)
        if mx < 2 or not _gpu.multistep(self, "ppf_cahnhilliard_multistep", chunk, float(u["M"]),
                                         float(u["epsilon"]), float(self.dt), nactive=1, mx=mx):
            super()._capture_chunk(chunk)

    def simulation_loop(self):
        if self.user_data.get("solver", "explicit") != "explicit":
            _implicit.cahnhilliard_step(self)
            return
        u = self.user_data
        _gpu.stencil_step(self, "ppf_cahnhilliard_step", float(u["M"]), float(u["epsilon"]), float(self.dt),
                          nactive=1)

    def free_energy(self):
        """FP64 F = sum[c^2(1-c)^2 + eps^2/2 |grad c|^2] dx^d on the device."""
        import torch
        from .. import _lib
        if self._fields_gpu_device is None:
            self.send_fields_to_GPU()
        L = _lib.lib()
        dev = self._fields_gpu_device.device
        work = torch.empty(int(L.ppf_reduce_work_bytes(self._grid)) // 8 + 1, dtype=torch.float64, device=dev)
        out = torch.zeros(1, dtype=torch.float64, device=dev)
        _lib.check(L.ppf_reduce(self._grid, _lib.ptr(self._fields_gpu_device),
                                _lib.ptr(self._boundary_conditions_gpu_device), 1,
                                float(self.user_data["epsilon"]), _lib.ptr(out), _lib.ptr(work),
                                _lib.stream_handle()), "reduce")
        return float(out.item())
