"""Warren1995 binary-alloy solidification engine (Engines/Warren1995.py:133-207
of the reference), on the GPU as one fused (phi, c) stencil kernel.

phi = 0 is solid and phi = 1 liquid, as in the reference
(Warren1995.py:97-100).  Parameters come from user_data (D4 resolution);
derived parameters and dt follow Warren1995.py:197-203.  New optional
user_data keys for the initial condition: "seed_shape" ("diamond", the
reference's default, or "circle"), "seed_radius", "c_solid", "c_liquid".
"""

import numpy as np

from .. import _lib
from ..simulation import Simulation
from . import _gpu

_DEFAULTS = {
    "T_mA": 1728., "T_mB": 1358., "L_A": 2350000000., "L_B": 1728000000., "s_A": 0.37, "s_B": 0.29,
    "D_L": 1e-9, "D_S": 1e-13, "B_A": 0.0033, "B_B": 0.0039, "v_m": 0.00000742, "R": 8.314,
    "y_e": 0.04, "T": 1574., "alpha": 0.3, "diamond_size": 15,
}  # Warren1995.py:138-169


def warren_derived(ud, dx):
    """Warren1995.py:197-203."""
    out = dict(ud)
    d = dx / 0.94
    out["dt"] = dx ** 2 / 5. / ud["D_L"]
    out["d"] = d
    out["ebar"] = np.sqrt(6 * np.sqrt(2) * ud["s_A"] * d / ud["T_mA"])
    out["W_A"] = 3 * ud["s_A"] / (np.sqrt(2) * ud["T_mA"] * d)
    out["W_B"] = 3 * ud["s_B"] / (np.sqrt(2) * ud["T_mB"] * d)
    out["M_A"] = (ud["T_mA"] ** 2) * ud["B_A"] / (6 * np.sqrt(2) * ud["L_A"] * d)
    out["M_B"] = (ud["T_mB"] ** 2) * ud["B_B"] / (6 * np.sqrt(2) * ud["L_B"] * d)
    return out


class Warren1995(Simulation):
    _DEFAULT_USER_DATA = _DEFAULTS
    # steps per captured HIP graph; the noise key comes from a device counter
    _GRAPH_CHUNK = 20
    _DEVICE_STEP = True

    def init_fields(self):
        dim = self.dimensions
        ud = self.user_data
        shape = ud.get("seed_shape", "diamond")
        phi = np.zeros(dim) + 1.
        if shape == "circle":
            r = float(ud.get("seed_radius", 8))
            yy, xx = np.meshgrid(np.arange(dim[0]), np.arange(dim[1]), indexing="ij")
            cy, cx = dim[0] / 2, dim[1] / 2
            solid = (yy - cy) ** 2 + (xx - cx) ** 2 < r * r
            phi[solid] = 0.
        else:
            # diamond of solid (Warren1995.py:177-181)
            ds = ud["diamond_size"]
            for i in range(ds):
                phi[int(dim[0] / 2 - i):int(dim[0] / 2 + i),
                    int(dim[1] / 2 - (ds - i)):int(dim[1] / 2 + (ds - i))] = 0
            solid = phi == 0
        c = np.zeros(dim) + ud.get("c_liquid", 0.40831)
        c[solid] = ud.get("c_solid", ud.get("c_liquid", 0.40831))
        self.add_field(phi, "phi", colormap="gray")
        self.add_field(c, "c", colormap="GnBu")

    def just_before_simulating(self):
        if self.dx is None:
            self.dx = 4.6e-8  # legacy init_Warren1995 cell spacing (Warren1995.py:122)
        super().just_before_simulating()
        self.user_data.update({k: v for k, v in warren_derived(self.user_data, self.dx).items()
                               if k not in ("dt",)})
        self.set_time_step_length(self.get_cell_spacing() ** 2 / 5. / self.user_data["D_L"])
        self._params = self._pack()

    def _pack(self):
        u = self.user_data
        p = _lib.WarrenParams()
        for name in ("T", "T_mA", "T_mB", "L_A", "L_B", "D_L", "D_S", "v_m", "R", "y_e", "alpha", "ebar",
                     "W_A", "W_B", "M_A", "M_B"):
            setattr(p, name, float(u[name]))
        p.dt = float(self.dt)
        return p

    def _graph_key(self):
        return (bytes(self._params), int(self.seed))

    def simulation_loop(self):
        L = _lib.lib()
        args = (self._grid, _lib.ptr(self._fields_gpu_device), _lib.ptr(self._fields_out_gpu_device),
                _lib.ptr(self._boundary_conditions_gpu_device), self._params, self.seed)
        src = self._step_source()
        if src[0] == "host":
            rc = L.ppf_warren_step(*args, src[1], _lib.stream_handle())
        else:
            rc = L.ppf_warren_step_dev(*args, src[1], src[2], _lib.stream_handle())
        _lib.check(rc, "warren_step")
        _gpu.carry_passive(self, 2)
        self._swap_buffers()
