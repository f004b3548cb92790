"""AnisoDorrGPU engine (Engines/AnisoDorrGPU.py of the reference) on the GPU.

Binary (Ni-Cu) anisotropic polycrystalline solidification with fields
phi, q1, q4 (orientation quaternion) and c_CU, using Warren1995's ideal
solution thermodynamics.  Per step a vertex pass (replaces
AnisoDorr_helper_kernel, :262-375) writes the staggered anisotropy and
diffusion terms for every padded cell, and the step kernel (replaces
AnisoDorr_kernel, :54-245) updates the four fields.  By default both run as
one tiled launch that keeps those terms in LDS; user_data["aniso_two_pass"]
selects the two launches through a transfer stack (same arithmetic).

The reference hard-codes every material constant inside both kernels
(:63-105); here they are user_data entries with those values as defaults
(T, M_qmax, H, T_mA, T_mB, L_A, L_B, s_A, s_B, D_L, D_S, B_A, B_B, v_m, R,
y_e, beta, alpha, d_ratio).  dx defaults to the reference's 4.6e-6 and
dt = dx^2/5/D_L/8 as there.  Deliberate differences:

* the step swaps the field buffers (the reference ``simulation_loop``,
  :592-619, never does, so its fields never advance);
* noise comes from the counter RNG keyed by (seed, step, cell) instead of
  xoroshiro states indexed by the thread id;
* ``sim_type == "seeds"`` reads ``number_of_seeds`` from user_data and draws
  positions and angles from a seeded generator (the reference reads an
  undefined name there and silently falls back to the seedless start).
"""

import numpy as np

from .. import _lib
from ..simulation import Simulation
from ..utils import make_seed


class AnisoDorrGPU(Simulation):
    _DEFAULT_USER_DATA = {"d_ratio": 1. / 0.94, "T": 1574., "M_qmax": 80000000., "H": 1e-11,
                          "T_mA": 1728., "T_mB": 1358., "L_A": 2350., "L_B": 1728., "s_A": 0.000037,
                          "s_B": 0.000029, "D_L": 1e-5, "D_S": 1e-9, "B_A": 0.33, "B_B": 0.39, "v_m": 7.42,
                          "R": 8.314, "y_e": 0.12, "beta": 1.5, "alpha": 0.3}

    def __init__(self, **kwargs):
        super().__init__(**kwargs)
        self._framework = "GPU_SERIAL"  # AnisoDorrGPU.py:452

    def init_fields(self):
        """AnisoDorrGPU.py:463-583: phi, q1, q4 and c_CU."""
        dim = self.dimensions
        if len(dim) != 2:
            raise ValueError("AnisoDorrGPU is a 2D engine")
        self._num_transfer_arrays = 6
        u = self.user_data
        sim_type = u.get("sim_type")
        phi = np.zeros(dim)
        if sim_type == "seed":
            melt_angle = u.get("melt_angle", 0 * np.pi / 8)
            q1 = np.zeros(dim) + np.cos(0.5 * melt_angle)
            q4 = np.zeros(dim) + np.sin(0.5 * melt_angle)
            seed_angle = u.get("seed_angle", 1 * np.pi / 4)
            phi, q1, q4 = make_seed(phi, q1, q4, dim[1] / 2, dim[0] / 2, seed_angle, 5)
        elif sim_type == "seeds":
            q1 = np.zeros(dim) + 1.0
            q4 = np.zeros(dim)
            rng = np.random.default_rng(self.seed)
            for _ in range(int(u.get("number_of_seeds", 1))):
                seed_angle = (rng.random() - 0.5) * np.pi / 2
                x_pos = int(rng.random() * dim[1])
                y_pos = int(rng.random() * dim[0])
                phi, q1, q4 = make_seed(phi, q1, q4, x_pos, y_pos, seed_angle, 5)
        else:
            q1 = np.zeros(dim) + 1.0
            q4 = np.zeros(dim)
        self.add_field(phi, "phi")
        self.add_field(q1, "q1")
        self.add_field(q4, "q4")
        ica = u.get("initial_concentration_array")
        c0 = ica[0] if ica is not None and len(ica) == 1 else 0.5
        self.add_field(np.zeros(dim) + c0, "c_CU")

    def aniso_params(self):
        """The constants of AnisoDorrGPU.py:63-105 from user_data and dx."""
        u = self.user_data
        dx = float(self.dx)
        d = dx * u["d_ratio"]
        s_A, s_B, T_mA, T_mB = u["s_A"], u["s_B"], u["T_mA"], u["T_mB"]
        ebar = np.sqrt(6. * np.sqrt(2.) * s_A * d / T_mA)
        p = _lib.AnisoParams()
        for name in ("T", "M_qmax", "H", "T_mA", "T_mB", "L_A", "L_B", "D_L", "D_S", "v_m", "R", "y_e", "beta",
                     "alpha"):
            setattr(p, name, float(u[name]))
        p.ebar = float(ebar)
        p.eqbar = float(0.5 * ebar)
        p.W_A = float(3. * s_A / (np.sqrt(2.) * T_mA * d))
        p.W_B = float(3. * s_B / (np.sqrt(2.) * T_mB * d))
        p.M_A = float(T_mA * T_mA * u["B_A"] / (6. * np.sqrt(2.) * u["L_A"] * d))
        p.M_B = float(T_mB * T_mB * u["B_B"] / (6. * np.sqrt(2.) * u["L_B"] * d))
        p.dt = float(self.dt)
        return p

    def just_before_simulating(self):
        if self.dx is None:
            self.dx = 0.0000046  # AnisoDorrGPU.py:88
        self.set_time_step_length(self.dx * self.dx / 5. / self.user_data["D_L"] / 8.)
        super().just_before_simulating()
        self._params = self.aniso_params()
        dev = self._fields_gpu_device.device
        self._fields_transfer_gpu_device = _lib.alloc_stack(6, list(self.dimensions), self._torch_dtype(), dev)
        self._tgrid = _lib.make_grid(self._fields_transfer_gpu_device, 2, self._boundary_conditions_type, self.dx)

    # the fused step replays from captured HIP graphs (step counter on the device)
    _GRAPH_CHUNK = 20
    _DEVICE_STEP = True

    def _graph_key(self):
        return (bytes(self._params), int(self.seed))

    def _graph_chunk(self):
        if self.user_data.get("aniso_two_pass", False):
            return 0
        return super()._graph_chunk()

    def simulation_loop(self):
        L = _lib.lib()
        s = _lib.stream_handle()
        bc = _lib.ptr(self._boundary_conditions_gpu_device)
        if not self.user_data.get("aniso_two_pass", False):
            # one tiled launch; the two-pass path below keeps the reference's
            # helper/kernel split (same arithmetic)
            args = (self._grid, _lib.ptr(self._fields_gpu_device), _lib.ptr(self._fields_out_gpu_device), bc,
                    self._params, self.seed)
            src = self._step_source()
            if src[0] == "host":
                rc = L.ppf_aniso_fused_step(*args, src[1], s)
            else:
                rc = L.ppf_aniso_fused_step_dev(*args, src[1], src[2], s)
            _lib.check(rc, "aniso_fused_step")
            self._swap_buffers()
            return
        _lib.check(L.ppf_aniso_helper(self._grid, _lib.ptr(self._fields_gpu_device), bc, self._tgrid,
                                      _lib.ptr(self._fields_transfer_gpu_device), self._params, s), "aniso_helper")
        _lib.check(L.ppf_aniso_step(self._grid, _lib.ptr(self._fields_gpu_device), _lib.ptr(self._fields_out_gpu_device),
                                    bc, self._tgrid, _lib.ptr(self._fields_transfer_gpu_device), self._params,
                                    self.seed, int(self.time_step_counter), s), "aniso_step")
        self._swap_buffers()
