"""N-component alloy solidification engine (Engines/NCGPU_new.py:340-554 of
the reference) on the GPU.

Per step: a thermodynamic pass (replaces NComponent_helper_kernel,
NCGPU_new.py:256-313) writes G_S - G_L, the solute mobilities and the
chemical potentials for every padded cell; the multi-field stencil (replaces
NComponent_kernel, NCGPU_new.py:37-250) updates phi, the orientation (q1, q4)
and the N-1 compositions.  Boundary conditions are folded into the loads and
the LINEAR_GRADIENT temperature is evaluated inside the kernels.

Thermodynamics come from a ThermoModel (pyphasefield_amd.thermo) passed as
``tdb_container``; user_data["dg_mode"] selects the derivative:
"forward_difference" (the reference's h=1e-7 difference, FP64 only; default
in FP64) or "analytic" (default and only choice in FP32).  Seeds use a
seeded generator and user_data["seed_radius"] (the reference hard-codes 5
and uses unseeded np.random, D8).
"""

import numpy as np

from .. import _lib
from ..simulation import Simulation
from ..utils import make_seed


class NCGPU_new(Simulation):
    # NCGPU_new.py:345-348 (there assigned unconditionally, overriding user values)
    _DEFAULT_USER_DATA = {"d_ratio": 4., "noise_phi": 1., "noise_c": 1., "noise_q": 1.}

    def __init__(self, **kwargs):
        super().__init__(**kwargs)
        self._framework = "GPU_SERIAL"

    def init_tdb_params(self):
        """Engine parameters from the model's TDB-style functions
        (NCGPU_new.py:358-383)."""
        super().init_tdb_params()
        if self._tdb is None:
            raise ValueError("NCGPU_new needs a thermodynamic model (tdb_container)")
        tdb = self._tdb
        comps = self._tdb_components
        u = self.user_data
        u["d"] = self.get_cell_spacing() * u["d_ratio"]
        u["R"] = 8.314
        for key in ("L", "T_M", "S", "B", "W", "M"):
            u[key] = []
        for i, cn in enumerate(comps):
            u["L"].append(tdb.value(cn + "_L"))
            u["T_M"].append(tdb.value(cn + "_TM"))
            u["S"].append(tdb.value(cn + "_S"))
            u["B"].append(tdb.value(cn + "_B"))
            u["W"].append(3 * u["S"][i] / (np.sqrt(2) * u["T_M"][i] * u["d"]))
            u["M"].append((u["T_M"][i] ** 2) * u["B"][i] / (6 * np.sqrt(2) * u["L"][i] * u["d"]) / 1574.)
        u["D_S"] = tdb.value("D_S")
        u["D_L"] = tdb.value("D_L")
        u["v_m"] = tdb.value("V_M")
        u["M_qmax"] = tdb.value("M_Q")
        u["H"] = tdb.value("H")
        u["y_e"] = tdb.value("Y_E")
        u["ebar"] = np.sqrt(6 * np.sqrt(2) * u["S"][1] * u["d"] / u["T_M"][1])
        u["eqbar"] = 0.1 * u["ebar"]
        self.set_time_step_length(self.get_cell_spacing() ** 2 / 5. / u["D_L"] / 8)
        u["beta"] = 1.5

    def init_fields(self):
        """NCGPU_new.py:389-483."""
        ncomp = len(self._tdb_components)
        self._num_transfer_arrays = 2 * ncomp - 1
        dim = self.dimensions
        u = self.user_data
        rng = np.random.default_rng(self.seed)
        melt_angle = u.get("melt_angle", 0.)
        q1 = np.zeros(dim) + np.cos(0.5 * melt_angle)
        q4 = np.zeros(dim) + np.sin(0.5 * melt_angle)
        phi = np.zeros(dim)
        radius = u.get("seed_radius", 5)
        sim_type = u.get("sim_type")
        if sim_type == "seed":
            seed_angle = u.get("seed_angle", 1 * np.pi / 4)
            phi, q1, q4 = make_seed(phi, q1, q4, dim[1] / 2, dim[0] / 2, seed_angle, radius)
        elif sim_type == "seeds":
            number_of_seeds = u.get("number_of_seeds", int(np.prod(dim) // 10000))
            for _ in range(number_of_seeds):
                seed_angle = (rng.random() - 0.5) * np.pi / 2 + melt_angle
                x_pos = int(rng.random() * dim[1])
                y_pos = int(rng.random() * dim[0])
                phi, q1, q4 = make_seed(phi, q1, q4, x_pos, y_pos, seed_angle, radius)
        self.add_field(phi, "phi", colormap="gray")
        self.add_field(q1, "q1")
        self.add_field(q4, "q4")
        ica = u.get("initial_concentration_array")
        if ica is not None and len(ica) + 1 == ncomp:
            for i in range(len(ica)):
                self.add_field(np.zeros(dim) + ica[i], "c_" + self._tdb_components[i])
        else:
            for i in range(ncomp - 1):
                self.add_field(np.zeros(dim) + 1. / ncomp, "c_" + self._tdb_components[i])

    def _pack_params(self):
        """user_data["params"] / ["c_params"] in the order of NCGPU_new.py:497-517
        and the matching ppf_ncomp_params (host only)."""
        u = self.user_data
        for k in ("noise_phi", "noise_c", "noise_q"):
            u.setdefault(k, 1.)
        ncomp = len(self._tdb_components)
        if len(self.dimensions) != 2:
            raise NotImplementedError("the N-component engine runs in 2D")
        if self.temperature is None:
            raise ValueError("the N-component engine needs a temperature field")
        p = _lib.NcompParams()
        p.ncomp = ncomp
        p.dx = float(self.dx)
        for name in ("d", "v_m", "M_qmax", "H", "y_e", "beta", "D_L", "D_S", "ebar", "eqbar",
                     "noise_phi", "noise_c", "noise_q"):
            setattr(p, name, float(u[name]))
        p.dt = float(self.dt)
        for name in ("L", "T_M", "S", "B", "W", "M"):
            arr = getattr(p, name)
            for i in range(ncomp):
                arr[i] = float(u[name][i])
        u["params"] = np.array([self.dx, u["d"], u["v_m"], u["M_qmax"], u["H"], u["y_e"], u["beta"], u["D_L"],
                                u["D_S"], self.dt, u["ebar"], u["eqbar"], u["noise_phi"], u["noise_c"],
                                u["noise_q"]])
        u["c_params"] = np.array([u["L"], u["T_M"], u["S"], u["B"], u["W"], u["M"]])
        self._ncparams = p

    def just_before_simulating(self):
        super().just_before_simulating()
        self._pack_params()
        u = self.user_data
        ncomp = len(self._tdb_components)
        mode = u.get("dg_mode", "forward_difference" if self._precision == "FP64" else "analytic")
        if mode not in ("forward_difference", "analytic"):
            raise ValueError("dg_mode must be 'forward_difference' or 'analytic'")
        self._thermo = self._tdb.to_ctypes(_lib.PPF_DG_FORWARD_DIFF if mode == "forward_difference"
                                           else _lib.PPF_DG_ANALYTIC)
        dev = self._fields_gpu_device.device
        self._fields_transfer_gpu_device = _lib.alloc_stack(2 * ncomp - 1, list(self.dimensions),
                                                            self._torch_dtype(), dev)
        self._tgrid = _lib.make_grid(self._fields_transfer_gpu_device, 2, self._boundary_conditions_type, self.dx)
        self._surrogate = None
        self._thermo_dev = None
        if u.get("thermo_eval", "plain") == "surrogate":
            self._setup_surrogate()
            # cells outside the trained box: the plain FP64 evaluation inside
            # the network kernel ("inline", reading this device copy of the
            # model) or in a second kernel ("kernel")
            fb = u.get("surrogate_fallback", "inline")
            if fb not in ("inline", "kernel"):
                raise ValueError("surrogate_fallback must be 'inline' or 'kernel'")
            if fb == "inline":
                import torch
                raw = bytes(self._thermo)
                self._thermo_dev = torch.frombuffer(bytearray(raw), dtype=torch.uint8).to(dev)

    def _temperature_range(self):
        if self._thermal is not None:
            return self._thermal.range()
        t = self.temperature_descriptor()
        if t.kind == _lib.PPF_T_CONST:
            return t.T0, t.T0
        T = self.temperature._data[1:-1, 1:-1]
        return float(T.min()), float(T.max())

    def _setup_surrogate(self):
        """Use user_data["surrogate"] (a trained SurrogateThermo, or the path of
        one saved with SurrogateThermo.save) or train one on 1e6 samples of the
        plain evaluation over this run's (T, c) box."""
        if self._precision != "FP32":
            raise ValueError("the surrogate thermodynamics run in FP32")
        if len(self._tdb_components) != 3:
            raise ValueError("the surrogate kernel is built for 3 components")
        from ..surrogate import SurrogateThermo
        sur = self.user_data.get("surrogate")
        if isinstance(sur, str):
            sur = SurrogateThermo.load(sur, self._tdb)
            self.user_data["surrogate"] = sur
        if sur is None:
            lo, hi = self._temperature_range()
            pad = max(5.0, 0.02 * (hi - lo))
            sur = SurrogateThermo(self._tdb, (lo - pad, hi + pad), c_lo=0.02, seed=self.seed)
            sur.train(n_samples=int(self.user_data.get("surrogate_samples", 1_000_000)),
                      epochs=int(self.user_data.get("surrogate_epochs", 20)))
            self.user_data["surrogate"] = sur
        self._surrogate = sur

    # steps per captured HIP graph; noise and temperature ramp read the step
    # counter from the device inside a capture
    _GRAPH_CHUNK = 20
    _DEVICE_STEP = True

    def _kernel_mode(self):
        """"fused" (default): one launch per step (ppf_ncomp_fused_step);
        "two_pass": the reference's helper + kernel structure
        (ppf_ncomp_thermo + ppf_ncomp_step, same results); the surrogate
        runs its network pass before the step."""
        if getattr(self, "_surrogate", None) is not None:
            return "surrogate"
        mode = self.user_data.get("ncomp_kernel", "fused")
        if mode not in ("fused", "two_pass"):
            raise ValueError("ncomp_kernel must be 'fused' or 'two_pass'")
        return mode

    def _graph_key(self):
        sur = getattr(self, "_surrogate", None)
        if sur is not None:
            sur.packed()[0].thermo = self._thermo_dev.data_ptr() if self._thermo_dev is not None else None
        net = bytes(sur.packed()[0]) if sur is not None else b""
        return (self._kernel_mode(), bytes(self._ncparams), bytes(self._thermo), bytes(self.temperature_descriptor()),
                int(self.seed), net, self._fields_transfer_gpu_device.data_ptr())

    def kernel_names(self):
        t = "float" if self._precision == "FP32" else "double"
        n = len(self._tdb_components)
        mode = self._kernel_mode()
        # the step from the transfer stack: FP32 2D grids of 1024 or more
        # 62 x 4 tiles run the register-tile kernel (csrc/ppf_ncomp.inl)
        ny, nx = self.dimensions[0], self.dimensions[-1]
        wave = t == "float" and len(self.dimensions) == 2 and -(-nx // 62) * -(-ny // 4) >= 1024
        step = "k_ncomp_step_w<%d>" % n if wave else "k_ncomp_step<%s,%d,false>" % (t, n)
        if mode == "surrogate":
            return ["k_ncomp_thermo_mlp", step]
        tdb = getattr(self._tdb, "device_tables", None) is not None and self._tdb.device_tables() is not None
        if mode == "fused":
            return ["k_ncomp_step_tdb<%s,%d>" % (t, n) if tdb else "k_ncomp_step<%s,%d,true>" % (t, n)]
        return ["k_ncomp_thermo%s<%s,%d>" % ("_tdb" if tdb else "", t, n), step]

    def thermo_pass(self):
        L = _lib.lib()
        src = self._step_source()
        head = (self._grid, _lib.ptr(self._fields_gpu_device), _lib.ptr(self._boundary_conditions_gpu_device),
                self.temperature_descriptor(), self._tgrid, _lib.ptr(self._fields_transfer_gpu_device),
                self._ncparams, self._thermo)
        if self._surrogate is not None:
            net, _keep = self._surrogate.packed()
            net.thermo = self._thermo_dev.data_ptr() if self._thermo_dev is not None else None
            if src[0] == "host":
                rc = L.ppf_ncomp_thermo_mlp(*head, net, self.seed, src[1], _lib.stream_handle())
            else:
                rc = L.ppf_ncomp_thermo_mlp_dev(*head, net, self.seed, src[1], src[2], _lib.stream_handle())
            _lib.check(rc, "ncomp_thermo_mlp")
            return
        if src[0] == "host":
            rc = L.ppf_ncomp_thermo(*head, self.seed, src[1], _lib.stream_handle())
        else:
            rc = L.ppf_ncomp_thermo_dev(*head, self.seed, src[1], src[2], _lib.stream_handle())
        _lib.check(rc, "ncomp_thermo")

    def simulation_loop(self):
        L = _lib.lib()
        src = self._step_source()
        if self._kernel_mode() == "fused":
            args = (self._grid, _lib.ptr(self._fields_gpu_device), _lib.ptr(self._fields_out_gpu_device),
                    _lib.ptr(self._boundary_conditions_gpu_device), self.temperature_descriptor(), self._ncparams,
                    self._thermo, self.seed)
            if src[0] == "host":
                rc = L.ppf_ncomp_fused_step(*args, src[1], _lib.stream_handle())
            else:
                rc = L.ppf_ncomp_fused_step_dev(*args, src[1], src[2], _lib.stream_handle())
            _lib.check(rc, "ncomp_fused_step")
            self._swap_buffers()
            return
        self.thermo_pass()
        args = (self._grid, _lib.ptr(self._fields_gpu_device), _lib.ptr(self._fields_out_gpu_device),
                _lib.ptr(self._boundary_conditions_gpu_device), self.temperature_descriptor(), self._tgrid,
                _lib.ptr(self._fields_transfer_gpu_device), self._ncparams, self.seed)
        if src[0] == "host":
            rc = L.ppf_ncomp_step(*args, src[1], _lib.stream_handle())
        else:
            rc = L.ppf_ncomp_step_dev(*args, src[1], src[2], _lib.stream_handle())
        _lib.check(rc, "ncomp_step")
        self._swap_buffers()
