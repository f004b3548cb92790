"""The legacy N-component engine (class NCGPU, Engines/NCGPU.py:329-532 of the
reference) on the GPU.

It shares NCGPU_new's fields, TDB parameters and kernels; its differences
are carried by ppf_ncomp_params (flags and values):

* the chemical potentials take one forward difference, along e_1 - e_N, for
  every component (NCGPU.py:287-290, 303 -- exact for two components only,
  the reference's defect D7, reproduced because it is this engine's
  behaviour), with one composition noise draw per cell shared by the
  components (:295-296, amplitude sqrt(2 R T / v_m));
* no phi noise (:185) and no orientation noise (:221-224);
* ebar^2 = 6 sqrt(2) S_1 d / T_M,1 and eqbar^2 = ebar^2 / 4 (:74-75), where
  ebar and eqbar in user_data are the square roots (:366-367);
* phi is clamped to [1e-6, 0.999999] (:229-232).

user_data["params"] has the reference's ten entries (:486-497).  Seeds:
"seed" as in NCGPU_new; "seeds" reads an undefined name in the reference
(:108), which sends it to the all-liquid start -- here user_data
["number_of_seeds"] seeds with random orientation and melt angle 0 are
placed, as the branch intends.  The noise streams come from the counter
RNG (the reference's xoroshiro states are indexed by thread id, :42,:253).
The surrogate thermodynamics are not offered for this engine.
"""

import numpy as np

from ..utils import make_seed
from .NCGPU_new import NCGPU_new

# PPF_NC_DERIV_FIRST | PPF_NC_SHARED_NOISE | PPF_NC_CLAMP_WIDE | PPF_NC_EBAR_FROM_S
_LEGACY_FLAGS = 1 | 2 | 4 | 8


class NCGPU(NCGPU_new):
    # NCGPU.py:333-334
    _DEFAULT_USER_DATA = {"d_ratio": 4.}

    def init_tdb_params(self):
        """NCGPU.py:336-373: NCGPU_new's parameters with eqbar = ebar / 2."""
        super().init_tdb_params()
        u = self.user_data
        u["eqbar"] = 0.5 * u["ebar"]

    def init_fields(self):
        """NCGPU.py:375-474."""
        u = self.user_data
        if u.get("sim_type") != "seeds":
            super().init_fields()
            return
        ncomp = len(self._tdb_components)
        self._num_transfer_arrays = 2 * ncomp - 1
        dim = self.dimensions
        rng = np.random.default_rng(self.seed)
        q1 = np.ones(dim)
        q4 = np.zeros(dim)
        phi = np.zeros(dim)
        radius = u.get("seed_radius", 5)
        for _ in range(int(u.get("number_of_seeds", int(np.prod(dim) // 10000)))):
            seed_angle = (rng.random() - 0.5) * np.pi / 2
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
        """user_data["params"] / ["c_params"] as NCGPU.py:476-497 packs them, and
        the ppf_ncomp_params of the legacy kernels."""
        u = self.user_data
        u["noise_phi"], u["noise_c"], u["noise_q"] = 0., 1., 0.
        super()._pack_params()
        u["params"] = u["params"][:10].copy()
        p = self._ncparams
        p.flags = _LEGACY_FLAGS  # ebar^2, eqbar^2 from S, T_M, d in the kernels
        p.ebar = float(np.sqrt(6. * np.sqrt(2.) * u["S"][1] * u["d"] / u["T_M"][1]))
        p.eqbar = 0.5 * p.ebar

    def just_before_simulating(self):
        if self.user_data.get("thermo_eval", "plain") != "plain":
            raise NotImplementedError("the legacy NCGPU engine evaluates the thermodynamics directly; "
                                      "the surrogate is offered for NCGPU_new")
        super().just_before_simulating()

    def _kernel_mode(self):
        mode = super()._kernel_mode()
        if mode == "surrogate":
            raise NotImplementedError("the surrogate is offered for NCGPU_new only")
        return mode
