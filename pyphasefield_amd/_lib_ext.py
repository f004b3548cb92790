"""ctypes signatures of the surrogate entry points (include/ppf_hip.h)."""

import ctypes

from . import _lib as L

c_i32, c_i64, c_u64, c_f32, c_vp = ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64, ctypes.c_float, ctypes.c_void_p


PPF_MLP_SPLIT3 = 1
PPF_MLP_SPLIT2 = 2


class MlpNet(ctypes.Structure):
    _fields_ = [("weights", c_vp * 2), ("bias", c_vp * 2), ("in_lo", c_f32 * 3), ("in_scale", c_f32 * 3),
                ("out_mu", (c_f32 * 3) * 2), ("out_sigma", (c_f32 * 3) * 2), ("T_lo", c_f32), ("T_hi", c_f32),
                ("c_lo", c_f32), ("flags", c_i32), ("marks", c_vp), ("thermo", c_vp)]


P = ctypes.POINTER
SIGNATURES = {
    "ppf_mlp_gemm": ([ctypes.c_int, ctypes.c_int, ctypes.c_int, c_vp, c_i64, c_i64, c_vp, c_i64, c_i64, c_vp, c_i64,
                      c_f32, c_f32, ctypes.c_int, ctypes.c_int, c_vp], c_i32),
    "ppf_mlp_bias_act": ([c_vp, c_vp, c_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, c_vp], c_i32),
    "ppf_mlp_relu_bwd": ([c_vp, c_vp, c_vp, c_i64, c_vp], c_i32),
    "ppf_mlp_rowsum": ([c_vp, c_vp, ctypes.c_int, ctypes.c_int, c_vp], c_i32),
    "ppf_mlp_mse_partials": ([c_i64], c_i64),
    "ppf_mlp_mse": ([c_vp, c_vp, c_vp, c_vp, c_i64, c_vp], c_i32),
    "ppf_mlp_adam": ([c_vp, c_vp, c_vp, c_vp, c_i64, c_f32, c_f32, c_f32, c_f32, c_i64, c_vp], c_i32),
    "ppf_ncomp_thermo_mlp": ([P(L.Grid), c_vp, c_vp, P(L.Temperature), P(L.Grid), c_vp, P(L.NcompParams),
                              P(L.Thermo), P(MlpNet), c_u64, c_i64, c_vp], c_i32),
    "ppf_ncomp_thermo_mlp_dev": ([P(L.Grid), c_vp, c_vp, P(L.Temperature), P(L.Grid), c_vp, P(L.NcompParams),
                                  P(L.Thermo), P(MlpNet), c_u64, c_vp, c_i64, c_vp], c_i32),
    "ppf_ensemble_num_configs": ([], ctypes.c_int),
    "ppf_ensemble_config": ([ctypes.c_int, P(c_i32)], ctypes.c_int),
    "ppf_ensemble_step": ([c_vp, c_i32, c_vp, c_i32, P(c_i32), c_i32, c_i32, c_vp, c_vp, c_vp], c_i32),
}
