// FP32 instantiation of every engine launcher (bulk stencil arithmetic in
// single precision; FP64 reductions inside the reduce kernels).
#define PPF_FP32_UNIT 1
#include "ppf_ncomp.inl"
#include "ppf_warren.inl"
#include "ppf_aniso.inl"

namespace ppf {
template int apply_bc<float>(const ppf_grid&, void*, const void*, hipStream_t);
template int apply_bc_temperature<float>(const ppf_grid&, void*, hipStream_t);
template int fill_temperature<float>(const ppf_grid&, void*, const ppf_temperature&, int64_t, hipStream_t);
template int diffusion_step<float>(const ppf_grid&, const void*, void*, const void*, double, double, hipStream_t);
template int diffusion_multistep<float>(const ppf_grid&, const void*, void*, const void*, double, double, int32_t,
                                      hipStream_t);
template int cahnhilliard_multistep<float>(const ppf_grid&, const void*, void*, const void*, double, double, double,
                                        int32_t, hipStream_t);
template int allencahn_multistep<float>(const ppf_grid&, const void*, void*, const void*, double, double, double,
                                     double, int32_t, hipStream_t);
template int allencahn_step<float>(const ppf_grid&, const void*, void*, const void*, double, double, double, double,
                                   hipStream_t);
template int cahnhilliard_step<float>(const ppf_grid&, const void*, void*, const void*, double, double, double,
                                      hipStream_t);
template int reduce<float>(const ppf_grid&, const void*, const void*, int, double, double*, void*, hipStream_t);
template int warren_step<float>(const ppf_grid&, const void*, void*, const void*, const ppf_warren_params&, uint64_t,
                            StepRef, hipStream_t);
template int ncomp_thermo<float>(const ppf_grid&, const void*, const void*, const ppf_temperature&, const ppf_grid&,
                             void*, const ppf_ncomp_params&, const ppf_thermo&, uint64_t, StepRef, hipStream_t);
template int aniso_helper<float>(const ppf_grid&, const void*, const void*, const ppf_grid&, void*,
                             const ppf_aniso_params&, hipStream_t);
template int aniso_fused_step<float>(const ppf_grid&, const void*, void*, const void*, const ppf_aniso_params&,
                                 uint64_t, StepRef, hipStream_t);
template int aniso_step<float>(const ppf_grid&, const void*, void*, const void*, const ppf_grid&, const void*,
                           const ppf_aniso_params&, uint64_t, int64_t, hipStream_t);
template int ncomp_step<float>(const ppf_grid&, const void*, void*, const void*, const ppf_temperature&,
                           const ppf_grid&, const void*, const ppf_ncomp_params&, uint64_t, StepRef, hipStream_t);
template int ncomp_fused_step<float>(const ppf_grid&, const void*, void*, const void*, const ppf_temperature&,
                                 const ppf_ncomp_params&, const ppf_thermo&, uint64_t, StepRef, hipStream_t);
}  // namespace ppf
