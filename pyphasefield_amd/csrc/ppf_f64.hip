// FP64 instantiation of every engine launcher.  This translation unit is
// compiled with -ffp-contract=off so that FP64 runs keep the operation order
// of the reference formulas (oracle parity mode).
#include "ppf_ncomp.inl"
#include "ppf_warren.inl"
#include "ppf_aniso.inl"

namespace ppf {
template int apply_bc<double>(const ppf_grid&, void*, const void*, hipStream_t);
template int apply_bc_temperature<double>(const ppf_grid&, void*, hipStream_t);
template int fill_temperature<double>(const ppf_grid&, void*, const ppf_temperature&, int64_t, hipStream_t);
template int diffusion_step<double>(const ppf_grid&, const void*, void*, const void*, double, double, hipStream_t);
template int diffusion_multistep<double>(const ppf_grid&, const void*, void*, const void*, double, double, int32_t,
                                      hipStream_t);
template int cahnhilliard_multistep<double>(const ppf_grid&, const void*, void*, const void*, double, double, double,
                                        int32_t, hipStream_t);
template int allencahn_multistep<double>(const ppf_grid&, const void*, void*, const void*, double, double, double,
                                     double, int32_t, hipStream_t);
template int allencahn_step<double>(const ppf_grid&, const void*, void*, const void*, double, double, double, double,
                                   hipStream_t);
template int cahnhilliard_step<double>(const ppf_grid&, const void*, void*, const void*, double, double, double,
                                      hipStream_t);
template int reduce<double>(const ppf_grid&, const void*, const void*, int, double, double*, void*, hipStream_t);
template int warren_step<double>(const ppf_grid&, const void*, void*, const void*, const ppf_warren_params&, uint64_t,
                            StepRef, hipStream_t);
template int ncomp_thermo<double>(const ppf_grid&, const void*, const void*, const ppf_temperature&, const ppf_grid&,
                             void*, const ppf_ncomp_params&, const ppf_thermo&, uint64_t, StepRef, hipStream_t);
template int aniso_helper<double>(const ppf_grid&, const void*, const void*, const ppf_grid&, void*,
                             const ppf_aniso_params&, hipStream_t);
template int aniso_fused_step<double>(const ppf_grid&, const void*, void*, const void*, const ppf_aniso_params&,
                                 uint64_t, StepRef, hipStream_t);
template int aniso_step<double>(const ppf_grid&, const void*, void*, const void*, const ppf_grid&, const void*,
                           const ppf_aniso_params&, uint64_t, int64_t, hipStream_t);
template int ncomp_step<double>(const ppf_grid&, const void*, void*, const void*, const ppf_temperature&,
                           const ppf_grid&, const void*, const ppf_ncomp_params&, uint64_t, StepRef, hipStream_t);
template int ncomp_fused_step<double>(const ppf_grid&, const void*, void*, const void*, const ppf_temperature&,
                                 const ppf_ncomp_params&, const ppf_thermo&, uint64_t, StepRef, hipStream_t);
// plain thermodynamics at scattered points (surrogate training data), FP64
template <int N>
__global__ __launch_bounds__(256) void k_thermo_eval(ppf_thermo th, int64_t n, const double* __restrict__ T,
                                                     const double* __restrict__ c, double* __restrict__ out) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double cc[N];
  double cN = 1.0;
#pragma unroll
  for (int l = 0; l < N - 1; ++l) {
    cc[l] = c[i * (N - 1) + l];
    cN -= cc[l];
  }
  cc[N - 1] = cN;
  double t = T[i];
  double lnT = plog(t);
  double id = ideal_sum<double, N>(cc);
  double dL[N], dS[N];
  double* o = out + i * 2 * N;
  with_phases<double, N>(th, t, lnT, [&](const auto& PL, const auto& PS) {
    gibbs_grad(cc, PL, dL);
    gibbs_grad(cc, PS, dS);
    o[0] = gibbs(cc, PL, id);
    o[N] = gibbs(cc, PS, id);
  });
#pragma unroll
  for (int l = 0; l < N - 1; ++l) {
    o[1 + l] = dL[l];
    o[N + 1 + l] = dS[l];
  }
}

int thermo_eval(const ppf_thermo& th, int64_t n, const double* T, const double* c, double* out, hipStream_t s) {
  dim3 grid(unsigned((n + 255) / 256));
  switch (th.liquid.ncomp) {
    case 2:
      hipLaunchKernelGGL(k_thermo_eval<2>, grid, dim3(256), 0, s, th, n, T, c, out);
      break;
    case 3:
      hipLaunchKernelGGL(k_thermo_eval<3>, grid, dim3(256), 0, s, th, n, T, c, out);
      break;
    default:
      hipLaunchKernelGGL(k_thermo_eval<4>, grid, dim3(256), 0, s, th, n, T, c, out);
      break;
  }
  return check_launch("thermo_eval");
}

}  // namespace ppf
