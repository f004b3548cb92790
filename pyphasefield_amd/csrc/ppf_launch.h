// Typed launchers (instantiated for float in ppf_f32.hip and for double in
// ppf_f64.hip) behind the extern "C" entry points of ppf_abi.hip.
#pragma once

#include "ppf_common.h"

namespace ppf {

template <typename R>
int apply_bc(const ppf_grid& g, void* fields, const void* bcvals, hipStream_t s);
template <typename R>
int apply_bc_temperature(const ppf_grid& g, void* T, hipStream_t s);
template <typename R>
int fill_temperature(const ppf_grid& g, void* out, const ppf_temperature& t, int64_t step, hipStream_t s);
template <typename R>
int diffusion_step(const ppf_grid& g, const void* in, void* out, const void* bcvals, double D, double dt,
                   hipStream_t s);
template <typename R>
int allencahn_step(const ppf_grid& g, const void* in, void* out, const void* bcvals, double M, double W,
                   double eps, double dt, hipStream_t s);
// several Diffusion / Allen-Cahn steps in one launch (2D grids whose halos
// are a coordinate remap: every side periodic, or Neumann without bc values)
#ifndef PPF_MULTI_MAX
#define PPF_MULTI_MAX 8
#endif
constexpr int kMultiMax = PPF_MULTI_MAX;  // most steps in one launch
// steps one launch can take on this grid (0: none, use the one-step kernels)
inline int32_t multistep_max(const ppf_grid& g, const void* bcvals) {
  if (g.ndim != 2 || bcvals || g.bc[0] == PPF_DIRICHLET || g.bc[1] == PPF_DIRICHLET) return 0;
  if ((g.n[1] + 2) * g.sy >= (int64_t(1) << 31)) return 0;
  const int64_t m = g.n[0] < g.n[1] ? g.n[0] : g.n[1];
  return int32_t(m < kMultiMax ? m : kMultiMax);
}
template <typename R>
int diffusion_multistep(const ppf_grid& g, const void* in, void* out, const void* bcvals, double D, double dt,
                        int32_t nsteps, hipStream_t s);
template <typename R>
int allencahn_multistep(const ppf_grid& g, const void* in, void* out, const void* bcvals, double M, double W,
                        double eps, double dt, int32_t nsteps, hipStream_t s);
template <typename R>
int cahnhilliard_step(const ppf_grid& g, const void* in, void* out, const void* bcvals, double M,
                      double eps, double dt, hipStream_t s);
// Cahn-Hilliard steps one launch can take: 2 on FP32 3D grids with every side
// periodic, at least 8 cells per axis (the 4-cell halo of two steps wraps
// once at most) and 32-bit plane offsets (the two-step register march); 1
// elsewhere
inline int32_t cahnhilliard_multistep_max(const ppf_grid& g) {
  if (g.ndim != 3 || g.dtype != PPF_F32) return 1;
  for (int a = 0; a < 3; ++a)
    if (g.bc[a] != PPF_PERIODIC || g.n[a] < 8) return 1;
  return g.sz * int64_t(sizeof(float)) < (int64_t(1) << 31) ? 2 : 1;
}
template <typename R>
int cahnhilliard_multistep(const ppf_grid& g, const void* in, void* out, const void* bcvals, double M, double eps,
                           double dt, int32_t nsteps, hipStream_t s);
template <typename R>
int reduce(const ppf_grid& g, const void* in, const void* bcvals, int kind, double eps, double* out,
           void* work, hipStream_t s);
template <typename R>
int warren_step(const ppf_grid& g, const void* in, void* out, const void* bcvals, const ppf_warren_params& p,
                uint64_t seed, StepRef step, hipStream_t s);
template <typename R>
int ncomp_thermo(const ppf_grid& g, const void* in, const void* bcvals, const ppf_temperature& t, const ppf_grid& tg,
                 void* transfer, const ppf_ncomp_params& p, const ppf_thermo& th, uint64_t seed, StepRef step,
                 hipStream_t s);
template <typename R>
int ncomp_step(const ppf_grid& g, const void* in, void* out, const void* bcvals, const ppf_temperature& t,
               const ppf_grid& tg, const void* transfer, const ppf_ncomp_params& p, uint64_t seed,
               StepRef step, hipStream_t s);
template <typename R>
int ncomp_fused_step(const ppf_grid& g, const void* in, void* out, const void* bcvals, const ppf_temperature& t,
                     const ppf_ncomp_params& p, const ppf_thermo& th, uint64_t seed, StepRef step, hipStream_t s);

template <typename R>
int aniso_helper(const ppf_grid& g, const void* in, const void* bcvals, const ppf_grid& tg, void* transfer,
                 const ppf_aniso_params& p, hipStream_t s);
template <typename R>
int aniso_fused_step(const ppf_grid& g, const void* in, void* out, const void* bcvals, const ppf_aniso_params& p,
                     uint64_t seed, StepRef step, hipStream_t s);
template <typename R>
int aniso_step(const ppf_grid& g, const void* in, void* out, const void* bcvals, const ppf_grid& tg,
               const void* transfer, const ppf_aniso_params& p, uint64_t seed, int64_t step, hipStream_t s);

int thermo_eval(const ppf_thermo& th, int64_t n, const double* T, const double* c, double* out, hipStream_t s);

// number of partial sums the reduction kernels write (per field): 8 blocks of
// 256 threads per CU on a 256-CU device
constexpr int kReduceBlocks = 2048;

}  // namespace ppf
