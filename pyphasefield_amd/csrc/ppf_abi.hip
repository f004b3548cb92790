// extern "C" entry points of libppf_hip.so: argument validation, dtype
// dispatch and thread-local error reporting.  No exception crosses the ABI.
#include <cstdio>
#include <string>

#include "ppf_launch.h"

namespace ppf {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg) {
  set_error(msg);
  return code;
}

int device_cus() {
  static int cached[64] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  if (cached[dev] == 0) {
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
    cached[dev] = n;
  }
  return cached[dev];
}

int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(int(e), std::string(what) + ": " + hipGetErrorString(e));
  return 0;
}

int check_grid(const ppf_grid* g, int lo, int hi, const char* what) {
  if (!g) return fail(PPF_EINVAL, std::string(what) + ": null grid");
  if (g->ndim < lo || g->ndim > hi)
    return fail(PPF_EINVAL, std::string(what) + ": unsupported ndim " + std::to_string(g->ndim));
  if (g->dtype != PPF_F32 && g->dtype != PPF_F64) return fail(PPF_EINVAL, std::string(what) + ": bad dtype");
  for (int a = 0; a < g->ndim; ++a) {
    if (g->n[a] < 1) return fail(PPF_EINVAL, std::string(what) + ": empty axis");
    if (g->bc[a] < PPF_PERIODIC || g->bc[a] > PPF_DIRICHLET)
      return fail(PPF_EINVAL, std::string(what) + ": bad boundary kind");
  }
  if (g->ndim >= 2 && g->sy < g->n[0] + 2) return fail(PPF_EINVAL, std::string(what) + ": row stride too small");
  if (g->ndim >= 3 && g->sz < g->sy * (g->n[1] + 2))
    return fail(PPF_EINVAL, std::string(what) + ": plane stride too small");
  if (g->nfields < 1) return fail(PPF_EINVAL, std::string(what) + ": no fields");
  if (!(g->dx > 0.0)) return fail(PPF_EINVAL, std::string(what) + ": dx must be positive");
  return 0;
}

int check_temperature(const ppf_temperature* t, const char* what) {
  if (!t) return fail(PPF_EINVAL, std::string(what) + ": null temperature");
  if (t->kind < PPF_T_CONST || t->kind > PPF_T_FRAMES)
    return fail(PPF_EINVAL, std::string(what) + ": bad temperature kind");
  if (t->kind == PPF_T_ARRAY && !t->data) return fail(PPF_EINVAL, std::string(what) + ": null T array");
  if (t->kind == PPF_T_FRAMES && (!t->data || !t->data1 || !(t->t_end > t->t_start)))
    return fail(PPF_EINVAL, std::string(what) + ": thermal history needs two frames with t_end > t_start");
  return 0;
}

// Advance a device-resident step counter by `inc` (one lane; the sum goes
// out through an ordinary vector store).
__global__ void k_step_advance(int64_t* __restrict__ counter, int64_t inc) {
  if (threadIdx.x == 0) counter[threadIdx.x] += inc;
}

}  // namespace ppf

using namespace ppf;

extern "C" {

int ppf_step_counter_add(int64_t* counter, int64_t inc, ppf_stream_t s) {
  if (!counter) return fail(PPF_EINVAL, "step_counter_add: null counter");
  hipLaunchKernelGGL(k_step_advance, dim3(1), dim3(64), 0, stream_of(s), counter, inc);
  return check_launch("step_counter_add");
}

int ppf_abi_version(void) { return PPF_ABI_VERSION; }

const char* ppf_last_error(void) { return g_last_error.c_str(); }

int ppf_apply_bc(const ppf_grid* g, void* fields, const void* bcvals, ppf_stream_t s) {
  if (int rc = check_grid(g, 1, 3, "apply_bc")) return rc;
  if (!fields) return fail(PPF_EINVAL, "apply_bc: null fields");
  return g->dtype == PPF_F32 ? apply_bc<float>(*g, fields, bcvals, stream_of(s))
                             : apply_bc<double>(*g, fields, bcvals, stream_of(s));
}

int ppf_apply_bc_temperature(const ppf_grid* g, void* T, ppf_stream_t s) {
  if (int rc = check_grid(g, 1, 3, "apply_bc_temperature")) return rc;
  if (!T) return fail(PPF_EINVAL, "apply_bc_temperature: null T");
  return g->dtype == PPF_F32 ? apply_bc_temperature<float>(*g, T, stream_of(s))
                             : apply_bc_temperature<double>(*g, T, stream_of(s));
}

int ppf_fill_temperature(const ppf_grid* g, void* out, const ppf_temperature* t, int64_t step, ppf_stream_t s) {
  if (int rc = check_grid(g, 1, 3, "fill_temperature")) return rc;
  if (!out) return fail(PPF_EINVAL, "fill_temperature: null argument");
  if (int rc = check_temperature(t, "fill_temperature")) return rc;
  return g->dtype == PPF_F32 ? fill_temperature<float>(*g, out, *t, step, stream_of(s))
                             : fill_temperature<double>(*g, out, *t, step, stream_of(s));
}

int ppf_diffusion_step(const ppf_grid* g, const void* in, void* out, const void* bcvals, double D, double dt,
                       ppf_stream_t s) {
  if (int rc = check_grid(g, 1, 3, "diffusion_step")) return rc;
  if (!in || !out || in == out) return fail(PPF_EINVAL, "diffusion_step: need distinct in/out buffers");
  return g->dtype == PPF_F32 ? diffusion_step<float>(*g, in, out, bcvals, D, dt, stream_of(s))
                             : diffusion_step<double>(*g, in, out, bcvals, D, dt, stream_of(s));
}

int ppf_allencahn_step(const ppf_grid* g, const void* in, void* out, const void* bcvals, double M, double W,
                       double epsilon, double dt, ppf_stream_t s) {
  if (int rc = check_grid(g, 1, 3, "allencahn_step")) return rc;
  if (!in || !out || in == out) return fail(PPF_EINVAL, "allencahn_step: need distinct in/out buffers");
  return g->dtype == PPF_F32 ? allencahn_step<float>(*g, in, out, bcvals, M, W, epsilon, dt, stream_of(s))
                             : allencahn_step<double>(*g, in, out, bcvals, M, W, epsilon, dt, stream_of(s));
}

int ppf_multistep_max(const ppf_grid* g, const void* bcvals) {
  if (int rc = check_grid(g, 1, 3, "multistep_max")) return rc;
  return multistep_max(*g, bcvals);
}

int ppf_diffusion_multistep(const ppf_grid* g, const void* in, void* out, const void* bcvals, double D, double dt,
                            int32_t nsteps, ppf_stream_t s) {
  if (int rc = check_grid(g, 1, 3, "diffusion_multistep")) return rc;
  if (!in || !out || in == out) return fail(PPF_EINVAL, "diffusion_multistep: need distinct in/out buffers");
  return g->dtype == PPF_F32 ? diffusion_multistep<float>(*g, in, out, bcvals, D, dt, nsteps, stream_of(s))
                             : diffusion_multistep<double>(*g, in, out, bcvals, D, dt, nsteps, stream_of(s));
}

int ppf_allencahn_multistep(const ppf_grid* g, const void* in, void* out, const void* bcvals, double M, double W,
                            double epsilon, double dt, int32_t nsteps, ppf_stream_t s) {
  if (int rc = check_grid(g, 1, 3, "allencahn_multistep")) return rc;
  if (!in || !out || in == out) return fail(PPF_EINVAL, "allencahn_multistep: need distinct in/out buffers");
  return g->dtype == PPF_F32
             ? allencahn_multistep<float>(*g, in, out, bcvals, M, W, epsilon, dt, nsteps, stream_of(s))
             : allencahn_multistep<double>(*g, in, out, bcvals, M, W, epsilon, dt, nsteps, stream_of(s));
}

int ppf_cahnhilliard_multistep_max(const ppf_grid* g) {
  if (int rc = check_grid(g, 1, 3, "cahnhilliard_multistep_max")) return rc;
  return cahnhilliard_multistep_max(*g);
}

int ppf_cahnhilliard_multistep(const ppf_grid* g, const void* in, void* out, const void* bcvals, double M,
                               double epsilon, double dt, int32_t nsteps, ppf_stream_t s) {
  if (int rc = check_grid(g, 1, 3, "cahnhilliard_multistep")) return rc;
  if (!in || !out || in == out) return fail(PPF_EINVAL, "cahnhilliard_multistep: need distinct in/out buffers");
  return g->dtype == PPF_F32
             ? cahnhilliard_multistep<float>(*g, in, out, bcvals, M, epsilon, dt, nsteps, stream_of(s))
             : cahnhilliard_multistep<double>(*g, in, out, bcvals, M, epsilon, dt, nsteps, stream_of(s));
}

int ppf_cahnhilliard_step(const ppf_grid* g, const void* in, void* out, const void* bcvals, double M,
                          double epsilon, double dt, ppf_stream_t s) {
  if (int rc = check_grid(g, 1, 3, "cahnhilliard_step")) return rc;
  if (!in || !out || in == out) return fail(PPF_EINVAL, "cahnhilliard_step: need distinct in/out buffers");
  return g->dtype == PPF_F32 ? cahnhilliard_step<float>(*g, in, out, bcvals, M, epsilon, dt, stream_of(s))
                             : cahnhilliard_step<double>(*g, in, out, bcvals, M, epsilon, dt, stream_of(s));
}

static int warren_dispatch(const ppf_grid* g, const void* in, void* out, const void* bcvals,
                           const ppf_warren_params* p, uint64_t seed, StepRef step, ppf_stream_t s) {
  if (int rc = check_grid(g, 2, 2, "warren_step")) return rc;
  if (g->nfields < 2) return fail(PPF_EINVAL, "warren_step: needs fields (phi, c)");
  if (!in || !out || in == out || !p) return fail(PPF_EINVAL, "warren_step: need distinct in/out and params");
  return g->dtype == PPF_F32 ? warren_step<float>(*g, in, out, bcvals, *p, seed, step, stream_of(s))
                             : warren_step<double>(*g, in, out, bcvals, *p, seed, step, stream_of(s));
}

int ppf_warren_step(const ppf_grid* g, const void* in, void* out, const void* bcvals, const ppf_warren_params* p,
                    uint64_t seed, int64_t step, ppf_stream_t s) {
  return warren_dispatch(g, in, out, bcvals, p, seed, step_host(step), s);
}

int ppf_warren_step_dev(const ppf_grid* g, const void* in, void* out, const void* bcvals,
                        const ppf_warren_params* p, uint64_t seed, const int64_t* step_base, int64_t step_offset,
                        ppf_stream_t s) {
  if (!step_base) return fail(PPF_EINVAL, "warren_step_dev: null step counter");
  return warren_dispatch(g, in, out, bcvals, p, seed, StepRef{step_base, step_offset}, s);
}

constexpr int32_t kNcFlags = PPF_NC_DERIV_FIRST | PPF_NC_SHARED_NOISE | PPF_NC_CLAMP_WIDE | PPF_NC_EBAR_FROM_S;

static int check_ncomp(const ppf_grid* g, const ppf_grid* tg, const ppf_ncomp_params* p, const char* what) {
  if (int rc = check_grid(g, 2, 2, what)) return rc;
  if (!p || !tg) return fail(PPF_EINVAL, std::string(what) + ": null params");
  if (p->ncomp < 2 || p->ncomp > PPF_MAX_COMP) return fail(PPF_EINVAL, std::string(what) + ": ncomp out of range");
  if (p->flags & ~kNcFlags) return fail(PPF_EINVAL, std::string(what) + ": unknown flags");
  if (g->nfields != p->ncomp + 2) return fail(PPF_EINVAL, std::string(what) + ": field stack must hold N+2 fields");
  if (tg->nfields != 2 * p->ncomp - 1 || tg->n[0] != g->n[0] || tg->n[1] != g->n[1] || tg->dtype != g->dtype ||
      tg->sy < g->n[0] + 2)
    return fail(PPF_EINVAL, std::string(what) + ": transfer grid does not match");
  return 0;
}

static int ncomp_thermo_dispatch(const ppf_grid* g, const void* in, const void* bcvals, const ppf_temperature* t,
                                 const ppf_grid* tg, void* transfer, const ppf_ncomp_params* p, const ppf_thermo* th,
                                 uint64_t seed, StepRef step, ppf_stream_t s) {
  if (int rc = check_ncomp(g, tg, p, "ncomp_thermo")) return rc;
  if (!in || !transfer || !t || !th) return fail(PPF_EINVAL, "ncomp_thermo: null argument");
  if (int rc = check_temperature(t, "ncomp_thermo")) return rc;
  if (th->liquid.ncomp != p->ncomp || th->solid.ncomp != p->ncomp)
    return fail(PPF_EINVAL, "ncomp_thermo: thermodynamic model has a different component count");
  if (g->dtype == PPF_F32 && th->dg_mode == PPF_DG_FORWARD_DIFF)
    return fail(PPF_EINVAL, "ncomp_thermo: the h=1e-7 forward difference needs FP64; use PPF_DG_ANALYTIC");
  return g->dtype == PPF_F32
             ? ncomp_thermo<float>(*g, in, bcvals, *t, *tg, transfer, *p, *th, seed, step, stream_of(s))
             : ncomp_thermo<double>(*g, in, bcvals, *t, *tg, transfer, *p, *th, seed, step, stream_of(s));
}

int ppf_ncomp_thermo(const ppf_grid* g, const void* in, const void* bcvals, const ppf_temperature* t,
                     const ppf_grid* tg, void* transfer, const ppf_ncomp_params* p, const ppf_thermo* th,
                     uint64_t seed, int64_t step, ppf_stream_t s) {
  return ncomp_thermo_dispatch(g, in, bcvals, t, tg, transfer, p, th, seed, step_host(step), s);
}

int ppf_ncomp_thermo_dev(const ppf_grid* g, const void* in, const void* bcvals, const ppf_temperature* t,
                         const ppf_grid* tg, void* transfer, const ppf_ncomp_params* p, const ppf_thermo* th,
                         uint64_t seed, const int64_t* step_base, int64_t step_offset, ppf_stream_t s) {
  if (!step_base) return fail(PPF_EINVAL, "ncomp_thermo_dev: null step counter");
  return ncomp_thermo_dispatch(g, in, bcvals, t, tg, transfer, p, th, seed, StepRef{step_base, step_offset}, s);
}

static int ncomp_step_dispatch(const ppf_grid* g, const void* in, void* out, const void* bcvals,
                               const ppf_temperature* t, const ppf_grid* tg, const void* transfer,
                               const ppf_ncomp_params* p, uint64_t seed, StepRef step, ppf_stream_t s) {
  if (int rc = check_ncomp(g, tg, p, "ncomp_step")) return rc;
  if (!in || !out || in == out || !transfer || !t) return fail(PPF_EINVAL, "ncomp_step: null argument");
  if (int rc = check_temperature(t, "ncomp_step")) return rc;
  return g->dtype == PPF_F32
             ? ncomp_step<float>(*g, in, out, bcvals, *t, *tg, transfer, *p, seed, step, stream_of(s))
             : ncomp_step<double>(*g, in, out, bcvals, *t, *tg, transfer, *p, seed, step, stream_of(s));
}

int ppf_ncomp_step(const ppf_grid* g, const void* in, void* out, const void* bcvals, const ppf_temperature* t,
                   const ppf_grid* tg, const void* transfer, const ppf_ncomp_params* p, uint64_t seed,
                   int64_t step, ppf_stream_t s) {
  return ncomp_step_dispatch(g, in, out, bcvals, t, tg, transfer, p, seed, step_host(step), s);
}

int ppf_ncomp_step_dev(const ppf_grid* g, const void* in, void* out, const void* bcvals, const ppf_temperature* t,
                       const ppf_grid* tg, const void* transfer, const ppf_ncomp_params* p, uint64_t seed,
                       const int64_t* step_base, int64_t step_offset, ppf_stream_t s) {
  if (!step_base) return fail(PPF_EINVAL, "ncomp_step_dev: null step counter");
  return ncomp_step_dispatch(g, in, out, bcvals, t, tg, transfer, p, seed, StepRef{step_base, step_offset}, s);
}

static int ncomp_fused_dispatch(const ppf_grid* g, const void* in, void* out, const void* bcvals,
                                const ppf_temperature* t, const ppf_ncomp_params* p, const ppf_thermo* th,
                                uint64_t seed, StepRef step, ppf_stream_t s) {
  if (int rc = check_grid(g, 2, 2, "ncomp_fused_step")) return rc;
  if (!p || !th || !t) return fail(PPF_EINVAL, "ncomp_fused_step: null argument");
  if (p->ncomp < 2 || p->ncomp > PPF_MAX_COMP) return fail(PPF_EINVAL, "ncomp_fused_step: ncomp out of range");
  if (p->flags & ~kNcFlags) return fail(PPF_EINVAL, "ncomp_fused_step: unknown flags");
  if (g->nfields != p->ncomp + 2) return fail(PPF_EINVAL, "ncomp_fused_step: field stack must hold N+2 fields");
  if (!in || !out || in == out) return fail(PPF_EINVAL, "ncomp_fused_step: need distinct in/out buffers");
  if (int rc = check_temperature(t, "ncomp_fused_step")) return rc;
  if (th->liquid.ncomp != p->ncomp || th->solid.ncomp != p->ncomp)
    return fail(PPF_EINVAL, "ncomp_fused_step: thermodynamic model has a different component count");
  if (g->dtype == PPF_F32 && th->dg_mode == PPF_DG_FORWARD_DIFF)
    return fail(PPF_EINVAL, "ncomp_fused_step: the h=1e-7 forward difference needs FP64; use PPF_DG_ANALYTIC");
  return g->dtype == PPF_F32 ? ncomp_fused_step<float>(*g, in, out, bcvals, *t, *p, *th, seed, step, stream_of(s))
                             : ncomp_fused_step<double>(*g, in, out, bcvals, *t, *p, *th, seed, step, stream_of(s));
}

int ppf_ncomp_fused_step(const ppf_grid* g, const void* in, void* out, const void* bcvals, const ppf_temperature* t,
                         const ppf_ncomp_params* p, const ppf_thermo* th, uint64_t seed, int64_t step,
                         ppf_stream_t s) {
  return ncomp_fused_dispatch(g, in, out, bcvals, t, p, th, seed, step_host(step), s);
}

int ppf_ncomp_fused_step_dev(const ppf_grid* g, const void* in, void* out, const void* bcvals,
                             const ppf_temperature* t, const ppf_ncomp_params* p, const ppf_thermo* th,
                             uint64_t seed, const int64_t* step_base, int64_t step_offset, ppf_stream_t s) {
  if (!step_base) return fail(PPF_EINVAL, "ncomp_fused_step_dev: null step counter");
  return ncomp_fused_dispatch(g, in, out, bcvals, t, p, th, seed, StepRef{step_base, step_offset}, s);
}

static int check_aniso(const ppf_grid* g, const ppf_grid* tg, const ppf_aniso_params* p, const char* what) {
  if (int rc = check_grid(g, 2, 2, what)) return rc;
  if (!p || !tg) return fail(PPF_EINVAL, std::string(what) + ": null params");
  if (g->nfields != 4) return fail(PPF_EINVAL, std::string(what) + ": field stack must hold phi, q1, q4, c");
  if (tg->nfields != 6 || tg->n[0] != g->n[0] || tg->n[1] != g->n[1] || tg->dtype != g->dtype ||
      tg->sy < g->n[0] + 2 || tg->sf < tg->sy * (g->n[1] + 2))
    return fail(PPF_EINVAL, std::string(what) + ": transfer grid does not match");
  if (!(p->T > 0.0) || !(p->R > 0.0) || !(p->beta > 0.0))
    return fail(PPF_EINVAL, std::string(what) + ": T, R and beta must be positive");
  return 0;
}

int ppf_aniso_helper(const ppf_grid* g, const void* in, const void* bcvals, const ppf_grid* tg, void* transfer,
                     const ppf_aniso_params* p, ppf_stream_t s) {
  if (int rc = check_aniso(g, tg, p, "aniso_helper")) return rc;
  if (!in || !transfer) return fail(PPF_EINVAL, "aniso_helper: null argument");
  return g->dtype == PPF_F32 ? aniso_helper<float>(*g, in, bcvals, *tg, transfer, *p, stream_of(s))
                             : aniso_helper<double>(*g, in, bcvals, *tg, transfer, *p, stream_of(s));
}

int ppf_aniso_step(const ppf_grid* g, const void* in, void* out, const void* bcvals, const ppf_grid* tg,
                   const void* transfer, const ppf_aniso_params* p, uint64_t seed, int64_t step, ppf_stream_t s) {
  if (int rc = check_aniso(g, tg, p, "aniso_step")) return rc;
  if (!in || !out || in == out || !transfer) return fail(PPF_EINVAL, "aniso_step: null or aliased argument");
  return g->dtype == PPF_F32
             ? aniso_step<float>(*g, in, out, bcvals, *tg, transfer, *p, seed, step, stream_of(s))
             : aniso_step<double>(*g, in, out, bcvals, *tg, transfer, *p, seed, step, stream_of(s));
}

static int aniso_fused_dispatch(const ppf_grid* g, const void* in, void* out, const void* bcvals,
                                const ppf_aniso_params* p, uint64_t seed, StepRef step, ppf_stream_t s) {
  if (int rc = check_grid(g, 2, 2, "aniso_fused_step")) return rc;
  if (!p) return fail(PPF_EINVAL, "aniso_fused_step: null params");
  if (g->nfields != 4) return fail(PPF_EINVAL, "aniso_fused_step: field stack must hold phi, q1, q4, c");
  if (!(p->T > 0.0) || !(p->R > 0.0) || !(p->beta > 0.0))
    return fail(PPF_EINVAL, "aniso_fused_step: T, R and beta must be positive");
  if (!in || !out || in == out) return fail(PPF_EINVAL, "aniso_fused_step: null or aliased argument");
  return g->dtype == PPF_F32 ? aniso_fused_step<float>(*g, in, out, bcvals, *p, seed, step, stream_of(s))
                             : aniso_fused_step<double>(*g, in, out, bcvals, *p, seed, step, stream_of(s));
}

int ppf_aniso_fused_step(const ppf_grid* g, const void* in, void* out, const void* bcvals,
                         const ppf_aniso_params* p, uint64_t seed, int64_t step, ppf_stream_t s) {
  return aniso_fused_dispatch(g, in, out, bcvals, p, seed, step_host(step), s);
}

int ppf_aniso_fused_step_dev(const ppf_grid* g, const void* in, void* out, const void* bcvals,
                             const ppf_aniso_params* p, uint64_t seed, const int64_t* step_base,
                             int64_t step_offset, ppf_stream_t s) {
  if (!step_base) return fail(PPF_EINVAL, "aniso_fused_step_dev: null step counter");
  return aniso_fused_dispatch(g, in, out, bcvals, p, seed, StepRef{step_base, step_offset}, s);
}

int64_t ppf_reduce_work_bytes(const ppf_grid* g) {
  if (!g) return 0;
  return int64_t(kReduceBlocks) * int64_t(g->nfields > 0 ? g->nfields : 1) * int64_t(sizeof(double));
}

int ppf_reduce(const ppf_grid* g, const void* in, const void* bcvals, int kind, double epsilon, double* out,
               void* work, ppf_stream_t s) {
  if (int rc = check_grid(g, 1, 3, "reduce")) return rc;
  if (kind != 0 && kind != 1) return fail(PPF_EINVAL, "reduce: kind must be 0 (sum) or 1 (free energy)");
  if (!in || !out || !work) return fail(PPF_EINVAL, "reduce: null argument");
  return g->dtype == PPF_F32 ? reduce<float>(*g, in, bcvals, kind, epsilon, out, work, stream_of(s))
                             : reduce<double>(*g, in, bcvals, kind, epsilon, out, work, stream_of(s));
}

int ppf_thermo_eval(const ppf_thermo* th, int64_t n, const double* T, const double* c, double* out,
                    ppf_stream_t s) {
  if (!th || !T || !c || !out) return fail(PPF_EINVAL, "thermo_eval: null argument");
  if (n < 0) return fail(PPF_EINVAL, "thermo_eval: negative count");
  if (n == 0) return 0;
  int N = th->liquid.ncomp;
  if (N != th->solid.ncomp) return fail(PPF_EINVAL, "thermo_eval: phases disagree on component count");
  if (N < 2 || N > PPF_MAX_COMP) return fail(PPF_EINVAL, "thermo_eval: ncomp must be 2..4");
  return thermo_eval(*th, n, T, c, out, stream_of(s));
}

}  // extern "C"
