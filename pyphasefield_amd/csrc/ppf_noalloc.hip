// ppf_reduce_noalloc: the FP64 conservation reductions of ppf_reduce for
// callers that cannot allocate (code under HIP graph capture, tight serving
// loops, harnesses that forbid hipMalloc): the partial sums live in a
// file-scope device array instead of a caller's workspace.
#include <mutex>
#include <string>

#include "ppf_launch.h"

namespace ppf {

// Partial sums of ppf_reduce_noalloc: a file-scope device array instead of a
// caller's workspace (each GPU holds its own copy until the process ends).
// k_reduce_sum / k_reduce_energy overwrite every slot the final pass reads,
// so nothing needs clearing between calls; calls that overlap on one GPU
// share the array and race.
constexpr int kNoAllocFields = PPF_REDUCE_NOALLOC_MAX_FIELDS;
static __device__ __align__(256) double scratch[kNoAllocFields * kReduceBlocks];

// The array's name is a host-side stand-in, not a device address: its address
// on each device is looked up once, with that device current, and kept.
static std::mutex g_scratch_mu;
static void* g_scratch_dev[64];

static int scratch_address(void** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(int(e), std::string("reduce_noalloc: hipGetDevice: ") + hipGetErrorString(e));
  if (dev < 0 || dev >= 64) return fail(PPF_EINVAL, "reduce_noalloc: device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(e != hipSuccess ? int(e) : PPF_EINVAL,
                  std::string("reduce_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e));
    g_scratch_dev[dev] = p;
  }
  *out = g_scratch_dev[dev];
  return 0;
}

}  // namespace ppf

using namespace ppf;

extern "C" {

int ppf_reduce_noalloc(const ppf_grid* g, const void* in, const void* bcvals, int kind, double epsilon,
                       double* out, ppf_stream_t s) {
  if (int rc = check_grid(g, 1, 3, "reduce_noalloc")) return rc;
  if (kind != 0 && kind != 1) return fail(PPF_EINVAL, "reduce_noalloc: kind must be 0 (sum) or 1 (free energy)");
  if (!in || !out) return fail(PPF_EINVAL, "reduce_noalloc: null argument");
  const int nf = kind == 0 ? g->nfields : 1;
  if (nf > kNoAllocFields)
    return fail(PPF_EINVAL, "reduce_noalloc: at most " + std::to_string(kNoAllocFields) +
                                " fields per call (got " + std::to_string(nf) + "); use ppf_reduce with a workspace");
  void* work = nullptr;
  if (int rc = scratch_address(&work)) return rc;
  return g->dtype == PPF_F32 ? reduce<float>(*g, in, bcvals, kind, epsilon, out, work, stream_of(s))
                             : reduce<double>(*g, in, bcvals, kind, epsilon, out, work, stream_of(s));
}

}  // extern "C"
