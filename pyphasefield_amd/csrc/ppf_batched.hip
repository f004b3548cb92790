// Batched ensemble of independent 2D Allen-Cahn / Cahn-Hilliard instances in
// one persistent FP32 launch per time step.
//
// Every instance keeps the single-instance layout (halo-padded, pitched rows)
// inside two shared arenas (in / out, swapped between steps).  The host builds
// a work-list of output tiles (instance, x0, y0) for one tile configuration,
// ordered by instance and row, and splits it into 8 contiguous segments of
// about equal cost, one per XCD.  Block b runs on XCD b % 8 (round-robin
// dispatch) and walks its XCD's segment with stride "blocks on that XCD", so
// neighbouring tiles of an instance are in flight on one XCD at the same time
// and share its L2 for their halos.  The tile bodies are the single-instance
// ones (point_tile, ch_tile), so a batched step equals the per-instance
// ppf_allencahn_step / ppf_cahnhilliard_step FP32 result bit for bit.
#include <hip/hip_runtime.h>

#include "ppf_stencil.inl"

namespace ppf {

constexpr int kEnsXcds = PPF_ENS_XCDS;

struct EnsSeg {
  int32_t start[kEnsXcds + 1];
};

template <int ATX, int ATY, int CTX, int CTY>
struct EnsShape {
  static constexpr int kLdsAC = (ATY + 2) * (ATX + 2);
  static constexpr int kLdsCH = (CTY + 4) * (CTX + 4) + (CTY + 2) * (CTX + 2);
  static constexpr int kLds = kLdsAC > kLdsCH ? kLdsAC : kLdsCH;
};

template <int ATX, int ATY, int CTX, int CTY>
__global__ __launch_bounds__(kThreads) void k_ensemble(const ppf_ens_instance* __restrict__ inst, int32_t ninst,
                                                       const int4* __restrict__ tiles, EnsSeg seg,
                                                       const float* __restrict__ in, float* __restrict__ out) {
  __shared__ float smem[EnsShape<ATX, ATY, CTX, CTY>::kLds];
  // With 8 or more blocks, block b walks segment b % 8 only.  A smaller grid
  // has segments without a block of their own: block b then also walks
  // segments b + G, b + 2G, ... (G = grid size), alone on each.
  const int j = int(blockIdx.x / kEnsXcds);
  for (int xcd = int(blockIdx.x % kEnsXcds); xcd < kEnsXcds; xcd += int(gridDim.x)) {
    const int nb = gridDim.x >= unsigned(kEnsXcds) ? int((gridDim.x - 1 - xcd) / kEnsXcds) + 1 : 1;
    for (int t = seg.start[xcd] + j; t < seg.start[xcd + 1]; t += nb) {
      const int4 td = tiles[t];
      if (td.x < 0 || td.x >= ninst) continue;  // uniform per block
      const ppf_ens_instance I = inst[td.x];
      Dom d;
      d.ndim = 2;
      d.ax = Axis{I.n[0], 1, I.bc[0]};
      d.ay = Axis{I.n[1], I.pitch, I.bc[1]};
      d.az = Axis{1, 0, PPF_PERIODIC};
      d.sf = 0;
      const float* fin = in + I.offset;
      float* fout = out + I.offset;
      if (I.kind == PPF_ENS_AC) {
        const PointParams p{I.M, I.W, I.eps, I.dt};
        point_tile<float, 1, ATX, ATY>(smem, fin, fout, nullptr, d, float(I.dx), p, td.y, td.z);
      } else {
        ch_tile<float, CTX, CTY>(smem, fin, fout, nullptr, d, float(I.dx), float(I.eps), float(I.dt * I.M), td.y,
                                 td.z);
      }
      __syncthreads();  // the next tile overwrites the LDS
    }
  }
}

// Tile configurations (Allen-Cahn TX x TY, Cahn-Hilliard TX x TY) the
// autotuner chooses from.
constexpr int kEnsConfigs[][4] = {{64, 8, 64, 8},  {64, 16, 64, 8},  {64, 16, 64, 16}, {128, 8, 64, 8},
                                  {64, 32, 64, 16}, {128, 16, 128, 8}, {32, 16, 32, 16}, {256, 4, 128, 4}};
constexpr int kEnsNumConfigs = int(sizeof(kEnsConfigs) / sizeof(kEnsConfigs[0]));

template <int C>
void launch_ensemble(dim3 grid, hipStream_t s, const ppf_ens_instance* inst, int32_t ninst, const int4* tiles,
                     const EnsSeg& seg, const float* in, float* out) {
  constexpr int ATX = kEnsConfigs[C][0], ATY = kEnsConfigs[C][1], CTX = kEnsConfigs[C][2],
                CTY = kEnsConfigs[C][3];
  hipLaunchKernelGGL((k_ensemble<ATX, ATY, CTX, CTY>), grid, dim3(kThreads), 0, s, inst, ninst, tiles, seg, in,
                     out);
}

}  // namespace ppf

using namespace ppf;

extern "C" {

int ppf_ensemble_num_configs(void) { return kEnsNumConfigs; }

int ppf_ensemble_config(int config, int32_t* shape) {
  if (config < 0 || config >= kEnsNumConfigs || !shape) return fail(PPF_EINVAL, "ensemble_config: bad config");
  for (int i = 0; i < 4; ++i) shape[i] = kEnsConfigs[config][i];
  return 0;
}

int ppf_ensemble_step(const ppf_ens_instance* instances, int32_t n_instances, const int32_t* tiles,
                      int32_t n_tiles, const int32_t* segments, int32_t config, int32_t blocks, const float* in,
                      float* out, ppf_stream_t s) {
  if (!instances || !tiles || !segments || !in || !out || in == out)
    return fail(PPF_EINVAL, "ensemble_step: null or aliased buffer");
  if (n_instances < 1 || n_tiles < 1) return fail(PPF_EINVAL, "ensemble_step: empty ensemble");
  if (config < 0 || config >= kEnsNumConfigs) return fail(PPF_EINVAL, "ensemble_step: bad tile configuration");
  if (blocks < 1) return fail(PPF_EINVAL, "ensemble_step: need at least one block");
  EnsSeg seg;
  for (int i = 0; i <= kEnsXcds; ++i) {
    seg.start[i] = segments[i];
    if (seg.start[i] < 0 || seg.start[i] > n_tiles || (i > 0 && seg.start[i] < seg.start[i - 1]))
      return fail(PPF_EINVAL, "ensemble_step: segments must be non-decreasing within [0, n_tiles]");
  }
  if (seg.start[0] != 0 || seg.start[kEnsXcds] != n_tiles)
    return fail(PPF_EINVAL, "ensemble_step: segments must cover [0, n_tiles]");
  const dim3 grid{unsigned(blocks), 1u, 1u};
  const int4* t4 = reinterpret_cast<const int4*>(tiles);
  hipStream_t st = stream_of(s);
  switch (config) {
    case 0: launch_ensemble<0>(grid, st, instances, n_instances, t4, seg, in, out); break;
    case 1: launch_ensemble<1>(grid, st, instances, n_instances, t4, seg, in, out); break;
    case 2: launch_ensemble<2>(grid, st, instances, n_instances, t4, seg, in, out); break;
    case 3: launch_ensemble<3>(grid, st, instances, n_instances, t4, seg, in, out); break;
    case 4: launch_ensemble<4>(grid, st, instances, n_instances, t4, seg, in, out); break;
    case 5: launch_ensemble<5>(grid, st, instances, n_instances, t4, seg, in, out); break;
    case 6: launch_ensemble<6>(grid, st, instances, n_instances, t4, seg, in, out); break;
    default: launch_ensemble<7>(grid, st, instances, n_instances, t4, seg, in, out); break;
  }
  return check_launch("ensemble_step");
}

}  // extern "C"
