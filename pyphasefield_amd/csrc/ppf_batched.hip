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
// and share its L2 for their halos.  Two tile forms: block tiles through LDS
// (k_ensemble, the single-instance bodies point_tile / ch_tile) and wave tiles
// held in registers (k_ensemble_w, the same expressions in the same order);
// either way a batched step equals the per-instance ppf_allencahn_step /
// ppf_cahnhilliard_step FP32 result bit for bit.
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

// ------------------------------------------------------------ wave tiles --
// Each wave owns a tile of its own and holds it in registers: lane l holds
// column x0 - H + l of the rows y0 - H .. y0 + TY - 1 + H (halo H = 1 for
// Allen-Cahn, 2 for Cahn-Hilliard), so 64 lanes give 64 - 2H output columns.
// x neighbours come from the adjacent lanes (DPP lane shifts), y neighbours
// from the lane's own rows.  No LDS and no barrier: the four waves of a block
// walk the work-list independently, each with all its rows' loads in flight
// at once.  Halo rows and columns are remapped into the load addresses
// (periodic wrap, or the nearest interior cell for zero gradient); the
// Cahn-Hilliard mu ring takes the zero-flux copies on non-periodic sides.
// The arithmetic is point_tile_body's and ch_tile_body's, expression for
// expression, so results equal the LDS tiles and the single-instance kernels
// bit for bit.
constexpr int kWaveLanes = 64;

// f[row][lane's column]: uniform 64-bit row base, 32-bit lane byte offset
__device__ __forceinline__ float wave_row_load(const float* __restrict__ f, int32_t row, int64_t pitch,
                                               uint32_t xbytes) {
  const char* base = reinterpret_cast<const char*>(f + int64_t(row) * pitch);
  return *reinterpret_cast<const float*>(base + xbytes);
}

template <int TY>
__device__ __forceinline__ void ac_wave_tile(const float* __restrict__ f, float* __restrict__ o,
                                             const ppf_ens_instance& I, int32_t x0, int32_t y0, int lane) {
  const int32_t nx = int32_t(I.n[0]), ny = int32_t(I.n[1]);
  const int64_t pitch = I.pitch;
  const int32_t xr = x0 - 1 + lane;
  const uint32_t xb = uint32_t(remap_coord32(xr, nx, I.bc[0])) * uint32_t(sizeof(float));
  float v[TY + 2];  // rows y0 - 1 .. y0 + TY
#pragma unroll
  for (int r = 0; r < TY + 2; ++r) v[r] = wave_row_load(f, remap_coord32(y0 - 1 + r, ny, I.bc[1]), pitch, xb);
  const PointParams p{I.M, I.W, I.eps, I.dt};
  const float dx = float(I.dx);
  const float idx = 1.f / dx;
  const float idx2 = idx * idx;
  const bool out_lane = lane >= 1 && lane < kWaveLanes - 1 && xr <= nx;
  float* __restrict__ orow = o + int64_t(y0) * pitch + xr;
#pragma unroll
  for (int r = 0; r < TY; ++r) {
    const float c = v[r + 1];
    // point_tile_body's order: y (c[y+1] + c[y-1]), then x (c[x+1] + c[x-1])
    float lap = (v[r + 2] + v[r] - 2.f * c) * idx2;
    lap += (lane_shift<kDppWaveShl1>(c) + lane_shift<kDppWaveShr1>(c) - 2.f * c) * idx2;
    const float res = point_update<float, 1>(c, lap, p);
    if (out_lane && y0 + r <= ny) orow[r * pitch] = res;
  }
}

template <int TY>
__device__ __forceinline__ void ch_wave_tile(const float* __restrict__ f, float* __restrict__ o,
                                             const ppf_ens_instance& I, int32_t x0, int32_t y0, int lane) {
  const int32_t nx = int32_t(I.n[0]), ny = int32_t(I.n[1]);
  const int64_t pitch = I.pitch;
  const int32_t xr = x0 - 2 + lane;
  const uint32_t xb = uint32_t(remap_coord32(xr, nx, I.bc[0])) * uint32_t(sizeof(float));
  float v[TY + 4];  // c on rows y0 - 2 .. y0 + TY + 1
#pragma unroll
  for (int r = 0; r < TY + 4; ++r) v[r] = wave_row_load(f, remap_coord32(y0 - 2 + r, ny, I.bc[1]), pitch, xb);
  const float dx = float(I.dx), eps = float(I.eps), dtM = float(I.dt * I.M);
  const float idx = 1.f / dx;
  const float idx2 = idx * idx;
  const float eps2 = eps * eps;
  float m[TY + 2];  // mu on rows y0 - 1 .. y0 + TY
#pragma unroll
  for (int j = 0; j < TY + 2; ++j) {
    const float c = v[j + 1];
    float lap = (v[j + 2] + v[j] - 2.f * c) * idx2;
    lap += (lane_shift<kDppWaveShl1>(c) + lane_shift<kDppWaveShr1>(c) - 2.f * c) * idx2;
    m[j] = ch_mu(c, lap, eps2);
  }
  // zero flux across non-periodic sides: the ring cell outside the domain
  // takes its interior neighbour's mu (corners are never read)
  if (I.bc[0] != PPF_PERIODIC) {
    const bool lo = xr < 1, hi = xr > nx;
#pragma unroll
    for (int j = 0; j < TY + 2; ++j) {
      const float from_right = lane_shift<kDppWaveShl1>(m[j]), from_left = lane_shift<kDppWaveShr1>(m[j]);
      m[j] = lo ? from_right : (hi ? from_left : m[j]);
    }
  }
  if (I.bc[1] != PPF_PERIODIC) {
    if (y0 == 1) m[0] = m[1];
#pragma unroll
    for (int j = 1; j < TY + 2; ++j)
      if (y0 - 1 + j == ny + 1) m[j] = m[j - 1];
  }
  const bool out_lane = lane >= 2 && lane < kWaveLanes - 2 && xr <= nx;
  float* __restrict__ orow = o + int64_t(y0) * pitch + xr;
#pragma unroll
  for (int r = 0; r < TY; ++r) {
    const float mm = m[r + 1];
    // ch_tile_body's order: y (mu[y-1] + mu[y+1]), then x (mu[x-1] + mu[x+1])
    float dm = (m[r] + m[r + 2] - 2.f * mm);
    dm = dm + (lane_shift<kDppWaveShr1>(mm) + lane_shift<kDppWaveShl1>(mm) - 2.f * mm);
    dm = dm / (dx * dx);
    const float res = v[r + 2] + dtM * dm;
    if (out_lane && y0 + r <= ny) orow[r * pitch] = res;
  }
}

template <int ATY, int CTY>
__global__ __launch_bounds__(kThreads) void k_ensemble_w(const ppf_ens_instance* __restrict__ inst, int32_t ninst,
                                                         const int4* __restrict__ tiles, EnsSeg seg,
                                                         const float* __restrict__ in, float* __restrict__ out) {
  constexpr int kWaves = kThreads / kWaveLanes;
  const int wave = __builtin_amdgcn_readfirstlane(int(threadIdx.x) / kWaveLanes);
  const int lane = int(threadIdx.x) % kWaveLanes;
  // as k_ensemble, with the block's four waves taking consecutive tiles
  const int j = int(blockIdx.x / kEnsXcds);
  for (int xcd = int(blockIdx.x % kEnsXcds); xcd < kEnsXcds; xcd += int(gridDim.x)) {
    const int nb = gridDim.x >= unsigned(kEnsXcds) ? int((gridDim.x - 1 - xcd) / kEnsXcds) + 1 : 1;
    for (int t = seg.start[xcd] + kWaves * j + wave; t < seg.start[xcd + 1]; t += kWaves * nb) {
      const int4 td = tiles[t];
      if (td.x < 0 || td.x >= ninst) continue;  // uniform per wave
      const ppf_ens_instance I = inst[td.x];
      if (I.kind == PPF_ENS_AC)
        ac_wave_tile<ATY>(in + I.offset, out + I.offset, I, td.y, td.z, lane);
      else
        ch_wave_tile<CTY>(in + I.offset, out + I.offset, I, td.y, td.z, lane);
    }
  }
}

// Tile configurations (Allen-Cahn TX x TY, Cahn-Hilliard TX x TY) the
// autotuner chooses from: block tiles through LDS (k_ensemble), then wave
// tiles in registers (k_ensemble_w; TX = 64 lanes less the halo).
constexpr int kEnsConfigs[][4] = {{64, 8, 64, 8},   {64, 16, 64, 8},   {64, 16, 64, 16}, {128, 8, 64, 8},
                                  {64, 32, 64, 16},  {128, 16, 128, 8}, {32, 16, 32, 16}, {256, 4, 128, 4},
                                  {62, 16, 60, 8},   {62, 32, 60, 16},  {62, 32, 60, 24}, {62, 48, 60, 32}};
constexpr int kEnsNumConfigs = int(sizeof(kEnsConfigs) / sizeof(kEnsConfigs[0]));
constexpr int kEnsFirstWave = 8;

template <int C>
void launch_ensemble(dim3 grid, hipStream_t s, const ppf_ens_instance* inst, int32_t ninst, const int4* tiles,
                     const EnsSeg& seg, const float* in, float* out) {
  constexpr int ATX = kEnsConfigs[C][0], ATY = kEnsConfigs[C][1], CTX = kEnsConfigs[C][2],
                CTY = kEnsConfigs[C][3];
  if constexpr (C >= kEnsFirstWave) {
    static_assert(ATX == kWaveLanes - 2 && CTX == kWaveLanes - 4, "wave tile widths");
    hipLaunchKernelGGL((k_ensemble_w<ATY, CTY>), grid, dim3(kThreads), 0, s, inst, ninst, tiles, seg, in, out);
  } else {
    hipLaunchKernelGGL((k_ensemble<ATX, ATY, CTX, CTY>), grid, dim3(kThreads), 0, s, inst, ninst, tiles, seg, in,
                       out);
  }
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
  static_assert(kEnsNumConfigs == 12, "one case per configuration");
  switch (config) {
    case 0: launch_ensemble<0>(grid, st, instances, n_instances, t4, seg, in, out); break;
    case 1: launch_ensemble<1>(grid, st, instances, n_instances, t4, seg, in, out); break;
    case 2: launch_ensemble<2>(grid, st, instances, n_instances, t4, seg, in, out); break;
    case 3: launch_ensemble<3>(grid, st, instances, n_instances, t4, seg, in, out); break;
    case 4: launch_ensemble<4>(grid, st, instances, n_instances, t4, seg, in, out); break;
    case 5: launch_ensemble<5>(grid, st, instances, n_instances, t4, seg, in, out); break;
    case 6: launch_ensemble<6>(grid, st, instances, n_instances, t4, seg, in, out); break;
    case 7: launch_ensemble<7>(grid, st, instances, n_instances, t4, seg, in, out); break;
    case 8: launch_ensemble<8>(grid, st, instances, n_instances, t4, seg, in, out); break;
    case 9: launch_ensemble<9>(grid, st, instances, n_instances, t4, seg, in, out); break;
    case 10: launch_ensemble<10>(grid, st, instances, n_instances, t4, seg, in, out); break;
    default: launch_ensemble<11>(grid, st, instances, n_instances, t4, seg, in, out); break;
  }
  return check_launch("ensemble_step");
}

}  // extern "C"
