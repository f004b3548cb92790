// Surrogate thermodynamics: BF16-MFMA kernels for a small MLP (forward and
// backward GEMMs for training, elementwise ops, Adam) and the N-component
// thermodynamic pass that evaluates the trained networks per cell.
//
// All matrix products use v_mfma_f32_32x32x8_bf16 (FP32 accumulate).  Operand
// layout of that instruction, lane l (0..63), element t (0..3):
//   A (32 x 8):  A[l % 32][4 * (l / 32) + t]
//   B (8 x 32):  B[4 * (l / 32) + t][l % 32]
//   D (32 x 32): item v (0..15) holds D[8 * (v / 4) + 4 * (l / 32) + v % 4][l % 32]
// Networks are kept "features x samples" (transposed), so the D tile of one
// layer (hidden units as rows, cells as columns) is directly the B operand of
// the next: for the k-block s = 4h + j of the next layer, lane l holds rows
// 8s + 4(l/32) + t, exactly items 4j..4j+3 of the h-th 32-row D tile.
#include <hip/hip_runtime.h>

#include "ppf_ncomp.inl"

namespace ppf {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// round-to-nearest-even FP32 -> BF16 bits
__device__ __forceinline__ short f2bf(float x) {
  uint32_t u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return short(u >> 16);
}

__device__ __forceinline__ f32x16 mfma32(bf16x4 a, bf16x4 b, f32x16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, c, 0, 0, 0);
}

// ------------------------------------------------------------------ GEMM --
// C[M x N] = alpha * A[M x K] B[K x N] (+ beta C); A(i,k) = A[i*sai + k*sak],
// B(k,j) = B[k*sbk + j*sbj], C row-major with ldc.  One wave per 32x32 tile;
// gridDim.z splits K, partial tiles are atomically added (C must be zeroed by
// the caller when gridDim.z > 1).  A is rounded to BF16 (weights); with
// split_b, B enters as the BF16 pair hi + lo, hi = bf16(b), lo = bf16(b - hi)
// (about 16 significant bits: activations without BF16 jitter).
__device__ __forceinline__ float bf2f(short h) { return __uint_as_float(uint32_t(uint16_t(h)) << 16); }

__global__ __launch_bounds__(64) void k_gemm(int M, int N, int K, const float* __restrict__ A, int64_t sai,
                                             int64_t sak, const float* __restrict__ B, int64_t sbk, int64_t sbj,
                                             float* __restrict__ C, int64_t ldc, float alpha, float beta,
                                             int kchunk, int split_b) {
  const int lane = threadIdx.x;
  const int i0 = blockIdx.y * 32, j0 = blockIdx.x * 32;
  const int kb = blockIdx.z * kchunk;
  const int ke = min(K, kb + kchunk);
  const int r = lane & 31, u = lane >> 5;
  f32x16 acc;
#pragma unroll
  for (int v = 0; v < 16; ++v) acc[v] = 0.f;
  const int i = i0 + r, j = j0 + r;
  for (int k0 = kb; k0 < ke; k0 += 8) {
    bf16x4 a, b, blo;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int k = k0 + 4 * u + t;
      const bool kin = k < ke;
      a[t] = (kin && i < M) ? f2bf(A[i * sai + k * sak]) : short(0);
      const float bv = (kin && j < N) ? B[k * sbk + j * sbj] : 0.f;
      b[t] = f2bf(bv);
      blo[t] = f2bf(bv - bf2f(b[t]));
    }
    acc = mfma32(a, b, acc);
    if (split_b) acc = mfma32(a, blo, acc);
  }
#pragma unroll
  for (int v = 0; v < 16; ++v) {
    const int ii = i0 + 8 * (v >> 2) + 4 * u + (v & 3);
    const int jj = j0 + r;
    if (ii < M && jj < N) {
      float* c = C + ii * ldc + jj;
      if (gridDim.z > 1) {
        atomicAdd(c, alpha * acc[v]);
      } else {
        *c = alpha * acc[v] + (beta != 0.f ? beta * *c : 0.f);
      }
    }
  }
}

// ----------------------------------------------------------- elementwise --
// H = act(Z + b[row]) over a [M x N] row-major array (act 0: identity, 1: ReLU)
__global__ __launch_bounds__(256) void k_bias_act(const float* __restrict__ Z, const float* __restrict__ b,
                                                  float* __restrict__ H, int M, int N, int act) {
  int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= int64_t(M) * N) return;
  int row = int(t / N);
  float z = Z[t] + (b ? b[row] : 0.f);
  H[t] = act == 1 ? fmaxf(z, 0.f) : z;
}

// dZ = dH * (H > 0)
__global__ __launch_bounds__(256) void k_relu_bwd(const float* __restrict__ dH, const float* __restrict__ H,
                                                  float* __restrict__ dZ, int64_t n) {
  int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= n) return;
  dZ[t] = H[t] > 0.f ? dH[t] : 0.f;
}

// out[row] = sum_j X[row * N + j]   (one block per row)
__global__ __launch_bounds__(256) void k_rowsum(const float* __restrict__ X, float* __restrict__ out, int N) {
  const float* x = X + int64_t(blockIdx.x) * N;
  float s = 0.f;
  for (int j = threadIdx.x; j < N; j += 256) s += x[j];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  __shared__ float part[4];
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) out[blockIdx.x] = (part[0] + part[1]) + (part[2] + part[3]);
}

// mean squared error over [M x N]: dY = 2 (Y - T) / (M N); loss partial sums
// per block into partial[blockIdx.x]
__global__ __launch_bounds__(256) void k_mse(const float* __restrict__ Y, const float* __restrict__ T,
                                             float* __restrict__ dY, float* __restrict__ partial, int64_t n) {
  int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  float e = 0.f;
  if (t < n) {
    float d = Y[t] - T[t];
    dY[t] = 2.f * d / float(n);
    e = d * d;
  }
  for (int off = 32; off > 0; off >>= 1) e += __shfl_down(e, off, 64);
  __shared__ float part[4];
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = e;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = (part[0] + part[1]) + (part[2] + part[3]);
}

__global__ __launch_bounds__(256) void k_adam(float* __restrict__ p, const float* __restrict__ g,
                                              float* __restrict__ m, float* __restrict__ v, int64_t n, float lr,
                                              float b1, float b2, float eps, float c1, float c2) {
  int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= n) return;
  float gt = g[t];
  float mt = b1 * m[t] + (1.f - b1) * gt;
  float vt = b2 * v[t] + (1.f - b2) * gt * gt;
  m[t] = mt;
  v[t] = vt;
  p[t] -= lr * (mt * c1) / (sqrtf(vt * c2) + eps);
}

// ----------------------------------------------- N-component inference --
// Packed weights of one phase network (host prepares the lane layout):
//   w1[2][64] bf16x4   A operand of layer 1 (rows = hidden half h; k = 0..7 hold
//                      the weights of T, c1, c2, T, c1, c2, -, - matching the
//                      hi/lo input pairs)
//   w2[2][8][64]       A operand of layer 2 (output half m, k-block s)
//   w3[8][64]          A operand of layer 3 (rows = outputs 0..2, zero padded)
// followed by FP32 b1[64], b2[64], b3[4] (b3[3] unused).
constexpr int kMlpW1 = 2 * 64, kMlpW2 = 16 * 64, kMlpW3 = 8 * 64;
constexpr int kMlpWords = kMlpW1 + kMlpW2 + kMlpW3;  // bf16x4 words per phase

struct MlpK {
  const bf16x4* w[2];   // liquid, solid packed weights (device)
  const float* bias[2]; // b1[64], b2[64], b3[4] per phase (device)
  float in_lo[3], in_scale[3];        // x_hat = (x - lo) * scale - 1 over (T, c1, c2)
  float out_mu[2][3], out_sigma[2][3];
  float T_lo, T_hi, c_lo;             // range gate: outside, the plain FP64 evaluation is used
};

// D tile of a 32-row layer as B operands of the next layer: ReLU, then the
// BF16 pair hi + lo of each activation
__device__ __forceinline__ void relu_pack(const f32x16& d, int j, bf16x4& hi, bf16x4& lo) {
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const float v = fmaxf(d[4 * j + t], 0.f);
    hi[t] = f2bf(v);
    lo[t] = f2bf(v - bf2f(hi[t]));
  }
}

// bias of rows 8*(v/4) + 4*(l/32) + v%4 of a 32-row tile starting at row0
__device__ __forceinline__ f32x16 bias_tile(const float* b, int row0, int u) {
  f32x16 acc;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[4 * j + t] = b[row0 + 8 * j + 4 * u + t];
  }
  return acc;
}

// Evaluates one phase network for the 32 cells whose inputs the lanes hold
// (lanes l and l+32 hold cell l%32; lanes >= 32 pass zeros for k = 4..7).
// Returns, in lanes 0..31, items 0..2 = normalised outputs of cell l.
__device__ __forceinline__ f32x16 mlp_phase(const bf16x4* __restrict__ sw, const float* __restrict__ sb,
                                            bf16x4 x, int lane) {
  const int u = lane >> 5;
  f32x16 h1[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) h1[h] = mfma32(sw[h * 64 + lane], x, bias_tile(sb, 32 * h, u));
  f32x16 h2[2];
#pragma unroll
  for (int m = 0; m < 2; ++m) {
    f32x16 acc = bias_tile(sb + 64, 32 * m, u);
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      bf16x4 hi, lo;
      relu_pack(h1[s >> 2], s & 3, hi, lo);
      const bf16x4 w = sw[kMlpW1 + (m * 8 + s) * 64 + lane];
      acc = mfma32(w, hi, acc);
      acc = mfma32(w, lo, acc);
    }
    h2[m] = acc;
  }
  f32x16 y;
#pragma unroll
  for (int v = 0; v < 16; ++v) y[v] = 0.f;
  if (u == 0) {
    y[0] = sb[128];
    y[1] = sb[129];
    y[2] = sb[130];
  }
#pragma unroll
  for (int s = 0; s < 8; ++s) {
    bf16x4 hi, lo;
    relu_pack(h2[s >> 2], s & 3, hi, lo);
    const bf16x4 w = sw[kMlpW1 + kMlpW2 + s * 64 + lane];
    y = mfma32(w, hi, y);
    y = mfma32(w, lo, y);
  }
  return y;
}

// plain FP64 evaluation for cells outside the trained range (same outputs)
template <int N>
__device__ void plain_outputs(const float (&cin)[N - 1], float T, const ppf_thermo& th, double* yL, double* yS) {
  double c[N];
  double cN = 1.0;
#pragma unroll
  for (int l = 0; l < N - 1; ++l) {
    c[l] = double(cin[l]);
    cN -= c[l];
  }
  c[N - 1] = cN;
  double t = double(T);
  double lnT = plog(t);
  PhaseAtT<double, N> PL, PS;
  phase_at<double, N>(th.liquid, t, lnT, PL);
  phase_at<double, N>(th.solid, t, lnT, PS);
  double id = ideal_sum<double, N>(c);
  double gl = gibbs<double, N>(c, PL, id), gs = gibbs<double, N>(c, PS, id);
  double dl[N], ds[N];
  gibbs_grad<double, N>(c, PL, dl);
  gibbs_grad<double, N>(c, PS, ds);
  yL[0] = gl;
  yS[0] = gs - gl;
#pragma unroll
  for (int l = 0; l < N - 1; ++l) {
    yL[1 + l] = dl[l];
    yS[1 + l] = ds[l];
  }
}

// One wave per 64 padded cells of a row; 4 rows per 256-thread block.
__global__ __launch_bounds__(256) void k_ncomp_thermo_mlp(const float* __restrict__ in, const float* __restrict__ b,
                                                          Dom d, float dx, ppf_temperature t,
                                                          float* __restrict__ tr, int64_t tsy, int64_t tsf, NcK k,
                                                          MlpK net, uint64_t seed, int64_t step) {
  constexpr int N = 3;
  __shared__ bf16x4 sw[2][kMlpWords];
  __shared__ float sb[2][132];
  for (int i = threadIdx.x; i < 2 * kMlpWords; i += 256) sw[i / kMlpWords][i % kMlpWords] = net.w[i / kMlpWords][i % kMlpWords];
  for (int i = threadIdx.x; i < 2 * 132; i += 256) sb[i / 132][i % 132] = net.bias[i / 132][i % 132];
  __syncthreads();
  const int lane = threadIdx.x & 63, u = lane >> 5;
  const int64_t px = d.ax.n + 2, py = d.ay.n + 2;
  const uint32_t nbx = uint32_t((px + 63) / 64);
  const uint32_t blk = xcd_swizzle(blockIdx.x, gridDim.x);
  const int64_t xb = int64_t(blk % nbx) * 64, y = int64_t(blk / nbx) * 4 + (threadIdx.x >> 6);
  if (y >= py) return;  // whole wave leaves together
  const bool interior = xb >= 1 && xb + 63 <= d.ax.n && y >= 1 && y <= d.ay.n;
  const int64_t n3[3] = {d.ax.n, d.ay.n, 1};
  const int bc3[3] = {d.ax.bc, d.ay.bc, PPF_PERIODIC};
  // two groups of 32 cells: group g covers columns xb + 32g + (lane % 32)
  f32x16 yL[2], yS[2];
  float phig[2], Tg[2], cg[2][N - 1];
  bool gate[2];
#pragma unroll
  for (int g = 0; g < 2; ++g) {
    const int64_t x = xb + 32 * g + (lane & 31);
    const bool valid = x < px;
    float phi = 0.f, T = float(net.T_lo), cin[N - 1] = {0.f, 0.f};
    if (valid) {
      if (interior) {
        const int64_t o = y * d.ay.stride + x;
        phi = in[o];
#pragma unroll
        for (int l = 0; l < N - 1; ++l) cin[l] = in[o + (3 + l) * d.sf];
      } else {
        phi = fold(in, b, d, x, y, int64_t(0), dx);
#pragma unroll
        for (int l = 0; l < N - 1; ++l)
          cin[l] = fold(in + (3 + l) * d.sf, b ? b + (3 + l) * d.sf : nullptr, d, x, y, int64_t(0), dx);
      }
      T = temperature_at<float>(t, n3, bc3, x, y, 1, double(dx), step);
    }
    phig[g] = phi;
    Tg[g] = T;
#pragma unroll
    for (int l = 0; l < N - 1; ++l) cg[g][l] = cin[l];
    const float cN = 1.f - cin[0] - cin[1];
    gate[g] = T < net.T_lo || T > net.T_hi || cin[0] < net.c_lo || cin[1] < net.c_lo || cN < net.c_lo;
    // inputs as BF16 pairs in the 8 k slots of layer 1:
    // lanes < 32 (k 0..3): T_hi, c1_hi, c2_hi, T_lo; lanes >= 32 (k 4..7): c1_lo, c2_lo, 0, 0
    const float xv[3] = {(T - net.in_lo[0]) * net.in_scale[0] - 1.f, (cin[0] - net.in_lo[1]) * net.in_scale[1] - 1.f,
                         (cin[1] - net.in_lo[2]) * net.in_scale[2] - 1.f};
    short xh[3], xl[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      xh[q] = f2bf(xv[q]);
      xl[q] = f2bf(xv[q] - bf2f(xh[q]));
    }
    bf16x4 xin;
    xin[0] = u == 0 ? xh[0] : xl[1];
    xin[1] = u == 0 ? xh[1] : xl[2];
    xin[2] = u == 0 ? xh[2] : short(0);
    xin[3] = u == 0 ? xl[0] : short(0);
    yL[g] = mlp_phase(sw[0], sb[0], xin, lane);
    yS[g] = mlp_phase(sw[1], sb[1], xin, lane);
  }
  // lanes 0..31 finish group 0, lanes 32..63 group 1 (its outputs moved over)
  float oL[3], oS[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const float l1 = __shfl(yL[1][q], lane & 31, 64);
    const float s1 = __shfl(yS[1][q], lane & 31, 64);
    oL[q] = u == 0 ? yL[0][q] : l1;
    oS[q] = u == 0 ? yS[0][q] : s1;
  }
  const int g = u;
  const int64_t x = xb + 32 * g + (lane & 31);
  if (x >= px) return;
  const float phi = phig[g], T = Tg[g];
  const float cin[N - 1] = {cg[g][0], cg[g][1]};
  float GLn[N], GSn[N];  // non-ideal G and derivatives (liquid), G_S - G_L and solid derivatives
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    GLn[q] = oL[q] * net.out_sigma[0][q] + net.out_mu[0][q];
    GSn[q] = oS[q] * net.out_sigma[1][q] + net.out_mu[1][q];
  }
  const ppf_ncomp_params& p = k.p;
  const float cN = 1.f - cin[0] - cin[1];
  float dG, dGL[N - 1], dGS[N - 1];
  if (gate[g]) {
    double yl[3], ys[3];
    plain_outputs<N>(cin, T, k.th, yl, ys);
    dG = float(ys[0]);
#pragma unroll
    for (int l = 0; l < N - 1; ++l) {
      dGL[l] = float(yl[1 + l]);
      dGS[l] = float(ys[1 + l]);
    }
  } else {
    // ideal mixing R T sum c ln c is analytic: it cancels in G_S - G_L and adds
    // R T (ln c_l - ln c_N) to each derivative
    const float RT = float(k.th.liquid.R) * T;
    const float lnN = logf(fmaxf(cN, float(kCMin)));
    dG = GSn[0];
#pragma unroll
    for (int l = 0; l < N - 1; ++l) {
      const float idl = RT * (logf(fmaxf(cin[l], float(kCMin))) - lnN);
      dGL[l] = GLn[1 + l] + idl;
      dGS[l] = GSn[1 + l] + idl;
    }
  }
  const float g_ = (phi * phi) * ((1.f - phi) * (1.f - phi));
  const float h_ = (phi * phi * phi) * (6.f * (phi * phi) - 15.f * phi + 10.f);
  const float v_m = float(p.v_m), D_L = float(p.D_L), D_S = float(p.D_S);
  const uint64_t ncells = uint64_t(px * py), cell = uint64_t(y) * uint64_t(px) + uint64_t(x);
  const int64_t o = y * tsy + x;
  tr[o] = dG;
#pragma unroll
  for (int l = 0; l < N - 1; ++l) {
    tr[o + (1 + l) * tsf] = v_m * cin[l] * (D_L + h_ * (D_S - D_L)) / (8.314f * T);
    float noise = 0.f;
    if (p.noise_c != 0.0) {
      float xi = normal01<float>(hash64(seed, rng_counter(step, S_NC_C0 + l, ncells, cell)));
      noise = float(p.noise_c) * sqrtf(2.f * 8.314f * T / v_m) * xi;
    }
    tr[o + (N + l) * tsf] =
        (dGL[l] + h_ * (dGS[l] - dGL[l])) / v_m + (float(p.W[l]) - float(p.W[N - 1])) * g_ * T + noise;
  }
}

}  // namespace ppf

using namespace ppf;

extern "C" {

int ppf_mlp_gemm(int M, int N, int K, const float* A, int64_t sai, int64_t sak, const float* B, int64_t sbk,
                 int64_t sbj, float* C, int64_t ldc, float alpha, float beta, int ksplit, int split_b,
                 ppf_stream_t s) {
  if (M <= 0 || N <= 0 || K <= 0 || !A || !B || !C) return fail(PPF_EINVAL, "mlp_gemm: bad shape or null pointer");
  if (ksplit < 1) ksplit = 1;
  int kchunk = ((K + ksplit - 1) / ksplit + 7) / 8 * 8;
  ksplit = (K + kchunk - 1) / kchunk;
  if (ksplit > 1 && beta != 0.f) return fail(PPF_EINVAL, "mlp_gemm: split K needs beta = 0 and a zeroed C");
  dim3 grid(unsigned((N + 31) / 32), unsigned((M + 31) / 32), unsigned(ksplit));
  hipLaunchKernelGGL(k_gemm, grid, dim3(64), 0, stream_of(s), M, N, K, A, sai, sak, B, sbk, sbj, C, ldc, alpha, beta,
                     kchunk, split_b);
  return check_launch("mlp_gemm");
}

int ppf_mlp_bias_act(const float* Z, const float* b, float* H, int M, int N, int act, ppf_stream_t s) {
  if (M <= 0 || N <= 0 || !Z || !H) return fail(PPF_EINVAL, "mlp_bias_act: bad arguments");
  int64_t n = int64_t(M) * N;
  hipLaunchKernelGGL(k_bias_act, dim3(unsigned((n + 255) / 256)), dim3(256), 0, stream_of(s), Z, b, H, M, N, act);
  return check_launch("mlp_bias_act");
}

int ppf_mlp_relu_bwd(const float* dH, const float* H, float* dZ, int64_t n, ppf_stream_t s) {
  if (n <= 0 || !dH || !H || !dZ) return fail(PPF_EINVAL, "mlp_relu_bwd: bad arguments");
  hipLaunchKernelGGL(k_relu_bwd, dim3(unsigned((n + 255) / 256)), dim3(256), 0, stream_of(s), dH, H, dZ, n);
  return check_launch("mlp_relu_bwd");
}

int ppf_mlp_rowsum(const float* X, float* out, int M, int N, ppf_stream_t s) {
  if (M <= 0 || N <= 0 || !X || !out) return fail(PPF_EINVAL, "mlp_rowsum: bad arguments");
  hipLaunchKernelGGL(k_rowsum, dim3(unsigned(M)), dim3(256), 0, stream_of(s), X, out, N);
  return check_launch("mlp_rowsum");
}

int64_t ppf_mlp_mse_partials(int64_t n) { return (n + 255) / 256; }

int ppf_mlp_mse(const float* Y, const float* T, float* dY, float* partial, int64_t n, ppf_stream_t s) {
  if (n <= 0 || !Y || !T || !dY || !partial) return fail(PPF_EINVAL, "mlp_mse: bad arguments");
  hipLaunchKernelGGL(k_mse, dim3(unsigned((n + 255) / 256)), dim3(256), 0, stream_of(s), Y, T, dY, partial, n);
  return check_launch("mlp_mse");
}

int ppf_mlp_adam(float* p, const float* g, float* m, float* v, int64_t n, float lr, float b1, float b2, float eps,
                 int64_t t, ppf_stream_t s) {
  if (n <= 0 || !p || !g || !m || !v || t < 1) return fail(PPF_EINVAL, "mlp_adam: bad arguments");
  float c1 = 1.f / (1.f - powf(b1, float(t)));
  float c2 = 1.f / (1.f - powf(b2, float(t)));
  hipLaunchKernelGGL(k_adam, dim3(unsigned((n + 255) / 256)), dim3(256), 0, stream_of(s), p, g, m, v, n, lr, b1, b2,
                     eps, c1, c2);
  return check_launch("mlp_adam");
}

int ppf_ncomp_thermo_mlp(const ppf_grid* g, const void* in, const void* bcvals, const ppf_temperature* t,
                         const ppf_grid* tg, void* transfer, const ppf_ncomp_params* p, const ppf_thermo* th,
                         const ppf_mlp_net* net, uint64_t seed, int64_t step, ppf_stream_t s) {
  if (int rc = check_grid(g, 2, 2, "ncomp_thermo_mlp")) return rc;
  if (g->dtype != PPF_F32) return fail(PPF_EINVAL, "ncomp_thermo_mlp: the surrogate path is FP32");
  if (!p || p->ncomp != 3 || g->nfields != 5) return fail(PPF_EINVAL, "ncomp_thermo_mlp: needs 3 components");
  if (!tg || tg->nfields != 5 || tg->n[0] != g->n[0] || tg->n[1] != g->n[1] || tg->dtype != PPF_F32)
    return fail(PPF_EINVAL, "ncomp_thermo_mlp: transfer grid does not match");
  if (!in || !transfer || !t || !th || !net) return fail(PPF_EINVAL, "ncomp_thermo_mlp: null argument");
  for (int q = 0; q < 2; ++q)
    if (!net->weights[q] || !net->bias[q]) return fail(PPF_EINVAL, "ncomp_thermo_mlp: missing network");
  Dom d = make_dom(*g);
  NcK k{*p, *th};
  MlpK m;
  for (int q = 0; q < 2; ++q) {
    m.w[q] = static_cast<const bf16x4*>(net->weights[q]);
    m.bias[q] = net->bias[q];
    for (int j = 0; j < 3; ++j) {
      m.out_mu[q][j] = net->out_mu[q][j];
      m.out_sigma[q][j] = net->out_sigma[q][j];
    }
  }
  for (int j = 0; j < 3; ++j) {
    m.in_lo[j] = net->in_lo[j];
    m.in_scale[j] = net->in_scale[j];
  }
  m.T_lo = net->T_lo;
  m.T_hi = net->T_hi;
  m.c_lo = net->c_lo;
  dim3 grid(unsigned(((g->n[0] + 2 + 63) / 64) * ((g->n[1] + 2 + 3) / 4)));
  hipLaunchKernelGGL(k_ncomp_thermo_mlp, grid, dim3(256), 0, stream_of(s), static_cast<const float*>(in),
                     static_cast<const float*>(bcvals), d, float(g->dx), *t, static_cast<float*>(transfer), tg->sy,
                     tg->sf, k, m, seed, step);
  return check_launch("ncomp_thermo_mlp");
}

}  // extern "C"
