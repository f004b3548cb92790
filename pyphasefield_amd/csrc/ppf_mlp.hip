// Surrogate thermodynamics: MFMA kernels for a small MLP (forward and
// backward GEMMs for training, elementwise ops, Adam) and the N-component
// thermodynamic pass that evaluates the trained networks per cell.
//
// The training GEMMs use v_mfma_f32_32x32x8_bf16 / _f16 (FP32 accumulate).
// Operand layout of those instructions, lane l (0..63), element t (0..3):
//   A (32 x 8):  A[l % 32][4 * (l / 32) + t]
//   B (8 x 32):  B[4 * (l / 32) + t][l % 32]
//   D (32 x 32): item v (0..15) holds D[8 * (v / 4) + 4 * (l / 32) + v % 4][l % 32]
// Networks are kept "features x samples" (transposed), so the D tile of one
// layer (hidden units as rows, cells as columns) is directly the B operand of
// the next: for the k-block s = 4h + j of the next layer, lane l holds rows
// 8s + 4(l/32) + t, exactly items 4j..4j+3 of the h-th 32-row D tile.
#include <hip/hip_runtime.h>

#include <algorithm>

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
// the caller when gridDim.z > 1).  Operand formats (MODE):
//   0: A and B rounded to BF16;
//   1: A in BF16, B as the BF16 pair hi + lo, hi = bf16(b), lo = bf16(b - hi)
//      (about 16 significant bits: activations without BF16 jitter);
//   2: A in FP16, B as the FP16 pair hi + lo -- the formats of the inference
//      kernel's weights and inputs (forward products, weights as A).
__device__ __forceinline__ float bf2f(short h) { return __uint_as_float(uint32_t(uint16_t(h)) << 16); }

typedef _Float16 f16x4 __attribute__((ext_vector_type(4)));

template <int MODE>
__global__ __launch_bounds__(64) void k_gemm(int M, int N, int K, const float* __restrict__ A, int64_t sai,
                                             int64_t sak, const float* __restrict__ B, int64_t sbk, int64_t sbj,
                                             float* __restrict__ C, int64_t ldc, float alpha, float beta,
                                             int kchunk) {
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
    if constexpr (MODE == 2) {
      f16x4 a, b, blo;
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int k = k0 + 4 * u + t;
        const bool kin = k < ke;
        a[t] = (kin && i < M) ? _Float16(A[i * sai + k * sak]) : _Float16(0.f);
        const float bv = (kin && j < N) ? B[k * sbk + j * sbj] : 0.f;
        b[t] = _Float16(bv);
        blo[t] = _Float16(bv - float(b[t]));
      }
      acc = __builtin_amdgcn_mfma_f32_32x32x8f16(a, b, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x8f16(a, blo, acc, 0, 0, 0);
    } else {
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
      if (MODE == 1) acc = mfma32(a, blo, acc);
    }
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
// Inference uses v_mfma_f32_32x32x16_f16: lane l, element t (0..7) of the A
// and B operands pair up along K in the same (l / 32, t) slot, so any slot ->
// k assignment shared by A and B is valid.  The slot of hidden row k used here
// lets the D tile of one layer feed the next layer without data movement:
// D item v of lane l holds row 32h + 8(v/4) + 4(l/32) + v%4 of hidden half h,
// so items 8p..8p+7 form the B operand of K block kb = 2h + p with
//   slot (l/32, t) -> row 32h + 16p + 8(t/4) + 4(l/32) + t%4.
// Packed weights of one phase network (FP16 x 8 words, 64 lanes each; host
// packing in pyphasefield_amd.surrogate.pack_phase):
//   set 0..1    layer 1, hidden half h: slot (0, t) = weight of input t of
//               (T_hi, c1_hi, c2_hi, T_lo, c1_lo, c2_lo, 1, 1): slots 6 and 7
//               hold the layer-1 bias as an FP16 hi + lo pair; lanes >= 32 zero
//   set 2..9    layer 2, output half m = (set-2)/4, K block kb = (set-2)%4
//   set 10..13  layer 3, K block kb, as the A operands of the 4x4x4 MFMA
//               (mlp_eval): lane l element t = weight of output row l % 4
//               (row 3 zero) for the hidden row of slot (l / 32, t)
// followed by FP32 b1[64] (unused: packed above), b2[64], b3[4] (b3[3] unused).
// Activation scaling: the host divides layer 1's weights and bias by a power
// of two s1 and layer 2's bias by s2, multiplies layer 2's weights by s1/s2
// and layer 3's by s2, with s1, s2 at least the interval bounds of the
// layers' pre-activations over the trained input box.  Powers of two are
// exact in FP16 and FP32, so the network is unchanged, and every hidden value
// of an accepted cell lies in [0, 1]: ReLU(x) = clamp(x, 0, 1) there, and the
// clamp rides on the FP16 conversion for free (v_cvt_pk_f16_f32 ... clamp).
// The inputs enter layer 1 as FP16 pairs hi + lo; the activations entering
// layers 2 and 3 are either such pairs (two MFMAs with the same weights) or
// FP16 only (the default).
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
constexpr int kMlpSets = 14;
constexpr int kMlpWords = kMlpSets * 64;  // FP16 x 8 words per phase

__device__ __forceinline__ f32x16 mfma16(f16x8 a, f16x8 b, f32x16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
}

struct MlpK {
  const f16x8* w[2];    // liquid, solid packed weights (device)
  const float* bias[2]; // b1[64], b2[64], b3[4] per phase (device)
  float in_lo[3], in_scale[3];        // x_hat = (x - lo) * scale - 1 over (T, c1, c2)
  float out_mu[2][3], out_sigma[2][3];
  float T_lo, T_hi, c_lo;             // range gate: outside, the plain FP64 evaluation is used
  float R;                            // gas constant of the ideal-mixing term
  int32_t* marks;                     // [2] cells marked for the fallback, by step parity (may be null)
};

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

// two FP32 values to one register of FP16 (round to nearest even): low half
// from a, high half from b (one v_cvt_pk_f16_f32)
__device__ __forceinline__ uint32_t cvt_pk_f16(float a, float b) {
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2{a, b}, f16x2));
}

// the same with the result clamped to [0, 1]: the compiler folds the clamp
// into the conversion's clamp bit (one instruction)
__device__ __forceinline__ uint32_t act_pk_f16(float a, float b) {
  f16x2 p = __builtin_convertvector(f32x2{a, b}, f16x2);
  p = __builtin_elementwise_min(__builtin_elementwise_max(p, f16x2{0, 0}), f16x2{1, 1});
  return __builtin_bit_cast(uint32_t, p);
}

__device__ __forceinline__ float f16_lo(uint32_t w) { return float(__builtin_bit_cast(f16x2, w)[0]); }
__device__ __forceinline__ float f16_hi(uint32_t w) { return float(__builtin_bit_cast(f16x2, w)[1]); }

// max(x, 0) as a signed-integer max of the bits (one instruction, no NaN
// canonicalisation; negative floats are negative integers)
__device__ __forceinline__ float relu(float x) { return __int_as_float(max(__float_as_int(x), 0)); }

// ReLU of items 8p..8p+7 of a D tile as the FP16 pair (hi, lo)
__device__ __forceinline__ void relu_split(const f32x16& d, int p, f16x8& hi, f16x8& lo) {
  u32x4 wh, wl;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const float v0 = relu(d[8 * p + 2 * t]), v1 = relu(d[8 * p + 2 * t + 1]);
    const uint32_t h = cvt_pk_f16(v0, v1);
    wh[t] = h;
    wl[t] = cvt_pk_f16(v0 - f16_lo(h), v1 - f16_hi(h));
  }
  hi = __builtin_bit_cast(f16x8, wh);
  lo = __builtin_bit_cast(f16x8, wl);
}

// ReLU of items 8p..8p+7 of a D tile in FP16 (clamp to [0, 1], see above)
__device__ __forceinline__ void relu_f16(const f32x16& d, int p, f16x8& hi) {
  u32x4 w;
#pragma unroll
  for (int t = 0; t < 4; ++t) w[t] = act_pk_f16(d[8 * p + 2 * t], d[8 * p + 2 * t + 1]);
  hi = __builtin_bit_cast(f16x8, w);
}

// next-layer B operands (K blocks kb = 0..3) from the two D tiles of a layer
template <bool SPLIT>
__device__ __forceinline__ void pack_layer(const f32x16 (&h)[2], f16x8 (&bh)[4], f16x8 (&bl)[4]) {
#pragma unroll
  for (int kb = 0; kb < 4; ++kb) {
    if (SPLIT) {
      relu_split(h[kb >> 1], kb & 1, bh[kb], bl[kb]);
    } else {
      relu_f16(h[kb >> 1], kb & 1, bh[kb]);
    }
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

typedef float f32x4 __attribute__((ext_vector_type(4)));

// One phase network on the two groups of 32 cells of a tile; the weight
// operands read from LDS are shared by the groups.  x[g]: layer-1 B operand
// of group g.  out[q] = normalised output q of the lane's own cell: cell
// lane % 32 of group lane / 32.  S2 / S3: the inputs of layer 2 / layer 3
// enter as FP16 hi + lo pairs (two MFMAs per K block) instead of FP16 only.
// Layer 3 (three outputs) runs on the 16-block 4x4x4 MFMA: block b is lanes
// 4b..4b+3; lane 4b + r supplies weight row r (row 3 is zero padding) and
// lane 4b + j its own cell's hidden values, so item r of lane l ends as
// output r of cell lane % 32 summed over the hidden rows lane l holds (half
// l / 32 of them); one lane-half swap adds the two halves.  A quarter of the
// 32x32x16 MFMA's work for 3 of its 32 rows.
template <bool S2, bool S3>
__device__ __forceinline__ void mlp_eval(const f16x8* __restrict__ w, const float* __restrict__ b,
                                         const f16x8 (&x)[2], int lane, float (&out)[3]) {
  constexpr int G = 2;
  const int u = lane >> 5;
  f32x16 h[G][2];
  const f32x16 zero = {};
#pragma unroll
  for (int hh = 0; hh < 2; ++hh) {
    // the layer-1 bias comes in through the packed words (constant-1 inputs)
    const f16x8 a = w[hh * 64 + lane];
#pragma unroll
    for (int g = 0; g < G; ++g) h[g][hh] = mfma16(a, x[g], zero);
  }
  f16x8 bh[G][4], bl[G][4];
#pragma unroll
  for (int g = 0; g < G; ++g) pack_layer<S2>(h[g], bh[g], bl[g]);
#pragma unroll
  for (int m = 0; m < 2; ++m) {
    const f32x16 bt = bias_tile(b + 64, 32 * m, u);
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) {
      const f16x8 a = w[(2 + 4 * m + kb) * 64 + lane];
      // the bias tile enters as the first accumulator operand (no copy)
#pragma unroll
      for (int g = 0; g < G; ++g) h[g][m] = mfma16(a, bh[g][kb], kb == 0 ? bt : h[g][m]);
      if (S2) {
#pragma unroll
        for (int g = 0; g < G; ++g) h[g][m] = mfma16(a, bl[g][kb], h[g][m]);
      }
    }
  }
#pragma unroll
  for (int g = 0; g < G; ++g) pack_layer<S3>(h[g], bh[g], bl[g]);
  f32x4 y[G] = {};
#pragma unroll
  for (int kb = 0; kb < 4; ++kb) {
    const f16x8 a = w[(10 + kb) * 64 + lane];
    const f16x4 a0 = __builtin_shufflevector(a, a, 0, 1, 2, 3), a1 = __builtin_shufflevector(a, a, 4, 5, 6, 7);
#pragma unroll
    for (int g = 0; g < G; ++g) {
      y[g] = __builtin_amdgcn_mfma_f32_4x4x4f16(a0, __builtin_shufflevector(bh[g][kb], bh[g][kb], 0, 1, 2, 3), y[g],
                                                0, 0, 0);
      y[g] = __builtin_amdgcn_mfma_f32_4x4x4f16(a1, __builtin_shufflevector(bh[g][kb], bh[g][kb], 4, 5, 6, 7), y[g],
                                                0, 0, 0);
      if (S3) {
        y[g] = __builtin_amdgcn_mfma_f32_4x4x4f16(a0, __builtin_shufflevector(bl[g][kb], bl[g][kb], 0, 1, 2, 3),
                                                  y[g], 0, 0, 0);
        y[g] = __builtin_amdgcn_mfma_f32_4x4x4f16(a1, __builtin_shufflevector(bl[g][kb], bl[g][kb], 4, 5, 6, 7),
                                                  y[g], 0, 0, 0);
      }
    }
  }
  // swap the lane halves between the groups' partial sums: r[0] = (group 0
  // low, group 1 low), r[1] = (group 0 high, group 1 high); their sum is
  // group 0's output in lanes 0..31 and group 1's in lanes 32..63
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(y[0][q]), __float_as_uint(y[1][q]), false, false);
    out[q] = (__uint_as_float(r[0]) + __uint_as_float(r[1])) + b[128 + q];
  }
}

struct Plain5 {
  float dG, dGL[2], dGS[2];
};

// plain FP64 evaluation of G_S - G_L and both phases' derivatives; the
// phases one after the other (each value depends on its own phase only, so
// the order changes no result and halves the coefficients held at a time)
__device__ __forceinline__ Plain5 plain_outputs(float c1, float c2, float T, const ppf_thermo& th) {
  constexpr int N = 3;
  double c[N] = {double(c1), double(c2), 1.0 - double(c1) - double(c2)};
  double t = double(T);
  double lnT = plog(t);
  double id = ideal_sum<double, N>(c);
  double gl = 0.0, gs = 0.0;
  double dl[N], ds[N];
  if (th.tables) {
    {
      PhaseAtT<double, N, true> P;
      phase_at_tables<double, N>(&th.tables->liquid, th.liquid.R, t, lnT, P);
      gl = gibbs(c, P, id);
      gibbs_grad(c, P, dl);
    }
    {
      PhaseAtT<double, N, true> P;
      phase_at_tables<double, N>(&th.tables->solid, th.solid.R, t, lnT, P);
      gs = gibbs(c, P, id);
      gibbs_grad(c, P, ds);
    }
  } else {
    {
      PhaseAtT<double, N, false> P;
      phase_at<double, N>(th.liquid, t, lnT, P);
      gl = gibbs(c, P, id);
      gibbs_grad(c, P, dl);
    }
    {
      PhaseAtT<double, N, false> P;
      phase_at<double, N>(th.solid, t, lnT, P);
      gs = gibbs(c, P, id);
      gibbs_grad(c, P, ds);
    }
  }
  Plain5 r;
  r.dG = float(gs - gl);
  for (int l = 0; l < N - 1; ++l) {
    r.dGL[l] = float(dl[l]);
    r.dGS[l] = float(ds[l]);
  }
  return r;
}

// the same, out of line, for the surrogate kernel's cells outside the trained
// box: a call keeps the FP64 evaluation's registers out of the tile loop's
// allocation (the model is read through a device pointer)
__device__ __attribute__((noinline)) Plain5 plain_outputs_call(float c1, float c2, float T, const ppf_thermo* th) {
  return plain_outputs(c1, c2, T, *th);
}

__device__ __forceinline__ float fast_ln(float x) { return __builtin_amdgcn_logf(x) * 0.693147180559945309f; }

// Scalars of the N-component pass the surrogate kernel needs (the plain
// evaluation's tables stay in global memory for the fallback).
struct NcMlpK {
  float v_m, inv_v_m, D_L, dD, noise_c, W0, W1;  // W0, W1: W_l - W_N
};

// transfer outputs of one cell (layout of ppf_ncomp_thermo): G_S - G_L,
// mobilities M_c,l, chemical potentials dF/dc_l (+ noise)
__device__ __forceinline__ void write_transfer(float* __restrict__ tr, uint32_t o, uint32_t tsf, const NcMlpK& k,
                                               float phi, float T, const float (&cl)[2], float dG,
                                               const float (&dGL)[2], const float (&dGS)[2], uint64_t seed,
                                               int64_t step, uint32_t ncell, uint32_t cell) {
  constexpr int N = 3;
  const float g_ = (phi * phi) * ((1.f - phi) * (1.f - phi));
  const float h_ = (phi * phi * phi) * (6.f * (phi * phi) - 15.f * phi + 10.f);
  tr[o] = dG;
  const float mob = k.v_m * (k.D_L + h_ * k.dD) * __builtin_amdgcn_rcpf(8.314f * T);
  const float Wd[N - 1] = {k.W0, k.W1};
#pragma unroll
  for (int l = 0; l < N - 1; ++l) {
    tr[o + (1 + l) * tsf] = mob * cl[l];
    float noise = 0.f;
    if (k.noise_c != 0.f) {
      float xi = normal01<float>(hash64(seed, rng_counter(step, S_NC_C0 + l, uint64_t(ncell), uint64_t(cell))));
      noise = k.noise_c * sqrtf(2.f * 8.314f * T * k.inv_v_m) * xi;
    }
    tr[o + (N + l) * tsf] = (dGL[l] + h_ * (dGS[l] - dGL[l])) * k.inv_v_m + Wd[l] * g_ * T + noise;
  }
}

// waves per SIMD the register budget is set for (3: no spills)
#ifndef PPF_MLP_WAVES
#define PPF_MLP_WAVES 3
#endif
// persistent blocks per CU of the inference kernel
#ifndef PPF_MLP_BLOCKS_PER_CU
#define PPF_MLP_BLOCKS_PER_CU PPF_MLP_WAVES
#endif

// The surrogate's thermodynamic pass.  Persistent: each block stages both
// networks in LDS once and its waves walk 64-cell tiles of the padded domain
// in row-major order (wave w of block b takes tiles 4b + w, 4b + w + 4
// gridDim.x, ...).  Lane l owns cell 64 tile + l; MFMA group 0 is the tile's
// first 32 cells, group 1 the next 32, so the weight operands are read once
// per 64 cells.
//  * Tile bookkeeping on the scalar unit: the row and column of a tile's
//    first cell advance by the host-computed (stride / px, stride % px); a
//    lane adds its index and wraps at most once (px >= 64; narrower grids
//    divide).
//  * Fields and transfer arrays through buffer instructions with 32-bit byte
//    offsets (the array offsets ride in scalar registers).
//  * One v_permlane32_swap per input gives both MFMA groups their cells:
//    swapping a register with itself broadcasts its low half (cells 0..31,
//    group 0) into one result and its high half (cells 32..63, group 1) into
//    the other.  The K slots of lanes 32..63 then hold a copy of the same
//    cell's inputs, which the zero weights of those slots cancel exactly.
//  * Layer 3's lane-half sum (mlp_eval) leaves group 0's outputs in lanes
//    0..31 and group 1's in lanes 32..63, each lane's own cell.
//  * NOISE: composition noise on (a separate instantiation, so the noise
//    path's arithmetic is never speculated into the common one).
//  * Cells outside the trained box: with a device copy of the thermodynamic
//    model (thermo), the plain FP64 evaluation runs here, in the lanes of
//    those cells (its coefficients are read through the pointer inside the
//    branch, so none of them is held in registers across the tile loop);
//    without one they are marked for k_ncomp_thermo_fallback.
struct MlpLean {
  uint32_t px, py, nx, ny, sy, sf4, tsy, tsf4;  // padded extents, strides (elements; *4: bytes)
  uint32_t ntiles, srow, scol;                  // tiles; wave stride of cells as rows and columns
  float T_lo, T_hi, c_lo, R;
  int32_t* marks;
  const ppf_thermo* thermo;  // device copy of the model, or null
};
static_assert(sizeof(ppf_mlp_net) == 136, "ppf_mlp_net layout (mirrored by _lib_ext.MlpNet)");

__device__ __forceinline__ void lean_store(float v, __amdgpu_buffer_rsrc_t rs, uint32_t voff, uint32_t soff) {
  __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), rs, voff, int32_t(soff), 0);
}

template <bool S2, bool S3, bool NOISE>
__global__ __launch_bounds__(256, PPF_MLP_WAVES) void k_ncomp_thermo_mlp(
    const float* __restrict__ in, const float* __restrict__ b, Dom d, float dx, ppf_temperature t,
    float* __restrict__ tr, NcMlpK k, MlpK net, MlpLean g, uint64_t seed, StepRef sr) {
  constexpr int N = 3;
  const int64_t step = sr.get();
  __shared__ f16x8 sw[2][kMlpWords];
  __shared__ float sb[2][132];
  for (int i = threadIdx.x; i < 2 * kMlpWords; i += 256) sw[i / kMlpWords][i % kMlpWords] = net.w[i / kMlpWords][i % kMlpWords];
  for (int i = threadIdx.x; i < 2 * 132; i += 256) sb[i / 132][i % 132] = net.bias[i / 132][i % 132];
  __syncthreads();
  // this step's cells outside the box are counted in marks[step & 1]; the
  // next step's counter starts at zero (cleared here when no second kernel runs)
  if (g.thermo && g.marks && blockIdx.x == 0 && threadIdx.x == 0) g.marks[(step + 1) & 1] = 0;
  const int lane = threadIdx.x & 63;
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint32_t tstride = gridDim.x * 4;
  uint32_t tile = blockIdx.x * 4 + wave;
  // row and column of the tile's first cell (wave-uniform)
  uint32_t y0 = __builtin_amdgcn_readfirstlane((tile * 64) / g.px);
  uint32_t x0 = __builtin_amdgcn_readfirstlane(tile * 64 - y0 * g.px);
  const __amdgpu_buffer_rsrc_t rin = __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(in), int16_t(0),
                                                                        int32_t(-1), 0x00020000);
  const __amdgpu_buffer_rsrc_t rtr = __builtin_amdgcn_make_buffer_rsrc(tr, int16_t(0), int32_t(-1), 0x00020000);
  const int64_t n3[3] = {d.ax.n, d.ay.n, 1};
  const int bc3[3] = {d.ax.bc, d.ay.bc, PPF_PERIODIC};
  const uint64_t ncell = uint64_t(g.px) * g.py;
  for (; tile < g.ntiles; tile += tstride) {
    uint32_t x = x0 + lane, y = y0;
    if (g.px >= 64) {
      if (x >= g.px) {
        x -= g.px;
        ++y;
      }
    } else {
      y += x / g.px;
      x = x % g.px;
    }
    const bool valid = y < g.py;
    if (!valid) x = y = 1;
    float phi, c1, c2, T;
    if (x - 1 < g.nx && y - 1 < g.ny) {
      const uint32_t o = (y * g.sy + x) * 4u;
      phi = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rin, o, 0, 0));
      c1 = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rin, o, int32_t(3 * g.sf4), 0));
      c2 = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rin, o, int32_t(4 * g.sf4), 0));
      T = temperature_interior<float>(t, n3, bc3, int64_t(x), int64_t(y), double(dx), step);
    } else {
      phi = fold(in, b, d, int64_t(x), int64_t(y), int64_t(0), dx);
      c1 = fold(in + 3 * d.sf, b ? b + 3 * d.sf : nullptr, d, int64_t(x), int64_t(y), int64_t(0), dx);
      c2 = fold(in + 4 * d.sf, b ? b + 4 * d.sf : nullptr, d, int64_t(x), int64_t(y), int64_t(0), dx);
      T = temperature_at<float>(t, n3, bc3, int64_t(x), int64_t(y), 1, double(dx), step);
    }
    const float cN = 1.f - c1 - c2;
    const bool gate = T < g.T_lo || T > g.T_hi || c1 < g.c_lo || c2 < g.c_lo || cN < g.c_lo;
    const float xv[3] = {(T - net.in_lo[0]) * net.in_scale[0] - 1.f, (c1 - net.in_lo[1]) * net.in_scale[1] - 1.f,
                         (c2 - net.in_lo[2]) * net.in_scale[2] - 1.f};
    // xs[0]: cell lane % 32 (group 0), xs[1]: cell 32 + lane % 32 (group 1)
    float xs[2][3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(xv[q]), __float_as_uint(xv[q]), false, false);
      xs[0][q] = __uint_as_float(r[0]);
      xs[1][q] = __uint_as_float(r[1]);
    }
    // slots (T_hi, c1_hi, c2_hi, T_lo, c1_lo, c2_lo, 1, 1), hi = f16(x), lo = f16(x - hi)
    f16x8 xin[2];
#pragma unroll
    for (int gg = 0; gg < 2; ++gg) {
      const uint32_t p01 = cvt_pk_f16(xs[gg][0], xs[gg][1]);
      const float t_lo = xs[gg][0] - f16_lo(p01);
      const float c1_lo = xs[gg][1] - f16_hi(p01);
      const uint32_t p23 = cvt_pk_f16(xs[gg][2], t_lo);
      const float c2_lo = xs[gg][2] - f16_lo(p23);
      const uint32_t p45 = cvt_pk_f16(c1_lo, c2_lo);
      xin[gg] = __builtin_bit_cast(f16x8, u32x4{p01, p23, p45, 0x3c003c00u});
    }
    // the two networks one after the other (not unrolled: one network's
    // operands and activations live at a time)
    float GL[3], GS[3];
#pragma unroll 1
    for (int ph = 0; ph < 2; ++ph) {
      float o[3];
      mlp_eval<S2, S3>(sw[ph], sb[ph], xin, lane, o);
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const float v = o[q] * net.out_sigma[ph][q] + net.out_mu[ph][q];
        GL[q] = ph == 0 ? v : GL[q];
        GS[q] = ph == 1 ? v : GS[q];
      }
    }
    const uint32_t cell = y * g.px + x;
    const uint32_t ob = (y * g.tsy + x) * 4u;
    x0 += g.scol;
    y0 += g.srow;
    if (x0 >= g.px) {
      x0 -= g.px;
      ++y0;
    }
    if (!valid) continue;
    const float cl[N - 1] = {c1, c2};
    float dG, dGL[N - 1], dGS[N - 1];
    if (gate) {  // outside the trained box
      if (g.marks) atomicAdd(&g.marks[step & 1], 1);
      if (!g.thermo) {  // marked for k_ncomp_thermo_fallback
        lean_store(__builtin_nanf(""), rtr, ob, 0);
        continue;
      }
      const Plain5 r = plain_outputs_call(c1, c2, T, g.thermo);
      dG = r.dG;
#pragma unroll
      for (int l = 0; l < N - 1; ++l) {
        dGL[l] = r.dGL[l];
        dGS[l] = r.dGS[l];
      }
    } else {
      // ideal mixing R T sum c ln c is analytic: it cancels in G_S - G_L and
      // adds R T (ln c_l - ln c_N) to each derivative
      const float RT = g.R * T;
      const float lnN = fast_ln(fmaxf(cN, float(kCMin)));
      dG = GS[0];
#pragma unroll
      for (int l = 0; l < N - 1; ++l) {
        const float idl = RT * (fast_ln(fmaxf(cl[l], float(kCMin))) - lnN);
        dGL[l] = GL[1 + l] + idl;
        dGS[l] = GS[1 + l] + idl;
      }
    }
    const float g_ = (phi * phi) * ((1.f - phi) * (1.f - phi));
    const float h_ = (phi * phi * phi) * (6.f * (phi * phi) - 15.f * phi + 10.f);
    lean_store(dG, rtr, ob, 0);
    const float mob = k.v_m * (k.D_L + h_ * k.dD) * __builtin_amdgcn_rcpf(8.314f * T);
    const float Wd[N - 1] = {k.W0, k.W1};
#pragma unroll
    for (int l = 0; l < N - 1; ++l) {
      lean_store(mob * cl[l], rtr, ob, (1 + l) * g.tsf4);
      float noise = 0.f;
      if constexpr (NOISE) {
        float xi = normal01<float>(hash64(seed, rng_counter(step, S_NC_C0 + l, ncell, uint64_t(cell))));
        noise = k.noise_c * sqrtf(2.f * 8.314f * T * k.inv_v_m) * xi;
      }
      lean_store((dGL[l] + h_ * (dGS[l] - dGL[l])) * k.inv_v_m + Wd[l] * g_ * T + noise, rtr, ob, (N + l) * g.tsf4);
    }
  }
}

// Cells the surrogate kernel marked (NaN in G_S - G_L): the plain FP64
// evaluation, same outputs.  A grid-stride scan of the padded cells (a few
// blocks per CU: when the step's mark counter is zero every block returns
// after one load).
__global__ __launch_bounds__(256) void k_ncomp_thermo_fallback(const float* __restrict__ in,
                                                               const float* __restrict__ b, Dom d, float dx,
                                                               ppf_temperature t, float* __restrict__ tr,
                                                               int64_t tsy, int64_t tsf, NcMlpK k, ppf_thermo th,
                                                               const int32_t* marks_in, uint64_t seed,
                                                               StepRef sr) {
  const int64_t step = sr.get();
  int32_t* marks = const_cast<int32_t*>(marks_in);
  const uint32_t px = uint32_t(d.ax.n + 2), py = uint32_t(d.ay.n + 2), ncell = px * py;
  const int64_t n3[3] = {d.ax.n, d.ay.n, 1};
  const int bc3[3] = {d.ax.bc, d.ay.bc, PPF_PERIODIC};
  if (marks) {
    // this step's count (zero: nothing to recompute); the next step's
    // counter is cleared here, before the next surrogate pass runs
    if (blockIdx.x == 0 && threadIdx.x == 0) marks[(step + 1) & 1] = 0;
    if (marks[step & 1] == 0) return;
  }
  for (uint32_t cell = blockIdx.x * 256 + threadIdx.x; cell < ncell; cell += gridDim.x * 256) {
    const uint32_t y = cell / px, x = cell - y * px;
    const uint32_t o = y * uint32_t(tsy) + x;
    if (!__builtin_isnan(tr[o])) continue;
    const float phi = fold(in, b, d, int64_t(x), int64_t(y), int64_t(0), dx);
    const float c1 = fold(in + 3 * d.sf, b ? b + 3 * d.sf : nullptr, d, int64_t(x), int64_t(y), int64_t(0), dx);
    const float c2 = fold(in + 4 * d.sf, b ? b + 4 * d.sf : nullptr, d, int64_t(x), int64_t(y), int64_t(0), dx);
    const float T = temperature_at<float>(t, n3, bc3, int64_t(x), int64_t(y), 1, double(dx), step);
    const Plain5 r = plain_outputs(c1, c2, T, th);
    const float cl[2] = {c1, c2};
    write_transfer(tr, o, uint32_t(tsf), k, phi, T, cl, r.dG, r.dGL, r.dGS, seed, step, ncell, cell);
  }
}
}  // namespace ppf

using namespace ppf;

extern "C" {

int ppf_mlp_gemm(int M, int N, int K, const float* A, int64_t sai, int64_t sak, const float* B, int64_t sbk,
                 int64_t sbj, float* C, int64_t ldc, float alpha, float beta, int ksplit, int split_b,
                 ppf_stream_t s) {
  if (M <= 0 || N <= 0 || K <= 0 || !A || !B || !C) return fail(PPF_EINVAL, "mlp_gemm: bad shape or null pointer");
  if (split_b < 0 || split_b > 2) return fail(PPF_EINVAL, "mlp_gemm: operand format must be 0, 1 or 2");
  if (ksplit < 1) ksplit = 1;
  int kchunk = ((K + ksplit - 1) / ksplit + 7) / 8 * 8;
  ksplit = (K + kchunk - 1) / kchunk;
  if (ksplit > 1 && beta != 0.f) return fail(PPF_EINVAL, "mlp_gemm: split K needs beta = 0 and a zeroed C");
  dim3 grid(unsigned((N + 31) / 32), unsigned((M + 31) / 32), unsigned(ksplit));
  auto kern = split_b == 2 ? k_gemm<2> : (split_b == 1 ? k_gemm<1> : k_gemm<0>);
  hipLaunchKernelGGL(kern, grid, dim3(64), 0, stream_of(s), M, N, K, A, sai, sak, B, sbk, sbj, C, ldc, alpha, beta,
                     kchunk);
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

static int ncomp_thermo_mlp_dispatch(const ppf_grid* g, const void* in, const void* bcvals,
                                     const ppf_temperature* t, const ppf_grid* tg, void* transfer,
                                     const ppf_ncomp_params* p, const ppf_thermo* th, const ppf_mlp_net* net,
                                     uint64_t seed, StepRef step, ppf_stream_t s) {
  if (int rc = check_grid(g, 2, 2, "ncomp_thermo_mlp")) return rc;
  if (g->dtype != PPF_F32) return fail(PPF_EINVAL, "ncomp_thermo_mlp: the surrogate path is FP32");
  if (!p || p->ncomp != 3 || g->nfields != 5) return fail(PPF_EINVAL, "ncomp_thermo_mlp: needs 3 components");
  if (p->flags & (PPF_NC_DERIV_FIRST | PPF_NC_SHARED_NOISE))
    return fail(PPF_EINVAL, "ncomp_thermo_mlp: the surrogate has no legacy-engine derivative or noise form");
  if (!tg || tg->nfields != 5 || tg->n[0] != g->n[0] || tg->n[1] != g->n[1] || tg->dtype != PPF_F32)
    return fail(PPF_EINVAL, "ncomp_thermo_mlp: transfer grid does not match");
  if (!in || !transfer || !t || !th || !net) return fail(PPF_EINVAL, "ncomp_thermo_mlp: null argument");
  if (int rc = check_temperature(t, "ncomp_thermo_mlp")) return rc;
  for (int q = 0; q < 2; ++q)
    if (!net->weights[q] || !net->bias[q]) return fail(PPF_EINVAL, "ncomp_thermo_mlp: missing network");
  Dom d = make_dom(*g);
  NcMlpK k;
  k.v_m = float(p->v_m);
  k.inv_v_m = float(1.0 / p->v_m);
  k.D_L = float(p->D_L);
  k.dD = float(p->D_S - p->D_L);
  k.noise_c = float(p->noise_c);
  k.W0 = float(p->W[0]) - float(p->W[2]);
  k.W1 = float(p->W[1]) - float(p->W[2]);
  MlpK m;
  for (int q = 0; q < 2; ++q) {
    m.w[q] = static_cast<const f16x8*>(net->weights[q]);
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
  m.R = float(th->liquid.R);
  m.marks = net->marks;
  const int64_t ncell = (g->n[0] + 2) * (g->n[1] + 2);
  // byte offsets of every element the kernel addresses (field stack and
  // transfer array) stay below 2^32
  const int64_t fbytes = (4 * g->sf + (g->n[1] + 2) * g->sy) * 4, tbytes = (4 * tg->sf + (g->n[1] + 2) * tg->sy) * 4;
  if (ncell >= (int64_t(1) << 31) || fbytes >= (int64_t(1) << 32) || tbytes >= (int64_t(1) << 32))
    return fail(PPF_EINVAL, "ncomp_thermo_mlp: grid too large for 32-bit offsets");
  const uint32_t ntiles = uint32_t((ncell + 63) / 64);
  const unsigned blocks = unsigned(std::min<int64_t>((ntiles + 3) / 4, int64_t(PPF_MLP_BLOCKS_PER_CU) * device_cus()));
  const bool s2 = (net->flags & PPF_MLP_SPLIT2) != 0, s3 = (net->flags & PPF_MLP_SPLIT3) != 0;
  MlpLean lg;
  lg.px = uint32_t(g->n[0] + 2);
  lg.py = uint32_t(g->n[1] + 2);
  lg.nx = uint32_t(g->n[0]);
  lg.ny = uint32_t(g->n[1]);
  lg.sy = uint32_t(g->sy);
  lg.sf4 = uint32_t(g->sf * 4);
  lg.tsy = uint32_t(tg->sy);
  lg.tsf4 = uint32_t(tg->sf * 4);
  lg.ntiles = ntiles;
  const uint64_t stride_cells = uint64_t(blocks) * 4 * 64;
  lg.srow = uint32_t(stride_cells / lg.px);
  lg.scol = uint32_t(stride_cells % lg.px);
  lg.T_lo = m.T_lo;
  lg.T_hi = m.T_hi;
  lg.c_lo = m.c_lo;
  lg.R = m.R;
  lg.marks = m.marks;
  lg.thermo = net->thermo;
  const bool nz = k.noise_c != 0.f;
  auto kern = nz ? (s2 ? (s3 ? k_ncomp_thermo_mlp<true, true, true> : k_ncomp_thermo_mlp<true, false, true>)
                       : (s3 ? k_ncomp_thermo_mlp<false, true, true> : k_ncomp_thermo_mlp<false, false, true>))
                 : (s2 ? (s3 ? k_ncomp_thermo_mlp<true, true, false> : k_ncomp_thermo_mlp<true, false, false>)
                       : (s3 ? k_ncomp_thermo_mlp<false, true, false> : k_ncomp_thermo_mlp<false, false, false>));
  hipLaunchKernelGGL(kern, dim3(blocks), dim3(256), 0, stream_of(s), static_cast<const float*>(in),
                     static_cast<const float*>(bcvals), d, float(g->dx), *t, static_cast<float*>(transfer), k, m, lg,
                     seed, step);
  if (int rc = check_launch("ncomp_thermo_mlp")) return rc;
  if (net->thermo) return 0;  // cells outside the box already evaluated in the kernel
  const unsigned fblocks = unsigned(std::min<int64_t>((ncell + 255) / 256, int64_t(4) * device_cus()));
  hipLaunchKernelGGL(k_ncomp_thermo_fallback, dim3(fblocks), dim3(256), 0, stream_of(s),
                     static_cast<const float*>(in), static_cast<const float*>(bcvals), d, float(g->dx), *t,
                     static_cast<float*>(transfer), tg->sy, tg->sf, k, *th, net->marks, seed, step);
  return check_launch("ncomp_thermo_mlp");
}

int ppf_ncomp_thermo_mlp(const ppf_grid* g, const void* in, const void* bcvals, const ppf_temperature* t,
                         const ppf_grid* tg, void* transfer, const ppf_ncomp_params* p, const ppf_thermo* th,
                         const ppf_mlp_net* net, uint64_t seed, int64_t step, ppf_stream_t s) {
  return ncomp_thermo_mlp_dispatch(g, in, bcvals, t, tg, transfer, p, th, net, seed, step_host(step), s);
}

int ppf_ncomp_thermo_mlp_dev(const ppf_grid* g, const void* in, const void* bcvals, const ppf_temperature* t,
                             const ppf_grid* tg, void* transfer, const ppf_ncomp_params* p, const ppf_thermo* th,
                             const ppf_mlp_net* net, uint64_t seed, const int64_t* step_base, int64_t step_offset,
                             ppf_stream_t s) {
  if (!step_base) return fail(PPF_EINVAL, "ncomp_thermo_mlp_dev: null step counter");
  return ncomp_thermo_mlp_dispatch(g, in, bcvals, t, tg, transfer, p, th, net, seed,
                                   StepRef{step_base, step_offset}, s);
}

}  // extern "C"
