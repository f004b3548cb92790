// AnisoDorr binary polycrystalline solidification (2D; fields phi, q1, q4, c):
// the vertex pass replacing AnisoDorr_helper_kernel
// (Engines/AnisoDorrGPU.py:262-375) and the step replacing AnisoDorr_kernel
// (:54-245, device helpers :25-52), as two kernels through a transfer stack
// (the reference's structure) or as one tiled kernel (k_aniso_fused) that
// keeps the transfer terms in LDS; both evaluate the same device functions.  Expressions keep the reference's
// operation order (FP64 parity); its hard-coded constants arrive in
// ppf_aniso_params.
//
// Transfer stack (6 arrays, every padded cell):
//   [0] pf_comp_xmm, [1] pf_comp_ymm, [2] t1_temp, [3] t4_temp  -- at the
//       vertex (x-1/2, y-1/2); zero on padded row 0 / column 0, which the
//       step never reads (the reference computes them from index -1)
//   [4] temp = D_C v_m c (1-c) (H_B - H_A) / R,  [5] D_C = D_S + (1-h)(D_L - D_S)
#pragma once

#include "ppf_stencil.inl"

namespace ppf {

constexpr int S_ANISO_PHI = 8;

struct AnisoK {
  double T, Mq0, dMq, twoHT, W_A, W_B, kA, kB, D_S, dLS, v_m, Rg, ye4, k16, beta, alpha, ebar2, eqbar2,
      one3ye, M_A, M_B, dt;
};

inline AnisoK aniso_constants(const ppf_aniso_params& p) {
  AnisoK k;
  k.T = p.T;
  k.Mq0 = 1e-6;
  k.dMq = p.M_qmax - 1e-6;
  k.twoHT = 2 * p.H * p.T;
  k.W_A = p.W_A;
  k.W_B = p.W_B;
  k.kA = 30 * p.L_A * (1 / p.T - 1 / p.T_mA);
  k.kB = 30 * p.L_B * (1 / p.T - 1 / p.T_mB);
  k.D_S = p.D_S;
  k.dLS = p.D_L - p.D_S;
  k.v_m = p.v_m;
  k.Rg = p.R;
  k.ye4 = 4 * p.y_e;
  k.k16 = 16 * p.ebar * p.ebar * p.y_e;
  k.beta = p.beta;
  k.alpha = p.alpha;
  k.ebar2 = p.ebar * p.ebar;
  k.eqbar2 = p.eqbar * p.eqbar;
  k.one3ye = 1 - 3 * p.y_e;
  k.M_A = p.M_A;
  k.M_B = p.M_B;
  k.dt = p.dt;
  return k;
}

template <typename R>
__device__ __forceinline__ R ad_g(R p) {
  return p * p * (R(1) - p) * (R(1) - p);
}
template <typename R>
__device__ __forceinline__ R ad_gprime(R p) {
  return R(4) * p * p * p - R(6) * p * p + R(2) * p;
}
template <typename R>
__device__ __forceinline__ R ad_h(R p) {
  return p * p * p * (R(10) - R(15) * p + R(6) * p * p);
}

template <typename R>
struct AdLoad {
  const R* f;
  const R* b;
  const Dom* d;
  R dx;
  __device__ __forceinline__ R operator()(int fi, int64_t x, int64_t y) const {
    return fold(f + fi * d->sf, b ? b + fi * d->sf : nullptr, *d, x, y, int64_t(0), dx);
  }
};

// Pointwise terms of one cell (AnisoDorrGPU.py:338-345).
template <typename R>
__device__ __forceinline__ void aniso_point(const AnisoK& k, R phi, R c, R& temp, R& D_C) {
  const R g = ad_g(phi), gp = ad_gprime(phi);
  const R m = R(1) - ad_h(phi);
  const R H_A = R(k.W_A) * gp - R(k.kA) * g;
  const R H_B = R(k.W_B) * gp - R(k.kB) * g;
  D_C = R(k.D_S) + m * R(k.dLS);
  if constexpr (sizeof(R) == 8) {
    temp = D_C * R(k.v_m) * c * (R(1) - c) * (H_B - H_A) / R(k.Rg);
  } else {  // FP32: v_m / R as one constant instead of a division per point
    temp = D_C * R(k.v_m / k.Rg) * c * (R(1) - c) * (H_B - H_A);
  }
}

// Vertex terms at (x-1/2, y-1/2) (AnisoDorrGPU.py:346-375) from the cells
// (x,y), (x-1,y), (x,y-1), (x-1,y-1) -- the reference's [i][j], [i][j-1],
// [i-1][j], [i-1][j-1] with i = y, j = x.
template <typename R>
__device__ __forceinline__ void aniso_vertex(const AnisoK& k, R idx, R phi, R pl, R pu, R pul, R q1, R q1u, R q1l,
                                             R q1ul, R q4, R q4u, R q4l, R q4ul, R& pfx, R& pfy, R& t1t, R& t4t) {
  const R gx = R(0.5) * (phi - pl + pu - pul) * idx;
  const R gy = R(0.5) * (phi - pu + pl - pul) * idx;
  const R q1m = R(0.25) * (q1 + q1u + q1l + q1ul);
  const R q4m = R(0.25) * (q4 + q4u + q4l + q4ul);
  const R a2_b2 = q1m * q1m - q4m * q4m;
  const R ab2 = R(2) * q1m * q4m;
  const R sx = a2_b2 * gx - ab2 * gy;
  const R sy = a2_b2 * gy + ab2 * gx;
  const R psi_xxy = sx * sx * sy;
  const R psi_xyy = sx * sy * sy;
  R mg2 = sx * sx + sy * sy;
  mg2 = mg2 > R(0.000000001) ? mg2 : R(0.000000001);
  const R sx3 = sx * sx * sx, sy3 = sy * sy * sy;
  const R ye4 = R(k.ye4);
  R kk;
  if constexpr (sizeof(R) == 8) {
    pfx = ye4 * ((R(2) * a2_b2 * sx3 + R(2) * ab2 * sy3) / mg2 - gx * (sx3 * gx + sy3 * gy) / (mg2 * mg2));
    pfy = ye4 * ((R(2) * a2_b2 * sy3 - R(2) * ab2 * sx3) / mg2 - gy * (sx3 * gx + sy3 * gy) / (mg2 * mg2));
    kk = R(k.k16) / mg2;
  } else {
    // FP32: one reciprocal of the clamped |psi|^2 (>= 1e-9, a normal float)
    const R img2 = __builtin_amdgcn_rcpf(mg2);
    const R t = (sx3 * gx + sy3 * gy) * (img2 * img2);
    pfx = ye4 * (R(2) * (a2_b2 * sx3 + ab2 * sy3) * img2 - gx * t);
    pfy = ye4 * (R(2) * (a2_b2 * sy3 - ab2 * sx3) * img2 - gy * t);
    kk = R(k.k16) * img2;
  }
  const R q1px = q1m * gx, q1py = q1m * gy, q4px = q4m * gx, q4py = q4m * gy;
  t1t = kk * (psi_xyy * (q1px - q4py) + psi_xxy * (q1py + q4px));
  t4t = kk * (psi_xyy * (-q4px - q1py) + psi_xxy * (-q4py + q1px));
}

// Five-point neighbourhood: centre, x+1, x-1, y+1, y-1.
template <typename R>
struct Ad5 {
  R c, xp, xm, yp, ym;
};
// Four vertices of a cell: (x,y), (x+1,y), (x,y+1), (x+1,y+1).
template <typename R>
struct Ad4 {
  R v00, v10, v01, v11;
};

// One cell of AnisoDorr_kernel (AnisoDorrGPU.py:136-245).  The noise
// (uniform [0,1), counter RNG) multiplies 16 g, so it is drawn only where
// g != 0.
template <typename R>
__device__ __forceinline__ void aniso_cell(const AnisoK& k, R idx, const Ad5<R>& P, const Ad5<R>& A,
                                           const Ad5<R>& B, const Ad5<R>& C, const Ad5<R>& T4, const Ad5<R>& DC,
                                           const Ad4<R>& X, const Ad4<R>& Y, const Ad4<R>& T1, const Ad4<R>& T4v,
                                           uint64_t seed, int64_t step, uint64_t ncell, uint64_t cell, R& nphi,
                                           R& nq1, R& nq4, R& nc) {
  const R idx2 = idx * idx;
  const R hidx2 = R(0.5) * idx * idx;
  const R p = P.c, a = A.c, b = B.c, c = C.c;
  const R g = ad_g(p), h = ad_h(p);
  const R m = R(1) - h;
  const R M_q = R(k.Mq0) + R(k.dMq) * m;
  const R lq1 = (A.xp + A.xm + A.yp + A.ym - R(4) * a) * idx2;
  const R lq4 = (B.xp + B.xm + B.yp + B.ym - R(4) * b) * idx2;
  const R pp = R(2) * p;
  const R gp = ad_gprime(p);
  const R H_A = R(k.W_A) * gp - R(k.kA) * g;
  const R H_B = R(k.W_B) * gp - R(k.kB) * g;
  // quaternion gradient magnitudes per face
  const R g1xm = (a - A.xm) * idx, g1ym = (a - A.ym) * idx, g1xp = (a - A.xp) * idx, g1yp = (a - A.yp) * idx;
  const R g4xm = (b - B.xm) * idx, g4ym = (b - B.ym) * idx, g4xp = (b - B.xp) * idx, g4yp = (b - B.yp) * idx;
  R sxm = g1xm * g1xm + g4xm * g4xm;
  R sym = g1ym * g1ym + g4ym * g4ym;
  R sxp = g1xp * g1xp + g4xp * g4xp;
  R syp = g1yp * g1yp + g4yp * g4yp;
  const R beta = R(k.beta);
  R rg0;
  if constexpr (sizeof(R) == 8) {
    rg0 = sqrt(R(0.5) * (sxm + sym + sxp + syp));
  } else {
    rg0 = __builtin_amdgcn_sqrtf(R(0.5) * (sxm + sym + sxp + syp));
  }
  sxm = sxm > beta ? sxm : beta;
  sym = sym > beta ? sym : beta;
  sxp = sxp > beta ? sxp : beta;
  syp = syp > beta ? syp : beta;

  // dc/dt: div(temp grad phi) + div(D_C grad c)
  R dc = hidx2 * ((T4.xp + T4.c) * (P.xp - p) - (T4.c + T4.xm) * (p - P.xm) + (T4.yp + T4.c) * (P.yp - p) -
                  (T4.c + T4.ym) * (p - P.ym));
  dc = dc + hidx2 * ((DC.xp + DC.c) * (C.xp - c) - (DC.c + DC.xm) * (c - C.xm) + (DC.yp + DC.c) * (C.yp - c) -
                     (DC.c + DC.ym) * (c - C.ym));

  // dphi/dt
  const R lphi = (P.xp + P.xm + P.yp + P.ym - R(4) * p) * idx2;
  const R pfx = R(0.5) * idx * (X.v10 + X.v11 - X.v00 - X.v01);
  const R pfy = R(0.5) * idx * (Y.v01 + Y.v11 - Y.v00 - Y.v10);
  const R M_phi = (R(1) - c) * R(k.M_A) + c * R(k.M_B);
  R dphi = M_phi * (R(k.ebar2) * (R(k.one3ye) * lphi + pfx + pfy) - (R(1) - c) * H_A - c * H_B -
                    R(k.twoHT) * pp * rg0);
  if (g != R(0)) {
    const R rnd = uniform01<R>(hash64(seed, rng_counter(step, S_ANISO_PHI, ncell, cell)));
    dphi = dphi + M_phi * R(k.alpha) * rnd * (R(16) * g) * ((R(1) - c) * H_A + c * H_B);
  }

  // dq/dt
  R gaq1, gaq4;
  if constexpr (sizeof(R) == 8) {
    const R rxm = sqrt(sxm), rym = sqrt(sym), rxp = sqrt(sxp), ryp = sqrt(syp);
    const R base = R(k.twoHT);
    const R Dq = base * (p * p);
    const R Dxp = base * P.xp * P.xp, Dxm = base * P.xm * P.xm, Dyp = base * P.yp * P.yp, Dym = base * P.ym * P.ym;
    gaq1 = hidx2 * ((Dq + Dxp) * (A.xp - a) / rxp - (Dq + Dxm) * (a - A.xm) / rxm +
                    (Dq + Dyp) * (A.yp - a) / ryp - (Dq + Dym) * (a - A.ym) / rym);
    gaq4 = hidx2 * ((Dq + Dxp) * (B.xp - b) / rxp - (Dq + Dxm) * (b - B.xm) / rxm +
                    (Dq + Dyp) * (B.yp - b) / ryp - (Dq + Dym) * (b - B.ym) / rym);
  } else {
    // FP32: face weights (D + D_nb) / |grad q|_face = 2HT (p^2 + p_nb^2) / sqrt(s),
    // one reciprocal square root per face (s >= beta), 2HT in the common
    // factor, the weights shared by q1 and q4
    const R p2 = p * p, hk = hidx2 * R(k.twoHT);
    const R wxp = (p2 + P.xp * P.xp) * __builtin_amdgcn_rsqf(sxp), wxm = (p2 + P.xm * P.xm) * __builtin_amdgcn_rsqf(sxm);
    const R wyp = (p2 + P.yp * P.yp) * __builtin_amdgcn_rsqf(syp), wym = (p2 + P.ym * P.ym) * __builtin_amdgcn_rsqf(sym);
    gaq1 = hk * (wxp * (A.xp - a) - wxm * (a - A.xm) + wyp * (A.yp - a) - wym * (a - A.ym));
    gaq4 = hk * (wxp * (B.xp - b) - wxm * (b - B.xm) + wyp * (B.yp - b) - wym * (b - B.ym));
  }
  const R cc1 = R(0.25) * (T1.v00 + T1.v01 + T1.v10 + T1.v11);
  const R cc4 = R(0.25) * (T4v.v00 + T4v.v01 + T4v.v10 + T4v.v11);
  const R t1 = R(k.eqbar2) * lq1 + gaq1 + cc1;
  const R t4 = R(k.eqbar2) * lq4 + gaq4 + cc4;
  const R lm = a * t1 + b * t4;
  const R dq1 = M_q * (t1 - a * lm);
  const R dq4 = M_q * (t4 - b * lm);

  const R dt = R(k.dt);
  nphi = p + dphi * dt;
  const R q1n = a + dq1 * dt;
  const R q4n = b + dq4 * dt;
  nc = c + dc * dt;
  if constexpr (sizeof(R) == 8) {
    const R rn = sqrt(q1n * q1n + q4n * q4n);
    nq1 = q1n / rn;
    nq4 = q4n / rn;
  } else {
    const R irn = __builtin_amdgcn_rsqf(q1n * q1n + q4n * q4n);
    nq1 = q1n * irn;
    nq4 = q4n * irn;
  }
}

// ------------------------------------------------------- two-pass path --
// one thread per padded cell, 64 x 4 cells per block
template <typename R>
__global__ __launch_bounds__(256) void k_aniso_helper(const R* __restrict__ in, const R* __restrict__ bv, Dom d,
                                                      R dx, R* __restrict__ tr, int64_t tsy, int64_t tsf,
                                                      AnisoK k) {
  const int64_t px = d.ax.n + 2, py = d.ay.n + 2;
  const uint32_t nbx = uint32_t((px + 63) / 64);
  const uint32_t blk = xcd_swizzle(blockIdx.x, gridDim.x);
  const int64_t x = int64_t(blk % nbx) * 64 + (threadIdx.x & 63);
  const int64_t y = int64_t(blk / nbx) * 4 + (threadIdx.x >> 6);
  if (x >= px || y >= py) return;
  const AdLoad<R> F{in, bv, &d, dx};
  const R idx = R(1) / dx;
  const R phi = F(0, x, y), c = F(3, x, y);
  R temp, D_C;
  aniso_point(k, phi, c, temp, D_C);
  R pfx = R(0), pfy = R(0), t1t = R(0), t4t = R(0);
  if (x >= 1 && y >= 1)
    aniso_vertex(k, idx, phi, F(0, x - 1, y), F(0, x, y - 1), F(0, x - 1, y - 1), F(1, x, y), F(1, x, y - 1),
                 F(1, x - 1, y), F(1, x - 1, y - 1), F(2, x, y), F(2, x, y - 1), F(2, x - 1, y),
                 F(2, x - 1, y - 1), pfx, pfy, t1t, t4t);
  const int64_t o = y * tsy + x;
  tr[o] = pfx;
  tr[o + tsf] = pfy;
  tr[o + 2 * tsf] = t1t;
  tr[o + 3 * tsf] = t4t;
  tr[o + 4 * tsf] = temp;
  tr[o + 5 * tsf] = D_C;
}

// one thread per interior cell, 64 x 4 cells per block
template <typename R>
__global__ __launch_bounds__(256) void k_aniso_step(const R* __restrict__ in, R* __restrict__ out,
                                                    const R* __restrict__ bv, Dom d, R dx,
                                                    const R* __restrict__ tr, int64_t tsy, int64_t tsf, AnisoK k,
                                                    uint64_t seed, int64_t step) {
  const int64_t nx = d.ax.n, ny = d.ay.n;
  const uint32_t nbx = uint32_t((nx + 63) / 64);
  const uint32_t blk = xcd_swizzle(blockIdx.x, gridDim.x);
  const int64_t x = 1 + int64_t(blk % nbx) * 64 + (threadIdx.x & 63);
  const int64_t y = 1 + int64_t(blk / nbx) * 4 + (threadIdx.x >> 6);
  if (x > nx || y > ny) return;
  const AdLoad<R> F{in, bv, &d, dx};
  auto five = [&](int f) { return Ad5<R>{F(f, x, y), F(f, x + 1, y), F(f, x - 1, y), F(f, x, y + 1), F(f, x, y - 1)}; };
  const int64_t o = y * tsy + x;
  auto tfive = [&](int f) {
    const R* t = tr + f * tsf;
    return Ad5<R>{t[o], t[o + 1], t[o - 1], t[o + tsy], t[o - tsy]};
  };
  auto tfour = [&](int f) {
    const R* t = tr + f * tsf;
    return Ad4<R>{t[o], t[o + 1], t[o + tsy], t[o + tsy + 1]};
  };
  R nphi, nq1, nq4, nc;
  aniso_cell(k, R(1) / dx, five(0), five(1), five(2), five(3), tfive(4), tfive(5), tfour(0), tfour(1), tfour(2),
             tfour(3), seed, step, uint64_t(nx + 2) * uint64_t(ny + 2), uint64_t(y) * uint64_t(nx + 2) + uint64_t(x),
             nphi, nq1, nq4, nc);
  const int64_t of = y * d.ay.stride + x;
  out[of] = nphi;
  out[of + d.sf] = nq1;
  out[of + 2 * d.sf] = nq4;
  out[of + 3 * d.sf] = nc;
}

// --------------------------------------------------------- fused path --
// One 64 x 8 output tile per 256-thread block.  The (64+2) x (8+2) cells of
// the four fields are loaded once (boundary tiles fold the rules in), the
// pointwise temp / D_C and the (64+1) x (8+1) vertex terms are computed into
// LDS, and each thread then updates two cells: the six transfer arrays of
// the two-pass path never touch HBM.
#ifndef PPF_AD_TY
#define PPF_AD_TY 8
#endif
constexpr int kAdTX = 64, kAdTY = PPF_AD_TY;
// 256 threads cover 64 columns x 4 row groups; each thread walks kAdTY / 4 rows
static_assert(kAdTY % 4 == 0 && kAdTY >= 4, "PPF_AD_TY must be a positive multiple of 4");

template <typename R>
__global__ __launch_bounds__(256) void k_aniso_fused(const R* __restrict__ in, R* __restrict__ out,
                                                     const R* __restrict__ bv, Dom d, R dx, AnisoK k, uint64_t seed,
                                                     StepRef sr) {
  const int64_t step = sr.get();
  constexpr int CW = kAdTX + 2, CH = kAdTY + 2, VW = kAdTX + 1, VH = kAdTY + 1;
  __shared__ R sf[4][CH][CW];
  __shared__ R sp[2][CH][CW];  // temp, D_C
  __shared__ R sv[4][VH][VW];  // pf_comp_x, pf_comp_y, t1_temp, t4_temp
  const int64_t nx = d.ax.n, ny = d.ay.n;
  const uint32_t nbx = uint32_t((nx + kAdTX - 1) / kAdTX);
  const uint32_t blk = xcd_swizzle(blockIdx.x, gridDim.x);
  const int64_t x0 = 1 + int64_t(blk % nbx) * kAdTX, y0 = 1 + int64_t(blk / nbx) * kAdTY;
  const bool interior = x0 - 1 >= 1 && x0 + kAdTX <= nx && y0 - 1 >= 1 && y0 + kAdTY <= ny;
  const R idx = R(1) / dx;
  // 1) cells (x0-1 .. x0+64, y0-1 .. y0+8) and their pointwise terms; an
  //    interior tile issues all its loads before the first LDS store
  constexpr int NC = CW * CH, NP = (NC + 255) / 256;
  if (interior) {
    const R* __restrict__ base = in + ((y0 - 1) * d.ay.stride + (x0 - 1));
    R f[NP][4];
#pragma unroll
    for (int j = 0; j < NP; ++j) {
      const int t = threadIdx.x + j * 256;
      if (t < NC) {
        const int ly = t / CW, lx = t - ly * CW;
        const int64_t o = int64_t(ly) * d.ay.stride + lx;
#pragma unroll
        for (int i = 0; i < 4; ++i) f[j][i] = base[o + i * d.sf];
      }
    }
#pragma unroll
    for (int j = 0; j < NP; ++j) {
      const int t = threadIdx.x + j * 256;
      if (t < NC) {
        const int ly = t / CW, lx = t - ly * CW;
#pragma unroll
        for (int i = 0; i < 4; ++i) sf[i][ly][lx] = f[j][i];
        aniso_point(k, f[j][0], f[j][3], sp[0][ly][lx], sp[1][ly][lx]);
      }
    }
  } else {
    for (int t = threadIdx.x; t < NC; t += 256) {
      const int ly = t / CW, lx = t - ly * CW;
      const int64_t gx = x0 - 1 + lx, gy = y0 - 1 + ly;
      R f[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) f[i] = fold(in + i * d.sf, bv ? bv + i * d.sf : nullptr, d, gx, gy, int64_t(0), dx);
#pragma unroll
      for (int i = 0; i < 4; ++i) sf[i][ly][lx] = f[i];
      aniso_point(k, f[0], f[3], sp[0][ly][lx], sp[1][ly][lx]);
    }
  }
  __syncthreads();
  // 2) vertices (x0 .. x0+64, y0 .. y0+8): vertex (vx, vy) belongs to cell
  //    (vx+1, vy+1) of the tile
  for (int t = threadIdx.x; t < VW * VH; t += 256) {
    const int vy = t / VW, vx = t - vy * VW;
    const int cx = vx + 1, cy = vy + 1;
    aniso_vertex(k, idx, sf[0][cy][cx], sf[0][cy][cx - 1], sf[0][cy - 1][cx], sf[0][cy - 1][cx - 1], sf[1][cy][cx],
                 sf[1][cy - 1][cx], sf[1][cy][cx - 1], sf[1][cy - 1][cx - 1], sf[2][cy][cx], sf[2][cy - 1][cx],
                 sf[2][cy][cx - 1], sf[2][cy - 1][cx - 1], sv[0][vy][vx], sv[1][vy][vx], sv[2][vy][vx],
                 sv[3][vy][vx]);
  }
  __syncthreads();
  // 3) two output rows per thread
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const uint64_t ncell = uint64_t(nx + 2) * uint64_t(ny + 2);
#pragma unroll
  for (int r = 0; r < kAdTY / 4; ++r) {
    const int oy = ty + 4 * r;
    const int64_t gx = x0 + tx, gy = y0 + oy;
    if (gx > nx || gy > ny) continue;
    const int lx = tx + 1, ly = oy + 1;
    auto five = [&](const R (&a)[CH][CW]) {
      return Ad5<R>{a[ly][lx], a[ly][lx + 1], a[ly][lx - 1], a[ly + 1][lx], a[ly - 1][lx]};
    };
    auto four = [&](const R (&v)[VH][VW]) { return Ad4<R>{v[oy][tx], v[oy][tx + 1], v[oy + 1][tx], v[oy + 1][tx + 1]}; };
    R nphi, nq1, nq4, nc;
    aniso_cell(k, idx, five(sf[0]), five(sf[1]), five(sf[2]), five(sf[3]), five(sp[0]), five(sp[1]), four(sv[0]),
               four(sv[1]), four(sv[2]), four(sv[3]), seed, step, ncell,
               uint64_t(gy) * uint64_t(nx + 2) + uint64_t(gx), nphi, nq1, nq4, nc);
    const int64_t of = gy * d.ay.stride + gx;
    out[of] = nphi;
    out[of + d.sf] = nq1;
    out[of + 2 * d.sf] = nq4;
    out[of + 3 * d.sf] = nc;
  }
}

// ------------------------------------------------------------ wave tiles --
// FP32 grids with enough tiles: each wave of a 256-thread block owns a 62 x
// kAwTY tile.  Lane l holds column x0 - 1 + l of the four fields on the rows
// y0 - 1 .. y0 + kAwTY in registers (all loads issued together; halos a
// coordinate remap when there are no bc values and no Dirichlet side, folded
// otherwise), computes temp / D_C on those rows and the vertex terms on the
// vertex rows y0 .. y0 + kAwTY (vertex (x, y) at (x - 1/2, y - 1/2) from the
// cells of lanes l and l - 1 on rows y and y - 1), then updates its rows.  x
// neighbours come from DPP lane shifts, y neighbours from the lane's rows:
// no LDS and no barrier.  The device functions are the block path's.
#ifndef PPF_AD_WTY
#define PPF_AD_WTY 4
#endif
constexpr int kAwTX = 62, kAwTY = PPF_AD_WTY;

template <typename R>
__global__ __launch_bounds__(256) void k_aniso_w(const R* __restrict__ in, R* __restrict__ out,
                                                 const R* __restrict__ bv, Dom d, R dx, AnisoK k, uint64_t seed,
                                                 StepRef sr, uint32_t nstrips, uint32_t ntiles) {
  const int wave = __builtin_amdgcn_readfirstlane(int(threadIdx.x) / 64);
  const int lane = int(threadIdx.x) % 64;
  const uint32_t t = xcd_swizzle(blockIdx.x, gridDim.x) * 4u + uint32_t(wave);
  if (t >= ntiles) return;  // uniform per wave
  const int32_t x0 = 1 + int32_t(t % nstrips) * kAwTX, y0 = 1 + int32_t(t / nstrips) * kAwTY;
  const int32_t nx = int32_t(d.ax.n), ny = int32_t(d.ay.n);
  const int32_t xr = x0 - 1 + lane;
  constexpr int NR = kAwTY + 2;  // cell rows y0 - 1 .. y0 + kAwTY
  R F[4][NR];
  if (!bv && d.ax.bc != PPF_DIRICHLET && d.ay.bc != PPF_DIRICHLET) {
    const __amdgpu_buffer_rsrc_t rs = plane_rsrc(in);
    const uint32_t xo = uint32_t(remap_coord32(xr, nx, d.ax.bc)) * uint32_t(sizeof(R));
    const int32_t sy = int32_t(d.ay.stride * int64_t(sizeof(R))), sfb = int32_t(d.sf * int64_t(sizeof(R)));
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const int32_t yo = remap_coord32(y0 - 1 + r, ny, d.ay.bc) * sy;
#pragma unroll
      for (int i = 0; i < 4; ++i) F[i][r] = buf_load<R>(rs, xo, yo + i * sfb);
    }
  } else {
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const int64_t y = y0 - 1 + r;
      const bool ok = xr <= nx + 1 && y <= ny + 1;  // beyond: never used
#pragma unroll
      for (int i = 0; i < 4; ++i)
        F[i][r] = ok ? fold(in + i * d.sf, bv ? bv + i * d.sf : nullptr, d, int64_t(xr), y, int64_t(0), dx) : R(0);
    }
  }
  const R idx = R(1) / dx;
  R Tp[NR], DC[NR];  // temp, D_C
#pragma unroll
  for (int r = 0; r < NR; ++r) aniso_point(k, F[0][r], F[3][r], Tp[r], DC[r]);
  // vertex rows y0 .. y0 + kAwTY: row j from cell rows j (y - 1) and j + 1 (y)
  R V[4][kAwTY + 1];
#pragma unroll
  for (int j = 0; j < kAwTY + 1; ++j) {
    R L[3][2];  // lane l - 1 of rows y and y - 1
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      L[i][0] = lane_shift<kDppWaveShr1>(F[i][j + 1]);
      L[i][1] = lane_shift<kDppWaveShr1>(F[i][j]);
    }
    aniso_vertex(k, idx, F[0][j + 1], L[0][0], F[0][j], L[0][1], F[1][j + 1], F[1][j], L[1][0], L[1][1], F[2][j + 1],
                 F[2][j], L[2][0], L[2][1], V[0][j], V[1][j], V[2][j], V[3][j]);
  }
  const int64_t step = sr.get();
  const uint64_t ncell = uint64_t(nx + 2) * uint64_t(ny + 2);
  const bool out_lane = lane >= 1 && lane < 63 && xr <= nx;
  const __amdgpu_buffer_rsrc_t ws = plane_rsrc(out);
  const uint32_t oo = uint32_t(int64_t(y0) * d.ay.stride + xr) * uint32_t(sizeof(R));
  const int32_t osy = int32_t(d.ay.stride * int64_t(sizeof(R))), osf = int32_t(d.sf * int64_t(sizeof(R)));
#pragma unroll
  for (int r = 0; r < kAwTY; ++r) {
    const int c = r + 1;  // cell row of the output
    auto five = [&](const R (&a)[NR]) {
      return Ad5<R>{a[c], lane_shift<kDppWaveShl1>(a[c]), lane_shift<kDppWaveShr1>(a[c]), a[c + 1], a[c - 1]};
    };
    auto four = [&](const R (&v)[kAwTY + 1]) {
      return Ad4<R>{v[r], lane_shift<kDppWaveShl1>(v[r]), v[r + 1], lane_shift<kDppWaveShl1>(v[r + 1])};
    };
    const int32_t y = y0 + r;
    R nphi, nq1, nq4, nc;
    aniso_cell(k, idx, five(F[0]), five(F[1]), five(F[2]), five(F[3]), five(Tp), five(DC), four(V[0]), four(V[1]),
               four(V[2]), four(V[3]), seed, step, ncell, uint64_t(y) * uint64_t(nx + 2) + uint64_t(xr), nphi, nq1,
               nq4, nc);
    if (out_lane && y <= ny) {
      buf_store<R>(nphi, ws, oo, r * osy);
      buf_store<R>(nq1, ws, oo, r * osy + osf);
      buf_store<R>(nq4, ws, oo, r * osy + 2 * osf);
      buf_store<R>(nc, ws, oo, r * osy + 3 * osf);
    }
  }
}

// whether the wave-tile kernel runs the grid: FP32, enough wave tiles to fill
// the GPU, and the four fields' bytes within the 32-bit buffer offsets
// (PPF_AD_WAVE_TILES=0 keeps the block tiles everywhere)
#ifndef PPF_AD_WAVE_TILES
#define PPF_AD_WAVE_TILES 1
#endif
inline bool aniso_wave_fits(const ppf_grid& g, size_t elem) {
  if (!PPF_AD_WAVE_TILES || elem != sizeof(float) || g.ndim != 2) return false;
  const int64_t tiles = ((g.n[0] + kAwTX - 1) / kAwTX) * ((g.n[1] + kAwTY - 1) / kAwTY);
  return tiles >= 1024 && (3 * g.sf + (g.n[1] + 2) * g.sy) * int64_t(elem) < (int64_t(1) << 31);
}

template <typename R>
int aniso_helper(const ppf_grid& g, const void* in, const void* bcvals, const ppf_grid& tg, void* transfer,
                 const ppf_aniso_params& p, hipStream_t s) {
  Dom d = make_dom(g);
  dim3 grid(unsigned(((g.n[0] + 2 + 63) / 64) * ((g.n[1] + 2 + 3) / 4)));
  hipLaunchKernelGGL(k_aniso_helper<R>, grid, dim3(256), 0, s, static_cast<const R*>(in),
                     static_cast<const R*>(bcvals), d, R(g.dx), static_cast<R*>(transfer), tg.sy, tg.sf,
                     aniso_constants(p));
  return check_launch("aniso_helper");
}

template <typename R>
int aniso_fused_step(const ppf_grid& g, const void* in, void* out, const void* bcvals, const ppf_aniso_params& p,
                     uint64_t seed, StepRef step, hipStream_t s) {
  Dom d = make_dom(g);
  if (sizeof(R) == 4 && aniso_wave_fits(g, sizeof(R))) {
    const uint32_t nstrips = uint32_t((g.n[0] + kAwTX - 1) / kAwTX);
    const uint32_t ntiles = nstrips * uint32_t((g.n[1] + kAwTY - 1) / kAwTY);
    hipLaunchKernelGGL(k_aniso_w<R>, dim3((ntiles + 3) / 4), dim3(256), 0, s, static_cast<const R*>(in),
                       static_cast<R*>(out), static_cast<const R*>(bcvals), d, R(g.dx), aniso_constants(p), seed,
                       step, nstrips, ntiles);
    return check_launch("aniso_fused_step");
  }
  dim3 grid(unsigned(((g.n[0] + kAdTX - 1) / kAdTX) * ((g.n[1] + kAdTY - 1) / kAdTY)));
  hipLaunchKernelGGL(k_aniso_fused<R>, grid, dim3(256), 0, s, static_cast<const R*>(in), static_cast<R*>(out),
                     static_cast<const R*>(bcvals), d, R(g.dx), aniso_constants(p), seed, step);
  return check_launch("aniso_fused_step");
}

template <typename R>
int aniso_step(const ppf_grid& g, const void* in, void* out, const void* bcvals, const ppf_grid& tg,
               const void* transfer, const ppf_aniso_params& p, uint64_t seed, int64_t step, hipStream_t s) {
  Dom d = make_dom(g);
  dim3 grid(unsigned(((g.n[0] + 63) / 64) * ((g.n[1] + 3) / 4)));
  hipLaunchKernelGGL(k_aniso_step<R>, grid, dim3(256), 0, s, static_cast<const R*>(in), static_cast<R*>(out),
                     static_cast<const R*>(bcvals), d, R(g.dx), static_cast<const R*>(transfer), tg.sy, tg.sf,
                     aniso_constants(p), seed, step);
  return check_launch("aniso_step");
}

}  // namespace ppf
