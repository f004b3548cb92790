// AnisoDorr binary polycrystalline solidification (2D; fields phi, q1, q4, c):
// the vertex pass replacing AnisoDorr_helper_kernel
// (Engines/AnisoDorrGPU.py:262-375) and the step replacing AnisoDorr_kernel
// (:54-245, device helpers :25-52).  Expressions keep the reference's
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
  const R g = ad_g(phi), gp = ad_gprime(phi);
  const R m = R(1) - ad_h(phi);
  const R H_A = R(k.W_A) * gp - R(k.kA) * g;
  const R H_B = R(k.W_B) * gp - R(k.kB) * g;
  const R D_C = R(k.D_S) + m * R(k.dLS);
  const R temp = D_C * R(k.v_m) * c * (R(1) - c) * (H_B - H_A) / R(k.Rg);
  R pfx = R(0), pfy = R(0), t1t = R(0), t4t = R(0);
  if (x >= 1 && y >= 1) {
    // cells [i][j], [i][j-1], [i-1][j], [i-1][j-1] of the reference (i = y, j = x)
    const R pl = F(0, x - 1, y), pu = F(0, x, y - 1), pul = F(0, x - 1, y - 1);
    const R gx = R(0.5) * (phi - pl + pu - pul) * idx;
    const R gy = R(0.5) * (phi - pu + pl - pul) * idx;
    const R q1m = R(0.25) * (F(1, x, y) + F(1, x, y - 1) + F(1, x - 1, y) + F(1, x - 1, y - 1));
    const R q4m = R(0.25) * (F(2, x, y) + F(2, x, y - 1) + F(2, x - 1, y) + F(2, x - 1, y - 1));
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
    pfx = ye4 * ((R(2) * a2_b2 * sx3 + R(2) * ab2 * sy3) / mg2 - gx * (sx3 * gx + sy3 * gy) / (mg2 * mg2));
    pfy = ye4 * ((R(2) * a2_b2 * sy3 - R(2) * ab2 * sx3) / mg2 - gy * (sx3 * gx + sy3 * gy) / (mg2 * mg2));
    const R q1px = q1m * gx, q1py = q1m * gy, q4px = q4m * gx, q4py = q4m * gy;
    const R kk = R(k.k16) / mg2;
    t1t = kk * (psi_xyy * (q1px - q4py) + psi_xxy * (q1py + q4px));
    t4t = kk * (psi_xyy * (-q4px - q1py) + psi_xxy * (-q4py + q1px));
  }
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
  const R idx = R(1) / dx;
  const R idx2 = idx * idx;
  const R hidx2 = R(0.5) * idx * idx;
  // five-point neighbourhoods: centre, x+, x-, y+, y-
  const R p = F(0, x, y), pxp = F(0, x + 1, y), pxm = F(0, x - 1, y), pyp = F(0, x, y + 1), pym = F(0, x, y - 1);
  const R a = F(1, x, y), axp = F(1, x + 1, y), axm = F(1, x - 1, y), ayp = F(1, x, y + 1), aym = F(1, x, y - 1);
  const R b = F(2, x, y), bxp = F(2, x + 1, y), bxm = F(2, x - 1, y), byp = F(2, x, y + 1), bym = F(2, x, y - 1);
  const R c = F(3, x, y), cxp = F(3, x + 1, y), cxm = F(3, x - 1, y), cyp = F(3, x, y + 1), cym = F(3, x, y - 1);
  const int64_t o = y * tsy + x;
  const R* T4 = tr + 4 * tsf;
  const R* DC = tr + 5 * tsf;

  const R g = ad_g(p), h = ad_h(p);
  const R m = R(1) - h;
  const R M_q = R(k.Mq0) + R(k.dMq) * m;
  const R lq1 = (axp + axm + ayp + aym - R(4) * a) * idx2;
  const R lq4 = (bxp + bxm + byp + bym - R(4) * b) * idx2;
  const R pp = R(2) * p;
  const R gp = ad_gprime(p);
  const R H_A = R(k.W_A) * gp - R(k.kA) * g;
  const R H_B = R(k.W_B) * gp - R(k.kB) * g;
  // quaternion gradient magnitudes per face
  const R g1xm = (a - axm) * idx, g1ym = (a - aym) * idx, g1xp = (a - axp) * idx, g1yp = (a - ayp) * idx;
  const R g4xm = (b - bxm) * idx, g4ym = (b - bym) * idx, g4xp = (b - bxp) * idx, g4yp = (b - byp) * idx;
  R sxm = g1xm * g1xm + g4xm * g4xm;
  R sym = g1ym * g1ym + g4ym * g4ym;
  R sxp = g1xp * g1xp + g4xp * g4xp;
  R syp = g1yp * g1yp + g4yp * g4yp;
  const R rg0 = sqrt(R(0.5) * (sxm + sym + sxp + syp));
  const R beta = R(k.beta);
  sxm = sxm > beta ? sxm : beta;
  sym = sym > beta ? sym : beta;
  sxp = sxp > beta ? sxp : beta;
  syp = syp > beta ? syp : beta;
  const R rxm = sqrt(sxm), rym = sqrt(sym), rxp = sqrt(sxp), ryp = sqrt(syp);

  // dc/dt: div(temp grad phi) + div(D_C grad c)
  const R tc = T4[o], txp = T4[o + 1], txm = T4[o - 1], typ = T4[o + tsy], tym = T4[o - tsy];
  R dc = hidx2 * ((txp + tc) * (pxp - p) - (tc + txm) * (p - pxm) + (typ + tc) * (pyp - p) - (tc + tym) * (p - pym));
  const R dcc = DC[o], dxp = DC[o + 1], dxm = DC[o - 1], dyp = DC[o + tsy], dym = DC[o - tsy];
  dc = dc + hidx2 * ((dxp + dcc) * (cxp - c) - (dcc + dxm) * (c - cxm) + (dyp + dcc) * (cyp - c) -
                     (dcc + dym) * (c - cym));

  // dphi/dt
  const R lphi = (pxp + pxm + pyp + pym - R(4) * p) * idx2;
  const R* X = tr;
  const R* Y = tr + tsf;
  const R pfx = R(0.5) * idx * (X[o + 1] + X[o + tsy + 1] - X[o] - X[o + tsy]);
  const R pfy = R(0.5) * idx * (Y[o + tsy] + Y[o + tsy + 1] - Y[o] - Y[o + 1]);
  const R M_phi = (R(1) - c) * R(k.M_A) + c * R(k.M_B);
  R dphi = M_phi * (R(k.ebar2) * (R(k.one3ye) * lphi + pfx + pfy) - (R(1) - c) * H_A - c * H_B -
                    R(k.twoHT) * pp * rg0);
  const uint64_t ncell = uint64_t(nx + 2) * uint64_t(ny + 2);
  const uint64_t cell = uint64_t(y) * uint64_t(nx + 2) + uint64_t(x);
  const R rnd = uniform01<R>(hash64(seed, rng_counter(step, S_ANISO_PHI, ncell, cell)));
  dphi = dphi + M_phi * R(k.alpha) * rnd * (R(16) * g) * ((R(1) - c) * H_A + c * H_B);

  // dq/dt
  const R base = R(k.twoHT);
  const R Dq = base * (p * p);
  const R Dxp = base * pxp * pxp, Dxm = base * pxm * pxm, Dyp = base * pyp * pyp, Dym = base * pym * pym;
  const R gaq1 = hidx2 * ((Dq + Dxp) * (axp - a) / rxp - (Dq + Dxm) * (a - axm) / rxm +
                          (Dq + Dyp) * (ayp - a) / ryp - (Dq + Dym) * (a - aym) / rym);
  const R gaq4 = hidx2 * ((Dq + Dxp) * (bxp - b) / rxp - (Dq + Dxm) * (b - bxm) / rxm +
                          (Dq + Dyp) * (byp - b) / ryp - (Dq + Dym) * (b - bym) / rym);
  const R* T1 = tr + 2 * tsf;
  const R* T4t = tr + 3 * tsf;
  const R cc1 = R(0.25) * (T1[o] + T1[o + tsy] + T1[o + 1] + T1[o + tsy + 1]);
  const R cc4 = R(0.25) * (T4t[o] + T4t[o + tsy] + T4t[o + 1] + T4t[o + tsy + 1]);
  const R t1 = R(k.eqbar2) * lq1 + gaq1 + cc1;
  const R t4 = R(k.eqbar2) * lq4 + gaq4 + cc4;
  const R lm = a * t1 + b * t4;
  const R dq1 = M_q * (t1 - a * lm);
  const R dq4 = M_q * (t4 - b * lm);

  const R dt = R(k.dt);
  const int64_t of = y * d.ay.stride + x;
  out[of] = p + dphi * dt;
  const R q1n = a + dq1 * dt;
  const R q4n = b + dq4 * dt;
  out[of + 3 * d.sf] = c + dc * dt;
  const R rn = sqrt(q1n * q1n + q4n * q4n);
  out[of + d.sf] = q1n / rn;
  out[of + 2 * d.sf] = q4n / rn;
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
