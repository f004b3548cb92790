// Warren1995 binary-alloy solidification step (phi, c), one fused kernel.
// Restates Engines/Warren1995.py:50-92 (interpolants :17-28, derivatives
// :30-48).  phi = 0 is solid, as in the reference (Warren1995.py:97-100).
#pragma once

#include "ppf_stencil.inl"

namespace ppf {

struct WarrenK {
  double W_A, W_B, kA, kB;  // kA = (30 L_A)(1/T - 1/T_mA), likewise kB
  double D_L, D_S, v_m, R, y_e, alpha, ebar2, M_A, M_B, dt;
};

// The reference names axis 0 "x" (gradx/gradxx roll along axis 0) and axis 1
// "y"; its gradx/grady are the negated centred differences
// (phi[i-1]-phi[i+1])/2dx.  Below, "r" is array axis 0 (rows, our y) and "k"
// is axis 1 (columns, our x), kept with the reference's sign convention.
template <typename R>
__global__ __launch_bounds__(kThreads) void k_warren(const R* __restrict__ in, R* __restrict__ out,
                                                     const R* __restrict__ b, Dom d, R dx, WarrenK p,
                                                     uint64_t seed, int64_t step) {
  constexpr int H = 1, W = kTX + 2 * H, HH = kTY + 2 * H;
  __shared__ R sphi[HH][W];
  __shared__ R sc[HH][W];
  __shared__ R sD[HH][W];
  __shared__ R stemp[HH][W];
  int64_t x0, y0;
  tile_coords(d, &x0, &y0);
  bool interior = tile_interior(d, x0, y0, H);
  load_tile<R, W, HH>(sphi, in, b, d, dx, x0 - H, y0 - H, interior);
  load_tile<R, W, HH>(sc, in + d.sf, b ? b + d.sf : nullptr, d, dx, x0 - H, y0 - H, interior);
  __syncthreads();
  const R D_L = R(p.D_L), D_S = R(p.D_S), v_m = R(p.v_m), Rg = R(p.R);
  const R W_A = R(p.W_A), W_B = R(p.W_B), kA = R(p.kA), kB = R(p.kB);
  // pointwise D_C and temp on the whole tile (Warren1995.py:69-70)
  for (int t = threadIdx.x; t < W * HH; t += kThreads) {
    int ly = t / W, lx = t - ly * W;
    R ph = sphi[ly][lx], c = sc[ly][lx];
    R pp = ph * ph * ph * (R(10) - R(15) * ph + R(6) * ph * ph);
    R g = (ph * ph * (R(1) - ph) * (R(1) - ph));
    R gp = (R(4) * ph * ph * ph - R(6) * ph * ph + R(2) * ph);
    R H_A = W_A * gp + kA * g;
    R H_B = W_B * gp + kB * g;
    R D_C = D_S + pp * (D_L - D_S);
    sD[ly][lx] = D_C;
    stemp[ly][lx] = D_C * v_m * c * (R(1) - c) * (H_B - H_A) / Rg;
  }
  __syncthreads();
  const R i2dx = R(1) / (R(2) * dx);
  const R idx2 = R(1) / (dx * dx);
  const R y_e = R(p.y_e), ebar2 = R(p.ebar2), M_A = R(p.M_A), M_B = R(p.M_B), dt = R(p.dt);
  const R alpha = R(p.alpha);
  const uint64_t ncells = uint64_t(d.ax.n + 2) * uint64_t(d.ay.n + 2);
  int tx = threadIdx.x % kTX, ty = threadIdx.x / kTX;
  int64_t x = x0 + tx;
  if (x > d.ax.n) return;
  const int lx = tx + 1;
#pragma unroll 2
  for (int k = 0; k < kTY / 4; ++k) {
    int ly = ty + 4 * k + 1;
    int64_t y = y0 + ly - 1;
    if (y > d.ay.n) break;
    // rows: ly-1 = "i-1" (roll +1 along axis 0), ly+1 = "i+1"
    R ph = sphi[ly][lx], c = sc[ly][lx];
    R g = (ph * ph * (R(1) - ph) * (R(1) - ph));
    R gp = (R(4) * ph * ph * ph - R(6) * ph * ph + R(2) * ph);
    R H_A = W_A * gp + kA * g;
    R H_B = W_B * gp + kB * g;
    R phixx = (sphi[ly - 1][lx] + sphi[ly + 1][lx] - R(2) * ph) * idx2;
    R phiyy = (sphi[ly][lx - 1] + sphi[ly][lx + 1] - R(2) * ph) * idx2;
    R lphi = phixx + phiyy;
    R phix = (sphi[ly - 1][lx] - sphi[ly + 1][lx]) * i2dx;
    R phiy = (sphi[ly][lx - 1] - sphi[ly][lx + 1]) * i2dx;
    R phix_km = (sphi[ly - 1][lx - 1] - sphi[ly + 1][lx - 1]) * i2dx;
    R phix_kp = (sphi[ly - 1][lx + 1] - sphi[ly + 1][lx + 1]) * i2dx;
    R phixy = (phix_km - phix_kp) * i2dx;
    // change in c (Warren1995.py:69-71)
    R D_C = sD[ly][lx];
    R temp = stemp[ly][lx];
    R cxx = (sc[ly - 1][lx] + sc[ly + 1][lx] - R(2) * c) * idx2;
    R cyy = (sc[ly][lx - 1] + sc[ly][lx + 1] - R(2) * c) * idx2;
    R cx = (sc[ly - 1][lx] - sc[ly + 1][lx]) * i2dx;
    R cy = (sc[ly][lx - 1] - sc[ly][lx + 1]) * i2dx;
    R Dx = (sD[ly - 1][lx] - sD[ly + 1][lx]) * i2dx;
    R Dy = (sD[ly][lx - 1] - sD[ly][lx + 1]) * i2dx;
    R tmx = (stemp[ly - 1][lx] - stemp[ly + 1][lx]) * i2dx;
    R tmy = (stemp[ly][lx - 1] - stemp[ly][lx + 1]) * i2dx;
    R deltac = D_C * (cxx + cyy) + (Dx * cx + Dy * cy) + temp * (lphi) + (tmx * phix + tmy * phiy);
    // change in phi (Warren1995.py:76-86); theta = atan2(phiy, phix) evaluated
    // algebraically: cos/sin of 2theta and 4theta from the gradient, with the
    // arctan2(0, 0) = 0 case giving cos = 1, sin = 0.
    R r2 = phix * phix + phiy * phiy;
    R c2 = R(1), s2 = R(0);
    if (r2 > R(0)) {
      R ir2 = R(1) / r2;
      c2 = (phix * phix - phiy * phiy) * ir2;
      s2 = R(2) * phix * phiy * ir2;
    }
    R c4 = c2 * c2 - s2 * s2;
    R s4 = R(2) * s2 * c2;
    R eta = R(1) + y_e * c4;
    R etap = R(-4) * y_e * s4;
    R etapp = R(-16) * (eta - R(1));
    R M_phi = (R(1) - c) * M_A + c * M_B;
    R deltaphi = M_phi * ((ebar2 * eta * eta * lphi - (R(1) - c) * H_A - c * H_B) +
                          ebar2 * eta * etap * (s2 * (phiyy - phixx) + R(2) * c2 * phixy) +
                          R(0.5) * ebar2 * (etap * etap + eta * etapp) *
                              (R(-2) * s2 * phixy + lphi + c2 * (phiyy - phixx)));
    if (p.alpha != 0.0) {
      uint64_t cell = uint64_t(y) * uint64_t(d.ax.n + 2) + uint64_t(x);
      R u = uniform01<R>(hash64(seed, rng_counter(step, S_WARREN_PHI, ncells, cell)));
      R rnd = R(2) * u - R(1);
      deltaphi = deltaphi + M_phi * alpha * rnd * (R(16) * g) * ((R(1) - c) * H_A + c * H_B);
    }
    int64_t o = y * d.ay.stride + x;
    out[o] = ph + deltaphi * dt;
    out[o + d.sf] = c + deltac * dt;
  }
}

template <typename R>
int warren_step(const ppf_grid& g, const void* in, void* out, const void* bcvals, const ppf_warren_params& p,
                uint64_t seed, int64_t step, hipStream_t s) {
  Dom d = make_dom(g);
  WarrenK k;
  k.W_A = p.W_A;
  k.W_B = p.W_B;
  k.kA = 30 * p.L_A * (1 / p.T - 1 / p.T_mA);
  k.kB = 30 * p.L_B * (1 / p.T - 1 / p.T_mB);
  k.D_L = p.D_L;
  k.D_S = p.D_S;
  k.v_m = p.v_m;
  k.R = p.R;
  k.y_e = p.y_e;
  k.alpha = p.alpha;
  k.ebar2 = p.ebar * p.ebar;
  k.M_A = p.M_A;
  k.M_B = p.M_B;
  k.dt = p.dt;
  hipLaunchKernelGGL(k_warren<R>, dim3(tiles2d(g)), dim3(kThreads), 0, s, static_cast<const R*>(in),
                     static_cast<R*>(out), static_cast<const R*>(bcvals), d, R(g.dx), k, seed, step);
  return check_launch("warren_step");
}

}  // namespace ppf
