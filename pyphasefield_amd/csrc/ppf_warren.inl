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
//
// Layout of the work: a 64 x 16 tile per 256-thread block.  Each thread loads
// its share of the (64+2) x (16+2) phi/c tile, derives D_C and temp there
// (pointwise, Warren1995.py:69-70) and stores all four in LDS; then each
// thread walks 4 consecutive rows of one column keeping a 3-row register
// window, so a cell costs 12 LDS reads.
template <typename R>
__device__ __forceinline__ R recip(R x) {
  if constexpr (sizeof(R) == 4) {
    return __builtin_amdgcn_rcpf(x);
  } else {
    return R(1) / x;
  }
}

template <typename R>
__global__ __launch_bounds__(kThreads) void k_warren(const R* __restrict__ in, R* __restrict__ out,
                                                     const R* __restrict__ b, Dom d, R dx, WarrenK p,
                                                     uint64_t seed, int64_t step) {
  constexpr int W = kTX + 2, HH = kTY + 2;
  __shared__ R sphi[HH][W];
  __shared__ R sc[HH][W];
  __shared__ R sD[HH][W];
  __shared__ R stemp[HH][W];
  int64_t x0, y0;
  tile_coords(d, &x0, &y0);
  const bool interior = tile_interior(d, x0, y0, 1);
  const int tx = threadIdx.x % kTX, ty = threadIdx.x / kTX;
  const R D_L = R(p.D_L), D_S = R(p.D_S), v_m = R(p.v_m), Rg = R(p.R);
  const R W_A = R(p.W_A), W_B = R(p.W_B), kA = R(p.kA), kB = R(p.kB);
  const R vmR = v_m / Rg;
  const int64_t xs = x0 - 1, ys = y0 - 1;
  const int32_t sy32 = int32_t(d.ay.stride);
  const R* __restrict__ gphi = in + (ys * d.ay.stride + xs);
  const R* __restrict__ gc = gphi + d.sf;
  // load + pointwise derivation (rows ty, ty+4, ...; columns tx and 64+tx)
#pragma unroll
  for (int k = 0; k < (HH + 3) / 4; ++k) {
    const int ly = ty + 4 * k;
    if (ly < HH) {
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int lx = tx + kTX * h;
        if (lx < W) {
          R ph, c;
          if (interior) {
            const int32_t o = ly * sy32 + lx;
            ph = gphi[o];
            c = gc[o];
          } else {
            const int64_t x = xs + lx, y = ys + ly;
            if (x > d.ax.n + 1 || y > d.ay.n + 1) {
              ph = R(1);
              c = R(0);
            } else {
              ph = fold(in, b, d, x, y, int64_t(0), dx);
              c = fold(in + d.sf, b ? b + d.sf : nullptr, d, x, y, int64_t(0), dx);
            }
          }
          const R pp = ph * ph * ph * (R(10) - R(15) * ph + R(6) * ph * ph);
          const R g = (ph * ph * (R(1) - ph) * (R(1) - ph));
          const R gp = (R(4) * ph * ph * ph - R(6) * ph * ph + R(2) * ph);
          const R H_A = W_A * gp + kA * g;
          const R H_B = W_B * gp + kB * g;
          const R D_C = D_S + pp * (D_L - D_S);
          R temp;
          if constexpr (sizeof(R) == 8) {
            temp = D_C * v_m * c * (R(1) - c) * (H_B - H_A) / Rg;
          } else {
            temp = D_C * vmR * c * (R(1) - c) * (H_B - H_A);
          }
          sphi[ly][lx] = ph;
          sc[ly][lx] = c;
          sD[ly][lx] = D_C;
          stemp[ly][lx] = temp;
        }
      }
    }
  }
  __syncthreads();

  const int64_t x = x0 + tx;
  if (x > d.ax.n) return;
  const R i2dx = R(1) / (R(2) * dx);
  const R idx2 = R(1) / (dx * dx);
  const R y_e = R(p.y_e), ebar2 = R(p.ebar2), M_A = R(p.M_A), M_B = R(p.M_B), dt = R(p.dt);
  const R alpha = R(p.alpha);
  const uint64_t ncells = uint64_t(d.ax.n + 2) * uint64_t(d.ay.n + 2);
  const int lx = tx + 1;
  const int ly0 = 1 + 4 * ty;
  // register window: phi rows ly-1, ly, ly+1 (3 columns each); c, D_C, temp:
  // centre of row ly-1, 3 columns of row ly, centre of row ly+1
  R pm0 = sphi[ly0 - 1][lx - 1], pm1 = sphi[ly0 - 1][lx], pm2 = sphi[ly0 - 1][lx + 1];
  R pc0 = sphi[ly0][lx - 1], pc1 = sphi[ly0][lx], pc2 = sphi[ly0][lx + 1];
  R cm = sc[ly0 - 1][lx], c0 = sc[ly0][lx - 1], c1 = sc[ly0][lx], c2v = sc[ly0][lx + 1];
  R Dm = sD[ly0 - 1][lx], D0 = sD[ly0][lx - 1], D1 = sD[ly0][lx], D2 = sD[ly0][lx + 1];
  R tm = stemp[ly0 - 1][lx], t0 = stemp[ly0][lx - 1], t1 = stemp[ly0][lx], t2 = stemp[ly0][lx + 1];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int ly = ly0 + r;
    const int64_t y = y0 + ly - 1;
    const R pp0 = sphi[ly + 1][lx - 1], pp1 = sphi[ly + 1][lx], pp2 = sphi[ly + 1][lx + 1];
    const R cp = sc[ly + 1][lx], Dp = sD[ly + 1][lx], tp = stemp[ly + 1][lx];
    if (y <= d.ay.n) {
      const R ph = pc1, c = c1;
      const R g = (ph * ph * (R(1) - ph) * (R(1) - ph));
      const R gp = (R(4) * ph * ph * ph - R(6) * ph * ph + R(2) * ph);
      const R H_A = W_A * gp + kA * g;
      const R H_B = W_B * gp + kB * g;
      // rows: pm = "i-1" (roll +1 along axis 0), pp = "i+1"
      const R phixx = (pm1 + pp1 - R(2) * ph) * idx2;
      const R phiyy = (pc0 + pc2 - R(2) * ph) * idx2;
      const R lphi = phixx + phiyy;
      const R phix = (pm1 - pp1) * i2dx;
      const R phiy = (pc0 - pc2) * i2dx;
      const R phix_km = (pm0 - pp0) * i2dx;
      const R phix_kp = (pm2 - pp2) * i2dx;
      const R phixy = (phix_km - phix_kp) * i2dx;
      // change in c (Warren1995.py:69-71)
      const R cxx = (cm + cp - R(2) * c) * idx2;
      const R cyy = (c0 + c2v - R(2) * c) * idx2;
      const R cx = (cm - cp) * i2dx;
      const R cy = (c0 - c2v) * i2dx;
      const R Dx = (Dm - Dp) * i2dx;
      const R Dy = (D0 - D2) * i2dx;
      const R tmx = (tm - tp) * i2dx;
      const R tmy = (t0 - t2) * i2dx;
      const R deltac = D1 * (cxx + cyy) + (Dx * cx + Dy * cy) + t1 * (lphi) + (tmx * phix + tmy * phiy);
      // change in phi (Warren1995.py:76-86); theta = atan2(phiy, phix) evaluated
      // algebraically: cos/sin of 2theta and 4theta from the gradient, with the
      // arctan2(0, 0) = 0 case giving cos = 1, sin = 0.
      const R r2 = phix * phix + phiy * phiy;
      R c2 = R(1), s2 = R(0);
      if (r2 > R(0)) {
        const R ir2 = recip(r2);
        c2 = (phix * phix - phiy * phiy) * ir2;
        s2 = R(2) * phix * phiy * ir2;
      }
      const R c4 = c2 * c2 - s2 * s2;
      const R s4 = R(2) * s2 * c2;
      const R eta = R(1) + y_e * c4;
      const R etap = R(-4) * y_e * s4;
      const R etapp = R(-16) * (eta - R(1));
      const R M_phi = (R(1) - c) * M_A + c * M_B;
      R deltaphi = M_phi * ((ebar2 * eta * eta * lphi - (R(1) - c) * H_A - c * H_B) +
                            ebar2 * eta * etap * (s2 * (phiyy - phixx) + R(2) * c2 * phixy) +
                            R(0.5) * ebar2 * (etap * etap + eta * etapp) *
                                (R(-2) * s2 * phixy + lphi + c2 * (phiyy - phixx)));
      // noise (Warren1995.py:85-86); its factor g vanishes in the bulk, where
      // the hash is skipped (the term is exactly zero there)
      if (p.alpha != 0.0 && g != R(0)) {
        const uint64_t cell = uint64_t(y) * uint64_t(d.ax.n + 2) + uint64_t(x);
        const R u = uniform01<R>(hash64(seed, rng_counter(step, S_WARREN_PHI, ncells, cell)));
        const R rnd = R(2) * u - R(1);
        deltaphi = deltaphi + M_phi * alpha * rnd * (R(16) * g) * ((R(1) - c) * H_A + c * H_B);
      }
      const int64_t o = y * d.ay.stride + x;
      out[o] = ph + deltaphi * dt;
      out[o + d.sf] = c + deltac * dt;
    }
    // slide the window down one row
    pm0 = pc0; pm1 = pc1; pm2 = pc2;
    pc0 = pp0; pc1 = pp1; pc2 = pp2;
    cm = c1; Dm = D1; tm = t1;
    c1 = cp; D1 = Dp; t1 = tp;
    if (r < 3) {
      c0 = sc[ly + 1][lx - 1]; c2v = sc[ly + 1][lx + 1];
      D0 = sD[ly + 1][lx - 1]; D2 = sD[ly + 1][lx + 1];
      t0 = stemp[ly + 1][lx - 1]; t2 = stemp[ly + 1][lx + 1];
    }
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
