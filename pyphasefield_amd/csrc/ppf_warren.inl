// Warren1995 binary-alloy solidification step (phi, c), one fused kernel.
// Restates Engines/Warren1995.py:50-92 (interpolants :17-28, derivatives
// :30-48).  phi = 0 is solid, as in the reference (Warren1995.py:97-100).
#pragma once

#include "ppf_stencil.inl"

namespace ppf {

// Tile shape knob of the LDS-tile kernel: rows per thread (the tile is
// 64 x 4*RPT) -- chosen by sweep.
#ifndef PPF_WARREN_RPT
#define PPF_WARREN_RPT 4
#endif
constexpr int kWRpt = PPF_WARREN_RPT;

constexpr int kWTY = 4 * kWRpt;

struct WarrenK {
  double W_A, W_B, kA, kB;  // kA = (30 L_A)(1/T - 1/T_mA), likewise kB
  double D_L, D_S, v_m, R, y_e, alpha, ebar2, M_A, M_B, dt;
};

// The reference names axis 0 "x" (gradx/gradxx roll along axis 0) and axis 1
// "y"; its gradx/grady are the negated centred differences
// (phi[i-1]-phi[i+1])/2dx.  Below, "r" is array axis 0 (rows, our y) and "k"
// is axis 1 (columns, our x), kept with the reference's sign convention.
//
// Layout of the work: a 64 x 16 tile per 256-thread block.  The threads load
// the (64+2) x (16+2) phi/c tile linearly, derive D_C and temp there
// (pointwise, Warren1995.py:69-70) and store all four in LDS; then each
// thread walks 4 consecutive rows of one column keeping a 3-row register
// window, so a cell costs 14 LDS reads.
template <typename R>
__device__ __forceinline__ R recip(R x) {
  if constexpr (sizeof(R) == 4) {
    return __builtin_amdgcn_rcpf(x);
  } else {
    return R(1) / x;
  }
}

// Per-launch constants of the cell update, in the arithmetic type.
template <typename R>
struct WarrenC {
  R D_L, D_S, v_m, Rg, vmR, W_A, W_B, kA, kB;
  R y_e, M_A, M_B, dt, A, dtc, alpha;
  R dD, dW, dk, dM;  // D_L - D_S, W_B - W_A, kB - kA, M_B - M_A (FP32 forms)
  bool noise;
  __host__ __device__ WarrenC(const WarrenK& p, R dx) {
    D_L = R(p.D_L); D_S = R(p.D_S); v_m = R(p.v_m); Rg = R(p.R); vmR = v_m / Rg;
    W_A = R(p.W_A); W_B = R(p.W_B); kA = R(p.kA); kB = R(p.kB);
    dD = R(p.D_L - p.D_S); dW = R(p.W_B - p.W_A); dk = R(p.kB - p.kA); dM = R(p.M_B - p.M_A);
    const R idx2 = R(1) / (dx * dx);
    y_e = R(p.y_e); M_A = R(p.M_A); M_B = R(p.M_B); dt = R(p.dt);
    A = R(p.ebar2) * idx2;  // ebar^2 / dx^2
    dtc = R(p.dt) * idx2;   // dt / dx^2 for the c update
    alpha = R(p.alpha);
    noise = p.alpha != 0.0;
  }
};

// pointwise D_C and temp (Warren1995.py:69-70)
template <typename R>
__device__ __forceinline__ void warren_point(const WarrenC<R>& K, R ph, R c, R& D_C, R& temp) {
  if constexpr (sizeof(R) == 8) {
    // FP64: the reference's expressions in its order
    const R pp = ph * ph * ph * (R(10) - R(15) * ph + R(6) * ph * ph);
    const R g = (ph * ph * (R(1) - ph) * (R(1) - ph));
    const R gp = (R(4) * ph * ph * ph - R(6) * ph * ph + R(2) * ph);
    const R H_A = K.W_A * gp + K.kA * g;
    const R H_B = K.W_B * gp + K.kB * g;
    D_C = K.D_S + pp * (K.D_L - K.D_S);
    temp = D_C * K.v_m * c * (R(1) - c) * (H_B - H_A) / K.Rg;
  } else {
    // FP32: Horner forms; H_B - H_A = (W_B - W_A) g' + (kB - kA) g
    const R ph2 = ph * ph;
    const R pp = ph2 * ph * __builtin_fmaf(ph, __builtin_fmaf(R(6), ph, R(-15)), R(10));
    const R q = ph - ph2;  // ph (1 - ph)
    const R g = q * q;
    const R gp = ph * __builtin_fmaf(ph, __builtin_fmaf(R(4), ph, R(-6)), R(2));
    D_C = __builtin_fmaf(pp, K.dD, K.D_S);
    temp = D_C * K.vmR * (c - c * c) * __builtin_fmaf(K.dW, gp, K.dk * g);
  }
}

// The cell update (Warren1995.py:69-86) from its stencil terms: raw grid
// differences with the dx scalings folded into the constants, phixx =
// Dxx/dx^2, phix = gx/(2dx), phixy = gxy/(4dx^2); products of two first
// differences carry 1/(4dx^2).  clap is the 5-point sum of c less 4c, cross
// the sum of the four products of first differences (D_C and c, temp and phi).
template <typename R>
__device__ __forceinline__ void warren_update(const WarrenC<R>& K, R ph, R c, R Dxx, R Dyy, R gx, R gy, R gxy,
                                              R clap, R cross, R D1, R t1, uint64_t seed, int64_t step,
                                              uint64_t ncells, uint64_t cell, R& nph, R& nc) {
  R g, gp, Hmix, M_phi;
  if constexpr (sizeof(R) == 8) {
    g = (ph * ph * (R(1) - ph) * (R(1) - ph));
    gp = (R(4) * ph * ph * ph - R(6) * ph * ph + R(2) * ph);
    Hmix = (R(1) - c) * (K.W_A * gp + K.kA * g) + c * (K.W_B * gp + K.kB * g);
    M_phi = (R(1) - c) * K.M_A + c * K.M_B;
  } else {
    // FP32: (1-c) H_A + c H_B = g' (W_A + c dW) + g (kA + c dk)
    const R q = ph - ph * ph;
    g = q * q;
    gp = ph * __builtin_fmaf(ph, __builtin_fmaf(R(4), ph, R(-6)), R(2));
    Hmix = __builtin_fmaf(gp, __builtin_fmaf(c, K.dW, K.W_A), g * __builtin_fmaf(c, K.dk, K.kA));
    M_phi = __builtin_fmaf(c, K.dM, K.M_A);
  }
  const R lap = Dxx + Dyy;
  // change in c (Warren1995.py:69-71):
  // D_C lap c + grad D_C . grad c + temp lap phi + grad temp . grad phi
  const R deltac = D1 * clap + t1 * lap + R(0.25) * cross;
  // change in phi (Warren1995.py:76-86); theta = atan2(phiy, phix) evaluated
  // algebraically from the gradient, with the arctan2(0, 0) = 0 case giving
  // cos = 1, sin = 0; eta'' = -16 (eta - 1) = -16 y_e cos(4 theta).  FP32:
  // the hardware reciprocal flushes a subnormal r2 to zero and returns
  // infinity, so gradients with r2 below the smallest normal float (phi
  // decaying to ~1e-22 inside the solid) take the zero-gradient case; the
  // terms they multiply are then below 1e-19.
  const R r2 = gx * gx + gy * gy;
  R c2 = R(1), s2 = R(0);
  if (r2 > (sizeof(R) == 4 ? R(0x1.0p-126f) : R(0))) {
    const R ir2 = recip(r2);
    c2 = (gx * gx - gy * gy) * ir2;
    s2 = R(2) * gx * gy * ir2;
  }
  const R c4 = c2 * c2 - s2 * s2;
  const R s4 = R(2) * s2 * c2;
  const R eta = R(1) + K.y_e * c4;
  const R etap = R(-4) * K.y_e * s4;
  const R etapp = R(-16) * K.y_e * c4;
  const R dd = Dyy - Dxx;
  const R B = eta * eta * lap + eta * etap * (s2 * dd + R(0.5) * c2 * gxy) +
              R(0.5) * (etap * etap + eta * etapp) * (lap - R(0.5) * s2 * gxy + c2 * dd);
  R deltaphi = M_phi * (K.A * B - Hmix);
  // noise (Warren1995.py:85-86); its factor g vanishes in the bulk, where
  // the hash is skipped (the term is exactly zero there)
  if (K.noise && g != R(0)) {
    const R u = uniform01<R>(hash64(seed, rng_counter(step, S_WARREN_PHI, ncells, cell)));
    const R rnd = R(2) * u - R(1);
    deltaphi = deltaphi + M_phi * K.alpha * rnd * (R(16) * g) * Hmix;
  }
  nph = ph + deltaphi * K.dt;
  nc = c + deltac * K.dtc;
}

// One cell from its 3x3 phi neighbourhood (rows "m" = i-1, "c" = i, "p" =
// i+1 along array axis 0; columns 0/1/2 = x-1/x/x+1) and the 5-point c, D_C,
// temp neighbourhoods.
template <typename R>
__device__ __forceinline__ void warren_cell(const WarrenC<R>& K, R pm0, R pm1, R pm2, R pc0, R ph, R pc2, R pp0,
                                            R pp1, R pp2, R cm, R c0, R c, R c2v, R cp, R Dm, R D0, R D1, R D2,
                                            R Dp, R tm, R t0, R t1, R t2, R tp, uint64_t seed, int64_t step,
                                            uint64_t ncells, uint64_t cell, R& nph, R& nc) {
  const R Dxx = pm1 + pp1 - R(2) * ph;
  const R Dyy = pc0 + pc2 - R(2) * ph;
  const R gx = pm1 - pp1;
  const R gy = pc0 - pc2;
  const R gxy = (pm0 - pp0) - (pm2 - pp2);
  const R clap = cm + cp + c0 + c2v - R(4) * c;
  const R cross = (Dm - Dp) * (cm - cp) + (D0 - D2) * (c0 - c2v) + (tm - tp) * gx + (t0 - t2) * gy;
  warren_update(K, ph, c, Dxx, Dyy, gx, gy, gxy, clap, cross, D1, t1, seed, step, ncells, cell, nph, nc);
}

// ---- pieces of the tile step
constexpr int kWW = kTX + 2, kWH = kWTY + 2;  // LDS tile: (64+2) x (16+2)
constexpr int kWNP = (kWW * kWH + kThreads - 1) / kThreads, kWNFULL = (kWW * kWH) / kThreads;

template <typename R>
struct WarrenLds {
  R phi[kWH][kWW];
  R c[kWH][kWW];
  R D[kWH][kWW];
  R temp[kWH][kWW];
};

// tile t (after the XCD swizzle over ntiles): output origin and whether its
// halo ring lies inside the interior (no boundary rule needed)
__device__ __forceinline__ void warren_tile(const Dom& d, uint32_t t, uint32_t ntiles, int64_t& x0, int64_t& y0,
                                            bool& interior) {
  const uint32_t nbx = uint32_t((d.ax.n + kTX - 1) / kTX);
  const uint32_t blk = xcd_swizzle(t, ntiles);
  x0 = 1 + int64_t(blk % nbx) * kTX;
  y0 = 1 + int64_t(blk / nbx) * kWTY;
  interior = x0 - 1 >= 1 && x0 + kTX <= d.ax.n && y0 - 1 >= 1 && y0 + kWTY <= d.ay.n;
}

// phi and c of an interior tile (halo included), kWNP per thread, walked
// linearly (LDS index t = ly * kWW + lx); only the last pass is partial
template <typename R>
__device__ __forceinline__ void warren_load(const R* __restrict__ in, const Dom& d, int64_t x0, int64_t y0,
                                            R (&vph)[kWNP], R (&vc)[kWNP]) {
  const int32_t sy32 = int32_t(d.ay.stride);
  const R* __restrict__ gphi = in + ((y0 - 1) * d.ay.stride + (x0 - 1));
  const R* __restrict__ gc = gphi + d.sf;
#pragma unroll
  for (int j = 0; j < kWNP; ++j) {
    const uint32_t t = threadIdx.x + uint32_t(j * kThreads);
    if (j < kWNFULL || t < uint32_t(kWW * kWH)) {
      const uint32_t ly = t / uint32_t(kWW), lx = t - ly * uint32_t(kWW);
      const int32_t o = int32_t(ly) * sy32 + int32_t(lx);
      vph[j] = gphi[o];
      vc[j] = gc[o];
    }
  }
}

template <typename R>
__device__ __forceinline__ void warren_put(WarrenLds<R>& S, const WarrenC<R>& K, int t, R ph, R c) {
  R D_C, temp;
  warren_point(K, ph, c, D_C, temp);
  (&S.phi[0][0])[t] = ph;
  (&S.c[0][0])[t] = c;
  (&S.D[0][0])[t] = D_C;
  (&S.temp[0][0])[t] = temp;
}

// pointwise D_C and temp (Warren1995.py:69-70) of a loaded interior tile
template <typename R>
__device__ __forceinline__ void warren_fill(WarrenLds<R>& S, const WarrenC<R>& K, const R (&vph)[kWNP],
                                            const R (&vc)[kWNP]) {
#pragma unroll
  for (int j = 0; j < kWNP; ++j) {
    const uint32_t t = threadIdx.x + uint32_t(j * kThreads);
    if (j < kWNFULL || t < uint32_t(kWW * kWH)) warren_put(S, K, int(t), vph[j], vc[j]);
  }
}

// the same for a tile touching the boundary: the rules are folded into the
// loads, in a loop kept out of line so the interior path stays short
template <typename R>
__device__ __forceinline__ void warren_fill_fold(WarrenLds<R>& S, const WarrenC<R>& K, const R* __restrict__ in,
                                                 const R* __restrict__ b, const Dom& d, R dx, int64_t x0,
                                                 int64_t y0) {
#pragma unroll 1
  for (int t = threadIdx.x; t < kWW * kWH; t += kThreads) {
    const int ly = t / kWW, lx = t - (t / kWW) * kWW;
    const int64_t x = x0 - 1 + lx, y = y0 - 1 + ly;
    R ph = R(1), c = R(0);
    if (x <= d.ax.n + 1 && y <= d.ay.n + 1) {
      ph = fold(in, b, d, x, y, int64_t(0), dx);
      c = fold(in + d.sf, b ? b + d.sf : nullptr, d, x, y, int64_t(0), dx);
    }
    warren_put(S, K, t, ph, c);
  }
}

// the cell updates of one tile from LDS: each thread walks kWRpt
// consecutive rows of one column keeping a 3-row register window, so a cell
// costs 14 LDS reads
template <typename R>
__device__ __forceinline__ void warren_rows(const WarrenLds<R>& S, const WarrenC<R>& K, R* __restrict__ out,
                                            const Dom& d, int64_t x0, int64_t y0, bool interior, uint64_t seed,
                                            int64_t step) {
  const int tx = threadIdx.x % kTX, ty = threadIdx.x / kTX;
  const int64_t x = x0 + tx;
  if (x > d.ax.n) return;
  const uint64_t ncells = uint64_t(d.ax.n + 2) * uint64_t(d.ay.n + 2);
  const int lx = tx + 1;
  const int ly0 = 1 + kWRpt * ty;
  R* __restrict__ orow = out + ((y0 + ly0 - 1) * d.ay.stride + x);
  // register window: phi rows ly-1, ly, ly+1 (3 columns each); c, D_C, temp:
  // centre of row ly-1, 3 columns of row ly, centre of row ly+1
  R pm0 = S.phi[ly0 - 1][lx - 1], pm1 = S.phi[ly0 - 1][lx], pm2 = S.phi[ly0 - 1][lx + 1];
  R pc0 = S.phi[ly0][lx - 1], pc1 = S.phi[ly0][lx], pc2 = S.phi[ly0][lx + 1];
  R cm = S.c[ly0 - 1][lx], c0 = S.c[ly0][lx - 1], c1 = S.c[ly0][lx], c2v = S.c[ly0][lx + 1];
  R Dm = S.D[ly0 - 1][lx], D0 = S.D[ly0][lx - 1], D1 = S.D[ly0][lx], D2 = S.D[ly0][lx + 1];
  R tm = S.temp[ly0 - 1][lx], t0 = S.temp[ly0][lx - 1], t1 = S.temp[ly0][lx], t2 = S.temp[ly0][lx + 1];
#pragma unroll
  for (int r = 0; r < kWRpt; ++r) {
    const int ly = ly0 + r;
    const int64_t y = y0 + ly - 1;
    const R pp0 = S.phi[ly + 1][lx - 1], pp1 = S.phi[ly + 1][lx], pp2 = S.phi[ly + 1][lx + 1];
    const R cp = S.c[ly + 1][lx], Dp = S.D[ly + 1][lx], tp = S.temp[ly + 1][lx];
    if (interior || y <= d.ay.n) {
      const uint64_t cell = uint64_t(y) * uint64_t(d.ax.n + 2) + uint64_t(x);
      R nph, nc;
      warren_cell(K, pm0, pm1, pm2, pc0, pc1, pc2, pp0, pp1, pp2, cm, c0, c1, c2v, cp, Dm, D0, D1, D2, Dp, tm, t0,
                  t1, t2, tp, seed, step, ncells, cell, nph, nc);
      R* __restrict__ o = orow + r * int32_t(d.ay.stride);
      o[0] = nph;
      o[d.sf] = nc;
    }
    // slide the window down one row
    pm0 = pc0; pm1 = pc1; pm2 = pc2;
    pc0 = pp0; pc1 = pp1; pc2 = pp2;
    cm = c1; Dm = D1; tm = t1;
    c1 = cp; D1 = Dp; t1 = tp;
    if (r < kWRpt - 1) {
      c0 = S.c[ly + 1][lx - 1]; c2v = S.c[ly + 1][lx + 1];
      D0 = S.D[ly + 1][lx - 1]; D2 = S.D[ly + 1][lx + 1];
      t0 = S.temp[ly + 1][lx - 1]; t2 = S.temp[ly + 1][lx + 1];
    }
  }
}

// One 64 x 16 tile per 256-thread block: load phi/c (halo included), derive
// D_C and temp there into LDS, then the row walk.
template <typename R>
__device__ __forceinline__ void warren_tile_step(const R* __restrict__ in, R* __restrict__ out,
                                                 const R* __restrict__ b, const Dom& d, R dx, const WarrenK& p,
                                                 uint64_t seed, StepRef sr) {
  __shared__ WarrenLds<R> S;
  const WarrenC<R> K(p, dx);
  const int64_t step = sr.get();
  int64_t x0, y0;
  bool interior;
  warren_tile(d, blockIdx.x, gridDim.x, x0, y0, interior);
  if (interior) {
    R vph[kWNP], vc[kWNP];
    warren_load(in, d, x0, y0, vph, vc);  // all loads first, then the arithmetic
    warren_fill(S, K, vph, vc);
  } else {
    warren_fill_fold(S, K, in, b, d, dx, x0, y0);
  }
  __syncthreads();
  warren_rows(S, K, out, d, x0, y0, interior, seed, step);
}

template <typename R>
__global__ __launch_bounds__(kThreads) void k_warren(const R* __restrict__ in, R* __restrict__ out,
                                                     const R* __restrict__ b, Dom d, R dx, WarrenK p,
                                                     uint64_t seed, StepRef sr) {
  warren_tile_step<R>(in, out, b, d, dx, p, seed, sr);
}

// ------------------------------------------------------------ wave tiles --
// Each wave of a 256-thread block owns a tile of its own: lane l holds column
// x0 - 1 + l of phi and c on the rows y0 - 1 .. y0 + kWvTY in registers (all
// its loads issued together), so the 64 lanes give 62 output columns.  The
// wave walks its rows computing D_C and temp one row ahead; x neighbours come
// from the adjacent lanes (DPP lane shifts), y neighbours from the lane's own
// rows.  No LDS and no barrier.  Halo rows and columns are a coordinate remap
// (periodic wrap, nearest interior cell for zero gradient): the kernel runs
// grids without bc values and without a Dirichlet side only.  Stencil
// terms: the x differences of a row are differences of shifted copies, and
// gxy = (pm0 - pp0) - (pm2 - pp2) is the difference of the x neighbours' gx,
// with the same operands and operations as warren_cell.
constexpr int kWvTX = 62;
#ifndef PPF_WARREN_WTY
#define PPF_WARREN_WTY 4
#endif
constexpr int kWvTY = PPF_WARREN_WTY;
// rows fetched ahead of the row being updated (kWvTY or more: the whole tile
// is fetched before the first update)
#ifndef PPF_WARREN_AHEAD
#define PPF_WARREN_AHEAD 64
#endif
constexpr int kWvAhead = PPF_WARREN_AHEAD;

// f(0), f(1), ... (compile-time row indices)
template <typename F, int... P>
__device__ __forceinline__ void warren_unrolled(F& f, std::integer_sequence<int, P...>) {
  (f(std::integral_constant<int, P>{}), ...);
}

template <typename R>
__device__ __forceinline__ void warren_wave_cell(const WarrenC<R>& K, R pm, R pc, R pp, R cm, R cc, R cp, R Dm,
                                                 R D1, R Dp, R tm, R t1, R tp, uint64_t seed, int64_t step,
                                                 uint64_t ncells, uint64_t cell, R& nph, R& nc) {
  const R pc0 = lane_shift<kDppWaveShr1>(pc), pc2 = lane_shift<kDppWaveShl1>(pc);
  const R Dxx = pm + pp - R(2) * pc;
  const R Dyy = pc0 + pc2 - R(2) * pc;
  const R gx = pm - pp;
  const R gy = pc0 - pc2;
  const R gxy = lane_shift<kDppWaveShr1>(gx) - lane_shift<kDppWaveShl1>(gx);
  const R c0 = lane_shift<kDppWaveShr1>(cc), c2v = lane_shift<kDppWaveShl1>(cc);
  const R clap = cm + cp + c0 + c2v - R(4) * cc;
  const R D0 = lane_shift<kDppWaveShr1>(D1), D2 = lane_shift<kDppWaveShl1>(D1);
  const R t0 = lane_shift<kDppWaveShr1>(t1), t2 = lane_shift<kDppWaveShl1>(t1);
  const R cross = (Dm - Dp) * (cm - cp) + (D0 - D2) * (c0 - c2v) + (tm - tp) * gx + (t0 - t2) * gy;
  warren_update(K, pc, cc, Dxx, Dyy, gx, gy, gxy, clap, cross, D1, t1, seed, step, ncells, cell, nph, nc);
}

// occupancy hint of the wave-tile kernel (waves per SIMD; 0 = compiler's choice)
#ifndef PPF_WARREN_W_WAVES
#define PPF_WARREN_W_WAVES 0
#endif
#if PPF_WARREN_W_WAVES
#define PPF_WARREN_W_ATTR __attribute__((amdgpu_waves_per_eu(PPF_WARREN_W_WAVES, 8)))
#else
#define PPF_WARREN_W_ATTR
#endif

template <typename R>
__global__ __launch_bounds__(kThreads) PPF_WARREN_W_ATTR void k_warren_w(const R* __restrict__ in, R* __restrict__ out,
                                                       const R* __restrict__ b, Dom d, R dx, WarrenC<R> K,
                                                       uint64_t seed, StepRef sr, uint32_t nstrips,
                                                       uint32_t ntiles) {
  constexpr int kWaves = kThreads / 64;
  const int wave = __builtin_amdgcn_readfirstlane(int(threadIdx.x) / 64);
  const int lane = int(threadIdx.x) % 64;
  const uint32_t t = xcd_swizzle(blockIdx.x, gridDim.x) * uint32_t(kWaves) + uint32_t(wave);
  if (t >= ntiles) return;  // uniform per wave
  const int32_t x0 = 1 + int32_t(t % nstrips) * kWvTX, y0 = 1 + int32_t(t / nstrips) * kWvTY;
  const int32_t nx = int32_t(d.ax.n), ny = int32_t(d.ay.n);
  const int32_t xr = x0 - 1 + lane;
  const int32_t sy = int32_t(d.ay.stride * int64_t(sizeof(R))), sfb = int32_t(d.sf * int64_t(sizeof(R)));
  const int64_t step = sr.get();
  const uint64_t ncells = uint64_t(nx + 2) * uint64_t(ny + 2);
  const bool out_lane = lane >= 1 && lane < 63 && xr <= nx;
  const __amdgpu_buffer_rsrc_t rs = plane_rsrc(in), ws = plane_rsrc(out);
  const uint32_t oo = uint32_t(int64_t(y0) * d.ay.stride + xr) * uint32_t(sizeof(R));
  // the tile's march: halo rows and columns are a coordinate remap (the
  // dispatcher sends grids with bc values or a Dirichlet side to the LDS
  // tiles, which fold them), and rows are fetched kWvAhead ahead of the
  // update (unconditional loads: a conditional one makes the compiler wait
  // for it at once)
  auto march = [&]() {
    R ph[kWvTY + 2], cc[kWvTY + 2];  // rows y0 - 1 .. y0 + kWvTY
    const uint32_t xo = uint32_t(remap_coord32(xr, nx, d.ax.bc)) * uint32_t(sizeof(R));
    auto fetch = [&](auto kc) {
      constexpr int k = decltype(kc)::value;
      if constexpr (k < kWvTY + 2) {
        const int32_t yo = remap_coord32(y0 - 1 + k, ny, d.ay.bc) * sy;
        ph[k] = buf_load<R>(rs, xo, yo);
        cc[k] = buf_load<R>(rs, xo, yo + sfb);
      }
    };
    warren_unrolled(fetch, std::make_integer_sequence<int, (kWvAhead + 3 < kWvTY + 2 ? kWvAhead + 3 : kWvTY + 2)>{});
    R Dm, tm, D1, t1;
    warren_point(K, ph[0], cc[0], Dm, tm);
    warren_point(K, ph[1], cc[1], D1, t1);
    auto row = [&](auto rc) {
      constexpr int r = decltype(rc)::value;
      // rows up to r + 2 are in flight; start row r + kWvAhead + 3
      if constexpr (kWvAhead + 3 < kWvTY + 2) {
        __builtin_amdgcn_sched_barrier(0);
        fetch(std::integral_constant<int, r + kWvAhead + 3>{});
      }
      R Dp, tp;
      warren_point(K, ph[r + 2], cc[r + 2], Dp, tp);
      const int32_t y = y0 + r;
      R nph, nc;
      warren_wave_cell(K, ph[r], ph[r + 1], ph[r + 2], cc[r], cc[r + 1], cc[r + 2], Dm, D1, Dp, tm, t1, tp,
                       seed, step, ncells, uint64_t(y) * uint64_t(nx + 2) + uint64_t(xr), nph, nc);
      if (out_lane && y <= ny) {
        buf_store<R>(nph, ws, oo, r * sy);
        buf_store<R>(nc, ws, oo, r * sy + sfb);
      }
      Dm = D1;
      tm = t1;
      D1 = Dp;
      t1 = tp;
    };
    warren_unrolled(row, std::make_integer_sequence<int, kWvTY>{});
  };
  march();
}

// whether the wave-tile kernel runs the grid: FP32 (FP64 keeps the LDS tiles,
// which are faster there), 2D, enough wave tiles to fill the GPU (smaller
// grids keep the 256-cell LDS tiles, four times as many work units), and both
// fields' bytes within the 32-bit buffer offsets.  PPF_WARREN_WAVE_TILES=0
// keeps the LDS tiles everywhere.
#ifndef PPF_WARREN_WAVE_TILES
#define PPF_WARREN_WAVE_TILES 1
#endif
constexpr int64_t kWvMinTiles = 1024;
inline bool warren_wave_fits(const ppf_grid& g, size_t elem, const void* bcvals) {
  if (!PPF_WARREN_WAVE_TILES || elem != sizeof(float) || g.ndim != 2) return false;
  if (bcvals || g.bc[0] == PPF_DIRICHLET || g.bc[1] == PPF_DIRICHLET) return false;
  const int64_t tiles = ((g.n[0] + kWvTX - 1) / kWvTX) * ((g.n[1] + kWvTY - 1) / kWvTY);
  return tiles >= kWvMinTiles && (g.sf + (g.n[1] + 2) * g.sy) * int64_t(elem) < (int64_t(1) << 31);
}

// FP32 occupancy hint (waves per SIMD the allocation must allow; 0 =
// compiler's choice).  Left to itself the compiler takes 102 SGPRs, which
// caps the kernel at 7 waves per SIMD (7 blocks per CU: a 2048^2 grid's 4096
// tiles then run in 2.3 rounds); with 8 it keeps 78 SGPRs, spilling the rest
// to VGPR lanes, and 4096 tiles fill exactly two rounds.  FP64 keeps the
// compiler's choice: its registers allow 4 waves either way.
#ifndef PPF_WARREN_WAVES
#define PPF_WARREN_WAVES 8
#endif
#if PPF_WARREN_WAVES
#define PPF_WARREN_ATTR __attribute__((amdgpu_waves_per_eu(PPF_WARREN_WAVES, 8)))
#else
#define PPF_WARREN_ATTR
#endif
// (defined in the FP32 translation unit only: an explicit specialization is
// an ordinary function)
#ifdef PPF_FP32_UNIT
template <>
__global__ __launch_bounds__(kThreads) PPF_WARREN_ATTR void k_warren<float>(const float* __restrict__ in,
                                                                            float* __restrict__ out,
                                                                            const float* __restrict__ b, Dom d,
                                                                            float dx, WarrenK p, uint64_t seed,
                                                                            StepRef sr) {
  warren_tile_step<float>(in, out, b, d, dx, p, seed, sr);
}
#endif

template <typename R>
int warren_step(const ppf_grid& g, const void* in, void* out, const void* bcvals, const ppf_warren_params& p,
                uint64_t seed, StepRef step, hipStream_t s) {
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
  if (sizeof(R) == 4 && warren_wave_fits(g, sizeof(R), bcvals)) {
    const WarrenC<R> K(k, R(g.dx));
    const uint32_t nstrips = uint32_t((g.n[0] + kWvTX - 1) / kWvTX);
    const uint32_t ntiles = nstrips * uint32_t((g.n[1] + kWvTY - 1) / kWvTY);
    const unsigned blocks = (ntiles + kThreads / 64 - 1) / (kThreads / 64);
    hipLaunchKernelGGL(k_warren_w<R>, dim3(blocks), dim3(kThreads), 0, s, static_cast<const R*>(in),
                       static_cast<R*>(out), static_cast<const R*>(bcvals), d, R(g.dx), K, seed, step, nstrips,
                       ntiles);
    return check_launch("warren_step");
  }
  const unsigned tiles = unsigned(((g.n[0] + kTX - 1) / kTX) * ((g.n[1] + kWTY - 1) / kWTY));
  hipLaunchKernelGGL(k_warren<R>, dim3(tiles), dim3(kThreads), 0, s, static_cast<const R*>(in),
                     static_cast<R*>(out), static_cast<const R*>(bcvals), d, R(g.dx), k, seed, step);
  return check_launch("warren_step");
}

}  // namespace ppf
