// Implicit and Crank-Nicolson diffusion on periodic grids
// (Engines/Diffusion.py:18-633).  The reference assembles dense cyclic
// matrices (diffusion_matrix_1d/2d/3d) and calls np.linalg.solve / inv /
// scipy gmres on them.  Here the same systems are solved without matrices:
//
//  * along one axis (1D solvers and every alternating-direction variant) the
//    system is cyclic tridiagonal with constant coefficients,
//      -a x[i-1] + (1+2a) x[i] - a x[i+1] = d[i]   (indices mod n),
//    solved per grid line by the Thomas recurrence on the matrix with its
//    corners removed plus a Sherman-Morrison rank-one correction.  Every
//    coefficient that does not depend on d (elimination factors, the
//    correction vector) is line-independent and comes from a host-built
//    table, so a line costs three streaming passes: forward elimination,
//    back substitution, correction.  One thread per line; lines along y and
//    z are read coalesced across x.
//  * the full 2D/3D systems (1 + 2*ndim*a) x - a*sum_nb x = rhs are symmetric
//    positive definite: matrix-free conjugate gradients, two kernels per
//    iteration, scalars reduced on the device (every block reduces the same
//    partial sums in the same order, so all blocks agree on alpha, beta and
//    the convergence test without a host round trip).
#include <cmath>

#include "ppf_common.h"

namespace ppf {
namespace {

constexpr int kLineThreads = 64;
constexpr int kCgThreads = 256;
constexpr int kCgBlocks = 1024;
constexpr int64_t kTableHead = 4;  // off, fac, r, reserved; then cp[n], den[n], q[n]

__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// deterministic block sum (256 threads); result valid in every thread
__device__ __forceinline__ double block_sum_all(double v) {
  __shared__ double part[kCgThreads / 64];
  v = wave_sum_d(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) part[w] = v;
  __syncthreads();
  return (part[0] + part[1]) + (part[2] + part[3]);
}

// sum of nb partials, identical in every block
__device__ __forceinline__ double reduce_partials(const double* __restrict__ p, int nb) {
  double acc = 0.0;
  for (int i = threadIdx.x; i < nb; i += kCgThreads) acc += p[i];
  return block_sum_all(acc);
}

// ------------------------------------------------------------ line solve --
struct Line {
  int64_t n;        // cells along the solve axis
  int64_t stride;   // element stride along the solve axis
  int64_t lines;    // number of lines
  int64_t m0, m1;   // sizes of the two other axes (line index = i0 + m0*i1)
  int64_t s0, s1;   // element strides of the two other axes
  int64_t base;     // padded offset of the first interior cell of line 0
  // explicit half step along `eaxis` (-1: none); when eaxis differs from the
  // solve axis it is one of the two other axes: eslot 0 or 1
  int eslot;        // -1: none or along the line; 0/1: other axis
  int ealong;       // 1: explicit along the solve axis
  int64_t en;       // cells along the explicit axis (other-axis case)
};

template <typename R>
__global__ __launch_bounds__(kLineThreads) void k_line_solve(const R* __restrict__ in, R* __restrict__ out, Line L,
                                                            const double* __restrict__ tab, double bexp) {
  const int64_t l = int64_t(blockIdx.x) * kLineThreads + threadIdx.x;
  if (l >= L.lines) return;
  const int64_t i0 = l % L.m0, i1 = l / L.m0;
  const int64_t o = L.base + i0 * L.s0 + i1 * L.s1;
  const int64_t n = L.n, st = L.stride;
  const R off = R(tab[0]);
  const double* cp = tab + kTableHead;
  const double* den = cp + n;
  const double* q = den + n;
  // explicit neighbours across lines (periodic wrap)
  int64_t dm = 0, dp = 0;
  if (L.eslot >= 0) {
    const int64_t c = L.eslot == 0 ? i0 : i1;
    const int64_t s = L.eslot == 0 ? L.s0 : L.s1;
    dm = c == 0 ? (L.en - 1) * s : -s;
    dp = c == L.en - 1 ? -(L.en - 1) * s : s;
  }
  const R b = R(bexp), c0 = R(1.0 - 2.0 * bexp);
  // pass 1: forward elimination on the corner-free matrix
  R prev = R(0);
  for (int64_t i = 0; i < n; ++i) {
    const int64_t p = o + i * st;
    R d = in[p];
    if (L.eslot >= 0) {
      d = c0 * d + b * (in[p + dm] + in[p + dp]);
    } else if (L.ealong) {
      const R xm = in[i == 0 ? o + (n - 1) * st : p - st];
      const R xp = in[i == n - 1 ? o : p + st];
      d = c0 * d + b * (xm + xp);
    }
    prev = (d - off * prev) * R(den[i]);
    out[p] = prev;
  }
  // pass 2: back substitution
  const R ylast = prev;
  R y = prev;
  for (int64_t i = n - 2; i >= 0; --i) {
    const int64_t p = o + i * st;
    y = out[p] - R(cp[i]) * y;
    out[p] = y;
  }
  // pass 3: rank-one correction  x = y - ((y0 + r*y_{n-1}) * fac) * q
  const R f = (y + R(tab[2]) * ylast) * R(tab[1]);
  for (int64_t i = 0; i < n; ++i) {
    const int64_t p = o + i * st;
    out[p] = out[p] - f * R(q[i]);
  }
}

// Allen-Cahn semi-implicit line solve: the diagonal of line element i is
// 1 + 2a + srck (4 d^2 - 6 d + 2), d = its right-hand side (the linearised
// double-well force, Engines/CahnAllen.py:146-148), so the elimination
// factors are per line.  rhs: the explicit half step as in k_line_solve, or
// with imexk != 0 d = c - imexk (4 c^3 - 6 c^2 + 2 c).  Scratch s1 holds the
// elimination factors, s2 the correction vector (same layout as the field).
template <typename R>
__global__ __launch_bounds__(kLineThreads) void k_line_solve_var(const R* __restrict__ in, R* __restrict__ out,
                                                                Line L, double a, double bexp, double srck,
                                                                double imexk, R* __restrict__ s1,
                                                                R* __restrict__ s2) {
  const int64_t l = int64_t(blockIdx.x) * kLineThreads + threadIdx.x;
  if (l >= L.lines) return;
  const int64_t i0 = l % L.m0, i1 = l / L.m0;
  const int64_t o = L.base + i0 * L.s0 + i1 * L.s1;
  const int64_t n = L.n, st = L.stride;
  int64_t dm = 0, dp = 0;
  if (L.eslot >= 0) {
    const int64_t c = L.eslot == 0 ? i0 : i1;
    const int64_t s = L.eslot == 0 ? L.s0 : L.s1;
    dm = c == 0 ? (L.en - 1) * s : -s;
    dp = c == L.en - 1 ? -(L.en - 1) * s : s;
  }
  const R off = R(-a), d0 = R(1.0 + 2.0 * a), K = R(srck), KI = R(imexk);
  const R b = R(bexp), c0 = R(1.0 - 2.0 * bexp);
  // pass 1: forward elimination of the corner-free matrix for the solution
  // (into out) and the correction vector (into s2)
  R gam = R(0), cp = R(0), dy = R(0), dq = R(0);
  for (int64_t i = 0; i < n; ++i) {
    const int64_t p = o + i * st;
    R d = in[p];
    if (L.eslot >= 0) {
      d = c0 * d + b * (in[p + dm] + in[p + dp]);
    } else if (L.ealong) {
      const R xm = in[i == 0 ? o + (n - 1) * st : p - st];
      const R xp = in[i == n - 1 ? o : p + st];
      d = c0 * d + b * (xm + xp);
    }
    if (imexk != 0.0) d = d - KI * (R(4) * d * d * d - R(6) * d * d + R(2) * d);
    R diag = d0 + K * (R(4) * d * d - R(6) * d + R(2));
    R u = R(0);
    if (i == 0) {
      gam = -diag;
      diag = diag - gam;
      u = gam;
    } else if (i == n - 1) {
      diag = diag - off * off / gam;
      u = off;
    }
    const R m = i == 0 ? diag : diag - off * cp;
    const R rm = R(1) / m;
    cp = off * rm;
    dy = (d - off * dy) * rm;
    dq = (u - off * dq) * rm;
    out[p] = dy;
    s1[p] = cp;
    s2[p] = dq;
  }
  // pass 2: back substitution of both
  const R ylast = dy, qlast = dq;
  R y = dy, q = dq;
  for (int64_t i = n - 2; i >= 0; --i) {
    const int64_t p = o + i * st;
    const R f = s1[p];
    y = out[p] - f * y;
    q = s2[p] - f * q;
    out[p] = y;
    s2[p] = q;
  }
  // pass 3: Sherman-Morrison correction x = y - (v.y / (1 + v.q)) q
  const R r = off / gam;
  const R f = (y + r * ylast) / (R(1) + q + r * qlast);
  for (int64_t i = 0; i < n; ++i) {
    const int64_t p = o + i * st;
    out[p] = out[p] - f * s2[p];
  }
}

// ----------------------------------------------- scan-based line solve --
// The same constant-coefficient cyclic systems, with parallelism along each
// line.  With r = ((1+2a) - sqrt(1+4a)) / (2a) < 1 (the root of
// a r^2 - (1+2a) r + a = 0) the matrix factors as
//   (1+2a) I - a (S + S^T) = (a/r) (I - r S)(I - r S^T),   (S x)_i = x_{i-1 mod n},
// so x = (I - r S^T)^-1 (I - r S)^-1 (r/a) d: a forward and a backward
// first-order recurrence, each closed around the ring (a recurrence started
// from zero is off by r^(i+1) C at element i, C = end value / (1 - r^n)).
// Each line is cut into chunks of m cells, one thread per (line, chunk):
//   A  forward recurrence inside each chunk from zero, chunk end value E
//   B  per line: carries between chunks and the ring closure
//        K_k = carry into chunk k + r^(start_k) C_u
//   C  forward fix-up u_i = v_i + r^(i+1-start) K_k, then the backward
//      recurrence inside the chunk, chunk end value E'
//   D  per line: backward carries and closure K'_k = carry + r^(n-end_k) C_x
//   E  backward fix-up x_i = w_i + r^(end-i) K'_k
// Table (host built, ppf_scan_table): r, r/a, r^n, m, then r^0 .. r^m.
constexpr int64_t kScanHead = 4;

struct Scan {
  int64_t lines, nch, m;
};

__device__ __forceinline__ int64_t line_base(const Line& L, int64_t l) {
  const int64_t i0 = l % L.m0, i1 = l / L.m0;
  return L.base + i0 * L.s0 + i1 * L.s1;
}

template <typename R>
__global__ __launch_bounds__(256) void k_scan_a(const R* __restrict__ in, R* __restrict__ out, Line L, Scan S,
                                                const double* __restrict__ tab, double bexp, double* __restrict__ E) {
  const int64_t t = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (t >= S.lines * S.nch) return;
  const int64_t l = t % S.lines, k = t / S.lines;
  const int64_t o = line_base(L, l), n = L.n, st = L.stride;
  int64_t dm = 0, dp = 0;
  if (L.eslot >= 0) {
    const int64_t i0 = l % L.m0, i1 = l / L.m0;
    const int64_t c = L.eslot == 0 ? i0 : i1;
    const int64_t s = L.eslot == 0 ? L.s0 : L.s1;
    dm = c == 0 ? (L.en - 1) * s : -s;
    dp = c == L.en - 1 ? -(L.en - 1) * s : s;
  }
  const R r = R(tab[0]), scale = R(tab[1]);
  const R b = R(bexp), c0 = R(1.0 - 2.0 * bexp);
  const int64_t i0 = k * S.m, i1 = i0 + S.m < n ? i0 + S.m : n;
  R acc = R(0);
  for (int64_t i = i0; i < i1; ++i) {
    const int64_t p = o + i * st;
    R d = in[p];
    if (L.eslot >= 0) {
      d = c0 * d + b * (in[p + dm] + in[p + dp]);
    } else if (L.ealong) {
      const R xm = in[i == 0 ? o + (n - 1) * st : p - st];
      const R xp = in[i == n - 1 ? o : p + st];
      d = c0 * d + b * (xm + xp);
    }
    acc = scale * d + r * acc;
    out[p] = acc;
  }
  E[t] = double(acc);
}

// B (forward) / D (backward): one wavefront per line.  Chunk k in the order
// of the recurrence acts on the carry as the affine map x -> r^(len_k) x + E_k;
// each lane composes the maps of its consecutive chunks, a shuffle scan
// composes across the lanes, and the lane's chunks are walked again to write
// K_k = (carry into chunk k) + (product of the maps' slopes before k) * C,
// C = (composition of all chunks applied to 0) / (1 - r^n) closing the ring.
__global__ __launch_bounds__(256) void k_scan_carry(int64_t n, Scan S, const double* __restrict__ tab,
                                                    const double* __restrict__ E, double* __restrict__ K,
                                                    int backward) {
  const int lane = threadIdx.x & 63;
  const int64_t l = int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
  if (l >= S.lines) return;  // uniform per wavefront
  const double rn = tab[2];
  const double* P = tab + kScanHead;
  const int64_t q = (S.nch + 63) / 64;
  const int64_t j0 = int64_t(lane) * q, j1 = j0 + q < S.nch ? j0 + q : S.nch;
  auto chunk = [&](int64_t j) { return backward ? S.nch - 1 - j : j; };
  auto slope = [&](int64_t k) { return P[(k + 1) * S.m < n ? S.m : n - k * S.m]; };
  double A = 1.0, B = 0.0;
  for (int64_t j = j0; j < j1; ++j) {
    const int64_t k = chunk(j);
    const double a = slope(k);
    B = B * a + E[k * S.lines + l];
    A *= a;
  }
#pragma unroll
  for (int sft = 1; sft < 64; sft <<= 1) {
    const double Ap = __shfl_up(A, sft, 64), Bp = __shfl_up(B, sft, 64);
    if (lane >= sft) {
      B = Bp * A + B;
      A = Ap * A;
    }
  }
  const double total = __shfl(B, 63, 64);
  double Ar = __shfl_up(A, 1, 64), Br = __shfl_up(B, 1, 64);
  if (lane == 0) {
    Ar = 1.0;
    Br = 0.0;
  }
  const double C = total / (1.0 - rn);
  for (int64_t j = j0; j < j1; ++j) {
    const int64_t k = chunk(j);
    const double e = E[k * S.lines + l];
    K[k * S.lines + l] = Br + Ar * C;
    const double a = slope(k);
    Br = Br * a + e;
    Ar *= a;
  }
}

template <typename R>
__global__ __launch_bounds__(256) void k_scan_c(R* __restrict__ out, Line L, Scan S, const double* __restrict__ tab,
                                                const double* __restrict__ K, double* __restrict__ E) {
  const int64_t t = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (t >= S.lines * S.nch) return;
  const int64_t l = t % S.lines, k = t / S.lines;
  const int64_t o = line_base(L, l), n = L.n, st = L.stride;
  const R r = R(tab[0]);
  const double* P = tab + kScanHead;
  const int64_t i0 = k * S.m, i1 = i0 + S.m < n ? i0 + S.m : n;
  const R kk = R(K[t]);
  R acc = R(0);
  for (int64_t i = i1 - 1; i >= i0; --i) {
    const int64_t p = o + i * st;
    const R u = out[p] + R(P[i + 1 - i0]) * kk;
    acc = u + r * acc;
    out[p] = acc;
  }
  E[t] = double(acc);
}

template <typename R>
__global__ __launch_bounds__(256) void k_scan_e(R* __restrict__ out, Line L, Scan S, const double* __restrict__ tab,
                                                const double* __restrict__ K) {
  const int64_t t = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (t >= S.lines * S.nch) return;
  const int64_t l = t % S.lines, k = t / S.lines;
  const int64_t o = line_base(L, l), n = L.n, st = L.stride;
  const double* P = tab + kScanHead;
  const int64_t i0 = k * S.m, i1 = i0 + S.m < n ? i0 + S.m : n;
  const R kk = R(K[t]);
  for (int64_t i = i0; i < i1; ++i) {
    const int64_t p = o + i * st;
    out[p] = out[p] + R(P[i1 - i]) * kk;
  }
}

// ---------------------------------------------------- conjugate gradients --
struct Dense {
  int nd;
  int64_t nx, ny, nz, sy, sz, N;
  int nb;  // blocks of every CG launch
};

struct Work {
  ppf_cg_state* st;
  double* prr;  // [2][kCgBlocks]: r.r of iteration k at (k & 1)
  double* ppq;  // [kCgBlocks]
  double* pbb;  // [kCgBlocks]
  void* r;
  void* p[2];
  void* q;
};

__host__ __device__ inline int64_t align256(int64_t b) { return (b + 255) / 256 * 256; }

inline Work carve(void* w, int64_t N, int64_t esize) {
  char* b = static_cast<char*>(w);
  Work k;
  k.st = reinterpret_cast<ppf_cg_state*>(b);
  b += 256;
  k.prr = reinterpret_cast<double*>(b);
  k.ppq = k.prr + 2 * kCgBlocks;
  k.pbb = k.ppq + kCgBlocks;
  b += 4 * kCgBlocks * sizeof(double);
  const int64_t v = align256(N * esize);
  k.r = b;
  k.p[0] = b + v;
  k.p[1] = b + 2 * v;
  k.q = b + 3 * v;
  return k;
}

// The CG kernels walk whole rows: block b takes rows (y, z) = b, b + nb, ...
// in a fixed order and its threads stride along x, so a row's periodic
// neighbour offsets are computed once and the per-block partial sums are
// deterministic.
struct Row {
  int64_t t0;          // dense index of (0, y, z)
  int64_t p0;          // padded offset of interior (0, y, z)
  int64_t ym, yp, zm, zp;      // dense offsets of the y / z neighbours
  int64_t pym, pyp, pzm, pzp;  // padded offsets of the same
};

__device__ __forceinline__ Row row_of(const Dense& d, int64_t rw) {
  Row r;
  const int64_t y = d.nd >= 2 ? rw % d.ny : 0, z = d.nd >= 3 ? rw / d.ny : 0;
  r.t0 = rw * d.nx;
  r.p0 = 1 + (d.nd >= 2 ? (y + 1) * d.sy : 0) + (d.nd >= 3 ? (z + 1) * d.sz : 0);
  const int64_t pl = d.nx * d.ny;
  r.ym = (y == 0 ? d.ny - 1 : -1) * d.nx;
  r.yp = (y == d.ny - 1 ? -(d.ny - 1) : 1) * d.nx;
  r.zm = (z == 0 ? d.nz - 1 : -1) * pl;
  r.zp = (z == d.nz - 1 ? -(d.nz - 1) : 1) * pl;
  r.pym = (y == 0 ? d.ny - 1 : -1) * d.sy;
  r.pyp = (y == d.ny - 1 ? -(d.ny - 1) : 1) * d.sy;
  r.pzm = (z == 0 ? d.nz - 1 : -1) * d.sz;
  r.pzp = (z == d.nz - 1 ? -(d.nz - 1) : 1) * d.sz;
  return r;
}

// x0 = in (written to out), rhs = in or the explicit half step of in,
// r = rhs - A x0; partial sums of r.r and rhs.rhs
template <typename R, bool VAR>
__global__ __launch_bounds__(kCgThreads) void k_cg_begin(const R* __restrict__ in, R* __restrict__ out, Dense d,
                                                        double a, double b, double srck, Work w) {
  R* r = static_cast<R*>(w.r);
  double srr = 0.0, sbb = 0.0;
  const double na = 2.0 * d.nd;
  const int64_t rows = d.ny * d.nz;
  for (int64_t rw = blockIdx.x; rw < rows; rw += d.nb) {
    const Row o = row_of(d, rw);
    for (int64_t x = threadIdx.x; x < d.nx; x += kCgThreads) {
      const int64_t p = o.p0 + x;
      // periodic neighbours in the padded array
      R s = in[p + (x == 0 ? d.nx - 1 : -1)] + in[p + (x == d.nx - 1 ? -(d.nx - 1) : 1)];
      if (d.nd >= 2) s = s + in[p + o.pym] + in[p + o.pyp];
      if (d.nd >= 3) s = s + in[p + o.pzm] + in[p + o.pzp];
      const R v = in[p];
      const R rhs = b != 0.0 ? R(1.0 - na * b) * v + R(b) * s : v;
      R diag = R(1.0 + na * a);
      if (VAR) diag = diag + R(srck) * (R(4) * v * v - R(6) * v + R(2));
      const R ax = diag * v - R(a) * s;
      const R rr = rhs - ax;
      out[p] = v;
      r[o.t0 + x] = rr;
      srr += double(rr) * double(rr);
      sbb += double(rhs) * double(rhs);
    }
  }
  srr = block_sum_all(srr);
  sbb = block_sum_all(sbb);
  if (threadIdx.x == 0) {
    w.prr[blockIdx.x] = srr;
    w.pbb[blockIdx.x] = sbb;
    if (blockIdx.x == 0) {
      w.st->bb = 0.0;
      w.st->rr = 0.0;
      w.st->iterations = -1;
      w.st->converged = 0;
    }
  }
}

// iteration k, first half: convergence test, p_k = r_k + beta p_{k-1},
// q = A p_k, partial sums of p.q
template <typename R, bool VAR>
__global__ __launch_bounds__(kCgThreads) void k_cg_dir(const R* __restrict__ in, Dense d, double a, double srck,
                                                      double tol2, int64_t k, Work w) {
  const double rr = reduce_partials(w.prr + (k & 1) * kCgBlocks, d.nb);
  const double bb = k == 0 ? reduce_partials(w.pbb, d.nb) : w.st->bb;
  if (w.st->converged) return;
  const bool done = rr <= tol2 * bb;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (k == 0) w.st->bb = bb;
    w.st->rr = rr;
    if (done) {
      w.st->converged = 1;
      w.st->iterations = int32_t(k);
    }
  }
  if (done) return;
  const double beta = k == 0 ? 0.0 : rr / reduce_partials(w.prr + ((k + 1) & 1) * kCgBlocks, d.nb);
  const R* r = static_cast<const R*>(w.r);
  const R* po = static_cast<const R*>(w.p[(k + 1) & 1]);
  R* pn = static_cast<R*>(w.p[k & 1]);
  R* q = static_cast<R*>(w.q);
  const R bt = R(beta);
  const R cdiag = R(1.0 + 2.0 * d.nd * a), ra = R(a);
  // direction at dense index j (k = 0: the residual)
  auto dir = [&](int64_t j) { return k == 0 ? r[j] : r[j] + bt * po[j]; };
  double spq = 0.0;
  const int64_t rows = d.ny * d.nz;
  for (int64_t rw = blockIdx.x; rw < rows; rw += d.nb) {
    const Row o = row_of(d, rw);
    for (int64_t x = threadIdx.x; x < d.nx; x += kCgThreads) {
      const int64_t t = o.t0 + x;
      const R pv = dir(t);
      R s = dir(t + (x == 0 ? d.nx - 1 : -1)) + dir(t + (x == d.nx - 1 ? -(d.nx - 1) : 1));
      if (d.nd >= 2) s = s + dir(t + o.ym) + dir(t + o.yp);
      if (d.nd >= 3) s = s + dir(t + o.zm) + dir(t + o.zp);
      R dg = cdiag;
      if (VAR) {
        const R v = in[o.p0 + x];
        dg = dg + R(srck) * (R(4) * v * v - R(6) * v + R(2));
      }
      const R qv = dg * pv - ra * s;
      pn[t] = pv;
      q[t] = qv;
      spq += double(pv) * double(qv);
    }
  }
  spq = block_sum_all(spq);
  if (threadIdx.x == 0) w.ppq[blockIdx.x] = spq;
}

// iteration k, second half: x += alpha p, r -= alpha q, partial sums of r.r
template <typename R>
__global__ __launch_bounds__(kCgThreads) void k_cg_step(R* __restrict__ out, Dense d, int64_t k, Work w) {
  if (w.st->converged) return;
  const double pq = reduce_partials(w.ppq, d.nb);
  const double alpha = w.st->rr / pq;
  R* r = static_cast<R*>(w.r);
  const R* p = static_cast<const R*>(w.p[k & 1]);
  const R* q = static_cast<const R*>(w.q);
  const R al = R(alpha);
  double srr = 0.0;
  const int64_t rows = d.ny * d.nz;
  for (int64_t rw = blockIdx.x; rw < rows; rw += d.nb) {
    const Row o = row_of(d, rw);
    for (int64_t x = threadIdx.x; x < d.nx; x += kCgThreads) {
      const int64_t t = o.t0 + x, pp = o.p0 + x;
      out[pp] = out[pp] + al * p[t];
      const R rv = r[t] - al * q[t];
      r[t] = rv;
      srr += double(rv) * double(rv);
    }
  }
  srr = block_sum_all(srr);
  if (threadIdx.x == 0) w.prr[((k + 1) & 1) * kCgBlocks + blockIdx.x] = srr;
}

Dense dense_of(const ppf_grid& g) {
  Dense d;
  d.nd = g.ndim;
  d.nx = g.n[0];
  d.ny = g.ndim >= 2 ? g.n[1] : 1;
  d.nz = g.ndim >= 3 ? g.n[2] : 1;
  d.sy = g.sy;
  d.sz = g.sz;
  d.N = d.nx * d.ny * d.nz;
  const int64_t rows = d.ny * d.nz;  // blocks take whole rows
  d.nb = int(rows < kCgBlocks ? rows : kCgBlocks);
  return d;
}

int check_periodic(const ppf_grid* g, const char* what) {
  if (int rc = check_grid(g, 1, 3, what)) return rc;
  for (int a = 0; a < g->ndim; ++a) {
    if (g->bc[a] != PPF_PERIODIC)
      return fail(PPF_EINVAL, std::string(what) +
                                  ": the implicit diffusion matrices are periodic (Diffusion.py:18-196); "
                                  "every axis must be PERIODIC");
    if (g->n[a] < 3) return fail(PPF_EINVAL, std::string(what) + ": every axis needs at least 3 cells");
  }
  return 0;
}

Line line_of(const ppf_grid& g, int axis, int eaxis) {
  const int64_t st[3] = {1, g.ndim >= 2 ? g.sy : 0, g.ndim >= 3 ? g.sz : 0};
  const int64_t nn[3] = {g.n[0], g.ndim >= 2 ? g.n[1] : 1, g.ndim >= 3 ? g.n[2] : 1};
  int other[2], k = 0;
  for (int a = 0; a < 3; ++a)
    if (a != axis) other[k++] = a;
  Line L;
  L.n = nn[axis];
  L.stride = st[axis];
  L.m0 = nn[other[0]];
  L.m1 = nn[other[1]];
  L.s0 = st[other[0]];
  L.s1 = st[other[1]];
  L.lines = L.m0 * L.m1;
  L.base = 1 + (g.ndim >= 2 ? g.sy : 0) + (g.ndim >= 3 ? g.sz : 0);
  L.eslot = -1;
  L.ealong = 0;
  L.en = 1;
  if (eaxis == axis) {
    L.ealong = 1;
  } else if (eaxis >= 0) {
    L.eslot = eaxis == other[0] ? 0 : 1;
    L.en = nn[eaxis];
  }
  return L;
}

template <typename R>
int line_solve_var(const ppf_grid& g, const void* in, void* out, int axis, double a, int eaxis, double b,
                   double srck, double imexk, void* work, hipStream_t s) {
  const Line L = line_of(g, axis, eaxis);
  R* s1 = static_cast<R*>(work);
  R* s2 = s1 + g.sf;
  const unsigned blocks = unsigned((L.lines + kLineThreads - 1) / kLineThreads);
  hipLaunchKernelGGL(k_line_solve_var<R>, dim3(blocks), dim3(kLineThreads), 0, s, static_cast<const R*>(in),
                     static_cast<R*>(out), L, a, b, srck, imexk, s1, s2);
  return check_launch("allencahn_line_solve");
}

template <typename R>
int line_solve_scan(const ppf_grid& g, const void* in, void* out, int axis, const double* table, int eaxis,
                    double b, void* work, int64_t m, hipStream_t s) {
  const Line L = line_of(g, axis, eaxis);
  Scan S;
  S.lines = L.lines;
  S.m = m;
  S.nch = (L.n + m - 1) / m;
  double* E = static_cast<double*>(work);
  double* K = E + S.lines * S.nch;
  const int64_t nt = S.lines * S.nch;
  const unsigned gb = unsigned((nt + 255) / 256), lb = unsigned((S.lines + 3) / 4);
  hipLaunchKernelGGL(k_scan_a<R>, dim3(gb), dim3(256), 0, s, static_cast<const R*>(in), static_cast<R*>(out), L, S,
                     table, b, E);
  if (int rc = check_launch("diffusion_line_solve_scan")) return rc;
  hipLaunchKernelGGL(k_scan_carry, dim3(lb), dim3(256), 0, s, L.n, S, table, static_cast<const double*>(E), K, 0);
  if (int rc = check_launch("diffusion_line_solve_scan")) return rc;
  hipLaunchKernelGGL(k_scan_c<R>, dim3(gb), dim3(256), 0, s, static_cast<R*>(out), L, S, table,
                     static_cast<const double*>(K), E);
  if (int rc = check_launch("diffusion_line_solve_scan")) return rc;
  hipLaunchKernelGGL(k_scan_carry, dim3(lb), dim3(256), 0, s, L.n, S, table, static_cast<const double*>(E), K, 1);
  if (int rc = check_launch("diffusion_line_solve_scan")) return rc;
  hipLaunchKernelGGL(k_scan_e<R>, dim3(gb), dim3(256), 0, s, static_cast<R*>(out), L, S, table,
                     static_cast<const double*>(K));
  return check_launch("diffusion_line_solve_scan");
}

template <typename R>
int line_solve(const ppf_grid& g, const void* in, void* out, int axis, const double* table, int eaxis, double b,
               hipStream_t s) {
  const Line L = line_of(g, axis, eaxis);
  const unsigned blocks = unsigned((L.lines + kLineThreads - 1) / kLineThreads);
  hipLaunchKernelGGL(k_line_solve<R>, dim3(blocks), dim3(kLineThreads), 0, s, static_cast<const R*>(in),
                     static_cast<R*>(out), L, table, b);
  return check_launch("diffusion_line_solve");
}

template <typename R, bool VAR>
int cg_begin(const ppf_grid& g, const void* in, void* out, double a, double b, double srck, void* work,
             hipStream_t s) {
  Dense d = dense_of(g);
  Work w = carve(work, d.N, sizeof(R));
  hipLaunchKernelGGL((k_cg_begin<R, VAR>), dim3(d.nb), dim3(kCgThreads), 0, s, static_cast<const R*>(in),
                     static_cast<R*>(out), d, a, b, srck, w);
  return check_launch("cg_begin");
}

template <typename R, bool VAR>
int cg_iterate(const ppf_grid& g, const void* in, void* out, double a, double srck, double tol, int64_t k0,
               int iters, void* work, hipStream_t s) {
  Dense d = dense_of(g);
  Work w = carve(work, d.N, sizeof(R));
  const double tol2 = tol * tol;
  for (int i = 0; i < iters; ++i) {
    const int64_t k = k0 + i;
    hipLaunchKernelGGL((k_cg_dir<R, VAR>), dim3(d.nb), dim3(kCgThreads), 0, s, static_cast<const R*>(in), d, a,
                       srck, tol2, k, w);
    if (int rc = check_launch("diffusion_cg_iterate")) return rc;
    hipLaunchKernelGGL(k_cg_step<R>, dim3(d.nb), dim3(kCgThreads), 0, s, static_cast<R*>(out), d, k, w);
    if (int rc = check_launch("diffusion_cg_iterate")) return rc;
  }
  return 0;
}

// ---- Cahn-Hilliard semi-implicit 1D: the periodic pentadiagonal system
// whose row i is (1 + a (2 s_i + 6 e)) x_i - a (s_i + 4 e)(x_{i-1} + x_{i+1})
// + a e (x_{i-2} + x_{i+2}) = r_i, s_i = 4 r_i^2 - 6 r_i + 2.  Split
// A = B + U V^T: B the same rows without the six corner entries (a plain
// band of width 2, eliminated without pivoting), U = [e_0, e_1, e_{n-2},
// e_{n-1}]; then x = z - Y (I + V^T Y)^-1 V^T z with B z = r and B Y = U
// (Sherman-Morrison-Woodbury).  Thread 0 factors B; threads 0..4 substitute
// the five right-hand sides; thread 0 solves the 4 x 4 system; all threads
// correct.  Arithmetic in FP64 for both storage types.
struct ChPenta {
  double *u0, *u1, *u2, *l1, *l2, *z, *y;  // n each; y holds 4 columns
};

template <typename R>
__global__ __launch_bounds__(64) void k_ch_penta(const R* __restrict__ in, R* __restrict__ out, int64_t n, double a,
                                                 double e, ChPenta w) {
  const int tid = threadIdx.x;
  auto s_of = [&](int64_t i) {
    const double c = double(in[1 + i]);
    return 4.0 * c * c - 6.0 * c + 2.0;
  };
  __shared__ double wsol[4];
  if (tid == 0) {
    // banded elimination of B without pivoting (row i: lower entries at
    // i-2, i-1 when >= 0, upper entries at i+1, i+2 when < n); the two
    // previous pivot rows stay in registers
    double p2[3] = {1.0, 0.0, 0.0}, p1[3] = {1.0, 0.0, 0.0};
    const double far = a * e;
    for (int64_t i = 0; i < n; ++i) {
      const double si = s_of(i);
      const double nb = -a * (si + 4.0 * e);
      double em1 = i >= 1 ? nb : 0.0;
      double e0 = 1.0 + a * (2.0 * si + 6.0 * e);
      double e1 = i + 1 < n ? nb : 0.0;
      const double e2 = i + 2 < n ? far : 0.0;
      double l2 = 0.0, l1 = 0.0;
      if (i >= 2) {
        l2 = far / p2[0];
        em1 -= l2 * p2[1];
        e0 -= l2 * p2[2];
      }
      if (i >= 1) {
        l1 = em1 / p1[0];
        e0 -= l1 * p1[1];
        e1 -= l1 * p1[2];
      }
      w.u0[i] = e0;
      w.u1[i] = e1;
      w.u2[i] = e2;
      w.l1[i] = l1;
      w.l2[i] = l2;
      for (int k = 0; k < 3; ++k) p2[k] = p1[k];
      p1[0] = e0;
      p1[1] = e1;
      p1[2] = e2;
    }
  }
  __syncthreads();
  if (tid < 5) {
    // right-hand side tid: 0 = the field, 1..4 = e_0, e_1, e_{n-2}, e_{n-1}
    double* x = tid == 0 ? w.z : w.y + (tid - 1) * n;
    const int64_t unit = tid == 1 ? 0 : tid == 2 ? 1 : tid == 3 ? n - 2 : n - 1;
    double ym2 = 0.0, ym1 = 0.0;
    for (int64_t i = 0; i < n; ++i) {
      const double r = tid == 0 ? double(in[1 + i]) : (i == unit ? 1.0 : 0.0);
      const double yi = r - w.l2[i] * ym2 - w.l1[i] * ym1;
      x[i] = yi;
      ym2 = ym1;
      ym1 = yi;
    }
    double xp1 = 0.0, xp2 = 0.0;
    for (int64_t i = n - 1; i >= 0; --i) {
      const double xi = (x[i] - w.u1[i] * xp1 - w.u2[i] * xp2) / w.u0[i];
      x[i] = xi;
      xp2 = xp1;
      xp1 = xi;
    }
  }
  __syncthreads();
  if (tid == 0) {
    // V^T applied to a vector v: the corner entries of rows 0, 1, n-2, n-1
    const double s0 = s_of(0), s1 = s_of(n - 1);
    const double far = a * e, nb0 = -a * (s0 + 4.0 * e), nbn = -a * (s1 + 4.0 * e);
    auto vt = [&](const double* v, double (&o)[4]) {
      o[0] = far * v[n - 2] + nb0 * v[n - 1];
      o[1] = far * v[n - 1];
      o[2] = far * v[0];
      o[3] = nbn * v[0] + far * v[1];
    };
    double m[4][5];
    for (int j = 0; j < 4; ++j) {
      double col[4];
      vt(w.y + j * n, col);
      for (int r = 0; r < 4; ++r) m[r][j] = col[r] + (r == j ? 1.0 : 0.0);
    }
    double rhs[4];
    vt(w.z, rhs);
    for (int r = 0; r < 4; ++r) m[r][4] = rhs[r];
    // Gaussian elimination with partial pivoting on the 4 x 5 system
    for (int k = 0; k < 4; ++k) {
      int p = k;
      for (int r = k + 1; r < 4; ++r)
        if (fabs(m[r][k]) > fabs(m[p][k])) p = r;
      if (p != k)
        for (int c = 0; c < 5; ++c) {
          const double t = m[k][c];
          m[k][c] = m[p][c];
          m[p][c] = t;
        }
      for (int r = k + 1; r < 4; ++r) {
        const double f = m[r][k] / m[k][k];
        for (int c = k; c < 5; ++c) m[r][c] -= f * m[k][c];
      }
    }
    for (int k = 3; k >= 0; --k) {
      double v = m[k][4];
      for (int c = k + 1; c < 4; ++c) v -= m[k][c] * wsol[c];
      wsol[k] = v / m[k][k];
    }
  }
  __syncthreads();
  for (int64_t i = tid; i < n; i += blockDim.x) {
    double v = w.z[i];
    for (int j = 0; j < 4; ++j) v -= w.y[j * n + i] * wsol[j];
    out[1 + i] = R(v);
  }
}

template <typename R>
int ch_penta(const ppf_grid& g, const void* in, void* out, double a, double e, void* work, hipStream_t s) {
  const int64_t n = g.n[0];
  double* b = static_cast<double*>(work);
  ChPenta w{b, b + n, b + 2 * n, b + 3 * n, b + 4 * n, b + 5 * n, b + 6 * n};
  hipLaunchKernelGGL(k_ch_penta<R>, dim3(1), dim3(64), 0, s, static_cast<const R*>(in), static_cast<R*>(out), n, a,
                     e, w);
  return check_launch("cahnhilliard_line_solve");
}

}  // namespace
}  // namespace ppf

using namespace ppf;

extern "C" {

int64_t ppf_cyclic_table_len(int64_t n) { return n < 3 ? -1 : kTableHead + 3 * n; }

int ppf_cyclic_table(int64_t n, double a, double* table) {
  if (n < 3) return fail(PPF_EINVAL, "cyclic_table: n must be at least 3");
  if (!(a >= 0.0) || !std::isfinite(a)) return fail(PPF_EINVAL, "cyclic_table: a must be finite and >= 0");
  if (!table) return fail(PPF_EINVAL, "cyclic_table: null table");
  const double diag = 1.0 + 2.0 * a, off = -a;
  const double gam = -diag;
  double* cp = table + kTableHead;
  double* den = cp + n;
  double* q = den + n;
  // corner-free matrix: diag' = diag except diag'_0 = diag - gam, diag'_{n-1} = diag - off^2/gam
  for (int64_t i = 0; i < n; ++i) {
    double di = diag;
    if (i == 0) di = diag - gam;
    if (i == n - 1) di = diag - off * off / gam;
    const double m = i == 0 ? di : di - off * cp[i - 1];
    den[i] = 1.0 / m;
    cp[i] = off / m;
  }
  // correction vector: corner-free matrix times q = (gam, 0, ..., 0, off)
  for (int64_t i = 0; i < n; ++i) {
    const double u = i == 0 ? gam : (i == n - 1 ? off : 0.0);
    q[i] = i == 0 ? u * den[0] : (u - off * q[i - 1]) * den[i];
  }
  for (int64_t i = n - 2; i >= 0; --i) q[i] = q[i] - cp[i] * q[i + 1];
  const double r = off / gam;
  table[0] = off;
  table[1] = 1.0 / (1.0 + q[0] + r * q[n - 1]);
  table[2] = r;
  table[3] = 0.0;
  return 0;
}

int64_t ppf_scan_table_len(int64_t m) { return m < 1 ? -1 : kScanHead + m + 1; }

int ppf_scan_table(int64_t n, double a, int64_t m, double* table) {
  if (n < 3) return fail(PPF_EINVAL, "scan_table: n must be at least 3");
  if (m < 1) return fail(PPF_EINVAL, "scan_table: chunk length must be positive");
  if (!(a >= 0.0) || !std::isfinite(a)) return fail(PPF_EINVAL, "scan_table: a must be finite and >= 0");
  if (!table) return fail(PPF_EINVAL, "scan_table: null table");
  // a = 0: the identity (the limit r -> 0 with r/a -> 1)
  const double r = a == 0.0 ? 0.0 : ((1.0 + 2.0 * a) - std::sqrt(1.0 + 4.0 * a)) / (2.0 * a);
  table[0] = r;
  table[1] = a == 0.0 ? 1.0 : r / a;
  table[2] = std::pow(r, double(n));
  table[3] = double(m);
  double pk = 1.0;
  for (int64_t k = 0; k <= m; ++k) {
    table[kScanHead + k] = pk;
    pk *= r;
  }
  return 0;
}

int64_t ppf_line_scan_work_bytes(const ppf_grid* g, int axis, int64_t m) {
  if (check_grid(g, 1, 3, "line_scan_work_bytes")) return -1;
  if (axis < 0 || axis >= g->ndim || m < 1) return -1;
  int64_t lines = 1;
  for (int a = 0; a < g->ndim; ++a)
    if (a != axis) lines *= g->n[a];
  const int64_t nch = (g->n[axis] + m - 1) / m;
  return 2 * lines * nch * int64_t(sizeof(double));
}

int ppf_diffusion_line_solve_scan(const ppf_grid* g, const void* in, void* out, int axis, const double* table,
                                  int64_t m, int explicit_axis, double b, void* work, ppf_stream_t s) {
  if (int rc = check_periodic(g, "diffusion_line_solve_scan")) return rc;
  if (axis < 0 || axis >= g->ndim) return fail(PPF_EINVAL, "diffusion_line_solve_scan: bad axis");
  if (explicit_axis < -1 || explicit_axis >= g->ndim)
    return fail(PPF_EINVAL, "diffusion_line_solve_scan: bad explicit axis");
  if (m < 1) return fail(PPF_EINVAL, "diffusion_line_solve_scan: chunk length must be positive");
  if (!in || !out || !table || !work) return fail(PPF_EINVAL, "diffusion_line_solve_scan: null pointer");
  if (in == out) return fail(PPF_EINVAL, "diffusion_line_solve_scan: in and out must differ");
  if (!(b >= 0.0) || !std::isfinite(b))
    return fail(PPF_EINVAL, "diffusion_line_solve_scan: b must be finite and >= 0");
  return g->dtype == PPF_F32
             ? line_solve_scan<float>(*g, in, out, axis, table, explicit_axis, b, work, m, stream_of(s))
             : line_solve_scan<double>(*g, in, out, axis, table, explicit_axis, b, work, m, stream_of(s));
}

int ppf_diffusion_line_solve(const ppf_grid* g, const void* in, void* out, int axis, const double* table,
                             int explicit_axis, double b, ppf_stream_t s) {
  if (int rc = check_periodic(g, "diffusion_line_solve")) return rc;
  if (axis < 0 || axis >= g->ndim) return fail(PPF_EINVAL, "diffusion_line_solve: bad axis");
  if (explicit_axis < -1 || explicit_axis >= g->ndim)
    return fail(PPF_EINVAL, "diffusion_line_solve: bad explicit axis");
  if (!in || !out || !table) return fail(PPF_EINVAL, "diffusion_line_solve: null pointer");
  if (in == out) return fail(PPF_EINVAL, "diffusion_line_solve: in and out must differ");
  if (!(b >= 0.0) || !std::isfinite(b)) return fail(PPF_EINVAL, "diffusion_line_solve: b must be finite and >= 0");
  return g->dtype == PPF_F32 ? line_solve<float>(*g, in, out, axis, table, explicit_axis, b, stream_of(s))
                             : line_solve<double>(*g, in, out, axis, table, explicit_axis, b, stream_of(s));
}

int64_t ppf_diffusion_cg_work_bytes(const ppf_grid* g) {
  if (check_grid(g, 1, 3, "diffusion_cg_work_bytes")) return -1;
  const int64_t N = g->n[0] * (g->ndim >= 2 ? g->n[1] : 1) * (g->ndim >= 3 ? g->n[2] : 1);
  const int64_t es = g->dtype == PPF_F32 ? 4 : 8;
  return 256 + 4 * kCgBlocks * int64_t(sizeof(double)) + 4 * align256(N * es);
}

int ppf_diffusion_cg_begin(const ppf_grid* g, const void* in, void* out, double a, double b, void* work,
                           ppf_stream_t s) {
  if (int rc = check_periodic(g, "diffusion_cg_begin")) return rc;
  if (!in || !out || !work) return fail(PPF_EINVAL, "diffusion_cg_begin: null pointer");
  if (in == out) return fail(PPF_EINVAL, "diffusion_cg_begin: in and out must differ");
  if (!(a >= 0.0) || !std::isfinite(a) || !(b >= 0.0) || !std::isfinite(b))
    return fail(PPF_EINVAL, "diffusion_cg_begin: a and b must be finite and >= 0");
  return g->dtype == PPF_F32 ? cg_begin<float, false>(*g, in, out, a, b, 0.0, work, stream_of(s))
                             : cg_begin<double, false>(*g, in, out, a, b, 0.0, work, stream_of(s));
}

int ppf_diffusion_cg_iterate(const ppf_grid* g, void* out, double a, double tol, int64_t first_iteration,
                             int iterations, void* work, ppf_stream_t s) {
  if (int rc = check_periodic(g, "diffusion_cg_iterate")) return rc;
  if (!out || !work) return fail(PPF_EINVAL, "diffusion_cg_iterate: null pointer");
  if (!(a >= 0.0) || !std::isfinite(a)) return fail(PPF_EINVAL, "diffusion_cg_iterate: a must be finite and >= 0");
  if (!(tol >= 0.0)) return fail(PPF_EINVAL, "diffusion_cg_iterate: tol must be >= 0");
  if (first_iteration < 0 || iterations < 0) return fail(PPF_EINVAL, "diffusion_cg_iterate: negative count");
  return g->dtype == PPF_F32
             ? cg_iterate<float, false>(*g, nullptr, out, a, 0.0, tol, first_iteration, iterations, work, stream_of(s))
             : cg_iterate<double, false>(*g, nullptr, out, a, 0.0, tol, first_iteration, iterations, work,
                                         stream_of(s));
}

int64_t ppf_allencahn_line_work_bytes(const ppf_grid* g) {
  if (check_grid(g, 1, 3, "allencahn_line_work_bytes")) return -1;
  return 2 * g->sf * (g->dtype == PPF_F32 ? 4 : 8);
}

int ppf_allencahn_line_solve(const ppf_grid* g, const void* in, void* out, int axis, double a, int explicit_axis,
                             double b, double src_k, double imex_k, void* work, ppf_stream_t s) {
  if (int rc = check_periodic(g, "allencahn_line_solve")) return rc;
  if (axis < 0 || axis >= g->ndim) return fail(PPF_EINVAL, "allencahn_line_solve: bad axis");
  if (explicit_axis < -1 || explicit_axis >= g->ndim)
    return fail(PPF_EINVAL, "allencahn_line_solve: bad explicit axis");
  if (!in || !out || !work) return fail(PPF_EINVAL, "allencahn_line_solve: null pointer");
  if (in == out) return fail(PPF_EINVAL, "allencahn_line_solve: in and out must differ");
  if (!(a >= 0.0) || !std::isfinite(a) || !(b >= 0.0) || !std::isfinite(b) || !std::isfinite(src_k) ||
      !std::isfinite(imex_k))
    return fail(PPF_EINVAL, "allencahn_line_solve: coefficients must be finite, a and b >= 0");
  return g->dtype == PPF_F32
             ? line_solve_var<float>(*g, in, out, axis, a, explicit_axis, b, src_k, imex_k, work, stream_of(s))
             : line_solve_var<double>(*g, in, out, axis, a, explicit_axis, b, src_k, imex_k, work, stream_of(s));
}

int ppf_allencahn_cg_begin(const ppf_grid* g, const void* in, void* out, double a, double src_k, void* work,
                           ppf_stream_t s) {
  if (int rc = check_periodic(g, "allencahn_cg_begin")) return rc;
  if (!in || !out || !work) return fail(PPF_EINVAL, "allencahn_cg_begin: null pointer");
  if (in == out) return fail(PPF_EINVAL, "allencahn_cg_begin: in and out must differ");
  if (!(a >= 0.0) || !std::isfinite(a) || !std::isfinite(src_k))
    return fail(PPF_EINVAL, "allencahn_cg_begin: a must be finite and >= 0, src_k finite");
  return g->dtype == PPF_F32 ? cg_begin<float, true>(*g, in, out, a, 0.0, src_k, work, stream_of(s))
                             : cg_begin<double, true>(*g, in, out, a, 0.0, src_k, work, stream_of(s));
}

int ppf_allencahn_cg_iterate(const ppf_grid* g, const void* in, void* out, double a, double src_k, double tol,
                             int64_t first_iteration, int iterations, void* work, ppf_stream_t s) {
  if (int rc = check_periodic(g, "allencahn_cg_iterate")) return rc;
  if (!in || !out || !work) return fail(PPF_EINVAL, "allencahn_cg_iterate: null pointer");
  if (!(a >= 0.0) || !std::isfinite(a) || !std::isfinite(src_k) || !(tol >= 0.0))
    return fail(PPF_EINVAL, "allencahn_cg_iterate: bad coefficients");
  if (first_iteration < 0 || iterations < 0) return fail(PPF_EINVAL, "allencahn_cg_iterate: negative count");
  return g->dtype == PPF_F32
             ? cg_iterate<float, true>(*g, in, out, a, src_k, tol, first_iteration, iterations, work, stream_of(s))
             : cg_iterate<double, true>(*g, in, out, a, src_k, tol, first_iteration, iterations, work,
                                        stream_of(s));
}

int64_t ppf_cahnhilliard_line_work_bytes(const ppf_grid* g) {
  if (check_grid(g, 1, 1, "cahnhilliard_line_work_bytes")) return -1;
  return 10 * g->n[0] * int64_t(sizeof(double));
}

int ppf_cahnhilliard_line_solve(const ppf_grid* g, const void* in, void* out, double alpha, double e, void* work,
                                ppf_stream_t s) {
  if (int rc = check_periodic(g, "cahnhilliard_line_solve")) return rc;
  if (g->ndim != 1) return fail(PPF_EINVAL, "cahnhilliard_line_solve: the pentadiagonal system is 1D");
  if (g->n[0] < 5) return fail(PPF_EINVAL, "cahnhilliard_line_solve: needs at least 5 cells");
  if (!in || !out || !work) return fail(PPF_EINVAL, "cahnhilliard_line_solve: null pointer");
  if (in == out) return fail(PPF_EINVAL, "cahnhilliard_line_solve: in and out must differ");
  return g->dtype == PPF_F32 ? ch_penta<float>(*g, in, out, alpha, e, work, stream_of(s))
                             : ch_penta<double>(*g, in, out, alpha, e, work, stream_of(s));
}

}  // extern "C"
