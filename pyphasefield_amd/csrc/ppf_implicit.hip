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

struct Cell {
  int64_t pad;      // padded offset
  int64_t nb[6];    // dense indices of the neighbours
  int cnt;
};

__device__ __forceinline__ Cell cell_of(const Dense& d, int64_t t) {
  Cell c;
  const int64_t x = t % d.nx;
  const int64_t y = d.nd >= 2 ? (t / d.nx) % d.ny : 0;
  const int64_t z = d.nd >= 3 ? t / (d.nx * d.ny) : 0;
  c.pad = (x + 1) + (d.nd >= 2 ? (y + 1) * d.sy : 0) + (d.nd >= 3 ? (z + 1) * d.sz : 0);
  const int64_t row = t - x;
  c.nb[0] = row + (x == 0 ? d.nx - 1 : x - 1);
  c.nb[1] = row + (x == d.nx - 1 ? 0 : x + 1);
  c.cnt = 2;
  if (d.nd >= 2) {
    const int64_t plane = t - (t % (d.nx * d.ny));
    const int64_t inrow = t - plane - y * d.nx;
    c.nb[2] = plane + (y == 0 ? d.ny - 1 : y - 1) * d.nx + inrow;
    c.nb[3] = plane + (y == d.ny - 1 ? 0 : y + 1) * d.nx + inrow;
    c.cnt = 4;
  }
  if (d.nd >= 3) {
    const int64_t pl = d.nx * d.ny;
    const int64_t inpl = t - z * pl;
    c.nb[4] = (z == 0 ? d.nz - 1 : z - 1) * pl + inpl;
    c.nb[5] = (z == d.nz - 1 ? 0 : z + 1) * pl + inpl;
    c.cnt = 6;
  }
  return c;
}

// x0 = in (written to out), rhs = in or the explicit half step of in,
// r = rhs - A x0; partial sums of r.r and rhs.rhs
template <typename R, bool VAR>
__global__ __launch_bounds__(kCgThreads) void k_cg_begin(const R* __restrict__ in, R* __restrict__ out, Dense d,
                                                        double a, double b, double srck, Work w) {
  R* r = static_cast<R*>(w.r);
  double srr = 0.0, sbb = 0.0;
  const double na = 2.0 * d.nd;
  for (int64_t t = int64_t(blockIdx.x) * kCgThreads + threadIdx.x; t < d.N; t += int64_t(d.nb) * kCgThreads) {
    const int64_t x = t % d.nx;
    const int64_t y = d.nd >= 2 ? (t / d.nx) % d.ny : 0;
    const int64_t z = d.nd >= 3 ? t / (d.nx * d.ny) : 0;
    const int64_t p = (x + 1) + (d.nd >= 2 ? (y + 1) * d.sy : 0) + (d.nd >= 3 ? (z + 1) * d.sz : 0);
    // periodic neighbours in the padded array
    // padded column of interior x' is x' + 1, so neighbour x' sits at p - x + x'
    R s = in[p - x + (x == 0 ? d.nx - 1 : x - 1)] + in[p - x + (x == d.nx - 1 ? 0 : x + 1)];
    if (d.nd >= 2) {
      s = s + in[p + (y == 0 ? (d.ny - 1) * d.sy : -d.sy)] + in[p + (y == d.ny - 1 ? -(d.ny - 1) * d.sy : d.sy)];
    }
    if (d.nd >= 3) {
      s = s + in[p + (z == 0 ? (d.nz - 1) * d.sz : -d.sz)] + in[p + (z == d.nz - 1 ? -(d.nz - 1) * d.sz : d.sz)];
    }
    const R v = in[p];
    const R rhs = b != 0.0 ? R(1.0 - na * b) * v + R(b) * s : v;
    R diag = R(1.0 + na * a);
    if (VAR) diag = diag + R(srck) * (R(4) * v * v - R(6) * v + R(2));
    const R ax = diag * v - R(a) * s;
    const R rr = rhs - ax;
    out[p] = v;
    r[t] = rr;
    srr += double(rr) * double(rr);
    sbb += double(rhs) * double(rhs);
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
  double spq = 0.0;
  for (int64_t t = int64_t(blockIdx.x) * kCgThreads + threadIdx.x; t < d.N; t += int64_t(d.nb) * kCgThreads) {
    const Cell c = cell_of(d, t);
    const R pv = k == 0 ? r[t] : r[t] + bt * po[t];
    R s = R(0);
    for (int j = 0; j < c.cnt; ++j) s = s + (k == 0 ? r[c.nb[j]] : r[c.nb[j]] + bt * po[c.nb[j]]);
    R dg = cdiag;
    if (VAR) {
      const R v = in[c.pad];
      dg = dg + R(srck) * (R(4) * v * v - R(6) * v + R(2));
    }
    const R qv = dg * pv - ra * s;
    pn[t] = pv;
    q[t] = qv;
    spq += double(pv) * double(qv);
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
  for (int64_t t = int64_t(blockIdx.x) * kCgThreads + threadIdx.x; t < d.N; t += int64_t(d.nb) * kCgThreads) {
    const int64_t x = t % d.nx;
    const int64_t y = d.nd >= 2 ? (t / d.nx) % d.ny : 0;
    const int64_t z = d.nd >= 3 ? t / (d.nx * d.ny) : 0;
    const int64_t pp = (x + 1) + (d.nd >= 2 ? (y + 1) * d.sy : 0) + (d.nd >= 3 ? (z + 1) * d.sz : 0);
    out[pp] = out[pp] + al * p[t];
    const R rv = r[t] - al * q[t];
    r[t] = rv;
    srr += double(rv) * double(rv);
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
  const int64_t want = (d.N + kCgThreads - 1) / kCgThreads;
  d.nb = int(want < kCgBlocks ? want : kCgBlocks);
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

}  // extern "C"
