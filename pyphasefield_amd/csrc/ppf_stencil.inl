// Stencil kernels of the Diffusion, Allen-Cahn and Cahn-Hilliard engines, the
// boundary-condition materialisation, the temperature fill and the FP64
// conservation reductions.  Included by ppf_f32.hip and ppf_f64.hip, which
// instantiate every launcher for one precision.
#pragma once

#include "ppf_launch.h"

namespace ppf {

struct Dom {
  Axis ax, ay, az;
  int ndim;
  int64_t sf;
};

inline Dom make_dom(const ppf_grid& g) {
  Dom d;
  d.ndim = g.ndim;
  d.ax = Axis{g.n[0], 1, g.bc[0]};
  d.ay = Axis{g.ndim >= 2 ? g.n[1] : 1, g.sy, g.ndim >= 2 ? g.bc[1] : PPF_PERIODIC};
  d.az = Axis{g.ndim >= 3 ? g.n[2] : 1, g.sz, g.ndim >= 3 ? g.bc[2] : PPF_PERIODIC};
  d.sf = g.sf;
  return d;
}

__device__ __forceinline__ bool inside(const Dom& d, int64_t x, int64_t y, int64_t z) {
  bool in = x >= 1 && x <= d.ax.n;
  if (d.ndim >= 2) in = in && y >= 1 && y <= d.ay.n;
  if (d.ndim >= 3) in = in && z >= 1 && z <= d.az.n;
  return in;
}

__device__ __forceinline__ int64_t offset(const Dom& d, int64_t x, int64_t y, int64_t z) {
  int64_t o = x;
  if (d.ndim >= 2) o += y * d.ay.stride;
  if (d.ndim >= 3) o += z * d.az.stride;
  return o;
}

// value of padded cell (x,y,z) with the boundary rules folded in
template <typename R>
__device__ __forceinline__ R fold(const R* __restrict__ f, const R* __restrict__ b, const Dom& d, int64_t x,
                                  int64_t y, int64_t z, R dx) {
  if (inside(d, x, y, z)) return f[offset(d, x, y, z)];
  if (d.ndim == 1) return fold_x(f, b, d.ax, x, 0, dx);
  if (d.ndim == 2) return fold_xy(f, b, d.ax, d.ay, x, y, 0, dx);
  return fold_xyz(f, b, d.ax, d.ay, d.az, x, y, z, dx);
}

// padded extents and decode of a linear padded index
__device__ __forceinline__ void decode(const Dom& d, int64_t t, int64_t* x, int64_t* y, int64_t* z) {
  int64_t px = d.ax.n + 2;
  int64_t py = d.ndim >= 2 ? d.ay.n + 2 : 1;
  *x = t % px;
  int64_t r = t / px;
  *y = d.ndim >= 2 ? r % py : 0;
  *z = d.ndim >= 3 ? r / py : 0;
}

inline int64_t padded_cells(const ppf_grid& g) {
  int64_t c = g.n[0] + 2;
  if (g.ndim >= 2) c *= g.n[1] + 2;
  if (g.ndim >= 3) c *= g.n[2] + 2;
  return c;
}

inline int64_t interior_cells(const ppf_grid& g) {
  int64_t c = g.n[0];
  if (g.ndim >= 2) c *= g.n[1];
  if (g.ndim >= 3) c *= g.n[2];
  return c;
}

// ------------------------------------------------------------------ halos --
template <typename R>
__global__ __launch_bounds__(256) void k_apply_bc(R* __restrict__ f, const R* __restrict__ b, Dom d, R dx,
                                                  int64_t total) {
  int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  int fi = blockIdx.y;
  if (t >= total) return;
  int64_t x, y, z;
  decode(d, t, &x, &y, &z);
  if (inside(d, x, y, z)) return;
  R* ff = f + fi * d.sf;
  const R* bb = b ? b + fi * d.sf : nullptr;
  R v;
  if (d.ndim == 1) v = fold_x(ff, bb, d.ax, x, 0, dx);
  else if (d.ndim == 2) v = fold_xy(ff, bb, d.ax, d.ay, x, y, 0, dx);
  else v = fold_xyz(ff, bb, d.ax, d.ay, d.az, x, y, z, dx);
  ff[offset(d, x, y, z)] = v;
}

template <typename R>
int apply_bc(const ppf_grid& g, void* fields, const void* bcvals, hipStream_t s) {
  Dom d = make_dom(g);
  int64_t total = padded_cells(g);
  dim3 grid(unsigned((total + 255) / 256), unsigned(g.nfields));
  hipLaunchKernelGGL(k_apply_bc<R>, grid, dim3(256), 0, s, static_cast<R*>(fields),
                     static_cast<const R*>(bcvals), d, R(g.dx), total);
  return check_launch("apply_bc");
}

template <typename R>
__global__ __launch_bounds__(256) void k_apply_bc_t(R* __restrict__ T, Dom d, int64_t total) {
  int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= total) return;
  int64_t x, y, z;
  decode(d, t, &x, &y, &z);
  if (inside(d, x, y, z)) return;
  int64_t xi = resolve_t(x, d.ax.n, d.ax.bc);
  int64_t yi = d.ndim >= 2 ? resolve_t(y, d.ay.n, d.ay.bc) : 0;
  int64_t zi = d.ndim >= 3 ? resolve_t(z, d.az.n, d.az.bc) : 0;
  T[offset(d, x, y, z)] = T[offset(d, xi, yi, zi)];
}

template <typename R>
int apply_bc_temperature(const ppf_grid& g, void* T, hipStream_t s) {
  Dom d = make_dom(g);
  int64_t total = padded_cells(g);
  hipLaunchKernelGGL(k_apply_bc_t<R>, dim3(unsigned((total + 255) / 256)), dim3(256), 0, s, static_cast<R*>(T),
                     d, total);
  return check_launch("apply_bc_temperature");
}

template <typename R>
__global__ __launch_bounds__(256) void k_fill_t(R* __restrict__ out, Dom d, ppf_temperature t, double dx,
                                                int64_t step, int64_t total) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= total) return;
  int64_t x, y, z;
  decode(d, i, &x, &y, &z);
  int64_t n[3] = {d.ax.n, d.ndim >= 2 ? d.ay.n : 1, d.ndim >= 3 ? d.az.n : 1};
  int bc[3] = {d.ax.bc, d.ay.bc, d.az.bc};
  // absent axes sit at interior coordinate 1
  int64_t yy = d.ndim >= 2 ? y : 1, zz = d.ndim >= 3 ? z : 1;
  out[offset(d, x, y, z)] = temperature_at<R>(t, n, bc, x, yy, zz, dx, step);
}

template <typename R>
int fill_temperature(const ppf_grid& g, void* out, const ppf_temperature& t, int64_t step, hipStream_t s) {
  Dom d = make_dom(g);
  int64_t total = padded_cells(g);
  hipLaunchKernelGGL(k_fill_t<R>, dim3(unsigned((total + 255) / 256)), dim3(256), 0, s, static_cast<R*>(out), d,
                     t, g.dx, step, total);
  return check_launch("fill_temperature");
}

// ------------------------------------------------------- 2D tile loading --
// Tile of TX x TY outputs, 256 threads as 64 columns x 4 row groups; each
// thread owns rows ty, ty+4, ... of column tx.
constexpr int kTX = 64;
constexpr int kTY = 16;
constexpr int kThreads = 256;

// Interior coordinate a halo coordinate reads under a periodic or zero-flux
// rule without bc values: resolve()'s interior coordinate.
__device__ __forceinline__ int32_t remap_coord32(int32_t p, int32_t n, int bc) {
  if (p >= 1 && p <= n) return p;
  if (bc == PPF_PERIODIC) {
    int32_t q = p < 1 ? p + n : p - n;
    q = q < 1 ? 1 : q;
    return q > n ? n : q;
  }
  return p < 1 ? 1 : n;
}

// register slots per thread of a W x HH tile
template <int W, int HH>
constexpr int tile_slots() {
  return (W * HH + kThreads - 1) / kThreads;
}

// whether a tile's halo is a pure coordinate remap (periodic wrap, or zero-flux
// Neumann copy when there are no bc values: the fold then adds -/+dx*0) that
// 32-bit offsets can address
template <typename R>
__device__ __forceinline__ bool tile_remappable(const R* __restrict__ b, const Dom& d) {
  return !b && d.ax.bc != PPF_DIRICHLET && d.ay.bc != PPF_DIRICHLET &&
         (d.ay.n + 2) * d.ay.stride < (int64_t(1) << 31);
}

// this thread's elements of the W x HH tile at padded (xs, ys) into registers,
// all loads issued together: interior tiles directly, other tiles through the
// remap (the caller checks tile_remappable)
template <typename R, int W, int HH, int NPX>
__device__ __forceinline__ void tile_fetch(R (&v)[NPX], const R* __restrict__ f, const Dom& d, int64_t xs,
                                           int64_t ys, bool interior) {
  constexpr int NP = tile_slots<W, HH>(), NFULL = (W * HH) / kThreads;
  static_assert(NP <= NPX, "register slots");
  if (interior) {
    const R* __restrict__ base = f + ys * d.ay.stride + xs;
    const int32_t sy = int32_t(d.ay.stride);
#pragma unroll
    for (int j = 0; j < NP; ++j) {
      const int t = threadIdx.x + j * kThreads;
      if (j < NFULL || t < W * HH) {
        const int ly = t / W, lx = t - ly * W;
        v[j] = base[ly * sy + lx];
      }
    }
    return;
  }
  const int32_t nx = int32_t(d.ax.n), ny = int32_t(d.ay.n), sy = int32_t(d.ay.stride);
  const int32_t xs32 = int32_t(xs), ys32 = int32_t(ys);
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    const int t = threadIdx.x + j * kThreads;
    if (j < NFULL || t < W * HH) {
      const int ly = t / W, lx = t - ly * W;
      const int32_t x = remap_coord32(xs32 + lx, nx, d.ax.bc), y = remap_coord32(ys32 + ly, ny, d.ay.bc);
      v[j] = f[uint32_t(y * sy + x)];
    }
  }
}

template <typename R, int W, int HH, int NPX>
__device__ __forceinline__ void tile_put(R* __restrict__ flat, const R (&v)[NPX]) {
  constexpr int NP = tile_slots<W, HH>(), NFULL = (W * HH) / kThreads;
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    const int t = threadIdx.x + j * kThreads;
    if (j < NFULL || t < W * HH) flat[t] = v[j];
  }
}

template <typename R, int W, int HH>
__device__ __forceinline__ void load_tile(R (*s)[W], const R* __restrict__ f, const R* __restrict__ b,
                                          const Dom& d, R dx, int64_t xs, int64_t ys, bool interior) {
  // s[ly][lx] = padded cell (xs + lx, ys + ly).  Interior tiles, and tiles
  // whose rules are a pure coordinate remap, read every element with its loads
  // issued before the first LDS store; other tiles fold element by element.
  if (interior || tile_remappable(b, d)) {
    constexpr int NP = tile_slots<W, HH>();
    R v[NP];
    tile_fetch<R, W, HH, NP>(v, f, d, xs, ys, interior);
    tile_put<R, W, HH, NP>(&s[0][0], v);
    return;
  }
  for (int t = threadIdx.x; t < W * HH; t += kThreads) {
    int ly = t / W;
    int lx = t - ly * W;
    int64_t x = xs + lx, y = ys + ly;
    R v;
    if (interior) {
      v = f[y * d.ay.stride + x];
    } else if (x > d.ax.n + 2 || y > d.ay.n + 2) {
      v = R(0);  // beyond a partial tile: never used
    } else {
      v = fold(f, b, d, x, y, int64_t(0), dx);
    }
    s[ly][lx] = v;
  }
}

template <int TX = kTX, int TY = kTY>
__device__ __forceinline__ void tile_coords(const Dom& d, int64_t* x0, int64_t* y0) {
  uint32_t nbx = uint32_t((d.ax.n + TX - 1) / TX);
  uint32_t b = xcd_swizzle(blockIdx.x, gridDim.x);
  *x0 = 1 + int64_t(b % nbx) * TX;
  *y0 = 1 + int64_t(b / nbx) * TY;
}

template <int TX = kTX, int TY = kTY>
__device__ __forceinline__ bool tile_interior(const Dom& d, int64_t x0, int64_t y0, int h) {
  return x0 - h >= 1 && x0 + TX - 1 + h <= d.ax.n && y0 - h >= 1 && y0 + TY - 1 + h <= d.ay.n;
}

inline unsigned tiles2d(const ppf_grid& g) {
  return unsigned(((g.n[0] + kTX - 1) / kTX) * ((g.n[1] + kTY - 1) / kTY));
}

// -------------------------------------------- Diffusion / Allen-Cahn (2D) --
// kind 0: c + dt*(D*lap c)                          (Diffusion.py:204-207)
// kind 1: phi + dt*M*(eps^2 lap - 16W(4p^3-6p^2+2p)) (CahnAllen.py:129-135)
struct PointParams {
  double a, b, c, dt;  // kind 0: a=D; kind 1: a=M, b=W, c=eps
};

template <typename R, int KIND>
__device__ __forceinline__ R point_update(R v, R lap, const PointParams& p) {
  if (KIND == 0) {
    return v + R(p.dt) * (R(p.a) * lap);
  } else {
    R eps2 = R(p.c) * R(p.c);
    R well = R(4) * v * v * v - R(6) * v * v + R(2) * v;
    return v + R(p.dt) * R(p.a) * (eps2 * lap - R(16) * R(p.b) * well);
  }
}

template <typename R, int KIND, int TX, int TY>
__device__ __forceinline__ void point_tile_body(const R* smem, R* __restrict__ out, const Dom& d, R dx,
                                                const PointParams& p, int64_t x0, int64_t y0);

// One TX x TY output tile at padded (x0, y0) of a 2D field: 1-cell halo
// staged in LDS (s holds (TY + 2) x (TX + 2) values), 256 threads as TX
// columns x (256 / TX) row groups.
template <typename R, int KIND, int TX, int TY>
__device__ __forceinline__ void point_tile(R* smem, const R* __restrict__ in, R* __restrict__ out,
                                           const R* __restrict__ b, const Dom& d, R dx, const PointParams& p,
                                           int64_t x0, int64_t y0) {
  constexpr int H = 1, W = TX + 2 * H, HH = TY + 2 * H, RP = kThreads / TX;
  static_assert(kThreads % TX == 0 && TY % RP == 0, "tile shape");
  R(*s)[W] = reinterpret_cast<R(*)[W]>(smem);
  load_tile<R, W, HH>(s, in, b, d, dx, x0 - H, y0 - H, tile_interior<TX, TY>(d, x0, y0, H));
  __syncthreads();
  point_tile_body<R, KIND, TX, TY>(smem, out, d, dx, p, x0, y0);
}

// the update of a point tile whose (TY + 2) x (TX + 2) values are in LDS
template <typename R, int KIND, int TX, int TY>
__device__ __forceinline__ void point_tile_body(const R* smem, R* __restrict__ out, const Dom& d, R dx,
                                                const PointParams& p, int64_t x0, int64_t y0) {
  constexpr int H = 1, W = TX + 2 * H, RP = kThreads / TX;
  const R(*s)[W] = reinterpret_cast<const R(*)[W]>(smem);
  const R idx = R(1) / dx;
  const R idx2 = idx * idx;
  int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
  int64_t x = x0 + tx;
  if (x > d.ax.n) return;
#pragma unroll
  for (int k = 0; k < TY / RP; ++k) {
    int ly = ty + RP * k;
    int64_t y = y0 + ly;
    if (y > d.ay.n) break;
    R c = s[ly + 1][tx + 1];
    // field.py:64-66 order: axis 0 (y) then axis 1 (x), each (c[+1]+c[-1]-2c)*idx2
    R lap = (s[ly + 2][tx + 1] + s[ly][tx + 1] - R(2) * c) * idx2;
    lap += (s[ly + 1][tx + 2] + s[ly + 1][tx] - R(2) * c) * idx2;
    out[y * d.ay.stride + x] = point_update<R, KIND>(c, lap, p);
  }
}

template <typename R, int KIND, int TY = kTY>
__global__ __launch_bounds__(kThreads) void k_point2d(const R* __restrict__ in, R* __restrict__ out,
                                                      const R* __restrict__ b, Dom d, R dx, PointParams p) {
  __shared__ R s[(TY + 2) * (kTX + 2)];
  int64_t x0, y0;
  tile_coords<kTX, TY>(d, &x0, &y0);
  point_tile<R, KIND, kTX, TY>(s, in, out, b, d, dx, p, x0, y0);
}

// grids with fewer 64 x 16 tiles than this take 64 x 4 tiles (four times the
// blocks; a 256^2 grid is otherwise 64 blocks on 256 CUs)
constexpr unsigned kSmallGridTiles = 512;

// generic 1D / 3D version (one cell per thread, folded global loads)
template <typename R, int KIND>
__global__ __launch_bounds__(256) void k_point_nd(const R* __restrict__ in, R* __restrict__ out,
                                                  const R* __restrict__ b, Dom d, R dx, PointParams p) {
  int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  int64_t nx = d.ax.n, ny = d.ndim >= 2 ? d.ay.n : 1, nz = d.ndim >= 3 ? d.az.n : 1;
  if (t >= nx * ny * nz) return;
  int64_t x = 1 + t % nx;
  int64_t y = d.ndim >= 2 ? 1 + (t / nx) % ny : 0;
  int64_t z = d.ndim >= 3 ? 1 + t / (nx * ny) : 0;
  const R idx = R(1) / dx;
  const R idx2 = idx * idx;
  R c = in[offset(d, x, y, z)];
  R lap;
  if (d.ndim == 3) {
    lap = (fold(in, b, d, x, y, z + 1, dx) + fold(in, b, d, x, y, z - 1, dx) - R(2) * c) * idx2;
    lap += (fold(in, b, d, x, y + 1, z, dx) + fold(in, b, d, x, y - 1, z, dx) - R(2) * c) * idx2;
    lap += (fold(in, b, d, x + 1, y, z, dx) + fold(in, b, d, x - 1, y, z, dx) - R(2) * c) * idx2;
  } else {
    lap = (fold(in, b, d, x + 1, y, z, dx) + fold(in, b, d, x - 1, y, z, dx) - R(2) * c) * idx2;
  }
  out[offset(d, x, y, z)] = point_update<R, KIND>(c, lap, p);
}

template <typename R, int KIND>
int point_step(const ppf_grid& g, const void* in, void* out, const void* bcvals, const PointParams& p,
               hipStream_t s, const char* what) {
  Dom d = make_dom(g);
  if (g.ndim == 2 && tiles2d(g) < kSmallGridTiles) {
    const unsigned blocks = unsigned(((g.n[0] + kTX - 1) / kTX) * ((g.n[1] + 3) / 4));
    hipLaunchKernelGGL((k_point2d<R, KIND, 4>), dim3(blocks), dim3(kThreads), 0, s, static_cast<const R*>(in),
                       static_cast<R*>(out), static_cast<const R*>(bcvals), d, R(g.dx), p);
  } else if (g.ndim == 2) {
    hipLaunchKernelGGL((k_point2d<R, KIND>), dim3(tiles2d(g)), dim3(kThreads), 0, s,
                       static_cast<const R*>(in), static_cast<R*>(out), static_cast<const R*>(bcvals), d,
                       R(g.dx), p);
  } else {
    int64_t total = interior_cells(g);
    hipLaunchKernelGGL((k_point_nd<R, KIND>), dim3(unsigned((total + 255) / 256)), dim3(256), 0, s,
                       static_cast<const R*>(in), static_cast<R*>(out), static_cast<const R*>(bcvals), d,
                       R(g.dx), p);
  }
  return check_launch(what);
}

// ----------------------------------------- several steps in one launch --
// Diffusion / Allen-Cahn on 2D grids whose halos are a coordinate remap
// (every side periodic, or zero gradient without bc values): a block loads
// its output tile plus a halo of ks cells once and advances ks steps in LDS,
// the valid region shrinking by one cell a step (temporal blocking), then
// writes the tile.  Periodic images in the halo evolve exactly as the cells
// they copy; on zero-gradient sides the cells outside the domain are reset to
// their nearest interior cell after every step, as the one-step kernels'
// loads read them.  Each update is point_update on the same five values, so
// the results equal ks one-step launches bit for bit.
template <typename R, int KIND, int TX, int TY>
__global__ __launch_bounds__(kThreads) void k_point2d_multi(const R* __restrict__ in, R* __restrict__ out, Dom d,
                                                            R dx, PointParams p, int32_t ks) {
  constexpr int H = kMultiMax, W = TX + 2 * H, HH = TY + 2 * H;
  __shared__ R s[2][HH][W];
  const int32_t nx = int32_t(d.ax.n), ny = int32_t(d.ay.n), sy = int32_t(d.ay.stride);
  const uint32_t nbx = uint32_t((nx + TX - 1) / TX);
  const uint32_t b = xcd_swizzle(blockIdx.x, gridDim.x);
  const int32_t x0 = 1 + int32_t(b % nbx) * TX, y0 = 1 + int32_t(b / nbx) * TY;
  // LDS (ly, lx) holds padded cell (x0 - H + lx, y0 - H + ly); the region
  // of step s is [H - ks + s, H + T + ks - s) on both axes
  const int lo0 = H - ks, wx = TX + 2 * ks, wy = TY + 2 * ks;
  for (int t = threadIdx.x; t < wx * wy; t += kThreads) {
    const int ry = t / wx, rx = t - ry * wx;
    const int32_t gx = remap_coord32(x0 - ks + rx, nx, d.ax.bc), gy = remap_coord32(y0 - ks + ry, ny, d.ay.bc);
    s[0][lo0 + ry][lo0 + rx] = in[gy * sy + gx];
  }
  const bool nx_edge = d.ax.bc != PPF_PERIODIC && (x0 - ks < 1 || x0 + TX - 1 + ks > nx);
  const bool ny_edge = d.ay.bc != PPF_PERIODIC && (y0 - ks < 1 || y0 + TY - 1 + ks > ny);
  const R idx = R(1) / dx;
  const R idx2 = idx * idx;
  for (int st = 1; st <= ks; ++st) {
    __syncthreads();
    const R(*src)[W] = s[(st - 1) & 1];
    R(*dst)[W] = s[st & 1];
    const int lo = lo0 + st, vx = TX + 2 * (ks - st), vy = TY + 2 * (ks - st);
    for (int t = threadIdx.x; t < vx * vy; t += kThreads) {
      const int ly = lo + t / vx, lx = lo + (t - (t / vx) * vx);
      const R c = src[ly][lx];
      // point_tile_body's order: y (c[y+1] + c[y-1]), then x (c[x+1] + c[x-1])
      R lap = (src[ly + 1][lx] + src[ly - 1][lx] - R(2) * c) * idx2;
      lap += (src[ly][lx + 1] + src[ly][lx - 1] - R(2) * c) * idx2;
      dst[ly][lx] = point_update<R, KIND>(c, lap, p);
    }
    if (nx_edge || ny_edge) {  // uniform per block
      __syncthreads();
      for (int t = threadIdx.x; t < vx * vy; t += kThreads) {
        const int ly = lo + t / vx, lx = lo + (t - (t / vx) * vx);
        const int32_t gx = x0 - H + lx, gy = y0 - H + ly;
        const bool ox = d.ax.bc != PPF_PERIODIC && (gx < 1 || gx > nx);
        const bool oy = d.ay.bc != PPF_PERIODIC && (gy < 1 || gy > ny);
        if (ox || oy) {
          const int sx = ox ? (gx < 1 ? 1 : nx) - (x0 - H) : lx;
          const int sy2 = oy ? (gy < 1 ? 1 : ny) - (y0 - H) : ly;
          dst[ly][lx] = dst[sy2][sx];  // an interior cell: never rewritten here
        }
      }
    }
  }
  __syncthreads();
  const R(*fin)[W] = s[ks & 1];
  for (int t = threadIdx.x; t < TX * TY; t += kThreads) {
    const int ty = t / TX, tx = t - ty * TX;
    const int32_t gx = x0 + tx, gy = y0 + ty;
    if (gx <= nx && gy <= ny) out[gy * sy + gx] = fin[H + ty][H + tx];
  }
}

// grids with at least this many 64 x 16 tiles take them; smaller grids take
// 16 x 8 tiles (more blocks: two per CU at 256^2)
#ifndef PPF_MULTI_BIG_TILES
#define PPF_MULTI_BIG_TILES 512
#endif
#ifndef PPF_MULTI_SMALL_TY
#define PPF_MULTI_SMALL_TY 8
#endif
#ifndef PPF_MULTI_SMALL_TX
#define PPF_MULTI_SMALL_TX 16
#endif
template <typename R, int KIND>
int point_multistep(const ppf_grid& g, const void* in, void* out, const void* bcvals, const PointParams& p,
                    int32_t nsteps, hipStream_t s, const char* what) {
  const int32_t mx = multistep_max(g, bcvals);
  if (nsteps < 1 || nsteps > mx)
    return fail(PPF_EINVAL, std::string(what) + ": nsteps must be 1.." + std::to_string(mx) +
                                " on this grid (2D, periodic or Neumann sides without bc values)");
  Dom d = make_dom(g);
  if (((g.n[0] + 63) / 64) * ((g.n[1] + 15) / 16) >= PPF_MULTI_BIG_TILES) {
    const unsigned blocks = unsigned(((g.n[0] + 63) / 64) * ((g.n[1] + 15) / 16));
    hipLaunchKernelGGL((k_point2d_multi<R, KIND, 64, 16>), dim3(blocks), dim3(kThreads), 0, s,
                       static_cast<const R*>(in), static_cast<R*>(out), d, R(g.dx), p, nsteps);
  } else {
    constexpr int STY = PPF_MULTI_SMALL_TY, STX = PPF_MULTI_SMALL_TX;
    const unsigned blocks = unsigned(((g.n[0] + STX - 1) / STX) * ((g.n[1] + STY - 1) / STY));
    hipLaunchKernelGGL((k_point2d_multi<R, KIND, STX, STY>), dim3(blocks), dim3(kThreads), 0, s,
                       static_cast<const R*>(in), static_cast<R*>(out), d, R(g.dx), p, nsteps);
  }
  return check_launch(what);
}

template <typename R>
int diffusion_multistep(const ppf_grid& g, const void* in, void* out, const void* bcvals, double D, double dt,
                        int32_t nsteps, hipStream_t s) {
  PointParams p{D, 0, 0, dt};
  return point_multistep<R, 0>(g, in, out, bcvals, p, nsteps, s, "diffusion_multistep");
}

template <typename R>
int allencahn_multistep(const ppf_grid& g, const void* in, void* out, const void* bcvals, double M, double W,
                        double eps, double dt, int32_t nsteps, hipStream_t s) {
  PointParams p{M, W, eps, dt};
  return point_multistep<R, 1>(g, in, out, bcvals, p, nsteps, s, "allencahn_multistep");
}

template <typename R>
int diffusion_step(const ppf_grid& g, const void* in, void* out, const void* bcvals, double D, double dt,
                   hipStream_t s) {
  PointParams p{D, 0, 0, dt};
  return point_step<R, 0>(g, in, out, bcvals, p, s, "diffusion_step");
}

template <typename R>
int allencahn_step(const ppf_grid& g, const void* in, void* out, const void* bcvals, double M, double W,
                   double eps, double dt, hipStream_t s) {
  PointParams p{M, W, eps, dt};
  return point_step<R, 1>(g, in, out, bcvals, p, s, "allencahn_step");
}

// ------------------------------------------------------ Cahn-Hilliard --
template <typename R>
__device__ __forceinline__ R ch_mu(R c, R lap, R eps2) {
  return -eps2 * lap + (R(4) * c * c * c - R(6) * c * c + R(2) * c);
}

template <typename R, int TX, int TY>
__device__ __forceinline__ void ch_tile_body(R* smem, R* __restrict__ out, const Dom& d, R dx, R eps, R dtM,
                                             int64_t x0, int64_t y0);

// One TX x TY Cahn-Hilliard output tile: 2-cell halo of c in LDS, mu on the
// tile + 1 ring in LDS (smem holds (TY + 4)(TX + 4) + (TY + 2)(TX + 2) values).
template <typename R, int TX, int TY>
__device__ __forceinline__ void ch_tile(R* smem, const R* __restrict__ in, R* __restrict__ out,
                                        const R* __restrict__ b, const Dom& d, R dx, R eps, R dtM, int64_t x0,
                                        int64_t y0) {
  constexpr int H = 2, W = TX + 2 * H, HH = TY + 2 * H, RP = kThreads / TX;
  static_assert(kThreads % TX == 0 && TY % RP == 0, "tile shape");
  R(*s)[W] = reinterpret_cast<R(*)[W]>(smem);
  load_tile<R, W, HH>(s, in, b, d, dx, x0 - H, y0 - H, tile_interior<TX, TY>(d, x0, y0, H));
  __syncthreads();
  ch_tile_body<R, TX, TY>(smem, out, d, dx, eps, dtM, x0, y0);
}

// mu ring and update of a Cahn-Hilliard tile whose (TY + 4) x (TX + 4) values
// of c are in LDS (the mu ring follows them there)
template <typename R, int TX, int TY>
__device__ __forceinline__ void ch_tile_body(R* smem, R* __restrict__ out, const Dom& d, R dx, R eps, R dtM,
                                             int64_t x0, int64_t y0) {
  constexpr int H = 2, W = TX + 2 * H, HH = TY + 2 * H, RP = kThreads / TX;
  constexpr int MW = TX + 2, MH = TY + 2;
  R(*s)[W] = reinterpret_cast<R(*)[W]>(smem);
  R(*mu)[MW] = reinterpret_cast<R(*)[MW]>(smem + W * HH);
  const R idx = R(1) / dx;
  const R idx2 = idx * idx;
  const R eps2 = eps * eps;
  // mu on the tile + 1 ring: mu[my][mx] at padded (x0-1+mx, y0-1+my)
  for (int t = threadIdx.x; t < MW * MH; t += kThreads) {
    int my = t / MW, mx = t - my * MW;
    int sy = my + 1, sx = mx + 1;
    R c = s[sy][sx];
    R lap = (s[sy + 1][sx] + s[sy - 1][sx] - R(2) * c) * idx2;
    lap += (s[sy][sx + 1] + s[sy][sx - 1] - R(2) * c) * idx2;
    mu[my][mx] = ch_mu(c, lap, eps2);
  }
  __syncthreads();
  // non-periodic sides: the ring outside the domain carries zero flux
  if (d.ax.bc != PPF_PERIODIC || d.ay.bc != PPF_PERIODIC) {
    for (int t = threadIdx.x; t < MW * MH; t += kThreads) {
      int my = t / MW, mx = t - my * MW;
      int64_t x = x0 - 1 + mx, y = y0 - 1 + my;
      bool xin = x >= 1 && x <= d.ax.n, yin = y >= 1 && y <= d.ay.n;
      if (xin == yin) continue;  // interior, or a corner (never read)
      R v = mu[my][mx];
      if (!xin && d.ax.bc != PPF_PERIODIC) v = mu[my][x < 1 ? mx + 1 : mx - 1];
      if (!yin && d.ay.bc != PPF_PERIODIC) v = mu[y < 1 ? my + 1 : my - 1][mx];
      // the cell read is always interior (never rewritten here): no hazard
      mu[my][mx] = v;
    }
  }
  __syncthreads();
  int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
  int64_t x = x0 + tx;
  if (x > d.ax.n) return;
#pragma unroll
  for (int k = 0; k < TY / RP; ++k) {
    int ly = ty + RP * k;
    int64_t y = y0 + ly;
    if (y > d.ay.n) break;
    R m = mu[ly + 1][tx + 1];
    // CahnHilliard.py:142-146 order: sum over axes of (mu[-1] + mu[+1] - 2mu), then /dx^2
    R dm = (mu[ly][tx + 1] + mu[ly + 2][tx + 1] - R(2) * m);
    dm = dm + (mu[ly + 1][tx] + mu[ly + 1][tx + 2] - R(2) * m);
    dm = dm / (dx * dx);
    out[y * d.ay.stride + x] = s[ly + 2][tx + 2] + dtM * dm;
  }
}

template <typename R>
__global__ __launch_bounds__(kThreads) void k_ch2d(const R* __restrict__ in, R* __restrict__ out,
                                                   const R* __restrict__ b, Dom d, R dx, R eps, R dtM) {
  __shared__ R s[(kTY + 4) * (kTX + 4) + (kTY + 2) * (kTX + 2)];
  int64_t x0, y0;
  tile_coords(d, &x0, &y0);
  ch_tile<R, kTX, kTY>(s, in, out, b, d, dx, eps, dtM, x0, y0);
}

// generic 1D / 3D version: mu at the centre and its 2*ndim neighbours
template <typename R>
__device__ __forceinline__ R ch_mu_at(const R* __restrict__ in, const R* __restrict__ b, const Dom& d, int64_t x,
                                      int64_t y, int64_t z, R dx, R idx2, R eps2) {
  R c = fold(in, b, d, x, y, z, dx);
  R lap;
  if (d.ndim == 3) {
    lap = (fold(in, b, d, x, y, z + 1, dx) + fold(in, b, d, x, y, z - 1, dx) - R(2) * c) * idx2;
    lap += (fold(in, b, d, x, y + 1, z, dx) + fold(in, b, d, x, y - 1, z, dx) - R(2) * c) * idx2;
    lap += (fold(in, b, d, x + 1, y, z, dx) + fold(in, b, d, x - 1, y, z, dx) - R(2) * c) * idx2;
  } else {
    lap = (fold(in, b, d, x + 1, y, z, dx) + fold(in, b, d, x - 1, y, z, dx) - R(2) * c) * idx2;
  }
  return ch_mu(c, lap, eps2);
}

// mu at a neighbour position: zero-flux ring on non-periodic sides
template <typename R>
__device__ __forceinline__ R ch_mu_nb(const R* __restrict__ in, const R* __restrict__ b, const Dom& d, int64_t x,
                                      int64_t y, int64_t z, int64_t cx, int64_t cy, int64_t cz, R dx, R idx2,
                                      R eps2) {
  bool out_x = x < 1 || x > d.ax.n;
  bool out_y = d.ndim >= 2 && (y < 1 || y > d.ay.n);
  bool out_z = d.ndim >= 3 && (z < 1 || z > d.az.n);
  if ((out_x && d.ax.bc != PPF_PERIODIC) || (out_y && d.ay.bc != PPF_PERIODIC) ||
      (out_z && d.az.bc != PPF_PERIODIC))
    return ch_mu_at(in, b, d, cx, cy, cz, dx, idx2, eps2);
  return ch_mu_at(in, b, d, x, y, z, dx, idx2, eps2);
}

template <typename R>
__global__ __launch_bounds__(256) void k_ch_nd(const R* __restrict__ in, R* __restrict__ out,
                                               const R* __restrict__ b, Dom d, R dx, R eps, R dtM) {
  int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  int64_t nx = d.ax.n, ny = d.ndim >= 2 ? d.ay.n : 1, nz = d.ndim >= 3 ? d.az.n : 1;
  if (t >= nx * ny * nz) return;
  int64_t x = 1 + t % nx;
  int64_t y = d.ndim >= 2 ? 1 + (t / nx) % ny : 0;
  int64_t z = d.ndim >= 3 ? 1 + t / (nx * ny) : 0;
  const R idx = R(1) / dx;
  const R idx2 = idx * idx;
  const R eps2 = eps * eps;
  R m = ch_mu_at(in, b, d, x, y, z, dx, idx2, eps2);
  R dm;
  if (d.ndim == 3) {
    dm = ch_mu_nb(in, b, d, x, y, z - 1, x, y, z, dx, idx2, eps2) +
         ch_mu_nb(in, b, d, x, y, z + 1, x, y, z, dx, idx2, eps2) - R(2) * m;
    dm = dm + (ch_mu_nb(in, b, d, x, y - 1, z, x, y, z, dx, idx2, eps2) +
               ch_mu_nb(in, b, d, x, y + 1, z, x, y, z, dx, idx2, eps2) - R(2) * m);
    dm = dm + (ch_mu_nb(in, b, d, x - 1, y, z, x, y, z, dx, idx2, eps2) +
               ch_mu_nb(in, b, d, x + 1, y, z, x, y, z, dx, idx2, eps2) - R(2) * m);
  } else {
    dm = ch_mu_nb(in, b, d, x - 1, y, z, x, y, z, dx, idx2, eps2) +
         ch_mu_nb(in, b, d, x + 1, y, z, x, y, z, dx, idx2, eps2) - R(2) * m;
  }
  dm = dm / (dx * dx);
  int64_t o = offset(d, x, y, z);
  out[o] = in[o] + dtM * dm;
}

// 3D Cahn-Hilliard, marching in z.  A block owns a C3X x C3Y column of
// outputs and a chunk of planes; planes of c (2-cell xy halo) and of mu (1-cell
// xy ring) rotate through LDS rings, so every c value is read from memory once
// per block (plus the xy halo and a few extra planes per chunk).  Iteration k
// of the march: store c(k+1) (fetched into registers during iteration k-2;
// halo cells outside the domain were written to LDS directly) and start
// fetching c(k+3) into the same registers, so two planes are in flight;
// barrier; mu(k) on the ring from c(k-1..k+1); barrier; output plane k-1 from
// mu(k-2..k) and c(k-1).  Six c slots and four mu slots keep every slot
// written in an iteration clear of every slot read in it or the one before,
// so two barriers per plane suffice.  Zero flux on
// non-periodic sides: a mu neighbour outside the domain is replaced by the
// centre value (as k_ch_nd and the oracle).
#ifndef PPF_CH3D_TY
#define PPF_CH3D_TY 8
#endif
constexpr int C3X = 64;
template <int C3Y>
struct Ch3dShape {
  static constexpr int W = C3X + 4, H = C3Y + 4, N = W * H;      // c plane, 2-cell halo
  static constexpr int MW = C3X + 2, MH = C3Y + 2, MN = MW * MH;  // mu plane, tile + 1 ring
  static constexpr int LOADS = (N + kThreads - 1) / kThreads;
  static constexpr int MUS = (MN + kThreads - 1) / kThreads;
  static constexpr size_t lds(size_t r) { return r * (6 * N + 4 * MN); }
};

// Per-thread constants of the march: the c-plane elements this thread
// fetches (in-plane offset, whether the (x, y) position is inside the domain)
// and the mu ring positions it evaluates.
template <int C3Y>
struct Ch3dLanes {
  int32_t off[Ch3dShape<C3Y>::LOADS];  // y * sy + x of fetched element j
  uint32_t xy_in;        // bit j: element j inside the domain in x and y
  uint32_t valid;        // bit j: element j exists (t < C3N)
};

// the halo cells of this thread's share of c plane z that lie outside the
// domain (bit j of `mask` clear), through the boundary rules straight into
// LDS
template <typename R, int C3Y>
__device__ __forceinline__ void ch3d_fold_halo(R* __restrict__ slot, const R* __restrict__ in,
                                            const R* __restrict__ b, const Dom& d, R dx, int64_t x0,
                                            int64_t y0, int64_t z, uint32_t mask) {
  constexpr int C3LOADS = Ch3dShape<C3Y>::LOADS, C3N = Ch3dShape<C3Y>::N, C3W = Ch3dShape<C3Y>::W;
#pragma unroll 1
  for (int j = 0; j < C3LOADS; ++j) {
    const int t = threadIdx.x + j * kThreads;
    if (t >= C3N || (mask >> j) & 1u) continue;
    const int ly = t / C3W, lx = t - ly * C3W;
    const int64_t x = x0 - 2 + lx, y = y0 - 2 + ly;
    // beyond a partial tile: never used
    slot[t] = (x > d.ax.n + 2 || y > d.ay.n + 2) ? R(0) : fold(in, b, d, x, y, z, dx);
  }
}

// this thread's share of c plane z: elements inside the domain into
// registers (bit j of the returned mask), the halo cells outside it straight
// into the plane's LDS slot through the boundary rules (one rolled loop, so
// the fold code is not replicated)
template <typename R, int C3Y>
__device__ __forceinline__ uint32_t ch3d_fetch(R (&v)[Ch3dShape<C3Y>::LOADS], R* __restrict__ slot,
                                               const R* __restrict__ in, const R* __restrict__ b, const Dom& d,
                                               R dx, int64_t x0, int64_t y0, int64_t z,
                                               const Ch3dLanes<C3Y>& L) {
  constexpr int C3LOADS = Ch3dShape<C3Y>::LOADS, C3N = Ch3dShape<C3Y>::N, C3W = Ch3dShape<C3Y>::W;
  const int64_t zw = wrap_periodic(z, d.az);
  const uint32_t mask = (zw >= 1 && zw <= d.az.n) ? L.xy_in : 0u;
  const R* plane = in + zw * d.az.stride;
#pragma unroll
  for (int j = 0; j < C3LOADS; ++j)
    if ((mask >> j) & 1u) v[j] = plane[L.off[j]];
  if (mask != L.valid) ch3d_fold_halo<R, C3Y>(slot, in, b, d, dx, x0, y0, z, mask);
  return mask;
}

template <typename R, int C3Y>
__device__ __forceinline__ void ch3d_store(R* __restrict__ dst, const R (&v)[Ch3dShape<C3Y>::LOADS],
                                           uint32_t mask) {
  constexpr int C3LOADS = Ch3dShape<C3Y>::LOADS;
#pragma unroll
  for (int j = 0; j < C3LOADS; ++j) {
    const int t = threadIdx.x + j * kThreads;
    if ((mask >> j) & 1u) dst[t] = v[j];
  }
}

// (sum of mu differences) / dx^2: FP64 keeps the reference's division, FP32
// multiplies by 1/dx^2
template <typename R>
__device__ __forceinline__ R ch3d_scale(R dm, R dx, R idx2) {
  if constexpr (sizeof(R) == 8) {
    return dm / (dx * dx);
  } else {
    return dm * idx2;
  }
}

template <typename R, int C3Y>
__global__ __launch_bounds__(kThreads) void k_ch3d(const R* __restrict__ in, R* __restrict__ out,
                                                   const R* __restrict__ b, Dom d, R dx, R eps, R dtM,
                                                   int32_t zchunk) {
  using S = Ch3dShape<C3Y>;
  constexpr int C3W = S::W, C3N = S::N, C3MW = S::MW, C3MN = S::MN, C3LOADS = S::LOADS, C3MUS = S::MUS;
  extern __shared__ __align__(16) unsigned char ch3d_lds[];
  R(*sc)[C3N] = reinterpret_cast<R(*)[C3N]>(ch3d_lds);
  R(*sm)[C3MN] = reinterpret_cast<R(*)[C3MN]>(ch3d_lds + 6 * C3N * sizeof(R));
  const uint32_t nbx = uint32_t((d.ax.n + C3X - 1) / C3X), nby = uint32_t((d.ay.n + C3Y - 1) / C3Y);
  const uint32_t blk = xcd_swizzle(blockIdx.x, gridDim.x);
  const uint32_t bxy = blk % (nbx * nby), bz = blk / (nbx * nby);
  const int64_t x0 = 1 + int64_t(bxy % nbx) * C3X, y0 = 1 + int64_t(bxy / nbx) * C3Y;
  const int64_t zb = 1 + int64_t(bz) * zchunk;
  const int64_t ze = zb + zchunk - 1 < d.az.n ? zb + zchunk - 1 : d.az.n;
  const R idx = R(1) / dx;
  const R idx2 = idx * idx;
  const R eps2 = eps * eps;
  const bool zflux = d.az.bc != PPF_PERIODIC, xflux = d.ax.bc != PPF_PERIODIC, yflux = d.ay.bc != PPF_PERIODIC;
  Ch3dLanes<C3Y> L;
  L.xy_in = 0;
  L.valid = 0;
#pragma unroll
  for (int j = 0; j < C3LOADS; ++j) {
    const int t = threadIdx.x + j * kThreads;
    const int ly = t / C3W, lx = t - ly * C3W;
    // periodic axes wrap here (as resolve()), so only non-periodic halo
    // cells take the fold path
    const int64_t x = wrap_periodic(x0 - 2 + lx, d.ax), y = wrap_periodic(y0 - 2 + ly, d.ay);
    const bool in_xy = t < C3N && x >= 1 && x <= d.ax.n && y >= 1 && y <= d.ay.n;
    L.off[j] = in_xy ? int32_t(y * d.ay.stride + x) : 0;
    L.xy_in |= in_xy ? 1u << j : 0u;
    L.valid |= t < C3N ? 1u << j : 0u;
  }
  int32_t mu_c[C3MUS];  // c-plane index of each mu ring position of this thread (-1: none)
#pragma unroll
  for (int j = 0; j < C3MUS; ++j) {
    const int t = threadIdx.x + j * kThreads;
    const int my = t / C3MW, mx = t - my * C3MW;
    mu_c[j] = t < C3MN ? (my + 1) * C3W + (mx + 1) : -1;
  }
  // LDS ring positions (rotated once per plane; plain 32-bit scalars)
  int c_mm = 5, c_m = 0, c_0 = 1, c_p = 2, c_n = 3, c_nn = 4;  // c(k-2) .. c(k+3)
  int m_mm = 0, m_m = 1, m_0 = 2, m_f = 3;                      // mu(k-2), mu(k-1), mu(k), free
  const int32_t zb32 = int32_t(zb), ze32 = int32_t(ze), nz32 = int32_t(d.az.n);
  // two register sets alternate between iterations, so two planes are in
  // flight from memory while one is computed
  R va[C3LOADS], vb[C3LOADS];
  // c(zb-2) -> c_m, c(zb-1) -> c_0; c(zb) (set a) and c(zb+1) (set b) in flight
  uint32_t ma = ch3d_fetch<R, C3Y>(va, sc[c_m], in, b, d, dx, x0, y0, zb - 2, L);
  ch3d_store<R, C3Y>(sc[c_m], va, ma);
  ma = ch3d_fetch<R, C3Y>(va, sc[c_0], in, b, d, dx, x0, y0, zb - 1, L);
  ch3d_store<R, C3Y>(sc[c_0], va, ma);
  ma = ch3d_fetch<R, C3Y>(va, sc[c_p], in, b, d, dx, x0, y0, zb, L);
  uint32_t mb = ch3d_fetch<R, C3Y>(vb, sc[c_n], in, b, d, dx, x0, y0, zb + 1, L);
  const int tx = threadIdx.x % C3X, ty = threadIdx.x / C3X;
  const int64_t x = x0 + tx;
  const bool x_in = x <= d.ax.n;
  const bool xlo = xflux && x == 1, xhi = xflux && x == d.ax.n;
  const bool any_flux = xflux || yflux || zflux;
  const int64_t orow = (y0 + ty) * d.ay.stride + x;  // in-plane offset of this thread's first output
  auto iteration = [&](int32_t k, R(&v)[C3LOADS], uint32_t& mask) {
    // iteration k: mu(k) and output plane k - 1.  c(k+1) -> c_p from
    // registers (fetched two iterations ago); c(k+3) is fetched into the same
    // registers (its halo cells straight into slot c_nn, which last held
    // c(k-3), read for the last time in iteration k-2)
    ch3d_store<R, C3Y>(sc[c_p], v, mask);
    if (k + 3 <= ze32 + 2) mask = ch3d_fetch<R, C3Y>(v, sc[c_nn], in, b, d, dx, x0, y0, int64_t(k) + 3, L);
    __syncthreads();
    // mu(k) into slot m_0 on the tile + 1 ring (corners are never read)
    {
      const R* cm = sc[c_m];
      const R* c0 = sc[c_0];
      const R* cp = sc[c_p];
      R* mu = sm[m_0];
#pragma unroll
      for (int j = 0; j < C3MUS; ++j) {
        const int i = mu_c[j];
        if (i < 0) break;
        const R c = c0[i];
        // axis 0 (z), then y, then x
        R lap = (cp[i] + cm[i] - R(2) * c) * idx2;
        lap += (c0[i + C3W] + c0[i - C3W] - R(2) * c) * idx2;
        lap += (c0[i + 1] + c0[i - 1] - R(2) * c) * idx2;
        mu[threadIdx.x + j * kThreads] = ch_mu(c, lap, eps2);
      }
    }
    __syncthreads();
    const int32_t z = k - 1;  // output plane, from mu(k-2) = m_mm, mu(k-1) = m_m, mu(k) = m_0, c(k-1) = c_m
    if (z >= zb32 && x_in) {
      const R* mm = sm[m_mm];
      const R* m0 = sm[m_m];
      const R* mp = sm[m_0];
      const R* c0 = sc[c_m];
      R* oz = out + int64_t(z) * d.az.stride + orow;
      if (!any_flux) {  // periodic on every axis (uniform branch): no boundary selects
#pragma unroll
        for (int r = 0; r < C3Y / 4; ++r) {
          const int ly = ty + 4 * r;
          if (y0 + ly > d.ay.n) break;
          const int i = (ly + 1) * C3MW + (tx + 1);
          const R m = m0[i];
          R dm = mm[i] + mp[i] - R(2) * m;
          dm = dm + (m0[i - C3MW] + m0[i + C3MW] - R(2) * m);
          dm = dm + (m0[i - 1] + m0[i + 1] - R(2) * m);
          dm = ch3d_scale(dm, dx, idx2);
          oz[int64_t(4 * r) * d.ay.stride] = c0[(ly + 2) * C3W + (tx + 2)] + dtM * dm;
        }
      } else {
        const bool zlo = zflux && z == 1, zhi = zflux && z == nz32;
#pragma unroll
        for (int r = 0; r < C3Y / 4; ++r) {
          const int ly = ty + 4 * r;
          const int64_t y = y0 + ly;
          if (y > d.ay.n) break;
          const int i = (ly + 1) * C3MW + (tx + 1);
          const R m = m0[i];
          const R mzm = zlo ? m : mm[i];
          const R mzp = zhi ? m : mp[i];
          const R mym = (yflux && y == 1) ? m : m0[i - C3MW];
          const R myp = (yflux && y == d.ay.n) ? m : m0[i + C3MW];
          const R mxm = xlo ? m : m0[i - 1];
          const R mxp = xhi ? m : m0[i + 1];
          R dm = mzm + mzp - R(2) * m;
          dm = dm + (mym + myp - R(2) * m);
          dm = dm + (mxm + mxp - R(2) * m);
          dm = ch3d_scale(dm, dx, idx2);
          oz[int64_t(4 * r) * d.ay.stride] = c0[(ly + 2) * C3W + (tx + 2)] + dtM * dm;
        }
      }
    }
    // rotate: c(k-1) becomes c(k-2) of the next iteration, ...
    const int c_old = c_mm;
    c_mm = c_m;
    c_m = c_0;
    c_0 = c_p;
    c_p = c_n;
    c_n = c_nn;
    c_nn = c_old;
    const int m_old = m_mm;
    m_mm = m_m;
    m_m = m_0;
    m_0 = m_f;
    m_f = m_old;
  };
  for (int32_t k = zb32 - 1; k <= ze32 + 1; k += 2) {
    iteration(k, va, ma);
    if (k + 1 <= ze32 + 1) iteration(k + 1, vb, mb);
  }
}

// Full-wave lane shifts (DPP wave_shr:1 / wave_shl:1): lane l receives the
// value of lane l-1 (resp. l+1); the end lane receives 0.  All 64 lanes must be
// active where these are used.
constexpr int kDppWaveShr1 = 0x138;
constexpr int kDppWaveShl1 = 0x130;
template <int CTRL, typename R>
__device__ __forceinline__ R lane_shift(R v) {
  if constexpr (sizeof(R) == 4) {
    return __builtin_bit_cast(R, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), CTRL, 0xF, 0xF, true));
  } else {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = uint32_t(__builtin_amdgcn_mov_dpp(int(uint32_t(u)), CTRL, 0xF, 0xF, true));
    const uint32_t hi = uint32_t(__builtin_amdgcn_mov_dpp(int(uint32_t(u >> 32)), CTRL, 0xF, 0xF, true));
    return __builtin_bit_cast(R, (uint64_t(hi) << 32) | lo);
  }
}

// 3D Cahn-Hilliard on a box whose halos are a coordinate remap (every side
// periodic, or zero-gradient Neumann without bc values), one wave per block, marching in z
// with the cells in registers.  Lane l owns column x = x0 - 2 + l of a 64-wide
// window (60 outputs and a 2-cell halo on each side) and the RPW rows
// y0 - 2 .. y0 + RPW - 3 (RPW - 4 output rows).  x neighbours come from the
// adjacent lanes (lane shifts), y neighbours from the lane's own rows, z
// neighbours from the register sets of the neighbouring planes, so the march
// reads no LDS and needs no barrier; halo coordinates are remapped once into
// the row offsets and the lane's column (the zero-flux mu ring of Neumann
// sides: lane shifts, row and plane copies), and planes are read and written
// with buffer instructions (plane base in the descriptor, row offset scalar,
// lane offset in one VGPR).  Iteration k: mu(k) on rows 1..RPW-2 from
// c(k-1..k+1); output plane k-1 on rows 2..RPW-3 from mu(k-2..k) and c(k-1);
// then c(k+3) is fetched into the registers c(k-1) held.  Four c sets (one
// plane always in flight) and three mu sets rotate by name: the march is
// unrolled twelve times so every set index is a compile-time constant.  FP64
// keeps k_ch3d's operation order; FP32 sums the six neighbours first and
// evaluates the well in Horner form.
constexpr int kCh3wTx = 60;
// planes fetched ahead of c(k+1) (each costs RPW registers)
#ifndef PPF_CH3W_AHEAD
#define PPF_CH3W_AHEAD 1
#endif

template <typename R>
__device__ __forceinline__ R buf_load(__amdgpu_buffer_rsrc_t rs, uint32_t voff, int32_t soff) {
  if constexpr (sizeof(R) == 4) {
    return __builtin_bit_cast(R, __builtin_amdgcn_raw_buffer_load_b32(rs, voff, soff, 0));
  } else {
    return __builtin_bit_cast(R, __builtin_amdgcn_raw_buffer_load_b64(rs, voff, soff, 0));
  }
}

template <typename R>
__device__ __forceinline__ void buf_store(R v, __amdgpu_buffer_rsrc_t rs, uint32_t voff, int32_t soff) {
  if constexpr (sizeof(R) == 4) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), rs, voff, soff, 0);
  } else {
    typedef uint32_t u32x2 __attribute__((__vector_size__(8)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), rs, voff, soff, 0);
  }
}

// raw buffer descriptor over [p, p + 4 GiB): dword 3 = 0x00020000 selects the
// 32-bit data format the untyped buffer instructions need on CDNA
__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(const void* p) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), int16_t(0), int32_t(-1), 0x00020000);
}

// Interior coordinate a halo coordinate reads: periodic wrap, or the nearest
// interior cell (zero gradient without bc values)
__device__ __forceinline__ int64_t remap_halo(int64_t p, const Axis& a) {
  if (a.bc == PPF_PERIODIC) return wrap_periodic(p, a);
  return p < 1 ? 1 : (p > a.n ? a.n : p);
}

template <typename R, int RPW>
__device__ __forceinline__ void ch3w_fetch(R (&v)[RPW], const R* __restrict__ in, const Dom& d,
                                           const int32_t (&yoff)[RPW], uint32_t xoff, int64_t z) {
  const __amdgpu_buffer_rsrc_t rs = plane_rsrc(in + remap_halo(z, d.az) * d.az.stride);
#pragma unroll
  for (int r = 0; r < RPW; ++r) v[r] = buf_load<R>(rs, xoff, yoff[r]);
}

// f(0), f(1), ... in order while they return true (compile-time phases)
template <typename F, int... P>
__device__ __forceinline__ bool ch3w_unrolled(F& f, std::integer_sequence<int, P...>) {
  return (f(std::integral_constant<int, P>{}) && ...);
}

template <typename R, int RPW>
__global__ __launch_bounds__(64) void k_ch3w(const R* __restrict__ in, R* __restrict__ out, Dom d, R dx,
                                             R eps, R dtM, int32_t zchunk) {
  constexpr int TYO = RPW - 4;
  const uint32_t nbx = uint32_t((d.ax.n + kCh3wTx - 1) / kCh3wTx), nby = uint32_t((d.ay.n + TYO - 1) / TYO);
  const uint32_t blk = xcd_swizzle(blockIdx.x, gridDim.x);
  const uint32_t bxy = blk % (nbx * nby), bz = blk / (nbx * nby);
  const int64_t x0 = 1 + int64_t(bxy % nbx) * kCh3wTx, y0 = 1 + int64_t(bxy / nbx) * TYO;
  const int64_t zb = 1 + int64_t(bz) * zchunk;
  const int64_t ze = zb + zchunk - 1 < d.az.n ? zb + zchunk - 1 : d.az.n;
  const R idx = R(1) / dx;
  const R idx2 = idx * idx;
  const R eps2 = eps * eps;
  const R k_mu = -eps2 * idx2;  // FP32 form: mu = well(c) + k_mu * (sum of neighbours - 6c)
  const R k_out = dtM * idx2;   // FP32 form: out = c + k_out * (sum of mu neighbours - 6 mu)
  const int lane = threadIdx.x;
  const int64_t xr = x0 - 2 + lane;
  const uint32_t xoff = uint32_t(remap_halo(xr, d.ax)) * uint32_t(sizeof(R));
  const uint32_t ooff = uint32_t(xr) * uint32_t(sizeof(R));  // used on output lanes only (xr >= 1 there)
  int32_t yoff[RPW];  // byte offset of row r in its plane (remapped)
#pragma unroll
  for (int r = 0; r < RPW; ++r) yoff[r] = int32_t(remap_halo(y0 - 2 + r, d.ay) * d.ay.stride * int64_t(sizeof(R)));
  // zero-flux sides: mu ring cells outside the domain take their interior
  // neighbour's mu (lanes xr = 0 / n + 1, rows y = 0 / n + 1, planes 0 / n + 1)
  const bool zfx = d.ax.bc != PPF_PERIODIC, zfy = d.ay.bc != PPF_PERIODIC, zfz = d.az.bc != PPF_PERIODIC;
  const bool xlo = xr < 1, xhi = xr > d.ax.n;
  const int32_t yrow0 = int32_t(y0 - 2);  // y of row 0
  const int32_t nyp1 = int32_t(d.ay.n + 1), nzp1 = int32_t(d.az.n + 1);
  const bool out_lane = lane >= 2 && lane < 2 + kCh3wTx && xr <= d.ax.n;
  const int32_t zb32 = int32_t(zb), ze32 = int32_t(ze);
  const int32_t ny_out = int32_t(d.ay.n - (y0 - 2));  // rows r <= ny_out are inside the domain
  const int32_t orow_bytes = int32_t(d.ay.stride * int64_t(sizeof(R)));
  constexpr int NC = 2 + PPF_CH3W_AHEAD;  // c sets: c(k-1), c(k), then AHEAD planes
  R C[NC + 1][RPW];  // c(k) lives in set (k - (zb - 1)) mod (NC + 1)
  R M[3][RPW];       // mu(k) lives in set (k - (zb - 1)) mod 3
  ch3w_fetch<R, RPW>(C[NC], in, d, yoff, xoff, zb - 2);
#pragma unroll
  for (int j = 0; j < NC; ++j) ch3w_fetch<R, RPW>(C[j], in, d, yoff, xoff, zb - 1 + j);
  auto iteration = [&](auto phase, int32_t k) {
    constexpr int p = decltype(phase)::value;
    __builtin_amdgcn_sched_barrier(0);  // one plane at a time (register pressure)
    R(&cm)[RPW] = C[(p + NC) % (NC + 1)];
    R(&c0)[RPW] = C[p % (NC + 1)];
    R(&cp)[RPW] = C[(p + 1) % (NC + 1)];
    R(&mm)[RPW] = M[(p + 1) % 3];
    R(&m0)[RPW] = M[(p + 2) % 3];
    R(&mp)[RPW] = M[p % 3];
    // mu(k); axis 0 (z), then y, then x
#pragma unroll
    for (int r = 1; r < RPW - 1; ++r) {
      const R c = c0[r];
      const R cxp = lane_shift<kDppWaveShl1>(c), cxm = lane_shift<kDppWaveShr1>(c);
      if constexpr (sizeof(R) == 8) {
        R lap = (cp[r] + cm[r] - R(2) * c) * idx2;
        lap += (c0[r + 1] + c0[r - 1] - R(2) * c) * idx2;
        lap += (cxp + cxm - R(2) * c) * idx2;
        mp[r] = ch_mu(c, lap, eps2);
      } else {
        const R sum = ((cp[r] + cm[r]) + (c0[r + 1] + c0[r - 1])) + (cxp + cxm);
        const R well = c * __builtin_fmaf(c, __builtin_fmaf(c, R(4), R(-6)), R(2));
        mp[r] = __builtin_fmaf(k_mu, __builtin_fmaf(R(-6), c, sum), well);
      }
    }
    if (zfx) {
#pragma unroll
      for (int r = 1; r < RPW - 1; ++r) {
        const R from_right = lane_shift<kDppWaveShl1>(mp[r]), from_left = lane_shift<kDppWaveShr1>(mp[r]);
        mp[r] = xlo ? from_right : (xhi ? from_left : mp[r]);
      }
    }
    if (zfy) {
#pragma unroll
      for (int r = 1; r < RPW - 1; ++r) {
        if (yrow0 + r == 0) mp[r] = mp[r + 1];
        if (yrow0 + r == nyp1) mp[r] = mp[r - 1];
      }
    }
    if (zfz) {
      if (k == 1) {  // mu(0) := mu(1)
#pragma unroll
        for (int r = 1; r < RPW - 1; ++r) m0[r] = mp[r];
      }
      if (k == nzp1) {  // mu(n + 1) := mu(n)
#pragma unroll
        for (int r = 1; r < RPW - 1; ++r) mp[r] = m0[r];
      }
    }
    // output plane k - 1 from mu(k-2) = mm, mu(k-1) = m0, mu(k) = mp, c(k-1) = cm
    const int32_t z = k - 1;
    if (z >= zb32) {
      const __amdgpu_buffer_rsrc_t ws = plane_rsrc(out + int64_t(z) * d.az.stride + (y0 - 2) * d.ay.stride);
      R o[RPW];
#pragma unroll
      for (int r = 2; r < RPW - 2; ++r) {
        const R m = m0[r];
        const R mxm = lane_shift<kDppWaveShr1>(m), mxp = lane_shift<kDppWaveShl1>(m);
        if constexpr (sizeof(R) == 8) {
          R dm = mm[r] + mp[r] - R(2) * m;
          dm = dm + (m0[r - 1] + m0[r + 1] - R(2) * m);
          dm = dm + (mxm + mxp - R(2) * m);
          o[r] = cm[r] + dtM * ch3d_scale(dm, dx, idx2);
        } else {
          const R sum = ((mm[r] + mp[r]) + (m0[r - 1] + m0[r + 1])) + (mxm + mxp);
          o[r] = __builtin_fmaf(k_out, __builtin_fmaf(R(-6), m, sum), cm[r]);
        }
      }
      // one divergent region for the whole plane's stores
      if (out_lane) {
#pragma unroll
        for (int r = 2; r < RPW - 2; ++r)
          if (r <= ny_out) buf_store<R>(o[r], ws, ooff, r * orow_bytes);
      }
    }
    // unconditional (the last iterations re-read plane ze + 2, whose values
    // are never used): a conditional fetch makes the compiler wait for it
    __builtin_amdgcn_sched_barrier(0);
    ch3w_fetch<R, RPW>(cm, in, d, yoff, xoff, int64_t(k + NC <= ze32 + 2 ? k + NC : ze32 + 2));
  };
  // the march, unrolled over one period of both rings
  constexpr int kPeriod = (NC + 1) % 3 == 0 ? NC + 1 : 3 * (NC + 1);
  for (int32_t k = zb32 - 1;; k += kPeriod) {
    auto step = [&](auto phase) {
      constexpr int q = decltype(phase)::value;
      if (k + q > ze32 + 1) return false;
      iteration(phase, k + q);
      return true;
    };
    if (!ch3w_unrolled(step, std::make_integer_sequence<int, kPeriod>{})) break;
  }
}

// Two explicit Cahn-Hilliard steps per pass of the register march (FP32,
// every side periodic).  Each wave owns 56 x (RPW - 8) output cells of a z
// chunk: a 4-cell ring of x and y halo (2 per step) in its lanes and rows.
// Plane k of the march computes mu(k), then c1(k-1) (the first step), then
// mu2(k-2) from c1, then c2(k-3) (the second step), which is written.  The
// halo cells are periodic images, which evolve exactly as the cells they
// copy, and every value is the same FP32 expression as in k_ch3w, so the
// result equals two k_ch3w launches bit for bit.
constexpr int kCh3w2Tx = 56;
#ifndef PPF_CH3W2_AHEAD
#define PPF_CH3W2_AHEAD 2  // planes of c fetched ahead of the march
#endif
template <int RPW>
__global__ __launch_bounds__(64) void k_ch3w2(const float* __restrict__ in, float* __restrict__ out, Dom d,
                                              float dx, float eps, float dtM, int32_t zchunk) {
  using R = float;
  constexpr int TYO = RPW - 8;
  const R idx = R(1) / dx;
  const R idx2 = idx * idx;
  const R eps2 = eps * eps;
  const R k_mu = -eps2 * idx2;  // as k_ch3w
  const R k_out = dtM * idx2;
  const uint32_t nbx = uint32_t((d.ax.n + kCh3w2Tx - 1) / kCh3w2Tx), nby = uint32_t((d.ay.n + TYO - 1) / TYO);
  const uint32_t blk = xcd_swizzle(blockIdx.x, gridDim.x);
  const uint32_t bxy = blk % (nbx * nby), bz = blk / (nbx * nby);
  const int64_t x0 = 1 + int64_t(bxy % nbx) * kCh3w2Tx, y0 = 1 + int64_t(bxy / nbx) * TYO;
  const int64_t zb = 1 + int64_t(bz) * zchunk;
  const int64_t ze = zb + zchunk - 1 < d.az.n ? zb + zchunk - 1 : d.az.n;
  const int lane = threadIdx.x;
  const int64_t xr = x0 - 4 + lane;
  const uint32_t xoff = uint32_t(remap_halo(xr, d.ax)) * uint32_t(sizeof(R));
  const uint32_t ooff = uint32_t(xr) * uint32_t(sizeof(R));  // used on output lanes only (xr >= 1 there)
  int32_t yoff[RPW];  // byte offset of row r in its plane (periodic image)
#pragma unroll
  for (int r = 0; r < RPW; ++r) yoff[r] = int32_t(remap_halo(y0 - 4 + r, d.ay) * d.ay.stride * int64_t(sizeof(R)));
  const bool out_lane = lane >= 4 && lane < 4 + kCh3w2Tx && xr <= d.ax.n;
  const int32_t zb32 = int32_t(zb), ze32 = int32_t(ze), k0 = zb32 - 3;
  const int32_t ny_out = int32_t(d.ay.n - (y0 - 4));  // rows r <= ny_out are inside the domain
  const int32_t orow_bytes = int32_t(d.ay.stride * int64_t(sizeof(R)));
  constexpr int NC = 2 + PPF_CH3W2_AHEAD;  // c sets: c(k-1), c(k), then AHEAD planes
  R C[NC + 1][RPW];  // c(k) lives in set (k - k0) mod (NC + 1)
  R M[3][RPW];       // mu(k): set (k - k0) mod 3, rows 1 .. RPW - 2
  R C1[3][RPW];      // c1(k): set (k - k0) mod 3, rows 2 .. RPW - 3
  R M2[3][RPW];      // mu2(k): set (k - k0) mod 3, rows 3 .. RPW - 4
  ch3w_fetch<R, RPW>(C[NC], in, d, yoff, xoff, k0 - 1);
#pragma unroll
  for (int j = 0; j < NC; ++j) ch3w_fetch<R, RPW>(C[j], in, d, yoff, xoff, k0 + j);
  // mu on rows lo .. hi - 1 of plane c0 (neighbour planes cm, cp)
  auto mu_rows = [&](R(&dst)[RPW], const R(&cm)[RPW], const R(&c0)[RPW], const R(&cp)[RPW], auto lo_c,
                     auto hi_c) {
    constexpr int lo = decltype(lo_c)::value, hi = decltype(hi_c)::value;
#pragma unroll
    for (int r = lo; r < hi; ++r) {
      const R c = c0[r];
      const R cxp = lane_shift<kDppWaveShl1>(c), cxm = lane_shift<kDppWaveShr1>(c);
      const R sum = ((cp[r] + cm[r]) + (c0[r + 1] + c0[r - 1])) + (cxp + cxm);
      const R well = c * __builtin_fmaf(c, __builtin_fmaf(c, R(4), R(-6)), R(2));
      dst[r] = __builtin_fmaf(k_mu, __builtin_fmaf(R(-6), c, sum), well);
    }
  };
  // c + dt M lap(mu) on rows lo .. hi - 1 (mu planes mm, m0, mp; c plane cc)
  auto step_rows = [&](R(&dst)[RPW], const R(&mm)[RPW], const R(&m0)[RPW], const R(&mp)[RPW],
                       const R(&cc)[RPW], auto lo_c, auto hi_c) {
    constexpr int lo = decltype(lo_c)::value, hi = decltype(hi_c)::value;
#pragma unroll
    for (int r = lo; r < hi; ++r) {
      const R m = m0[r];
      const R mxm = lane_shift<kDppWaveShr1>(m), mxp = lane_shift<kDppWaveShl1>(m);
      const R sum = ((mm[r] + mp[r]) + (m0[r - 1] + m0[r + 1])) + (mxm + mxp);
      dst[r] = __builtin_fmaf(k_out, __builtin_fmaf(R(-6), m, sum), cc[r]);
    }
  };
  auto iteration = [&](auto phase, int32_t k) {
    constexpr int p = decltype(phase)::value;
    __builtin_amdgcn_sched_barrier(0);  // one plane at a time (register pressure)
    R(&cm)[RPW] = C[(p + NC) % (NC + 1)];
    // mu(k) from c(k-1), c(k), c(k+1)
    mu_rows(M[p % 3], cm, C[p % (NC + 1)], C[(p + 1) % (NC + 1)], std::integral_constant<int, 1>{},
            std::integral_constant<int, RPW - 1>{});
    // c1(k-1) from mu(k-2), mu(k-1), mu(k) and c(k-1)
    step_rows(C1[(p + 2) % 3], M[(p + 1) % 3], M[(p + 2) % 3], M[p % 3], cm, std::integral_constant<int, 2>{},
              std::integral_constant<int, RPW - 2>{});
    // mu2(k-2) from c1(k-3), c1(k-2), c1(k-1)
    mu_rows(M2[(p + 1) % 3], C1[p % 3], C1[(p + 1) % 3], C1[(p + 2) % 3], std::integral_constant<int, 3>{},
            std::integral_constant<int, RPW - 3>{});
    // c2(k-3) from mu2(k-4), mu2(k-3), mu2(k-2) and c1(k-3): the output plane
    const int32_t z = k - 3;
    if (z >= zb32) {
      R o[RPW];
      step_rows(o, M2[(p + 2) % 3], M2[p % 3], M2[(p + 1) % 3], C1[p % 3], std::integral_constant<int, 4>{},
                std::integral_constant<int, RPW - 4>{});
      const __amdgpu_buffer_rsrc_t ws = plane_rsrc(out + int64_t(z) * d.az.stride + (y0 - 4) * d.ay.stride);
      if (out_lane) {
#pragma unroll
        for (int r = 4; r < RPW - 4; ++r)
          if (r <= ny_out) buf_store<R>(o[r], ws, ooff, r * orow_bytes);
      }
    }
    // unconditional (the last iterations re-read plane ze + 4, whose values
    // are never used): a conditional fetch makes the compiler wait for it
    __builtin_amdgcn_sched_barrier(0);
    ch3w_fetch<R, RPW>(cm, in, d, yoff, xoff, int64_t(k + NC <= ze32 + 4 ? k + NC : ze32 + 4));
  };
  constexpr int kPeriod = (NC + 1) % 3 == 0 ? NC + 1 : 3 * (NC + 1);
  for (int32_t k = k0;; k += kPeriod) {
    auto step = [&](auto phase) {
      constexpr int q = decltype(phase)::value;
      if (k + q > ze32 + 3) return false;
      iteration(phase, k + q);
      return true;
    };
    if (!ch3w_unrolled(step, std::make_integer_sequence<int, kPeriod>{})) break;
  }
}

#ifndef PPF_CH3W2_RPW
#define PPF_CH3W2_RPW 13
#endif
#ifndef PPF_CH3W2_ROUNDS
#define PPF_CH3W2_ROUNDS 4
#endif
inline int ch3w2_launch(const ppf_grid& g, const void* in, void* out, double M, double eps, double dt,
                        hipStream_t s) {
  // the step constants are formed in the kernel from these, as k_ch3w forms them
  constexpr int RPW = PPF_CH3W2_RPW;
  static int per_cu = 0;
  if (per_cu == 0) {
    int v = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&v, k_ch3w2<RPW>, 64, 0) != hipSuccess || v < 1) v = 1;
    per_cu = v;
  }
  Dom d = make_dom(g);
  const int64_t xy = ((g.n[0] + kCh3w2Tx - 1) / kCh3w2Tx) * ((g.n[1] + RPW - 9) / (RPW - 8));
  const int64_t resident = int64_t(device_cus()) * per_cu;
  const int64_t chunks = (PPF_CH3W2_ROUNDS * resident + xy - 1) / xy;
  int64_t zc = (g.n[2] + chunks - 1) / chunks;
  if (zc < 8) zc = 8;
  if (zc > g.n[2]) zc = g.n[2];
  const int64_t blocks = xy * ((g.n[2] + zc - 1) / zc);
  hipLaunchKernelGGL((k_ch3w2<RPW>), dim3(unsigned(blocks)), dim3(64), 0, s, static_cast<const float*>(in),
                     static_cast<float*>(out), d, float(g.dx), float(eps), float(dt * M), int32_t(zc));
  return check_launch("cahnhilliard_multistep");
}

#ifndef PPF_CH3D_WAVE
#define PPF_CH3D_WAVE 1
#endif
#ifndef PPF_CH3W_RPW
#define PPF_CH3W_RPW 16
#endif
// z chunks of the register march: as many blocks as PPF_CH3W_ROUNDS rounds of
// resident waves hold (so the last round is not a sliver), at least 8 planes
// per chunk
#ifndef PPF_CH3W_ROUNDS
#define PPF_CH3W_ROUNDS 4
#endif
template <typename R, int RPW>
int32_t ch3w_zchunk(const ppf_grid& g, int64_t xy_tiles) {
  static int per_cu = 0;
  if (per_cu == 0) {
    int v = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&v, k_ch3w<R, RPW>, 64, 0) != hipSuccess || v < 1) v = 1;
    per_cu = v;
  }
  const int64_t resident = int64_t(device_cus()) * per_cu;
  int64_t chunks = PPF_CH3W_ROUNDS * resident / xy_tiles;
  if (chunks < 1) chunks = 1;
  int64_t zc = (g.n[2] + chunks - 1) / chunks;
  if (zc < 8) zc = 8;
  if (zc > g.n[2]) zc = g.n[2];
  return int32_t(zc);
}

// z chunk: about PPF_CH3D_BLOCKS blocks (several rounds of resident blocks, so
// that blocks finishing early are replaced), at least 16 planes per chunk
#ifndef PPF_CH3D_BLOCKS
#define PPF_CH3D_BLOCKS 8192
#endif
inline int32_t ch3d_zchunk(const ppf_grid& g, int c3y) {
  const int64_t xy = ((g.n[0] + C3X - 1) / C3X) * ((g.n[1] + c3y - 1) / c3y);
  int64_t chunks = (PPF_CH3D_BLOCKS + xy - 1) / xy;
  int64_t zc = (g.n[2] + chunks - 1) / chunks;
  if (zc < 16) zc = 16;
  if (zc > g.n[2]) zc = g.n[2];
  return int32_t(zc);
}

// the register march reads halos as a coordinate remap: every side periodic,
// or zero-gradient Neumann without bc values (no Dirichlet side)
inline bool ch3w_remappable(const ppf_grid& g, const void* bcvals) {
  bool periodic = true, dirichlet = false;
  for (int a = 0; a < 3; ++a) {
    periodic = periodic && g.bc[a] == PPF_PERIODIC;
    dirichlet = dirichlet || g.bc[a] == PPF_DIRICHLET;
  }
  return periodic || (!bcvals && !dirichlet);
}

template <typename R>
int cahnhilliard_step(const ppf_grid& g, const void* in, void* out, const void* bcvals, double M, double eps,
                      double dt, hipStream_t s) {
  Dom d = make_dom(g);
  R dtM = R(dt * M);
  if (g.ndim == 2) {
    hipLaunchKernelGGL(k_ch2d<R>, dim3(tiles2d(g)), dim3(kThreads), 0, s, static_cast<const R*>(in),
                       static_cast<R*>(out), static_cast<const R*>(bcvals), d, R(g.dx), R(eps), dtM);
  } else if (g.ndim == 3 && g.sz * int64_t(sizeof(R)) < (int64_t(1) << 31) && PPF_CH3D_WAVE &&
             ch3w_remappable(g, bcvals)) {  // in-plane byte offsets (row offsets, buffer offsets) are 32-bit
    constexpr int RPW = sizeof(R) == 8 ? 8 : PPF_CH3W_RPW;
    const int64_t xy = ((g.n[0] + kCh3wTx - 1) / kCh3wTx) * ((g.n[1] + RPW - 5) / (RPW - 4));
    const int32_t zc = ch3w_zchunk<R, RPW>(g, xy);
    const int64_t blocks = xy * ((g.n[2] + zc - 1) / zc);
    hipLaunchKernelGGL((k_ch3w<R, RPW>), dim3(unsigned(blocks)), dim3(64), 0, s, static_cast<const R*>(in),
                       static_cast<R*>(out), d, R(g.dx), R(eps), dtM, zc);
  } else if (g.ndim == 3 && g.sz < (int64_t(1) << 31)) {  // in-plane offsets are 32-bit
    constexpr int Y = sizeof(R) == 8 ? 8 : PPF_CH3D_TY;  // FP64 tiles stay under 64 KB of LDS
    const size_t lds = Ch3dShape<Y>::lds(sizeof(R));
    const int32_t zc = ch3d_zchunk(g, Y);
    const int64_t blocks = ((g.n[0] + C3X - 1) / C3X) * ((g.n[1] + Y - 1) / Y) * ((g.n[2] + zc - 1) / zc);
    hipLaunchKernelGGL((k_ch3d<R, Y>), dim3(unsigned(blocks)), dim3(kThreads), lds, s, static_cast<const R*>(in),
                       static_cast<R*>(out), static_cast<const R*>(bcvals), d, R(g.dx), R(eps), dtM, zc);
  } else {
    int64_t total = interior_cells(g);
    hipLaunchKernelGGL(k_ch_nd<R>, dim3(unsigned((total + 255) / 256)), dim3(256), 0, s,
                       static_cast<const R*>(in), static_cast<R*>(out), static_cast<const R*>(bcvals), d,
                       R(g.dx), R(eps), dtM);
  }
  return check_launch("cahnhilliard_step");
}

// nsteps (1 .. cahnhilliard_multistep_max) successive explicit steps
template <typename R>
int cahnhilliard_multistep(const ppf_grid& g, const void* in, void* out, const void* bcvals, double M, double eps,
                           double dt, int32_t nsteps, hipStream_t s) {
  const int32_t mx = cahnhilliard_multistep_max(g);
  if (nsteps < 1 || nsteps > mx)
    return fail(PPF_EINVAL, "cahnhilliard_multistep: nsteps " + std::to_string(nsteps) + " outside 1.." +
                                std::to_string(mx) + " for this grid");
  if (nsteps == 1) return cahnhilliard_step<R>(g, in, out, bcvals, M, eps, dt, s);
  if constexpr (sizeof(R) == 4)
    return ch3w2_launch(g, in, out, M, eps, dt, s);
  else
    return fail(PPF_EINVAL, "cahnhilliard_multistep: FP64 grids take one step per launch");
}

// ------------------------------------------------------------ reductions --
// Two-pass FP64 reduction with a fixed summation order (deterministic):
// kReduceBlocks blocks write partial sums, one block adds them.
__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

__device__ __forceinline__ double block_sum(double v) {
  __shared__ double part[4];
  v = wave_sum(v);
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) part[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) r = (part[0] + part[1]) + (part[2] + part[3]);
  __syncthreads();
  return r;
}

// Partial sums of one field: block b walks interior rows b, b + G, ... (a
// row is a fixed (y, z)), kRRows of them at a time with all their loads
// issued before the arithmetic; threads walk x.  The (y, z) of the next row
// comes from an increment of G rows (no 64-bit division per row).  The two
// sums are separate kernels: the free energy's boundary folds would otherwise
// set the register allocation of the plain sum too.
#ifndef PPF_REDUCE_ROWS
#define PPF_REDUCE_ROWS 4
#endif
// the energy kernel's rows b, b + G, ... start from the XCD-swizzled block
// index, so rows r and r + 1 are read on one XCD and the y neighbour comes
// from its L2 (512^3: 251 -> 205 us)
#ifndef PPF_REDUCE_ESWZ
#define PPF_REDUCE_ESWZ 1
#endif
constexpr int kRRows = PPF_REDUCE_ROWS;

struct RowWalk {
  int64_t G, ny, dy, dz, cy, cz;
  __device__ RowWalk(int64_t ny_, int64_t start) : G(gridDim.x), ny(ny_) {
    dz = G / ny;
    dy = G - dz * ny;
    cz = start / ny;
    cy = start - cz * ny;
  }
  // interior (y, z) (1-based) of the current row, then advance by G rows
  __device__ __forceinline__ void next(int64_t& y, int64_t& z) {
    y = 1 + cy;
    z = 1 + cz;
    cy += dy;
    cz += dz;
    if (cy >= ny) {
      cy -= ny;
      ++cz;
    }
  }
};

// sum of the interior cells of field blockIdx.y
template <typename R>
__global__ __launch_bounds__(256) void k_reduce_sum(const R* __restrict__ in, Dom d, double* __restrict__ work) {
  const int fi = blockIdx.y;
  const R* __restrict__ f = in + fi * d.sf;
  const int64_t nx = d.ax.n, ny = d.ndim >= 2 ? d.ay.n : 1, nz = d.ndim >= 3 ? d.az.n : 1;
  const int64_t rows = ny * nz;
  const int64_t sy = d.ndim >= 2 ? d.ay.stride : 0, sz = d.ndim >= 3 ? d.az.stride : 0;
  RowWalk w(ny, blockIdx.x);
  double acc = 0.0;
  for (int64_t r = blockIdx.x; r < rows; r += kRRows * w.G) {
    const R* rowp[kRRows];
#pragma unroll
    for (int k = 0; k < kRRows; ++k) {
      int64_t y, z;
      w.next(y, z);
      rowp[k] = r + k * w.G < rows ? f + (d.ndim >= 3 ? z : 0) * sz + (d.ndim >= 2 ? y : 0) * sy : nullptr;
    }
    for (int64_t x = 1 + threadIdx.x; x <= nx; x += 256) {
      R v[kRRows];
#pragma unroll
      for (int k = 0; k < kRRows; ++k) v[k] = rowp[k] ? rowp[k][x] : R(0);
#pragma unroll
      for (int k = 0; k < kRRows; ++k) acc += double(v[k]);
    }
  }
  double s = block_sum(acc);
  if (threadIdx.x == 0) work[fi * kReduceBlocks + blockIdx.x] = s;
}

// free-energy density of one cell from its forward neighbours; the boundary
// rule is folded in where a neighbour lies outside the interior
template <typename R>
__device__ __forceinline__ double fe_fold(const R* __restrict__ f, const R* bb, const Dom& d, int64_t x, int64_t y,
                                          int64_t z, R dx, double idx, double ke) {
  const double c = double(f[x + (d.ndim >= 2 ? y * d.ay.stride : 0) + (d.ndim >= 3 ? z * d.az.stride : 0)]);
  const double gx = (double(fold(f, bb, d, x + 1, y, z, dx)) - c) * idx;
  double g2 = gx * gx;
  if (d.ndim >= 2) {
    const double gy = (double(fold(f, bb, d, x, y + 1, z, dx)) - c) * idx;
    g2 += gy * gy;
  }
  if (d.ndim >= 3) {
    const double gz = (double(fold(f, bb, d, x, y, z + 1, dx)) - c) * idx;
    g2 += gz * gz;
  }
  return c * c * (1.0 - c) * (1.0 - c) + ke * g2;
}

// sum of c^2 (1-c)^2 + eps^2/2 |grad c|^2 (forward differences) over the
// interior of field 0 (ND dimensions), in three parts so that the common
// case is a branch-free body whose loads are all in flight together:
//  1. cells x < nx of every row by direct loads; rows on the last y or z
//     layer (whose forward neighbour there is a halo cell, present in the
//     padded layout but not necessarily initialised) and the rows past the
//     end (pointed at row 0) are computed and dropped by a select;
//  2. the cells of the rows on the last y or z layer, in the folded form,
//     spread over the whole grid (with rows b, b + G, ... per block they
//     would all fall to the few blocks whose b matches ny - 1);
//  3. the last cell (x = nx) of the other rows, also spread over the grid.
template <typename R, int ND>
__global__ __launch_bounds__(256) void k_reduce_energy(const R* __restrict__ in, const R* __restrict__ b, Dom d,
                                                       R dx, double eps, double* __restrict__ work) {
  constexpr int kE = kRRows;
  const R* __restrict__ f = in;
  const int64_t nx = d.ax.n, ny = ND >= 2 ? d.ay.n : 1, nz = ND >= 3 ? d.az.n : 1;
  const int64_t rows = ny * nz;
  const int64_t sy = ND >= 2 ? d.ay.stride : 0, sz = ND >= 3 ? d.az.stride : 0;
  const double idx = 1.0 / double(dx);
  const double ke = 0.5 * eps * eps;
  auto last_layer = [&](int64_t y, int64_t z) { return (ND >= 2 && y == ny) || (ND >= 3 && z == nz); };
#if PPF_REDUCE_ESWZ
  // rows r and r + 1 on one XCD (the y neighbour from the same L2)
  const int64_t r0 = xcd_swizzle(blockIdx.x, gridDim.x);
#else
  const int64_t r0 = blockIdx.x;
#endif
  RowWalk w(ny, r0);
  double acc = 0.0;
  for (int64_t r = r0; r < rows; r += kE * w.G) {
    const R* rowp[kE];
    int64_t yy[kE], zz[kE];
    bool keep[kE];
#pragma unroll
    for (int k = 0; k < kE; ++k) {
      w.next(yy[k], zz[k]);
      if (ND < 2) yy[k] = 0;
      if (ND < 3) zz[k] = 0;
      const bool valid = r + k * w.G < rows;
      rowp[k] = f + (valid ? zz[k] * sz + yy[k] * sy : 0);
      keep[k] = valid && !last_layer(yy[k], zz[k]);
    }
    for (int64_t x = 1 + threadIdx.x; x < nx; x += 256) {
      R c[kE], xp[kE], yp[kE], zp[kE];
#pragma unroll
      for (int k = 0; k < kE; ++k) {
        c[k] = rowp[k][x];
        xp[k] = rowp[k][x + 1];
        if (ND >= 2) yp[k] = rowp[k][x + sy];
        if (ND >= 3) zp[k] = rowp[k][x + sz];
      }
#pragma unroll
      for (int k = 0; k < kE; ++k) {
        const double cc = double(c[k]);
        const double gx = (double(xp[k]) - cc) * idx;
        double g2 = gx * gx;
        if (ND >= 2) {
          const double gy = (double(yp[k]) - cc) * idx;
          g2 += gy * gy;
        }
        if (ND >= 3) {
          const double gz = (double(zp[k]) - cc) * idx;
          g2 += gz * gz;
        }
        const double e = cc * cc * (1.0 - cc) * (1.0 - cc) + ke * g2;
        acc += keep[k] ? e : 0.0;
      }
    }
  }
  const int64_t t0 = int64_t(blockIdx.x) * 256 + threadIdx.x, nt = w.G * 256;
  // slow rows: j < nz -> (y = ny, z = 1 + j); then (y = 1 + j - nz, z = nz), y < ny
  const int64_t nslow = ND == 1 ? 0 : ND == 2 ? 1 : nz + ny - 1;
  for (int64_t i = t0; i < nslow * nx; i += nt) {
    const int64_t j = i / nx, x = 1 + (i - j * nx);
    int64_t y = ny, z = ND >= 3 ? 1 + j : 0;
    if (ND >= 3 && j >= nz) {
      y = 1 + (j - nz);
      z = nz;
    }
    acc += fe_fold(f, b, d, x, y, z, dx, idx, ke);
  }
  for (int64_t r = t0; r < rows; r += nt) {
    const int64_t zq = r / ny;
    const int64_t y = ND >= 2 ? 1 + (r - zq * ny) : 0, z = ND >= 3 ? 1 + zq : 0;
    if (!last_layer(y, z)) acc += fe_fold(f, b, d, nx, y, z, dx, idx, ke);
  }
  double s = block_sum(acc);
  if (threadIdx.x == 0) work[blockIdx.x] = s;
}

template <typename R>
__global__ __launch_bounds__(256) void k_reduce_final(const double* __restrict__ work, double* __restrict__ out,
                                                      double scale) {
  int fi = blockIdx.x;
  double acc = 0.0;
  for (int i = threadIdx.x; i < kReduceBlocks; i += 256) acc += work[fi * kReduceBlocks + i];
  double s = block_sum(acc);
  if (threadIdx.x == 0) out[fi] = s * scale;
}

template <typename R>
int reduce(const ppf_grid& g, const void* in, const void* bcvals, int kind, double eps, double* out, void* work,
           hipStream_t s) {
  Dom d = make_dom(g);
  int nf = kind == 0 ? g.nfields : 1;
  double scale = 1.0;
  if (kind == 1) {
    for (int a = 0; a < g.ndim; ++a) scale *= g.dx;
  }
  if (kind == 0) {
    hipLaunchKernelGGL(k_reduce_sum<R>, dim3(kReduceBlocks, unsigned(nf)), dim3(256), 0, s, static_cast<const R*>(in),
                       d, static_cast<double*>(work));
  } else {
    auto kern = g.ndim == 1 ? k_reduce_energy<R, 1> : g.ndim == 2 ? k_reduce_energy<R, 2> : k_reduce_energy<R, 3>;
    hipLaunchKernelGGL(kern, dim3(kReduceBlocks), dim3(256), 0, s, static_cast<const R*>(in),
                       static_cast<const R*>(bcvals), d, R(g.dx), eps, static_cast<double*>(work));
  }
  int rc = check_launch("reduce_partial");
  if (rc) return rc;
  hipLaunchKernelGGL(k_reduce_final<R>, dim3(unsigned(nf)), dim3(256), 0, s, static_cast<const double*>(work), out,
                     scale);
  return check_launch("reduce_final");
}

}  // namespace ppf
