// Device-side helpers shared by the engine kernels: error reporting, the
// boundary-condition fold used by every halo load, and the counter RNG.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "../../include/ppf_hip.h"

namespace ppf {

// ---------------------------------------------------------------- errors --
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
int check_launch(const char* what);
// compute units of the current device (cached per device id)
int device_cus();
int check_temperature(const ppf_temperature* t, const char* what);
int check_grid(const ppf_grid* g, int want_ndim_min, int want_ndim_max, const char* what);

inline hipStream_t stream_of(ppf_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

// ----------------------------------------------------------- fold rules --
// Padded coordinate p along an axis with n interior cells (interior 1..n).
// The value of a halo cell is produced from interior cells by the rules of
// ppf_gpu_utils.py:11-96: PERIODIC wraps, NEUMANN copies the neighbour and
// adds -/+ dx*bc, DIRICHLET returns bc.  Corners compose the rules in the
// reference's kernel order: x innermost, then y, then z.

struct Axis {
  int64_t n;
  int64_t stride;
  int bc;
};

// Resolve one axis: returns the interior coordinate to read and, for a
// non-periodic halo, sets *halo to the halo coordinate (0 or n+1) whose bc
// value applies and *sign to -1 (low side) or +1 (high side).  For
// coordinates more than one cell outside on a non-periodic axis (second halo
// ring, values never used by the engines) the nearest halo cell is used.
__device__ __forceinline__ int64_t resolve(int64_t p, const Axis& a, int64_t* halo, int* sign) {
  *halo = -1;
  *sign = 0;
  if (p >= 1 && p <= a.n) return p;
  if (a.bc == PPF_PERIODIC) {
    int64_t q = p;
    if (q < 1) q += a.n;
    if (q > a.n) q -= a.n;
    if (q < 1) q = 1;           // domains thinner than the halo depth
    if (q > a.n) q = a.n;
    return q;
  }
  if (p < 1) {
    *halo = 0;
    *sign = -1;
    return 1;
  }
  *halo = a.n + 1;
  *sign = 1;
  return a.n;
}

// Coordinate p with a periodic axis wrapped into [1, n] exactly as resolve()
// does; other axes return p unchanged.
__device__ __forceinline__ int64_t wrap_periodic(int64_t p, const Axis& a) {
  if (a.bc != PPF_PERIODIC || (p >= 1 && p <= a.n)) return p;
  int64_t h;
  int s;
  return resolve(p, a, &h, &s);
}

// Value of padded cell (x, y, z) of a field with base pointer `f` (field
// offset applied) and bc-value pointer `b` (may be null = zeros).
template <typename R>
__device__ __forceinline__ R fold_x(const R* __restrict__ f, const R* __restrict__ b, const Axis& ax,
                                    int64_t x, int64_t row, R dx) {
  int64_t h;
  int s;
  int64_t xi = resolve(x, ax, &h, &s);
  if (h < 0) return f[row + xi];
  R bv = b ? b[row + h] : R(0);
  if (ax.bc == PPF_NEUMANN) return f[row + xi] + (s < 0 ? -dx : dx) * bv;
  return bv;
}

template <typename R>
__device__ __forceinline__ R fold_xy(const R* __restrict__ f, const R* __restrict__ b, const Axis& ax,
                                     const Axis& ay, int64_t x, int64_t y, int64_t plane, R dx) {
  int64_t h;
  int s;
  int64_t yi = resolve(y, ay, &h, &s);
  if (h < 0) return fold_x(f, b, ax, x, plane + yi * ay.stride, dx);
  if (ay.bc == PPF_NEUMANN) {
    // bc value sits at (x clamped into the padded range, halo row)
    int64_t xb = x < 0 ? 0 : (x > ax.n + 1 ? ax.n + 1 : x);
    R bv = b ? b[plane + h * ay.stride + xb] : R(0);
    return fold_x(f, b, ax, x, plane + yi * ay.stride, dx) + (s < 0 ? -dx : dx) * bv;
  }
  int64_t xb = x < 0 ? 0 : (x > ax.n + 1 ? ax.n + 1 : x);
  return b ? b[plane + h * ay.stride + xb] : R(0);
}

template <typename R>
__device__ __forceinline__ R fold_xyz(const R* __restrict__ f, const R* __restrict__ b, const Axis& ax,
                                      const Axis& ay, const Axis& az, int64_t x, int64_t y, int64_t z,
                                      R dx) {
  int64_t h;
  int s;
  int64_t zi = resolve(z, az, &h, &s);
  if (h < 0) return fold_xy(f, b, ax, ay, x, y, zi * az.stride, dx);
  int64_t xb = x < 0 ? 0 : (x > ax.n + 1 ? ax.n + 1 : x);
  int64_t yb = y < 0 ? 0 : (y > ay.n + 1 ? ay.n + 1 : y);
  R bv = b ? b[h * az.stride + yb * ay.stride + xb] : R(0);
  if (az.bc == PPF_NEUMANN) return fold_xy(f, b, ax, ay, x, y, zi * az.stride, dx) + (s < 0 ? -dx : dx) * bv;
  return bv;
}

// Temperature rule (ppf_gpu_utils.py:197-319): periodic wrap, else zero
// gradient; returns the interior coordinate to read.
__device__ __forceinline__ int64_t resolve_t(int64_t p, int64_t n, int bc) {
  if (p >= 1 && p <= n) return p;
  if (bc == PPF_PERIODIC) {
    int64_t q = p < 1 ? p + n : p - n;
    return q < 1 ? 1 : (q > n ? n : q);
  }
  return p < 1 ? 1 : n;
}

// Temperature at padded (x, y, z) for the engine step `step`.
template <typename R>
__device__ __forceinline__ R temperature_at(const ppf_temperature& t, const int64_t n[3], const int bc[3],
                                            int64_t x, int64_t y, int64_t z, double dx, int64_t step) {
  if (t.kind == PPF_T_CONST) return R(t.T0);
  int64_t xi = resolve_t(x, n[0], bc[0]);
  int64_t yi = resolve_t(y, n[1], bc[1]);
  int64_t zi = resolve_t(z, n[2], bc[2]);
  if (t.kind == PPF_T_ARRAY) {
    const R* d = static_cast<const R*>(t.data);
    return d[zi * t.sz + yi * t.sy + xi];
  }
  if (t.kind == PPF_T_FRAMES) {
    // two frames of a thermal history, interpolated in time in FP64 in the
    // order of update_thermal_file_*_kernel (ppf_gpu_utils.py:357-394)
    const int64_t o = zi * t.sz + yi * t.sy + xi;
    const double cur = double(step - 1) * t.dt;
    const double r1 = (cur - t.t_start) / (t.t_end - t.t_start);
    const double r0 = 1.0 - r1;
    return R(double(static_cast<const R*>(t.data)[o]) * r0 + double(static_cast<const R*>(t.data1)[o]) * r1);
  }
  // LINEAR: evaluated in double in the same order as the host/oracle formula
  double T = t.T0;
  if (t.dTdx != 0.0) T = T + double(xi - 1) * (dx * t.dTdx);
  if (t.dTdy != 0.0) T = T + double(yi - 1) * (dx * t.dTdy);
  if (t.dTdz != 0.0) T = T + double(zi - 1) * (dx * t.dTdz);
  if (t.dTdt != 0.0) T = T + double(step - 1) * (t.dTdt * t.dt);
  return R(T);
}

// temperature_at for an interior cell (x in [1, n_x], y in [1, n_y], 2D):
// the same operations in the same order, without the boundary resolution
template <typename R>
__device__ __forceinline__ R temperature_interior(const ppf_temperature& t, const int64_t n[3], const int bc[3],
                                                  int64_t x, int64_t y, double dx, int64_t step) {
  if (t.kind == PPF_T_CONST) return R(t.T0);
  if (t.kind == PPF_T_LINEAR) {
    double T = t.T0;
    if (t.dTdx != 0.0) T = T + double(x - 1) * (dx * t.dTdx);
    if (t.dTdy != 0.0) T = T + double(y - 1) * (dx * t.dTdy);
    if (t.dTdz != 0.0) T = T + double(1 - 1) * (dx * t.dTdz);
    if (t.dTdt != 0.0) T = T + double(step - 1) * (t.dTdt * t.dt);
    return R(T);
  }
  return temperature_at<R>(t, n, bc, x, y, 1, dx, step);
}

// ------------------------------------------------------------ step counter --
// Step number seen by a step kernel: `off` alone, or *base + off when `base`
// (a device int64) is set -- the form a captured HIP graph replays, with the
// counter advanced on the device between replays (k_step_advance).
struct StepRef {
  const int64_t* base;
  int64_t off;
  __device__ __forceinline__ int64_t get() const { return base ? *base + off : off; }
};
inline StepRef step_host(int64_t step) { return StepRef{nullptr, step}; }

// ------------------------------------------------------------ counter RNG --
// Stateless hash of (seed, step, stream, cell); see the oracle's rng module
// for the definition shared with the CPU side.
constexpr uint64_t kNStreams = 16;
enum { S_WARREN_PHI = 0, S_NC_PHI = 1, S_NC_Q1 = 2, S_NC_Q4 = 3, S_NC_C0 = 4 };

__device__ __forceinline__ uint64_t hash64(uint64_t seed, uint64_t counter) {
  uint64_t z = seed + counter * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__device__ __forceinline__ uint64_t rng_counter(int64_t step, int stream, uint64_t ncells, uint64_t cell) {
  return (uint64_t(step) * kNStreams + uint64_t(stream)) * ncells + cell;
}

template <typename R>
__device__ __forceinline__ R uniform01(uint64_t h);
template <>
__device__ __forceinline__ double uniform01<double>(uint64_t h) {
  return double(h >> 11) * 0x1.0p-53;
}
template <>
__device__ __forceinline__ float uniform01<float>(uint64_t h) {
  return float(uint32_t(h >> 40)) * 0x1.0p-24f;
}

// FP64 natural logarithm defined identically in the oracle: frexp range
// reduction to m in [sqrt(1/2), sqrt(2)), log m = 2 atanh((m-1)/(m+1)) by the
// odd power series to z^11 (z = s^2 <= 0.0295), exponent times ln 2 split in
// a 32-bit head and a tail.  The reference's h = 1e-7 forward difference
// turns a 1-ulp difference between log implementations into a 1e-8 relative
// error in the chemical potential; one shared definition keeps FP64 runs
// within 1e-10 of the oracle.  FP64 callers live in the translation unit compiled without FMA
// contraction.
__device__ __forceinline__ double plog(double x) {
  int ei;
  double m = frexp(x, &ei);
  if (m < 0x1.6a09e667f3bcdp-1) {
    m = m * 2.0;
    ei -= 1;
  }
  double e = double(ei);
  double f = m - 1.0;
  double s = f / (2.0 + f);
  double z = s * s;
  double p = 1.0 / 23.0;
  p = p * z + 1.0 / 21.0;
  p = p * z + 1.0 / 19.0;
  p = p * z + 1.0 / 17.0;
  p = p * z + 1.0 / 15.0;
  p = p * z + 1.0 / 13.0;
  p = p * z + 1.0 / 11.0;
  p = p * z + 1.0 / 9.0;
  p = p * z + 1.0 / 7.0;
  p = p * z + 1.0 / 5.0;
  p = p * z + 1.0 / 3.0;
  p = p * z + 1.0;
  return e * 0x1.62e42fee00000p-1 + (e * 0x1.a39ef35793c76p-33 + 2.0 * s * p);
}

// Standard normal by Marsaglia's polar method; attempt k > 0 re-hashes the
// previous word with k.  FP64 uses plog, so draws are bit-identical to the
// CPU side; FP32 uses logf.

template <typename R>
__device__ __forceinline__ R normal01(uint64_t h) {
  for (int k = 0; k < 64; ++k) {
    if (k > 0) h = hash64(h, uint64_t(k));
    R u, v;
    if (sizeof(R) == 8) {
      u = R(2.0 * ((double(h >> 32) + 0.5) * 0x1.0p-32) - 1.0);
      v = R(2.0 * ((double(h & 0xffffffffull) + 0.5) * 0x1.0p-32) - 1.0);
    } else {
      u = R(2) * ((R(uint32_t(h >> 40)) + R(0.5)) * R(0x1.0p-24)) - R(1);
      v = R(2) * ((R(uint32_t(h & 0xffffffffull) >> 8) + R(0.5)) * R(0x1.0p-24)) - R(1);
    }
    R s = u * u + v * v;
    if (s < R(1) && s > R(0)) {
      R l = sizeof(R) == 8 ? R(plog(double(s))) : R(logf(float(s)));
      return u * sqrt(R(-2) * l / s);
    }
  }
  return R(0);
}

// XCD-aware block order: the hardware deals workgroups round-robin over the
// 8 XCDs; remap so that consecutive logical tiles (which share halo rows in
// L2) are handled by the same XCD.
__device__ __forceinline__ uint32_t xcd_swizzle(uint32_t b, uint32_t nblocks) {
  constexpr uint32_t kXcd = 8;
  uint32_t per = nblocks / kXcd;
  if (per == 0 || b >= per * kXcd) return b;
  return (b % kXcd) * per + b / kXcd;
}

}  // namespace ppf
