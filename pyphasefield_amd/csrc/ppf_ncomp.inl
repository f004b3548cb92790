// N-component alloy solidification engine (2D): the thermodynamic pass that
// replaces NComponent_helper_kernel (Engines/NCGPU_new.py:256-313) and the
// multi-field stencil that replaces NComponent_kernel (NCGPU_new.py:37-250).
//
// Field stack: phi, q1, q4, c_1..c_{N-1} (NCGPU_new.py:63-66).
// Transfer stack written by the thermodynamic pass (2N-1 arrays, every padded
// cell including the halo ring, like the reference helper):
//   [0]        G_S - G_L
//   [1+l]      M_c,l  = v_m c_l (D_L + h (D_S - D_L)) / (8.314 T)      (:308)
//   [N+l]      dF/dc_l = (dGL_l + h (dGS_l - dGL_l))/v_m + (W_l - W_N) g T + noise  (:311)
#pragma once

#include "ppf_stencil.inl"

namespace ppf {

constexpr double kCMin = 1e-10;  // c ln c evaluated at max(c, kCMin)

template <typename R>
__device__ __forceinline__ R rlog(R x);
template <>
__device__ __forceinline__ double rlog<double>(double x) { return plog(x); }
template <>
__device__ __forceinline__ float rlog<float>(float x) { return logf(x); }

// phase coefficients evaluated at one temperature; TDB tables (EXT) add
// binary orders 3..4 (L34) and ternary terms (T3)
constexpr int nc_triples(int N) { return N * (N - 1) * (N - 2) / 6 > 0 ? N * (N - 1) * (N - 2) / 6 : 1; }

template <typename R, int N, bool EXT = false>
struct PhaseAtT {
  R gi[N];
  R L[N * (N - 1) / 2][3];
  R L34[EXT ? N * (N - 1) / 2 : 1][2];
  R T3[EXT ? nc_triples(N) : 1][3];
  R RT;
};

template <typename R, int N>
__device__ __forceinline__ void phase_at(const ppf_rk_phase& ph, R T, R lnT, PhaseAtT<R, N, false>& o) {
#pragma unroll
  for (int i = 0; i < N; ++i) {
    R a = R(ph.g0[i][0]), b = R(ph.g0[i][1]), c = R(ph.g0[i][2]), d = R(ph.g0[i][3]), e = R(ph.g0[i][4]);
    if (ph.g0[i][2] == 0.0 && ph.g0[i][3] == 0.0 && ph.g0[i][4] == 0.0) {
      o.gi[i] = a + b * T;  // uniform branch; same value as the full form
    } else {
      o.gi[i] = a + b * T + c * T * lnT + d * T * T + e / T;
    }
  }
#pragma unroll
  for (int p = 0; p < N * (N - 1) / 2; ++p) {
#pragma unroll
    for (int k = 0; k < 3; ++k) o.L[p][k] = R(ph.L[p][k][0]) + R(ph.L[p][k][1]) * T;
  }
  o.RT = R(ph.R) * T;
}

// piecewise TDB temperature function (operation order of ppf_tpoly)
template <typename R>
__device__ __forceinline__ R tpoly_eval(const ppf_tpoly* __restrict__ f, R T, R lnT) {
  const int n = f->nranges;
  if (n <= 0) return R(0);
  int r = 0;
  while (r < n - 1 && double(T) > f->T_hi[r]) ++r;
  const double* c = f->c[r];
  const R u = R(1) / T;
  const R T2 = T * T, T3 = T2 * T, T7 = (T3 * T3) * T;
  const R u3 = (u * u) * u, u9 = (u3 * u3) * u3;
  R v = R(c[0]) + R(c[1]) * T;
  v = v + R(c[2]) * (T * lnT);
  v = v + R(c[3]) * T2;
  v = v + R(c[4]) * T3;
  v = v + R(c[5]) * u;
  v = v + R(c[6]) * T7;
  v = v + R(c[7]) * u9;
  return v;
}

template <typename R, int N>
__device__ __forceinline__ void phase_at_tables(const ppf_tdb_phase* __restrict__ ph, R Rg, R T, R lnT,
                                                PhaseAtT<R, N, true>& o) {
#pragma unroll
  for (int i = 0; i < N; ++i) o.gi[i] = tpoly_eval<R>(&ph->g0[i], T, lnT);
#pragma unroll
  for (int p = 0; p < N * (N - 1) / 2; ++p) {
#pragma unroll
    for (int k = 0; k < 3; ++k) o.L[p][k] = tpoly_eval<R>(&ph->L[p][k], T, lnT);
#pragma unroll
    for (int k = 0; k < 2; ++k) o.L34[p][k] = tpoly_eval<R>(&ph->L[p][3 + k], T, lnT);
  }
#pragma unroll
  for (int t = 0; t < nc_triples(N); ++t) {
#pragma unroll
    for (int k = 0; k < 3; ++k) o.T3[t][k] = N >= 3 ? tpoly_eval<R>(&ph->T3[t][k], T, lnT) : R(0);
  }
  o.RT = Rg * T;
}

// f(PL, PS) with both phases at one temperature: from the TDB tables when
// present (binary orders 0..4, ternary terms), else from the inline
// Redlich-Kister coefficients (orders 0..2) -- two code paths, chosen once
template <typename R, int N, typename F>
__device__ __forceinline__ void with_phases(const ppf_thermo& th, R T, R lnT, F&& f) {
  if (th.tables) {
    PhaseAtT<R, N, true> PL, PS;
    phase_at_tables<R, N>(&th.tables->liquid, R(th.liquid.R), T, lnT, PL);
    phase_at_tables<R, N>(&th.tables->solid, R(th.solid.R), T, lnT, PS);
    f(PL, PS);
  } else {
    PhaseAtT<R, N, false> PL, PS;
    phase_at<R, N>(th.liquid, T, lnT, PL);
    phase_at<R, N>(th.solid, T, lnT, PS);
    f(PL, PS);
  }
}

// whether any coefficient multiplies T ln T (lnT is only needed then)
__device__ __forceinline__ bool needs_lnT(const ppf_thermo& th, int N) {
  if (th.tables) return true;
  for (int i = 0; i < N; ++i)
    if (th.liquid.g0[i][2] != 0.0 || th.solid.g0[i][2] != 0.0) return true;
  return false;
}

// sum_i c_i ln max(c_i, cmin), summed in component order
template <typename R, int N>
__device__ __forceinline__ R ideal_sum(const R (&c)[N]) {
  R s = R(0);
#pragma unroll
  for (int i = 0; i < N; ++i) {
    R ci = c[i] > R(kCMin) ? c[i] : R(kCMin);
    s = s + ci * rlog<R>(ci);
  }
  return s;
}

// binary Redlich-Kister polynomial sum_k L^k dd^k of pair p and its dd-derivative
template <typename R, int N, bool EXT>
__device__ __forceinline__ R rk_poly(const PhaseAtT<R, N, EXT>& ph, int p, R dd) {
  if constexpr (EXT)
    return ph.L[p][0] + dd * (ph.L[p][1] + dd * (ph.L[p][2] + dd * (ph.L34[p][0] + dd * ph.L34[p][1])));
  return ph.L[p][0] + dd * (ph.L[p][1] + dd * ph.L[p][2]);
}
template <typename R, int N, bool EXT>
__device__ __forceinline__ R rk_dpoly(const PhaseAtT<R, N, EXT>& ph, int p, R dd) {
  if constexpr (EXT)
    return ph.L[p][1] + dd * (R(2) * ph.L[p][2] + dd * (R(3) * ph.L34[p][0] + dd * (R(4) * ph.L34[p][1])));
  return ph.L[p][1] + R(2) * dd * ph.L[p][2];
}

// G of one phase (same operation order as the oracle's thermo.gibbs)
template <typename R, int N, bool EXT>
__device__ __forceinline__ R gibbs(const R (&c)[N], const PhaseAtT<R, N, EXT>& ph, R ideal) {
  R G = R(0);
#pragma unroll
  for (int i = 0; i < N; ++i) G = G + c[i] * ph.gi[i];
  int p = 0;
#pragma unroll
  for (int i = 0; i < N; ++i) {
#pragma unroll
    for (int j = i + 1; j < N; ++j, ++p) {
      R dd = c[i] - c[j];
      R P = rk_poly(ph, p, dd);
      G = G + c[i] * c[j] * P;
    }
  }
  if constexpr (EXT) {
    int t = 0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
#pragma unroll
      for (int j = i + 1; j < N; ++j) {
#pragma unroll
        for (int l = j + 1; l < N; ++l, ++t) {
          const R w = (R(1) - ((c[i] + c[j]) + c[l])) / R(3);
          const R Wt = ((c[i] + w) * ph.T3[t][0] + (c[j] + w) * ph.T3[t][1]) + (c[l] + w) * ph.T3[t][2];
          G = G + ((c[i] * c[j]) * c[l]) * Wt;
        }
      }
    }
  }
  return G + ph.RT * ideal;
}

// exact dG/dc_l along e_l - e_N
template <typename R, int N, bool EXT>
__device__ __forceinline__ void gibbs_grad(const R (&c)[N], const PhaseAtT<R, N, EXT>& ph, R (&dG)[N]) {
  R part[N];
#pragma unroll
  for (int i = 0; i < N; ++i) part[i] = ph.gi[i] + (c[i] > R(kCMin) ? ph.RT * (rlog<R>(c[i]) + R(1)) : R(0));
  int p = 0;
#pragma unroll
  for (int i = 0; i < N; ++i) {
#pragma unroll
    for (int j = i + 1; j < N; ++j, ++p) {
      R dd = c[i] - c[j];
      R P = rk_poly(ph, p, dd);
      R dP = rk_dpoly(ph, p, dd);
      part[i] = part[i] + c[j] * P + c[i] * c[j] * dP;
      part[j] = part[j] + c[i] * P - c[i] * c[j] * dP;
    }
  }
  if constexpr (EXT) {
    int t = 0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
#pragma unroll
      for (int j = i + 1; j < N; ++j) {
#pragma unroll
        for (int l = j + 1; l < N; ++l, ++t) {
          const R w = (R(1) - ((c[i] + c[j]) + c[l])) / R(3);
          const R Wt = ((c[i] + w) * ph.T3[t][0] + (c[j] + w) * ph.T3[t][1]) + (c[l] + w) * ph.T3[t][2];
          const R Vm = ((ph.T3[t][0] + ph.T3[t][1]) + ph.T3[t][2]) / R(3);
          const R prod = (c[i] * c[j]) * c[l];
          part[i] = part[i] + (c[j] * c[l]) * Wt + prod * (ph.T3[t][0] - Vm);
          part[j] = part[j] + (c[i] * c[l]) * Wt + prod * (ph.T3[t][1] - Vm);
          part[l] = part[l] + (c[i] * c[j]) * Wt + prod * (ph.T3[t][2] - Vm);
        }
      }
    }
  }
#pragma unroll
  for (int l = 0; l < N - 1; ++l) dG[l] = part[l] - part[N - 1];
}

// FP32 form of the inline Redlich-Kister phases, converted once on the host
// (the FP32 kernels then do no double arithmetic per cell): coefficients as
// float, and which temperature terms occur at all.
struct PhaseF {
  float g0[PPF_MAX_COMP][5];
  float L[6][3][2];
  float R;
};
struct ThermoF {
  PhaseF liquid, solid;
  int32_t has_TlnT;  // any c T ln T term
  int32_t has_more;  // any d T^2 or e / T term
  float inv_vm, D_L, dD, inv_R8314;  // 1/v_m, D_L, D_S - D_L, 1/8.314
};

inline PhaseF phase_f(const ppf_rk_phase& ph) {
  PhaseF o{};
  for (int i = 0; i < PPF_MAX_COMP; ++i)
    for (int j = 0; j < 5; ++j) o.g0[i][j] = float(ph.g0[i][j]);
  for (int p = 0; p < 6; ++p)
    for (int k = 0; k < 3; ++k)
      for (int j = 0; j < 2; ++j) o.L[p][k][j] = float(ph.L[p][k][j]);
  o.R = float(ph.R);
  return o;
}

inline ThermoF thermo_f(const ppf_thermo& th, const ppf_ncomp_params& p) {
  ThermoF f{};
  f.liquid = phase_f(th.liquid);
  f.solid = phase_f(th.solid);
  for (int i = 0; i < PPF_MAX_COMP; ++i) {
    for (const ppf_rk_phase* ph : {&th.liquid, &th.solid}) {
      if (ph->g0[i][2] != 0.0) f.has_TlnT = 1;
      if (ph->g0[i][3] != 0.0 || ph->g0[i][4] != 0.0) f.has_more = 1;
    }
  }
  f.inv_vm = float(1.0 / p.v_m);
  f.D_L = float(p.D_L);
  f.dD = float(p.D_S - p.D_L);
  f.inv_R8314 = float(1.0 / 8.314);
  return f;
}

struct NcK {
  ppf_ncomp_params p;
  ppf_thermo th;
  ThermoF tf;  // FP32 kernels without TDB tables
};

inline NcK make_nck(const ppf_ncomp_params& p, const ppf_thermo& th) {
  NcK k;
  k.p = p;
  k.th = th;
  k.tf = thermo_f(th, p);
  return k;
}

// FP32 phases at one temperature from the float coefficients
template <int N>
__device__ __forceinline__ void phase_at_f(const PhaseF& ph, float T, float lnT, bool tlnt, bool more,
                                           PhaseAtT<float, N, false>& o) {
#pragma unroll
  for (int i = 0; i < N; ++i) {
    float g = __builtin_fmaf(ph.g0[i][1], T, ph.g0[i][0]);
    if (tlnt) g = __builtin_fmaf(ph.g0[i][2], T * lnT, g);
    if (more) g = g + ph.g0[i][3] * T * T + ph.g0[i][4] / T;
    o.gi[i] = g;
  }
#pragma unroll
  for (int p = 0; p < N * (N - 1) / 2; ++p) {
#pragma unroll
    for (int k = 0; k < 3; ++k) o.L[p][k] = __builtin_fmaf(ph.L[p][k][1], T, ph.L[p][k][0]);
  }
  o.RT = ph.R * T;
}

// natural log of a positive normal float: the hardware log2 scaled by ln 2
// (the FP32 engines clamp compositions to >= 1e-10 and T is O(1e3))
__device__ __forceinline__ float nlogf(float x) { return __builtin_amdgcn_logf(x) * 0.693147180559945309f; }

// FP32 thermodynamics of one cell from the float coefficients (analytic
// derivatives; same quantities and order of terms as thermo_cell)
template <int N>
__device__ __forceinline__ void thermo_cell_f32(const float (&cin)[N - 1], float phi, float T, const NcK& k,
                                                int64_t step, uint64_t seed, uint64_t ncells, uint64_t cell,
                                                float* dG, float* Mc, float* dF) {
  const ThermoF& f = k.tf;
  float c[N];
  float cN = 1.f;
#pragma unroll
  for (int l = 0; l < N - 1; ++l) {
    c[l] = cin[l];
    cN = cN - cin[l];
  }
  c[N - 1] = cN;
  const float lnT = f.has_TlnT ? nlogf(T) : 0.f;
  PhaseAtT<float, N, false> PL, PS;
  phase_at_f<N>(f.liquid, T, lnT, f.has_TlnT != 0, f.has_more != 0, PL);
  phase_at_f<N>(f.solid, T, lnT, f.has_TlnT != 0, f.has_more != 0, PS);
  // ln max(c, 1e-10) once per component, shared by G and both gradients
  float cc[N], lc[N];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    cc[i] = fmaxf(c[i], float(kCMin));
    lc[i] = nlogf(cc[i]);
  }
  float ideal = 0.f;
#pragma unroll
  for (int i = 0; i < N; ++i) ideal = __builtin_fmaf(cc[i], lc[i], ideal);
  float GL = 0.f, GS = 0.f;
  float pL[N], pS[N];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    GL = __builtin_fmaf(c[i], PL.gi[i], GL);
    GS = __builtin_fmaf(c[i], PS.gi[i], GS);
    const float dmix = c[i] > float(kCMin) ? lc[i] + 1.f : 0.f;
    pL[i] = __builtin_fmaf(PL.RT, dmix, PL.gi[i]);
    pS[i] = __builtin_fmaf(PS.RT, dmix, PS.gi[i]);
  }
  int pr = 0;
#pragma unroll
  for (int i = 0; i < N; ++i) {
#pragma unroll
    for (int j = i + 1; j < N; ++j, ++pr) {
      const float dd = c[i] - c[j], cij = c[i] * c[j];
      const float PLv = __builtin_fmaf(dd, __builtin_fmaf(dd, PL.L[pr][2], PL.L[pr][1]), PL.L[pr][0]);
      const float PSv = __builtin_fmaf(dd, __builtin_fmaf(dd, PS.L[pr][2], PS.L[pr][1]), PS.L[pr][0]);
      const float dPL = __builtin_fmaf(2.f * dd, PL.L[pr][2], PL.L[pr][1]);
      const float dPS = __builtin_fmaf(2.f * dd, PS.L[pr][2], PS.L[pr][1]);
      GL = __builtin_fmaf(cij, PLv, GL);
      GS = __builtin_fmaf(cij, PSv, GS);
      pL[i] = pL[i] + __builtin_fmaf(cij, dPL, c[j] * PLv);
      pL[j] = pL[j] + __builtin_fmaf(-cij, dPL, c[i] * PLv);
      pS[i] = pS[i] + __builtin_fmaf(cij, dPS, c[j] * PSv);
      pS[j] = pS[j] + __builtin_fmaf(-cij, dPS, c[i] * PSv);
    }
  }
  // G_S - G_L: the ideal-mixing terms enter each phase as in gibbs()
  *dG = __builtin_fmaf(PS.RT, ideal, GS) - __builtin_fmaf(PL.RT, ideal, GL);
  const float phi2 = phi * phi;
  const float g = phi2 * ((1.f - phi) * (1.f - phi));
  const float h = phi2 * phi * __builtin_fmaf(phi, __builtin_fmaf(6.f, phi, -15.f), 10.f);
  const ppf_ncomp_params& p = k.p;
  const float mob = float(p.v_m) * __builtin_fmaf(h, f.dD, f.D_L) * (f.inv_R8314 * __builtin_amdgcn_rcpf(T));
  const float gT = g * T;
  const bool first = (p.flags & PPF_NC_DERIV_FIRST) != 0, shared = (p.flags & PPF_NC_SHARED_NOISE) != 0;
#pragma unroll
  for (int l = 0; l < N - 1; ++l) {
    Mc[l] = mob * cin[l];
    const int ld = first ? 0 : l;
    const float dGL = pL[ld] - pL[N - 1], dGS = pS[ld] - pS[N - 1];
    float noise = 0.f;
    if (p.noise_c != 0.0) {
      const float xi = normal01<float>(hash64(seed, rng_counter(step, S_NC_C0 + (shared ? 0 : l), ncells, cell)));
      noise = float(p.noise_c) * sqrtf(2.f * 8.314f * T * f.inv_vm) * xi;
    }
    dF[l] = __builtin_fmaf(__builtin_fmaf(h, dGS - dGL, dGL), f.inv_vm,
                           (float(p.W[l]) - float(p.W[N - 1])) * gT) + noise;
  }
}

// G_S - G_L and the chemical-potential derivatives of both phases at c:
// the h = 1e-7 forward differences of NCGPU_new.py:301-313 with its in-place
// perturb / restore sequence (fd; with `first` only its first pass, the
// legacy engine's NCGPU.py:287-290), or the exact derivatives
template <typename R, int N, bool EXT>
__device__ __forceinline__ void phase_thermo(const PhaseAtT<R, N, EXT>& PL, const PhaseAtT<R, N, EXT>& PS, bool fd,
                                             bool first, R (&c)[N], R* dG, R (&dGL)[N], R (&dGS)[N]) {
  R id0 = ideal_sum<R, N>(c);
  R G_L = gibbs(c, PL, id0);
  R G_S = gibbs(c, PS, id0);
  *dG = G_S - G_L;
  if (fd) {
    const R hstep = R(0.0000001);
    c[N - 1] = c[N - 1] - hstep;
#pragma unroll
    for (int l = 0; l < N - 1; ++l) {
      if (first && l > 0) break;
      c[l] = c[l] + hstep;
      R id = ideal_sum<R, N>(c);
      dGL[l] = R(10000000.) * (gibbs(c, PL, id) - G_L);
      dGS[l] = R(10000000.) * (gibbs(c, PS, id) - G_S);
      c[l] = c[l] - hstep;
    }
  } else {
    gibbs_grad(c, PL, dGL);
    gibbs_grad(c, PS, dGS);
  }
}

// TAB: the phases come from TDB tables (k.th.tables), else from the inline
// Redlich-Kister coefficients; the two are separate kernels, so neither
// path's registers are allocated for the other
template <typename R, int N, bool TAB>
__device__ __forceinline__ void thermo_cell(const R (&cin)[N - 1], R phi, R T, const NcK& k, int64_t step,
                                            uint64_t seed, uint64_t ncells, uint64_t cell, R* dG, R* Mc,
                                            R* dF) {
  if constexpr (sizeof(R) == 4 && !TAB) {
    thermo_cell_f32<N>(cin, phi, T, k, step, seed, ncells, cell, dG, Mc, dF);
    return;
  }
  const ppf_ncomp_params& p = k.p;
  R c[N];
  R cN = R(1);
#pragma unroll
  for (int l = 0; l < N - 1; ++l) {
    c[l] = cin[l];
    cN = cN - cin[l];
  }
  c[N - 1] = cN;
  R lnT = needs_lnT(k.th, N) ? rlog<R>(T) : R(0);
  // legacy engine: one derivative, along e_1 - e_N, for every component
  // (NCGPU.py:287-290, 303)
  const bool first = (p.flags & PPF_NC_DERIV_FIRST) != 0, shared = (p.flags & PPF_NC_SHARED_NOISE) != 0;
  R dGL[N], dGS[N];
  if constexpr (TAB) {
    PhaseAtT<R, N, true> PL, PS;
    phase_at_tables<R, N>(&k.th.tables->liquid, R(k.th.liquid.R), T, lnT, PL);
    phase_at_tables<R, N>(&k.th.tables->solid, R(k.th.solid.R), T, lnT, PS);
    phase_thermo(PL, PS, k.th.dg_mode == PPF_DG_FORWARD_DIFF, first, c, dG, dGL, dGS);
  } else {
    PhaseAtT<R, N, false> PL, PS;
    phase_at<R, N>(k.th.liquid, T, lnT, PL);
    phase_at<R, N>(k.th.solid, T, lnT, PS);
    phase_thermo(PL, PS, k.th.dg_mode == PPF_DG_FORWARD_DIFF, first, c, dG, dGL, dGS);
  }
  // grouping of NCGPU_new.py:298-299 (numba integer powers are products)
  R phi3 = phi * phi * phi;
  R g = (phi * phi) * ((R(1) - phi) * (R(1) - phi));
  R h = phi3 * (R(6) * (phi * phi) - R(15) * phi + R(10));
  if (first) {
#pragma unroll
    for (int l = 1; l < N - 1; ++l) {
      dGL[l] = dGL[0];
      dGS[l] = dGS[0];
    }
  }
  const R v_m = R(p.v_m), D_L = R(p.D_L), D_S = R(p.D_S);
#pragma unroll
  for (int l = 0; l < N - 1; ++l) {
    Mc[l] = v_m * cin[l] * (D_L + h * (D_S - D_L)) / (R(8.314) * T);
    R noise = R(0);
    if (p.noise_c != 0.0) {
      R xi = normal01<R>(hash64(seed, rng_counter(step, S_NC_C0 + (shared ? 0 : l), ncells, cell)));
      noise = R(p.noise_c) * sqrt(R(2.) * R(8.314) * T / v_m) * xi;
    }
    dF[l] = (dGL[l] + h * (dGS[l] - dGL[l])) / v_m + (R(p.W[l]) - R(p.W[N - 1])) * g * T + noise;
  }
}

// One thread per padded cell, 64 x 4 cells per block.  Blocks whose cells
// are all interior read the fields directly; the others fold the boundary
// rules in (the halo ring of the transfer arrays is computed from folded
// values, as the reference helper does from its BC-filled halos).
template <typename R, int N, bool TAB>
__device__ __forceinline__ void ncomp_thermo_body(const R* __restrict__ in, const R* __restrict__ b, Dom d, R dx,
                                                  ppf_temperature t, R* __restrict__ tr, int64_t tsy, int64_t tsf,
                                                  NcK k, uint64_t seed, StepRef sr) {
  const int64_t step = sr.get();
  const int64_t px = d.ax.n + 2, py = d.ay.n + 2;
  const uint32_t nbx = uint32_t((px + 63) / 64);
  const uint32_t blk = xcd_swizzle(blockIdx.x, gridDim.x);
  const int64_t xb = int64_t(blk % nbx) * 64, yb = int64_t(blk / nbx) * 4;
  const int64_t x = xb + (threadIdx.x & 63), y = yb + (threadIdx.x >> 6);
  if (x >= px || y >= py) return;
  const bool interior = xb >= 1 && xb + 63 <= d.ax.n && yb >= 1 && yb + 3 <= d.ay.n;
  R phi, cin[N - 1];
  if (interior) {
    const int64_t o = y * d.ay.stride + x;
    phi = in[o];
#pragma unroll
    for (int l = 0; l < N - 1; ++l) cin[l] = in[o + (3 + l) * d.sf];
  } else {
    phi = fold(in, b, d, x, y, int64_t(0), dx);
#pragma unroll
    for (int l = 0; l < N - 1; ++l)
      cin[l] = fold(in + (3 + l) * d.sf, b ? b + (3 + l) * d.sf : nullptr, d, x, y, int64_t(0), dx);
  }
  int64_t n3[3] = {d.ax.n, d.ay.n, 1};
  int bc3[3] = {d.ax.bc, d.ay.bc, PPF_PERIODIC};
  R T = temperature_at<R>(t, n3, bc3, x, y, 1, double(dx), step);
  R dG, Mc[N - 1], dF[N - 1];
  const uint64_t cell = uint64_t(y) * uint64_t(px) + uint64_t(x);
  thermo_cell<R, N, TAB>(cin, phi, T, k, step, seed, uint64_t(px * py), cell, &dG, Mc, dF);
  const int64_t o = y * tsy + x;
  tr[o] = dG;
#pragma unroll
  for (int l = 0; l < N - 1; ++l) {
    tr[o + (1 + l) * tsf] = Mc[l];
    tr[o + (N + l) * tsf] = dF[l];
  }
}

template <typename R, int N>
__global__ __launch_bounds__(256) void k_ncomp_thermo(const R* __restrict__ in, const R* __restrict__ b, Dom d,
                                                      R dx, ppf_temperature t, R* __restrict__ tr, int64_t tsy,
                                                      int64_t tsf, NcK k, uint64_t seed, StepRef sr) {
  ncomp_thermo_body<R, N, false>(in, b, d, dx, t, tr, tsy, tsf, k, seed, sr);
}
template <typename R, int N>
__global__ __launch_bounds__(256) void k_ncomp_thermo_tdb(const R* __restrict__ in, const R* __restrict__ b, Dom d,
                                                          R dx, ppf_temperature t, R* __restrict__ tr, int64_t tsy,
                                                          int64_t tsf, NcK k, uint64_t seed, StepRef sr) {
  ncomp_thermo_body<R, N, true>(in, b, d, dx, t, tr, tsy, tsf, k, seed, sr);
}

template <typename R>
int ncomp_thermo(const ppf_grid& g, const void* in, const void* bcvals, const ppf_temperature& t, const ppf_grid& tg,
                 void* transfer,
                 const ppf_ncomp_params& p, const ppf_thermo& th, uint64_t seed, StepRef step, hipStream_t s) {
  Dom d = make_dom(g);
  const NcK k = make_nck(p, th);
  dim3 grid(unsigned(((g.n[0] + 2 + 63) / 64) * ((g.n[1] + 2 + 3) / 4)));
  const R* fin = static_cast<const R*>(in);
  const R* bc = static_cast<const R*>(bcvals);
  R* tr = static_cast<R*>(transfer);
  switch (p.ncomp) {
    case 2:
      hipLaunchKernelGGL((th.tables ? k_ncomp_thermo_tdb<R, 2> : k_ncomp_thermo<R, 2>), grid, dim3(256), 0, s, fin,
                         bc, d, R(g.dx), t, tr, tg.sy, tg.sf, k, seed, step);
      break;
    case 3:
      hipLaunchKernelGGL((th.tables ? k_ncomp_thermo_tdb<R, 3> : k_ncomp_thermo<R, 3>), grid, dim3(256), 0, s, fin,
                         bc, d, R(g.dx), t, tr, tg.sy, tg.sf, k, seed, step);
      break;
    case 4:
      hipLaunchKernelGGL((th.tables ? k_ncomp_thermo_tdb<R, 4> : k_ncomp_thermo<R, 4>), grid, dim3(256), 0, s, fin,
                         bc, d, R(g.dx), t, tr, tg.sy, tg.sf, k, seed, step);
      break;
    default:
      return fail(PPF_EINVAL, "ncomp_thermo: ncomp must be 2..4");
  }
  return check_launch("ncomp_thermo");
}

// ------------------------------------------------------------ main step --
// NC tile: 64 x PPF_NC_TY outputs, PPF_NC_TY / 4 rows per thread (on 2048^2
// FP32 a 12-row tile measured 2.5 % faster than 8 rows and 5 % faster than 16)
#ifndef PPF_NC_TY
#define PPF_NC_TY 12
#endif
constexpr int kNcTY = PPF_NC_TY;
// 256 threads cover 64 columns x 4 row groups; each thread walks kNcTY / 4 rows
static_assert(kNcTY % 4 == 0 && kNcTY >= 4, "PPF_NC_TY must be a positive multiple of 4");

template <typename R>
__device__ __forceinline__ R divagradb(R a, R axp, R axm, R ayp, R aym, R b, R bxp, R bxm, R byp, R bym, R idx) {
  // NCGPU_new.py:28-30
  return R(0.5) * idx * idx *
         ((axp + a) * (bxp - b) - (a + axm) * (b - bxm) + (ayp + a) * (byp - b) - (a + aym) * (b - bym));
}

template <typename R>
__device__ __forceinline__ R f_ori_term(R D_q, R D_q_xp, R D_q_xm, R D_q_yp, R D_q_ym, R mgq_xp, R mgq_xm,
                                        R mgq_yp, R mgq_ym, R q, R q_xp, R q_xm, R q_yp, R q_ym, R idx) {
  // NCGPU_new.py:33-35
  return R(0.5) * idx * idx *
         ((D_q + D_q_xp) * (q_xp - q) / mgq_xp - (D_q + D_q_xm) * (q - q_xm) / mgq_xm +
          (D_q + D_q_yp) * (q_yp - q) / mgq_yp - (D_q + D_q_ym) * (q - q_ym) / mgq_ym);
}

template <typename R>
__device__ __forceinline__ R fsqrt(R x) {
  if constexpr (sizeof(R) == 4) {
    return __builtin_amdgcn_sqrtf(x);
  } else {
    return sqrt(x);
  }
}

template <typename R, int W>
struct Nb5 {
  // centre and neighbours of an LDS tile at (ly, lx)
  R c, xp, xm, yp, ym;
  __device__ __forceinline__ Nb5(const R (*s)[W], int ly, int lx)
      : c(s[ly][lx]), xp(s[ly][lx + 1]), xm(s[ly][lx - 1]), yp(s[ly + 1][lx]), ym(s[ly - 1][lx]) {}
};

// occupancy hint for the step (waves per SIMD the register allocation must
// allow; 0 = compiler's choice) -- chosen by sweep
#ifndef PPF_NC_STEP_WAVES
#define PPF_NC_STEP_WAVES 4
#endif
#if PPF_NC_STEP_WAVES
#define PPF_NC_STEP_ATTR __attribute__((amdgpu_waves_per_eu(PPF_NC_STEP_WAVES, 8)))
#else
#define PPF_NC_STEP_ATTR
#endif

// Per-launch constants of the step's cell update, in the arithmetic type.
template <typename R>
struct NcStepC {
  R idx, v_m, inv_v_m, M_qmax, Hh, y_e, beta, hidx, idx2, qidx2, Hidx, beta_g, dt, ebar2, eqbar2;
  __device__ __forceinline__ NcStepC(const ppf_ncomp_params& p, R dx) {
    idx = R(1) / dx;
    v_m = R(p.v_m);
    inv_v_m = R(1.0 / p.v_m);
    M_qmax = R(p.M_qmax);
    Hh = R(p.H);
    y_e = R(p.y_e);
    beta = R(p.beta);
    // FP32 scale factors: 0.5/dx, 1/dx^2, 0.25/dx^2, H/dx, and the clamp of
    // the face |grad q| in grid units (beta dx); FP64 clamps |grad q| itself
    hidx = R(0.5) * idx;
    idx2 = idx * idx;
    qidx2 = R(0.25) * idx2;
    Hidx = Hh * idx;
    beta_g = sizeof(R) == 8 ? beta : R(p.beta * double(dx));
    dt = R(p.dt);
    const R ebar = R(p.ebar), eqbar = R(p.eqbar);
    ebar2 = ebar * ebar;
    eqbar2 = eqbar * eqbar;
    if (p.flags & PPF_NC_EBAR_FROM_S) {
      // the legacy engine's values, in its operation order (NCGPU.py:74-75)
      ebar2 = R(6.0 * 1.4142135623730951 * p.S[1] * p.d / p.T_M[1]);
      eqbar2 = R(0.25) * ebar2;
    }
  }
};

// The neighbourhood one cell update reads, through an accessor evaluated at
// the point of use: ph(i, j) = phi at row y-1+i, column x-1+j; T(), q1(),
// q4(), c(l), Mc(l), dF(l) the 5-point stencils; dG() = G_S - G_L at the
// centre.
template <typename R>
struct Nb5v {
  R c, xp, xm, yp, ym;
};

template <typename R, int W>
__device__ __forceinline__ Nb5v<R> lds5(const R (*s)[W], int ly, int lx) {
  return Nb5v<R>{s[ly][lx], s[ly][lx + 1], s[ly][lx - 1], s[ly + 1][lx], s[ly - 1][lx]};
}

// One cell of NComponent_kernel (NCGPU_new.py:84-249) from its neighbourhood:
// new phi, q1, q4 (renormalised) and compositions.  `cell` keys the noise.
template <typename R, int N, typename Acc>
__device__ __forceinline__ void nc_cell(const Acc& v, const NcStepC<R>& K, const ppf_ncomp_params& p,
                                        uint64_t seed, int64_t step, uint64_t ncells, uint64_t cell, R& phi_out,
                                        R& q1_out, R& q4_out, R (&cn)[N - 1]) {
  const R idx = K.idx, v_m = K.v_m, inv_v_m = K.inv_v_m, M_qmax = K.M_qmax, Hh = K.Hh, y_e = K.y_e;
  const R hidx = K.hidx, idx2 = K.idx2, qidx2 = K.qidx2, Hidx = K.Hidx, beta_g = K.beta_g, dt = K.dt;
  const R ebar2 = K.ebar2, eqbar2 = K.eqbar2;
  R pc = v.ph(1, 1);
  // interpolating functions (NCGPU_new.py:84-87)
  R pc3 = pc * pc * pc;
  R g = (pc * pc) * ((R(1) - pc) * (R(1) - pc));
  R h, gprime;
  if constexpr (sizeof(R) == 8) {
    h = pc3 * (R(6) * (pc * pc) - R(15) * pc + R(10));
    gprime = R(4) * pc3 - R(6) * (pc * pc) + R(2) * pc;
  } else {  // Horner forms
    h = pc3 * (pc * (R(6) * pc - R(15)) + R(10));
    gprime = pc * (pc * (R(4) * pc - R(6)) + R(2));
  }
  R hprime = R(30) * g;
  // gradients (:89-107); FP32 takes the uniform scale factors as one
  // product each (0.5/dx, 1/dx^2, 0.25/dx^2), FP64 the reference's order
  R dphidx, dphidy, d2phidx2, d2phidy2, d2phidxy;
  if constexpr (sizeof(R) == 8) {
    dphidx = R(0.5) * (v.ph(1, 2) - v.ph(1, 0)) * idx;
    dphidy = R(0.5) * (v.ph(2, 1) - v.ph(0, 1)) * idx;
    d2phidx2 = (v.ph(1, 2) + v.ph(1, 0) - R(2) * pc) * idx * idx;
    d2phidy2 = (v.ph(2, 1) + v.ph(0, 1) - R(2) * pc) * idx * idx;
    d2phidxy = R(0.25) * (v.ph(2, 2) - v.ph(2, 0) - v.ph(0, 2) + v.ph(0, 0)) *
               idx * idx;
  } else {
    dphidx = hidx * (v.ph(1, 2) - v.ph(1, 0));
    dphidy = hidx * (v.ph(2, 1) - v.ph(0, 1));
    d2phidx2 = (v.ph(1, 2) + v.ph(1, 0) - R(2) * pc) * idx2;
    d2phidy2 = (v.ph(2, 1) + v.ph(0, 1) - R(2) * pc) * idx2;
    d2phidxy = qidx2 * (v.ph(2, 2) - v.ph(2, 0) - v.ph(0, 2) + v.ph(0, 0));
  }
  R dphidx2 = dphidx * dphidx;
  R dphidx3 = dphidx2 * dphidx;
  R dphidy2 = dphidy * dphidy;
  R dphidy3 = dphidy2 * dphidy;
  R Tc = v.T5().c;
  R dTdx = R(0.5) * idx * (v.T5().xp - v.T5().xm);
  R dTdy = R(0.5) * idx * (v.T5().yp - v.T5().ym);
  R lphi = d2phidx2 + d2phidy2;
  R mag_grad_phi2 = dphidx * dphidx + dphidy * dphidy;
  if (mag_grad_phi2 < R(1e-6)) mag_grad_phi2 = R(1e-6);
  R mag_grad_phi4 = mag_grad_phi2 * mag_grad_phi2;
  R mag_grad_phi8 = mag_grad_phi4 * mag_grad_phi4;
  // orientation gradients (:109-125)
  const Nb5v<R> Q1 = v.q1();
  const Nb5v<R> Q4 = v.q4();
  R dq1dx = R(0.5) * idx * (Q1.xp - Q1.xm);
  R dq1dy = R(0.5) * idx * (Q1.yp - Q1.ym);
  R dq4dx = R(0.5) * idx * (Q4.xp - Q4.xm);
  R dq4dy = R(0.5) * idx * (Q4.yp - Q4.ym);
  // face |grad q|; FP32 keeps them in grid units (r = dx |grad q|): the
  // 1/dx goes into mag_grad_q's factor and the face weights' common factor
  R mgq_xp = fsqrt((Q1.xp - Q1.c) * (Q1.xp - Q1.c) + (Q4.xp - Q4.c) * (Q4.xp - Q4.c));
  R mgq_xm = fsqrt((Q1.xm - Q1.c) * (Q1.xm - Q1.c) + (Q4.xm - Q4.c) * (Q4.xm - Q4.c));
  R mgq_yp = fsqrt((Q1.yp - Q1.c) * (Q1.yp - Q1.c) + (Q4.yp - Q4.c) * (Q4.yp - Q4.c));
  R mgq_ym = fsqrt((Q1.ym - Q1.c) * (Q1.ym - Q1.c) + (Q4.ym - Q4.c) * (Q4.ym - Q4.c));
  R mag_grad_q;
  if constexpr (sizeof(R) == 8) {
    mgq_xp = idx * mgq_xp;
    mgq_xm = idx * mgq_xm;
    mgq_yp = idx * mgq_yp;
    mgq_ym = idx * mgq_ym;
    mag_grad_q = R(0.5) * (mgq_xp + mgq_xm + mgq_yp + mgq_ym);
  } else {
    mag_grad_q = hidx * (mgq_xp + mgq_xm + mgq_yp + mgq_ym);
  }
  if (mgq_xp < beta_g) mgq_xp = beta_g;
  if (mgq_xm < beta_g) mgq_xm = beta_g;
  if (mgq_yp < beta_g) mgq_yp = beta_g;
  if (mgq_ym < beta_g) mgq_ym = beta_g;
  // psi terms (:127-146)
  R q1c = Q1.c, q4c = Q4.c;
  R q2q2 = (q1c * q1c - q4c * q4c);
  R qq2 = R(2) * q1c * q4c;
  R psix = q2q2 * dphidx - qq2 * dphidy;
  R psiy = qq2 * dphidx + q2q2 * dphidy;
  R psix2 = psix * psix;
  R psix3 = psix2 * psix;
  R psix4 = psix2 * psix2;
  R psiy2 = psiy * psiy;
  R psiy3 = psiy2 * psiy;
  R psiy4 = psiy2 * psiy2;
  R dq2q2dx = (R(2) * q1c * dq1dx - R(2) * q4c * dq4dx);
  R dqq2dx = (R(2) * q4c * dq1dx + R(2) * q1c * dq4dx);
  R dq2q2dy = (R(2) * q1c * dq1dy - R(2) * q4c * dq4dy);
  R dqq2dy = (R(2) * q4c * dq1dy + R(2) * q1c * dq4dy);
  R dpsixdx = dq2q2dx * dphidx + q2q2 * d2phidx2 - dqq2dx * dphidy - qq2 * d2phidxy;
  R dpsixdy = dq2q2dy * dphidx + q2q2 * d2phidxy - dqq2dy * dphidy - qq2 * d2phidy2;
  R dpsiydx = dq2q2dx * dphidy + q2q2 * d2phidxy + dqq2dx * dphidx + qq2 * d2phidx2;
  R dpsiydy = dq2q2dy * dphidy + q2q2 * d2phidy2 + dqq2dy * dphidx + qq2 * d2phidxy;
  R d_term_dx, d_term_dy;
  if constexpr (sizeof(R) == 8) {
    // reference form (oracle parity)
    R psi4 = psix4 + psiy4;
    // anisotropic terms (:148-160)
    d_term_dx = dTdx * ((R(2) * psix3 * q2q2 + R(2) * psiy3 * qq2) / mag_grad_phi2 - dphidx * psi4 / mag_grad_phi4);
    d_term_dx = d_term_dx + Tc *
                                (R(6) * psix2 * dpsixdx * q2q2 + R(2) * psix3 * dq2q2dx +
                                 R(6) * psiy2 * dpsiydx * qq2 + R(2) * psiy3 * dqq2dx) /
                                mag_grad_phi2;
    d_term_dx = d_term_dx - Tc * (R(2) * psix3 * q2q2 + R(2) * psiy3 * qq2) *
                                (R(2) * dphidx * d2phidx2 + R(2) * dphidy * d2phidxy) / mag_grad_phi4;
    d_term_dx = d_term_dx - Tc * (d2phidx2 * psi4 + dphidx * (R(4) * psix3 * dpsixdx + R(4) * psiy3 * dpsiydx)) /
                                mag_grad_phi4;
    d_term_dx = d_term_dx + Tc *
                                (dphidx * psi4 *
                                 (R(4) * dphidx3 * d2phidx2 + R(4) * dphidx * d2phidx2 * dphidy2 +
                                  R(4) * dphidy * d2phidxy * dphidx2 + R(4) * dphidy3 * d2phidxy)) /
                                mag_grad_phi8;
    d_term_dx = d_term_dx * (R(4) * y_e * ebar2);
    d_term_dy =
        dTdy * ((R(-2) * psix3 * qq2 + R(2) * psiy3 * q2q2) / mag_grad_phi2 - dphidy * psi4 / mag_grad_phi4);
    d_term_dy = d_term_dy + Tc *
                                (R(-6) * psix2 * dpsixdy * qq2 - R(2) * psix3 * dqq2dy +
                                 R(6) * psiy2 * dpsiydy * q2q2 + R(2) * psiy3 * dq2q2dy) /
                                mag_grad_phi2;
    d_term_dy = d_term_dy - Tc * (R(-2) * psix3 * qq2 + R(2) * psiy3 * q2q2) *
                                (R(2) * dphidx * d2phidxy + R(2) * dphidy * d2phidy2) / mag_grad_phi4;
    d_term_dy = d_term_dy - Tc * (d2phidy2 * psi4 + dphidy * (R(4) * psix3 * dpsixdy + R(4) * psiy3 * dpsiydy)) /
                                mag_grad_phi4;
    d_term_dy = d_term_dy + Tc *
                                (dphidy * psi4 *
                                 (R(4) * dphidx3 * d2phidxy + R(4) * dphidx * d2phidxy * dphidy2 +
                                  R(4) * dphidy * d2phidy2 * dphidx2 + R(4) * dphidy3 * d2phidy2)) /
                                mag_grad_phi8;
    d_term_dy = d_term_dy * (R(4) * y_e * ebar2);
  } else {
    // FP32: the same algebra with gradients and psi divided by m = sqrt of the
    // clamped |grad phi|^2 (every term is homogeneous in that scale), so no
    // |grad phi|^8 is formed; with dx ~ 5e-6 cm the reference form exceeds
    // the FP32 range.
    const R im = __builtin_amdgcn_rsqf(mag_grad_phi2);
    const R m = mag_grad_phi2 * im;
    // Common factors taken out: with s = X^2 + Y^2, u_x = X phi_xx + Y phi_xy
    // and u_y = X phi_xy + Y phi_yy, the last term of each component,
    // 4 X^3 phi_xx + 4 X phi_xx Y^2 + 4 Y phi_xy X^2 + 4 Y^3 phi_xy, is
    // 4 s u_x (and 4 s u_y), and the third is 2 A u.
    R X = dphidx * im, Yg = dphidy * im, Px = psix * im, Py = psiy * im;
    R Px2 = Px * Px, Py2 = Py * Py, Px3 = Px2 * Px, Py3 = Py2 * Py;
    R P4 = Px2 * Px2 + Py2 * Py2;
    R Ax = R(2) * (Px3 * q2q2 + Py3 * qq2);
    R Ay = R(2) * (Py3 * q2q2 - Px3 * qq2);
    const R s4 = R(4) * P4 * (X * X + Yg * Yg);
    const R ux = X * d2phidx2 + Yg * d2phidxy, uy = X * d2phidxy + Yg * d2phidy2;
    R tx = dTdx * m * (Ax - X * P4);
    tx += Tc * (R(6) * (Px2 * dpsixdx * q2q2 + Py2 * dpsiydx * qq2) + R(2) * m * (Px3 * dq2q2dx + Py3 * dqq2dx) -
                R(2) * Ax * ux - d2phidx2 * P4 - R(4) * X * (Px3 * dpsixdx + Py3 * dpsiydx) + X * s4 * ux);
    R ty = dTdy * m * (Ay - Yg * P4);
    ty += Tc * (R(6) * (Py2 * dpsiydy * q2q2 - Px2 * dpsixdy * qq2) + R(2) * m * (Py3 * dq2q2dy - Px3 * dqq2dy) -
                R(2) * Ay * uy - d2phidy2 * P4 - R(4) * Yg * (Px3 * dpsixdy + Py3 * dpsiydy) + Yg * s4 * uy);
    d_term_dx = tx * (R(4) * y_e * ebar2);
    d_term_dy = ty * (R(4) * y_e * ebar2);
  }
  // c_N, cW, M_phi (:162-173)
  R cW = R(0), c_N = R(1), M_phi = R(0);
#pragma unroll
  for (int l = 0; l < N - 1; ++l) {
    R ci = v.c(l).c;
    cW = cW + ci * R(p.W[l]);
    c_N = c_N - ci;
    M_phi = M_phi + ci * R(p.M[l]);
  }
  cW = cW + c_N * R(p.W[N - 1]);
  M_phi = M_phi + c_N * R(p.M[N - 1]);
  R M_q = M_qmax + (R(1e-6) - M_qmax) * h;
  // dphidt (:179-189)
  const R dG = v.dG();  // G_S - G_L
  R dphidt = ebar2 * (R(1) - R(3) * y_e) * (Tc * lphi + dTdx * dphidx + dTdy * dphidy);
  dphidt = dphidt + d_term_dx + d_term_dy;
  if constexpr (sizeof(R) == 8) {
    dphidt = dphidt - hprime * dG / v_m;
  } else {
    dphidt = dphidt - hprime * dG * inv_v_m;
  }
  dphidt = dphidt - gprime * Tc * cW;
  dphidt = dphidt - R(4) * Hh * Tc * pc * mag_grad_q;
  dphidt = dphidt * M_phi;
  if (p.noise_phi != 0.0) {
    R xi = normal01<R>(hash64(seed, rng_counter(step, S_NC_PHI, ncells, cell)));
    dphidt = dphidt + sqrt(R(2.) * R(8.314) * Tc * M_phi / v_m) * xi * R(p.noise_phi);
  }
  // dc_i/dt (:191-203) and update (:243-249)
#pragma unroll
  for (int l = 0; l < N - 1; ++l) {
    const Nb5v<R> Mc = v.Mc(l);
    const Nb5v<R> dF = v.dF(l);
    R acc = divagradb(Mc.c, Mc.xp, Mc.xm, Mc.yp, Mc.ym, dF.c, dF.xp, dF.xm, dF.yp, dF.ym, idx);
#pragma unroll
    for (int m = 0; m < N - 1; ++m) {
      const Nb5v<R> cj = v.c(m);
      const Nb5v<R> dFj = v.dF(m);
      acc = acc - divagradb(Mc.c * cj.c, Mc.xp * cj.xp, Mc.xm * cj.xm, Mc.yp * cj.yp, Mc.ym * cj.ym, dFj.c,
                            dFj.xp, dFj.xm, dFj.yp, dFj.ym, idx);
    }
    cn[l] = acc * dt + v.c(l).c;
  }
  // dq/dt (:205-232)
  R f_ori_1, f_ori_4;
  if constexpr (sizeof(R) == 8) {
    R D_q = R(2) * Hh * Tc * (pc * pc);
    R D_q_xp = R(2) * Hh * v.T5().xp * (v.ph(1, 2) * v.ph(1, 2));
    R D_q_xm = R(2) * Hh * v.T5().xm * (v.ph(1, 0) * v.ph(1, 0));
    R D_q_yp = R(2) * Hh * v.T5().yp * (v.ph(2, 1) * v.ph(2, 1));
    R D_q_ym = R(2) * Hh * v.T5().ym * (v.ph(0, 1) * v.ph(0, 1));
    f_ori_1 = f_ori_term(D_q, D_q_xp, D_q_xm, D_q_yp, D_q_ym, mgq_xp, mgq_xm, mgq_yp, mgq_ym, Q1.c, Q1.xp, Q1.xm,
                         Q1.yp, Q1.ym, idx);
    f_ori_4 = f_ori_term(D_q, D_q_xp, D_q_xm, D_q_yp, D_q_ym, mgq_xp, mgq_xm, mgq_yp, mgq_ym, Q4.c, Q4.xp, Q4.xm,
                         Q4.yp, Q4.ym, idx);
  } else {
    // FP32: D_q = 2 H T phi^2 and |grad q|_face = r / dx, so the face weight
    // (D_q + D_q,nb) / |grad q|_face is 2 H dx (e + e_nb) / r with e = T phi^2;
    // 2 H dx goes into the common factor 0.5 / dx^2 (giving H / dx), and
    // the four weights are shared by q1 and q4
    const R e = Tc * (pc * pc);
    const R wxp = (e + v.T5().xp * (v.ph(1, 2) * v.ph(1, 2))) * __builtin_amdgcn_rcpf(mgq_xp);
    const R wxm = (e + v.T5().xm * (v.ph(1, 0) * v.ph(1, 0))) * __builtin_amdgcn_rcpf(mgq_xm);
    const R wyp = (e + v.T5().yp * (v.ph(2, 1) * v.ph(2, 1))) * __builtin_amdgcn_rcpf(mgq_yp);
    const R wym = (e + v.T5().ym * (v.ph(0, 1) * v.ph(0, 1))) * __builtin_amdgcn_rcpf(mgq_ym);
    f_ori_1 = Hidx * (wxp * (Q1.xp - Q1.c) - wxm * (Q1.c - Q1.xm) + wyp * (Q1.yp - Q1.c) - wym * (Q1.c - Q1.ym));
    f_ori_4 = Hidx * (wxp * (Q4.xp - Q4.c) - wxm * (Q4.c - Q4.xm) + wyp * (Q4.yp - Q4.c) - wym * (Q4.c - Q4.ym));
  }
  R lq1, lq4;
  if constexpr (sizeof(R) == 8) {
    lq1 = (Q1.xp + Q1.xm + Q1.yp + Q1.ym - R(4) * Q1.c) * idx * idx;
    lq4 = (Q4.xp + Q4.xm + Q4.yp + Q4.ym - R(4) * Q4.c) * idx * idx;
  } else {
    lq1 = (Q1.xp + Q1.xm + Q1.yp + Q1.ym - R(4) * Q1.c) * idx2;
    lq4 = (Q4.xp + Q4.xm + Q4.yp + Q4.ym - R(4) * Q4.c) * idx2;
  }
  R noise_q1 = R(0), noise_q4 = R(0);
  if (p.noise_q != 0.0) {
    R amp = R(p.noise_q) * sqrt(R(0.0000000001) * R(8.314) * Tc / v_m);
    noise_q1 = amp * normal01<R>(hash64(seed, rng_counter(step, S_NC_Q1, ncells, cell)));
    noise_q4 = amp * normal01<R>(hash64(seed, rng_counter(step, S_NC_Q4, ncells, cell)));
  }
  R t1 = f_ori_1 + lq1 * eqbar2 + noise_q1;
  R t4 = f_ori_4 + lq4 * eqbar2 + noise_q4;
  R dq1dt = M_q * ((R(1) - q1c * q1c) * t1 - q1c * q4c * t4);
  R dq4dt = M_q * ((R(1) - q4c * q4c) * t4 - q1c * q4c * t1);
  R phin = pc + dt * dphidt;
  // NCGPU_new.py:234-237; the legacy engine's wider bounds, NCGPU.py:229-232
  const bool wide = (p.flags & PPF_NC_CLAMP_WIDE) != 0;
  const R plo = wide ? R(0.000001) : R(0.0001), phi_ = wide ? R(0.999999) : R(0.9999);
  if (phin < plo) phin = plo;
  if (phin > phi_) phin = phi_;
  R q1n = q1c + dt * dq1dt;
  R q4n = q4c + dt * dq4dt;
  phi_out = phin;
  if constexpr (sizeof(R) == 8) {
    R renorm = fsqrt(q1n * q1n + q4n * q4n);
    q1_out = q1n / renorm;
    q4_out = q4n / renorm;
  } else {
    const R irn = __builtin_amdgcn_rsqf(q1n * q1n + q4n * q4n);
    q1_out = q1n * irn;
    q4_out = q4n * irn;
  }
}

// Accessor of the LDS tile of ncomp_step_body at (ly, lx): every value is
// read where the update uses it.
template <typename R, int N, int W, int HH, bool FUSED>
struct NcLds {
  const R (*sf)[HH][W];
  const R (*sT)[W];
  const R (*st)[HH][W];
  const R* sG;  // FUSED: [kNcTY][kTX]
  const R* tr;
  int64_t tsy;
  int ly, lx;
  int64_t y, x;
  __device__ __forceinline__ R ph(int i, int j) const { return sf[0][ly - 1 + i][lx - 1 + j]; }
  __device__ __forceinline__ Nb5v<R> T5() const { return lds5(sT, ly, lx); }
  __device__ __forceinline__ Nb5v<R> q1() const { return lds5(sf[1], ly, lx); }
  __device__ __forceinline__ Nb5v<R> q4() const { return lds5(sf[2], ly, lx); }
  __device__ __forceinline__ Nb5v<R> c(int l) const { return lds5(sf[3 + l], ly, lx); }
  __device__ __forceinline__ Nb5v<R> Mc(int l) const { return lds5(st[l], ly, lx); }
  __device__ __forceinline__ Nb5v<R> dF(int l) const { return lds5(st[N - 1 + l], ly, lx); }
  __device__ __forceinline__ R dG() const {
    if constexpr (FUSED) {
      return sG[(ly - 1) * kTX + (lx - 1)];
    } else {
      return (tr + y * tsy)[x];
    }
  }
};

// The main step (NComponent_kernel) on a 64 x kNcTY tile.  The fields and T
// of the tile plus its +1 ring sit in LDS.  The thermodynamic quantities the
// step reads on that ring (M_c, dF/dc) and at its centre (G_S - G_L) come
// either from the transfer stack of ppf_ncomp_thermo (FUSED = false: the
// reference's two-kernel structure) or are evaluated here from the LDS tile
// by the same device function (FUSED = true: the thermodynamics is pointwise,
// so the ring needs no further halo and the transfer stack never reaches
// memory).
template <typename R, int N, bool FUSED, bool TAB>
__device__ __forceinline__ void ncomp_step_body(const R* __restrict__ in, R* __restrict__ out,
                                                const R* __restrict__ b, Dom d, R dx, ppf_temperature t,
                                                const R* __restrict__ tr, int64_t tsy, int64_t tsf, NcK k,
                                                uint64_t seed, StepRef sr) {
  constexpr int H = 1, W = kTX + 2 * H, HH = kNcTY + 2 * H;
  constexpr int NF = N + 2;         // phi, q1, q4, c_1..c_{N-1}
  constexpr int NT = 2 * (N - 1);   // M_c and dF/dc tiles
  __shared__ R sf[NF][HH][W];
  __shared__ R sT[HH][W];
  __shared__ R st[NT][HH][W];
  __shared__ R sG[FUSED ? kNcTY : 1][FUSED ? kTX : 1];  // G_S - G_L of the outputs (fused)
  const int64_t step = sr.get();

  uint32_t nbx = uint32_t((d.ax.n + kTX - 1) / kTX);
  uint32_t blk = xcd_swizzle(blockIdx.x, gridDim.x);
  int64_t x0 = 1 + int64_t(blk % nbx) * kTX;
  int64_t y0 = 1 + int64_t(blk / nbx) * kNcTY;
  bool interior = x0 - 1 >= 1 && x0 + kTX <= d.ax.n && y0 - 1 >= 1 && y0 + kNcTY <= d.ay.n;
  int64_t n3[3] = {d.ax.n, d.ay.n, 1};
  int bc3[3] = {d.ax.bc, d.ay.bc, PPF_PERIODIC};
  if (interior) {
    // every tile element of every array: one 32-bit offset per thread and slot,
    // computed once (the rows of the field and transfer stacks may differ in pitch)
    constexpr int NE = (W * HH + kThreads - 1) / kThreads;
    constexpr int NTL = FUSED ? 1 : NT;  // transfer arrays loaded
    int32_t eo[NE], et[NE];
#pragma unroll
    for (int j = 0; j < NE; ++j) {
      const int e = threadIdx.x + j * kThreads;
      const int ly = e / W, lx = e - ly * W;
      eo[j] = ly * int32_t(d.ay.stride) + lx;
      et[j] = ly * int32_t(tsy) + lx;
    }
    const R* fb = in + (y0 - 1) * d.ay.stride + (x0 - 1);
    const R* tb = FUSED ? nullptr : tr + (y0 - 1) * tsy + (x0 - 1);
    // all loads of the tile issued before the first LDS store (only the last
    // slot of each array is partial)
    constexpr int NFULL = (W * HH) / kThreads;
    R vf[NF][NE], vt[NTL][NE];
#pragma unroll
    for (int j = 0; j < NE; ++j) {
      const int e = threadIdx.x + j * kThreads;
      if (j < NFULL || e < W * HH) {
#pragma unroll
        for (int f = 0; f < NF; ++f) vf[f][j] = fb[f * d.sf + eo[j]];
        if constexpr (!FUSED) {
#pragma unroll
          for (int f = 0; f < NT; ++f) vt[f][j] = tb[(1 + f) * tsf + et[j]];
        }
      }
    }
#pragma unroll
    for (int j = 0; j < NE; ++j) {
      const int e = threadIdx.x + j * kThreads;
      if (j < NFULL || e < W * HH) {
#pragma unroll
        for (int f = 0; f < NF; ++f) (&sf[f][0][0])[e] = vf[f][j];
        if constexpr (!FUSED) {
#pragma unroll
          for (int f = 0; f < NT; ++f) (&st[f][0][0])[e] = vt[f][j];
        }
      }
    }
#pragma unroll
    for (int j = 0; j < NE; ++j) {
      const int e = threadIdx.x + j * kThreads;
      const int ly = e / W, lx = e - ly * W;
      if (e < W * HH) sT[ly][lx] = temperature_interior<R>(t, n3, bc3, x0 - 1 + lx, y0 - 1 + ly, double(dx), step);
    }
  } else {
#pragma unroll
    for (int f = 0; f < NF; ++f)
      load_tile<R, W, HH>(sf[f], in + f * d.sf, b ? b + f * d.sf : nullptr, d, dx, x0 - 1, y0 - 1, false);
    for (int i = threadIdx.x; i < W * HH; i += kThreads) {
      int ly = i / W, lx = i - ly * W;
      int64_t x = x0 - 1 + lx, y = y0 - 1 + ly;
      bool ok = x <= d.ax.n + 1 && y <= d.ay.n + 1;
      sT[ly][lx] = ok ? temperature_at<R>(t, n3, bc3, x, y, 1, double(dx), step) : R(1);
      if constexpr (!FUSED) {
#pragma unroll
        for (int f = 0; f < NT; ++f) st[f][ly][lx] = ok ? tr[(1 + f) * tsf + y * tsy + x] : R(0);
      }
    }
  }
  __syncthreads();
  if constexpr (FUSED) {
    // thermodynamics of every tile cell (NComponent_helper_kernel), from the
    // fields and T just staged; cells past the padded domain get finite
    // values that no output reads
    const uint64_t pxy = uint64_t(d.ax.n + 2) * uint64_t(d.ay.n + 2);
#pragma unroll 1
    for (int e = threadIdx.x; e < W * HH; e += kThreads) {
      const int ly = e / W, lx = e - ly * W;
      R cin[N - 1];
#pragma unroll
      for (int l = 0; l < N - 1; ++l) cin[l] = sf[3 + l][ly][lx];
      const uint64_t cell = uint64_t(y0 - 1 + ly) * uint64_t(d.ax.n + 2) + uint64_t(x0 - 1 + lx);
      R dGc, Mc[N - 1], dF[N - 1];
      thermo_cell<R, N, TAB>(cin, sf[0][ly][lx], sT[ly][lx], k, step, seed, pxy, cell, &dGc, Mc, dF);
#pragma unroll
      for (int l = 0; l < N - 1; ++l) {
        st[l][ly][lx] = Mc[l];
        st[N - 1 + l][ly][lx] = dF[l];
      }
      if (ly >= 1 && ly <= kNcTY && lx >= 1 && lx <= kTX) sG[ly - 1][lx - 1] = dGc;
    }
    __syncthreads();
  }

  const ppf_ncomp_params& p = k.p;
  const NcStepC<R> K(p, dx);
  const uint64_t ncells = uint64_t(d.ax.n + 2) * uint64_t(d.ay.n + 2);
  // kTX = 64: each wave owns one row of the tile per pass, so the row (and
  // every row base address) is wave-uniform
  static_assert(kTX == 64, "one wave per tile row");
  int tx = threadIdx.x % kTX, tyb = __builtin_amdgcn_readfirstlane(threadIdx.x / kTX);
  int64_t x = x0 + tx;
  if (x > d.ax.n) return;
  const int lx = tx + 1;
#pragma unroll 1
  for (int kk = 0; kk < kNcTY / 4; ++kk) {
    int ly = tyb + 4 * kk + 1;
    int64_t y = y0 + ly - 1;
    if (y > d.ay.n) break;
    const NcLds<R, N, W, HH, FUSED> v{sf, sT, st, &sG[0][0], tr, tsy, ly, lx, y, x};
    R nph, nq1, nq4, ncs[N - 1];
    nc_cell<R, N>(v, K, p, seed, step, ncells, uint64_t(y) * uint64_t(d.ax.n + 2) + uint64_t(x), nph, nq1, nq4, ncs);
    R* const orow = out + y * d.ay.stride;  // wave-uniform row base
#pragma unroll
    for (int l = 0; l < N - 1; ++l) orow[x + (3 + l) * d.sf] = ncs[l];
    orow[x] = nph;
    orow[x + d.sf] = nq1;
    orow[x + 2 * d.sf] = nq4;
  }
}

template <typename R, int N, bool FUSED>
__global__ __launch_bounds__(kThreads) PPF_NC_STEP_ATTR void k_ncomp_step(const R* __restrict__ in, R* __restrict__ out,
                                                         const R* __restrict__ b, Dom d, R dx, ppf_temperature t,
                                                         const R* __restrict__ tr, int64_t tsy, int64_t tsf,
                                                         NcK k, uint64_t seed, StepRef sr) {
  ncomp_step_body<R, N, FUSED, false>(in, out, b, d, dx, t, tr, tsy, tsf, k, seed, sr);
}
// the one-launch step with TDB-table phases
template <typename R, int N>
__global__ __launch_bounds__(kThreads) PPF_NC_STEP_ATTR void k_ncomp_step_tdb(const R* __restrict__ in,
                                                             R* __restrict__ out, const R* __restrict__ b, Dom d,
                                                             R dx, ppf_temperature t, const R* __restrict__ tr,
                                                             int64_t tsy, int64_t tsf, NcK k, uint64_t seed,
                                                             StepRef sr) {
  ncomp_step_body<R, N, true, true>(in, out, b, d, dx, t, tr, tsy, tsf, k, seed, sr);
}

// ------------------------------------------------------------ wave tiles --
// The step from the transfer stack (FP32): each wave of a 256-thread block
// owns a 62 x kNwTY tile.  Lane l holds column x0 - 1 + l of the five fields,
// T and the M_c / dF arrays on the rows y0 - 1 .. y0 + kNwTY in registers
// (all loads issued together; fields through the halo remap when there are
// no bc values and no Dirichlet side, else folded; the transfer stack and T
// at their padded coordinates), and G_S - G_L on the output rows.  The cell
// update is nc_cell with an accessor that takes x neighbours from DPP lane
// shifts and y neighbours from the lane's rows: no LDS and no barrier.
#ifndef PPF_NC_WTY
#define PPF_NC_WTY 4
#endif
constexpr int kNwTX = 62, kNwTY = PPF_NC_WTY;

template <int N, int NR, int R0>
struct NcWaveAcc {
  const float (&f)[N + 2][NR];       // phi, q1, q4, c_1..c_{N-1}
  const float (&T)[NR];
  const float (&tq)[2 * (N - 1)][NR];  // M_c, dF/dc
  float dg;
  __device__ __forceinline__ static Nb5v<float> five(const float (&a)[NR]) {
    return Nb5v<float>{a[R0], lane_shift<kDppWaveShl1>(a[R0]), lane_shift<kDppWaveShr1>(a[R0]), a[R0 + 1],
                       a[R0 - 1]};
  }
  __device__ __forceinline__ float ph(int i, int j) const {
    const float v = f[0][R0 - 1 + i];
    return j == 0 ? lane_shift<kDppWaveShr1>(v) : (j == 2 ? lane_shift<kDppWaveShl1>(v) : v);
  }
  __device__ __forceinline__ Nb5v<float> T5() const { return five(T); }
  __device__ __forceinline__ Nb5v<float> q1() const { return five(f[1]); }
  __device__ __forceinline__ Nb5v<float> q4() const { return five(f[2]); }
  __device__ __forceinline__ Nb5v<float> c(int l) const { return five(f[3 + l]); }
  __device__ __forceinline__ Nb5v<float> Mc(int l) const { return five(tq[l]); }
  __device__ __forceinline__ Nb5v<float> dF(int l) const { return five(tq[N - 1 + l]); }
  __device__ __forceinline__ float dG() const { return dg; }
};

// f(0), f(1), ... (compile-time row indices)
template <typename F, int... P>
__device__ __forceinline__ void nc_unrolled(F& f, std::integer_sequence<int, P...>) {
  (f(std::integral_constant<int, P>{}), ...);
}

template <int N>
__global__ __launch_bounds__(kThreads) void k_ncomp_step_w(const float* __restrict__ in, float* __restrict__ out,
                                                           const float* __restrict__ b, Dom d, float dx,
                                                           ppf_temperature t, const float* __restrict__ tr,
                                                           int64_t tsy, int64_t tsf, NcK k, uint64_t seed,
                                                           StepRef sr, uint32_t nstrips, uint32_t ntiles) {
  constexpr int NF = N + 2, NT = 2 * (N - 1), NR = kNwTY + 2;
  const int wave = __builtin_amdgcn_readfirstlane(int(threadIdx.x) / 64);
  const int lane = int(threadIdx.x) % 64;
  const uint32_t tile = xcd_swizzle(blockIdx.x, gridDim.x) * 4u + uint32_t(wave);
  if (tile >= ntiles) return;  // uniform per wave
  const int32_t x0 = 1 + int32_t(tile % nstrips) * kNwTX, y0 = 1 + int32_t(tile / nstrips) * kNwTY;
  const int32_t nx = int32_t(d.ax.n), ny = int32_t(d.ay.n);
  const int32_t xr = x0 - 1 + lane;
  const int64_t step = sr.get();
  float f[NF][NR], T[NR], tq[NT][NR], dg[kNwTY];
  // transfer stack and T at padded coordinates (lanes and rows past the
  // padded domain read its last cell: never used)
  const int32_t xp = xr <= nx + 1 ? xr : nx + 1;
  const __amdgpu_buffer_rsrc_t rt = plane_rsrc(tr);
  const uint32_t xto = uint32_t(xp) * 4u;
  const int32_t tsyb = int32_t(tsy * 4), tsfb = int32_t(tsf * 4);
  if (!b && d.ax.bc != PPF_DIRICHLET && d.ay.bc != PPF_DIRICHLET) {
    const __amdgpu_buffer_rsrc_t rs = plane_rsrc(in);
    const uint32_t xo = uint32_t(remap_coord32(xr, nx, d.ax.bc)) * 4u;
    const int32_t sy = int32_t(d.ay.stride * 4), sfb = int32_t(d.sf * 4);
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const int32_t yo = remap_coord32(y0 - 1 + r, ny, d.ay.bc) * sy;
#pragma unroll
      for (int i = 0; i < NF; ++i) f[i][r] = buf_load<float>(rs, xo, yo + i * sfb);
    }
  } else {
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const int64_t y = y0 - 1 + r;
      const bool ok = xr <= nx + 1 && y <= ny + 1;
#pragma unroll
      for (int i = 0; i < NF; ++i)
        f[i][r] = ok ? fold(in + i * d.sf, b ? b + i * d.sf : nullptr, d, int64_t(xr), y, int64_t(0), dx) : 0.f;
    }
  }
  const int64_t n3[3] = {d.ax.n, d.ay.n, 1};
  const int bc3[3] = {d.ax.bc, d.ay.bc, PPF_PERIODIC};
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    const int32_t yp = y0 - 1 + r <= ny + 1 ? y0 - 1 + r : ny + 1;
#pragma unroll
    for (int i = 0; i < NT; ++i) tq[i][r] = buf_load<float>(rt, xto, (1 + i) * tsfb + yp * tsyb);
    T[r] = temperature_at<float>(t, n3, bc3, xp, yp, 1, double(dx), step);
  }
#pragma unroll
  for (int r = 0; r < kNwTY; ++r) {
    const int32_t yp = y0 + r <= ny + 1 ? y0 + r : ny + 1;
    dg[r] = buf_load<float>(rt, xto, yp * tsyb);
  }
  const ppf_ncomp_params& p = k.p;
  const NcStepC<float> K(p, dx);
  const uint64_t ncells = uint64_t(nx + 2) * uint64_t(ny + 2);
  const bool out_lane = lane >= 1 && lane < 63 && xr <= nx;
  const __amdgpu_buffer_rsrc_t ws = plane_rsrc(out);
  const uint32_t oo = uint32_t(int64_t(y0) * d.ay.stride + xr) * 4u;
  const int32_t osy = int32_t(d.ay.stride * 4), osf = int32_t(d.sf * 4);
  auto row = [&](auto rc) {
    constexpr int r = decltype(rc)::value;
    const NcWaveAcc<N, NR, r + 1> acc{f, T, tq, dg[r]};
    const int32_t y = y0 + r;
    float nph, nq1, nq4, ncs[N - 1];
    nc_cell<float, N>(acc, K, p, seed, step, ncells, uint64_t(y) * uint64_t(nx + 2) + uint64_t(xr), nph, nq1, nq4,
                      ncs);
    if (out_lane && y <= ny) {
      buf_store<float>(nph, ws, oo, r * osy);
      buf_store<float>(nq1, ws, oo, r * osy + osf);
      buf_store<float>(nq4, ws, oo, r * osy + 2 * osf);
#pragma unroll
      for (int l = 0; l < N - 1; ++l) buf_store<float>(ncs[l], ws, oo, r * osy + (3 + l) * osf);
    }
  };
  nc_unrolled(row, std::make_integer_sequence<int, kNwTY>{});
}

// whether the wave-tile step runs the grid: FP32 from a transfer stack, 2D,
// enough wave tiles to fill the GPU, and the field and transfer stacks within
// the 32-bit buffer offsets (PPF_NC_WAVE_TILES=0 keeps the LDS tiles)
#ifndef PPF_NC_WAVE_TILES
#define PPF_NC_WAVE_TILES 1
#endif
inline bool ncomp_wave_fits(const ppf_grid& g, const ppf_grid* tg, int ncomp) {
  if (!PPF_NC_WAVE_TILES || !tg || g.dtype != PPF_F32 || g.ndim != 2) return false;
  const int64_t tiles = ((g.n[0] + kNwTX - 1) / kNwTX) * ((g.n[1] + kNwTY - 1) / kNwTY);
  const int64_t fbytes = (int64_t(ncomp + 1) * g.sf + (g.n[1] + 2) * g.sy) * 4;
  const int64_t tbytes = (int64_t(2 * ncomp - 2) * tg->sf + (g.n[1] + 2) * tg->sy) * 4;
  return tiles >= 1024 && fbytes < (int64_t(1) << 31) && tbytes < (int64_t(1) << 31);
}

template <typename R, bool FUSED>
int ncomp_step_launch(const ppf_grid& g, const void* in, void* out, const void* bcvals, const ppf_temperature& t,
                      const ppf_grid* tg, const void* transfer, const ppf_ncomp_params& p, const ppf_thermo* th,
                      uint64_t seed, StepRef step, hipStream_t s) {
  Dom d = make_dom(g);
  const NcK k = make_nck(p, th ? *th : ppf_thermo{});
  unsigned blocks = unsigned(((g.n[0] + kTX - 1) / kTX) * ((g.n[1] + kNcTY - 1) / kNcTY));
  const R* fin = static_cast<const R*>(in);
  R* fout = static_cast<R*>(out);
  const R* bc = static_cast<const R*>(bcvals);
  const R* tr = static_cast<const R*>(transfer);
  const int64_t tsy = tg ? tg->sy : 0, tsf = tg ? tg->sf : 0;
  if constexpr (sizeof(R) == 4 && !FUSED) {
    if (ncomp_wave_fits(g, tg, p.ncomp)) {
      const uint32_t nstrips = uint32_t((g.n[0] + kNwTX - 1) / kNwTX);
      const uint32_t ntiles = nstrips * uint32_t((g.n[1] + kNwTY - 1) / kNwTY);
      const dim3 wgrid((ntiles + 3) / 4);
      switch (p.ncomp) {
        case 2:
          hipLaunchKernelGGL(k_ncomp_step_w<2>, wgrid, dim3(kThreads), 0, s, fin, fout, bc, d, float(g.dx), t, tr, tsy,
                             tsf, k, seed, step, nstrips, ntiles);
          break;
        case 3:
          hipLaunchKernelGGL(k_ncomp_step_w<3>, wgrid, dim3(kThreads), 0, s, fin, fout, bc, d, float(g.dx), t, tr, tsy,
                             tsf, k, seed, step, nstrips, ntiles);
          break;
        case 4:
          hipLaunchKernelGGL(k_ncomp_step_w<4>, wgrid, dim3(kThreads), 0, s, fin, fout, bc, d, float(g.dx), t, tr, tsy,
                             tsf, k, seed, step, nstrips, ntiles);
          break;
        default:
          return fail(PPF_EINVAL, "ncomp_step: ncomp must be 2..4");
      }
      return check_launch("ncomp_step");
    }
  }
  switch (p.ncomp) {
    case 2:
      hipLaunchKernelGGL((FUSED && th && th->tables ? k_ncomp_step_tdb<R, 2> : k_ncomp_step<R, 2, FUSED>),
                         dim3(blocks), dim3(kThreads), 0, s, fin, fout, bc, d, R(g.dx), t, tr, tsy, tsf, k, seed,
                         step);
      break;
    case 3:
      hipLaunchKernelGGL((FUSED && th && th->tables ? k_ncomp_step_tdb<R, 3> : k_ncomp_step<R, 3, FUSED>),
                         dim3(blocks), dim3(kThreads), 0, s, fin, fout, bc, d, R(g.dx), t, tr, tsy, tsf, k, seed,
                         step);
      break;
    case 4:
      hipLaunchKernelGGL((FUSED && th && th->tables ? k_ncomp_step_tdb<R, 4> : k_ncomp_step<R, 4, FUSED>),
                         dim3(blocks), dim3(kThreads), 0, s, fin, fout, bc, d, R(g.dx), t, tr, tsy, tsf, k, seed,
                         step);
      break;
    default:
      return fail(PPF_EINVAL, "ncomp_step: ncomp must be 2..4");
  }
  return check_launch(FUSED ? "ncomp_fused_step" : "ncomp_step");
}

template <typename R>
int ncomp_step(const ppf_grid& g, const void* in, void* out, const void* bcvals, const ppf_temperature& t,
               const ppf_grid& tg, const void* transfer, const ppf_ncomp_params& p, uint64_t seed, StepRef step,
               hipStream_t s) {
  return ncomp_step_launch<R, false>(g, in, out, bcvals, t, &tg, transfer, p, nullptr, seed, step, s);
}

template <typename R>
int ncomp_fused_step(const ppf_grid& g, const void* in, void* out, const void* bcvals, const ppf_temperature& t,
                     const ppf_ncomp_params& p, const ppf_thermo& th, uint64_t seed, StepRef step, hipStream_t s) {
  return ncomp_step_launch<R, true>(g, in, out, bcvals, t, nullptr, nullptr, p, &th, seed, step, s);
}

}  // namespace ppf
