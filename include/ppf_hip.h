/*
 * ppf_hip.h -- C ABI of libppf_hip.so, the MI355X (gfx950) engine hot path of
 * the pyphasefield drop-in.
 *
 * Conventions
 *   - Every entry point returns 0 on success or a nonzero status (a hipError_t
 *     value, or PPF_EINVAL for a rejected argument); the message of the last
 *     failure on the calling thread is returned by ppf_last_error().
 *   - The caller owns every buffer (device pointers, typically from torch
 *     tensors).  No entry point allocates device memory or synchronises the
 *     device: all work is enqueued on the given stream, so a sequence of steps
 *     can be captured into a HIP graph.
 *   - Fields are stored halo-padded, +1 cell per axis, exactly as the
 *     reference Field does (field.py:6-20).  A field pointer addresses padded
 *     element (x=0, y=0, z=0) of field 0; element (f, z, y, x) lives at
 *     ptr + f*sf + z*sz + y*sy + x.  Step kernels never read halo cells of
 *     their input: boundary conditions are folded into the loads; output halos
 *     are left untouched (ppf_apply_bc materialises them).
 *   - "step" arguments are the 1-based time-step counter of the step being
 *     taken (the reference increments it before running the engine,
 *     simulation.py:263-264); it keys the noise and the temperature ramp.
 *
 * No C++ types, exceptions or torch types cross this boundary.
 */
#ifndef PPF_HIP_H
#define PPF_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t* ppf_stream_t; /* hipStream_t; NULL = default stream */

#define PPF_ABI_VERSION 1
#define PPF_EINVAL 1000

enum { PPF_F32 = 0, PPF_F64 = 1 };
/* boundary kinds, per axis; names as in simulation.py:110-117 */
enum { PPF_PERIODIC = 0, PPF_NEUMANN = 1, PPF_DIRICHLET = 2 };

/* Shape, layout and boundary rules of a stack of halo-padded fields. */
typedef struct ppf_grid {
  int32_t ndim;     /* 1, 2 or 3 */
  int32_t dtype;    /* PPF_F32 or PPF_F64 */
  int32_t nfields;  /* fields in the stack */
  int32_t bc[3];    /* per axis in the order x (last array axis), y, z
                       -- the order of ppf_gpu_utils.py:437-513 */
  int64_t n[3];     /* interior cells along x, y, z (1 for absent axes) */
  int64_t sy, sz;   /* element strides between y rows and z planes */
  int64_t sf;       /* element stride between fields */
  double dx;        /* cell spacing (Neumann halo = neighbour -/+ dx*bc) */
} ppf_grid;

/* Temperature seen by an engine (simulation.py:126-177, 303-332). */
enum { PPF_T_CONST = 0, PPF_T_LINEAR = 1, PPF_T_ARRAY = 2, PPF_T_FRAMES = 3 };
typedef struct ppf_temperature {
  int32_t kind;
  int32_t pad_;
  /* CONST: T0.  LINEAR: interior cell (i,j,k) at 0-based x,y,z index:
       T = ((T0 + i*(dx*dTdx)) + j*(dx*dTdy)) + k*(dx*dTdz) + (step-1)*(dTdt*dt)
     with halos per the temperature rule (periodic wrap, else zero gradient,
     ppf_gpu_utils.py:197-253). */
  double T0, dTdx, dTdy, dTdz, dTdt, dt;
  /* ARRAY: padded array, same dtype as the fields, strides sy/sz; only its
     interior is read (halos folded by the temperature rule). */
  const void* data;
  int64_t sy, sz;
  /* FRAMES (thermal history, update_thermal_file_*_kernel,
     ppf_gpu_utils.py:357-394): two padded frames like ARRAY (`data` at time
     t_start, `data1` at t_end); step s reads, at cur = (s-1)*dt,
       T = T_start*(1 - r) + T_end*r,  r = (cur - t_start)/(t_end - t_start)
     in FP64.  The host keeps t_start <= cur <= t_end by advancing frames. */
  const void* data1;
  double t_start, t_end;
} ppf_temperature;

/* Warren1995 parameters after Warren1995.py:197-203 derivation. */
typedef struct ppf_warren_params {
  double T, T_mA, T_mB, L_A, L_B, D_L, D_S, v_m, R, y_e, alpha;
  double ebar, W_A, W_B, M_A, M_B, dt;
} ppf_warren_params;

/* N-component engine parameters: params[15] and c_params[6][N] of
   Engines/NCGPU_new.py:495-519.  `flags` selects the legacy engine's
   differences (class NCGPU, Engines/NCGPU.py); 0 = NCGPU_new:
     PPF_NC_DERIV_FIRST  the free-energy derivative along e_1 - e_N (the first
                         component) is used for every component
                         (NCGPU.py:287-290, 303);
     PPF_NC_SHARED_NOISE one composition noise draw per cell, added to every
                         component (NCGPU.py:295-296);
     PPF_NC_CLAMP_WIDE   phi clamped to [1e-6, 0.999999] instead of
                         [1e-4, 0.9999] (NCGPU.py:229-232);
     PPF_NC_EBAR_FROM_S  ebar^2 = 6 sqrt(2) S[1] d / T_M[1] and
                         eqbar^2 = ebar^2 / 4 in that order (NCGPU.py:74-75)
                         instead of ebar^2, eqbar^2 from ebar, eqbar.
   The legacy engine's other differences are parameter values: no phi or
   orientation noise (noise_phi = noise_q = 0), composition noise amplitude 1. */
#define PPF_MAX_COMP 4
#define PPF_NC_DERIV_FIRST 1
#define PPF_NC_SHARED_NOISE 2
#define PPF_NC_CLAMP_WIDE 4
#define PPF_NC_EBAR_FROM_S 8
typedef struct ppf_ncomp_params {
  int32_t ncomp; /* N components, 2..PPF_MAX_COMP; field stack = N+2 */
  int32_t flags; /* PPF_NC_* (legacy engine), 0 for NCGPU_new */
  double dx, d, v_m, M_qmax, H, y_e, beta, D_L, D_S, dt, ebar, eqbar;
  double noise_phi, noise_c, noise_q;
  double L[PPF_MAX_COMP], T_M[PPF_MAX_COMP], S[PPF_MAX_COMP];
  double B[PPF_MAX_COMP], W[PPF_MAX_COMP], M[PPF_MAX_COMP];
} ppf_ncomp_params;

/* Redlich-Kister + ideal-mixing phase (pycalphad redlich_kister_sum +
   ideal_mixing_energy as used at simulation.py:212-214):
     G = sum c_i (a+bT+cT lnT+dT^2+e/T) + sum_{i<j} c_i c_j sum_k (A+BT)(c_i-c_j)^k
         + R T sum c_i ln max(c_i, 1e-10) */
typedef struct ppf_rk_phase {
  int32_t ncomp;
  int32_t pad_;
  double R;
  double g0[PPF_MAX_COMP][5];
  double L[6][3][2]; /* pairs (0,1),(0,2),(0,3),(1,2),(1,3),(2,3) as used */
} ppf_rk_phase;

/* derivative evaluation mode of the plain thermodynamics */
enum { PPF_DG_FORWARD_DIFF = 0, /* reference h=1e-7 forward difference */
       PPF_DG_ANALYTIC = 1 };

/* Piecewise temperature function read from a TDB file (FUNCTION /
   PARAMETER with temperature ranges): range r covers (T_hi[r-1], T_hi[r]]
   (below the first, the first; above the last, the last) with
     f = c0 + c1 T + c2 T ln T + c3 T^2 + c4 T^3 + c5/T + c6 T^7 + c7 T^-9
   summed in that order (T^7 = (T^3 T^3) T, T^-9 = (u^3 u^3) u^3, u = 1/T,
   u^3 = (u u) u).  nranges = 0: identically zero. */
#define PPF_TPOLY_TERMS 8
#define PPF_TPOLY_RANGES 6
typedef struct ppf_tpoly {
  int32_t nranges;
  int32_t pad_;
  double T_hi[PPF_TPOLY_RANGES];
  double c[PPF_TPOLY_RANGES][PPF_TPOLY_TERMS];
} ppf_tpoly; /* 440 bytes */

/* A TDB phase of one sublattice: unary G of each component (PARAMETER
   G(phase, i; 0)), binary Redlich-Kister terms L^k of each pair (i < j in
   component order, pairs as in ppf_rk_phase.L, orders 0..4):
     c_i c_j sum_k L^k_ij (c_i - c_j)^k,
   and ternary terms of each triple (i < j < l in component order: (0,1,2),
   (0,1,3), (0,2,3), (1,2,3)) in the Muggianu form pycalphad's
   redlich_kister_sum builds:
     c_i c_j c_l sum_k L^k_ijl v_k,  v_k = c_k + (1 - c_i - c_j - c_l)/3
   (k = 0, 1, 2 naming i, j, l; a file giving only L^0 means L^1 = L^2 = L^0,
   filled in on the host).  Evaluated as
     w = (1 - ((c_i + c_j) + c_l)) / 3,
     G += ((c_i c_j) c_l) (((c_i + w) L^0 + (c_j + w) L^1) + (c_l + w) L^2). */
#define PPF_RK_ORDERS 5
typedef struct ppf_tdb_phase {
  ppf_tpoly g0[PPF_MAX_COMP];
  ppf_tpoly L[6][PPF_RK_ORDERS];
  ppf_tpoly T3[4][3];
} ppf_tdb_phase;
typedef struct ppf_tdb_tables {
  ppf_tdb_phase liquid, solid;
} ppf_tdb_tables;

typedef struct ppf_thermo {
  ppf_rk_phase liquid; /* ufunc 1 in NCGPU_new.py:353 */
  ppf_rk_phase solid;  /* ufunc 0 (FCC_A1 sorts first), NCGPU_new.py:354 */
  int32_t dg_mode;
  int32_t pad_;
  /* device pointer or NULL: when set, the unary G and L^k of both phases come
     from these tables (the g0 and L fields above are ignored; R is used) */
  const ppf_tdb_tables* tables;
} ppf_thermo;

int ppf_abi_version(void);
const char* ppf_last_error(void);

/* Materialise the halo cells of a field stack: replaces apply_boundary_conditions
   (ppf_gpu_utils.py:430-514, kernels :11-195).  bcvals: NULL or an array with
   the same layout whose halo cells hold the bcarray values. */
int ppf_apply_bc(const ppf_grid* g, void* fields, const void* bcvals, ppf_stream_t s);

/* Temperature halo rule (ppf_gpu_utils.py:197-319): periodic wrap, else
   zero-gradient.  g->nfields must be 1. */
int ppf_apply_bc_temperature(const ppf_grid* g, void* T, ppf_stream_t s);

/* Write the temperature of step `step` (interior + halos) into a padded array
   (replaces init_temperature_field + update_temperature_field for the GPU
   path: simulation.py:126-156, ppf_gpu_utils.py:516-531). */
int ppf_fill_temperature(const ppf_grid* g, void* out, const ppf_temperature* t,
                         int64_t step, ppf_stream_t s);

/* One explicit-Euler step of c <- c + dt*D*lap(c)
   (Engines/Diffusion.py:198-208; replaces DiffusionGPU.py:18-75). */
int ppf_diffusion_step(const ppf_grid* g, const void* in, void* out, const void* bcvals,
                       double D, double dt, ppf_stream_t s);

/* phi <- phi + dt*M*(eps^2 lap(phi) - 16 W (4phi^3-6phi^2+2phi))
   (Engines/CahnAllen.py:128-135). */
int ppf_allencahn_step(const ppf_grid* g, const void* in, void* out, const void* bcvals,
                       double M, double W, double epsilon, double dt, ppf_stream_t s);

/* nsteps (1..ppf_multistep_max) successive ppf_diffusion_step /
   ppf_allencahn_step steps in one launch: out = step^nsteps(in), equal to the
   one-step calls bit for bit (no counterpart in the reference: in place of
   nsteps calls of its per-step kernel launch, Engines/Diffusion.py:198-208,
   Engines/CahnAllen.py:128-135).  Each block stages its tile with an nsteps
   halo and steps it in LDS.  ppf_multistep_max returns the largest nsteps
   the grid allows: 0 unless the grid is 2D and every side is periodic or
   Neumann without bc values (bcvals null), at most 8 and at most each side. */
int ppf_multistep_max(const ppf_grid* g, const void* bcvals);
int ppf_diffusion_multistep(const ppf_grid* g, const void* in, void* out, const void* bcvals,
                            double D, double dt, int32_t nsteps, ppf_stream_t s);
int ppf_allencahn_multistep(const ppf_grid* g, const void* in, void* out, const void* bcvals,
                            double M, double W, double epsilon, double dt, int32_t nsteps,
                            ppf_stream_t s);

/* mu = -eps^2 lap(phi) + 4phi^3-6phi^2+2phi;  phi <- phi + dt*M*lap(mu)
   (Engines/CahnHilliard.py:134-146; mu on the +1 ring: periodic wrap or
   zero-flux copy of the adjacent interior cell). */
int ppf_cahnhilliard_step(const ppf_grid* g, const void* in, void* out, const void* bcvals,
                          double M, double epsilon, double dt, ppf_stream_t s);

/* nsteps (1..ppf_cahnhilliard_multistep_max) successive ppf_cahnhilliard_step
   steps in one launch, equal to the one-step calls bit for bit (no
   counterpart in the reference: in place of nsteps calls of its per-step
   kernel launch, Engines/CahnHilliard.py:134-146).  Two steps take one pass
   of the 3D register march with a 4-cell halo.  ppf_cahnhilliard_multistep_max
   returns 2 for FP32 3D grids with every side periodic and at least 8 cells
   per axis, else 1. */
int ppf_cahnhilliard_multistep_max(const ppf_grid* g);
int ppf_cahnhilliard_multistep(const ppf_grid* g, const void* in, void* out, const void* bcvals,
                               double M, double epsilon, double dt, int32_t nsteps, ppf_stream_t s);

/* Warren1995 binary solidification step, fields (phi, c)
   (Engines/Warren1995.py:50-92).  Noise: counter RNG keyed by (seed, step). */
int ppf_warren_step(const ppf_grid* g, const void* in, void* out, const void* bcvals,
                    const ppf_warren_params* p, uint64_t seed, int64_t step, ppf_stream_t s);

/* N-component step in one launch: replaces NComponent_helper_kernel +
   NComponent_kernel (Engines/NCGPU_new.py:256-313 and :37-250, launched in
   that order by simulation_loop :521-554).  The thermodynamic quantities of
   the helper (G_S - G_L, M_c, dF/dc with its noise) are evaluated for each
   output tile and its +1 ring into LDS from the same fields the step loads,
   with the arithmetic of ppf_ncomp_thermo, so no transfer stack exists.
   Results equal ppf_ncomp_thermo followed by ppf_ncomp_step (bit for bit in
   FP64). */
int ppf_ncomp_fused_step(const ppf_grid* g, const void* in, void* out, const void* bcvals,
                         const ppf_temperature* t, const ppf_ncomp_params* p, const ppf_thermo* th,
                         uint64_t seed, int64_t step, ppf_stream_t s);

/* N-component thermodynamic pass: replaces NComponent_helper_kernel
   (Engines/NCGPU_new.py:256-313).  Writes, for every padded cell including the
   halo ring, transfer[0] = G_S - G_L, transfer[1+l] = M_c,l and
   transfer[N+l] = dF/dc_l (l < N-1).  `tg` describes transfer: nfields = 2N-1,
   same n/strides convention as g. */
int ppf_ncomp_thermo(const ppf_grid* g, const void* in, const void* bcvals, const ppf_temperature* t,
                     const ppf_grid* tg, void* transfer, const ppf_ncomp_params* p,
                     const ppf_thermo* th, uint64_t seed, int64_t step, ppf_stream_t s);

/* N-component main step: replaces NComponent_kernel (Engines/NCGPU_new.py:37-250)
   reading the transfer arrays ppf_ncomp_thermo wrote. */
int ppf_ncomp_step(const ppf_grid* g, const void* in, void* out, const void* bcvals,
                   const ppf_temperature* t, const ppf_grid* tg, const void* transfer,
                   const ppf_ncomp_params* p, uint64_t seed, int64_t step, ppf_stream_t s);

/* ---- device-resident step counter (HIP graph replay) ----
   The reference's loop increments a host counter before each step
   (simulation.py:262-264) and hands it to kernels that key their noise and
   temperature ramp on it.  The _dev variants below take the step from device
   memory instead: step = *step_base + step_offset (step_base: a device
   int64), so a chunk of k steps can be captured once (offsets 1..k) followed
   by ppf_step_counter_add(step_base, k) and replayed; each replay then takes
   the next k steps.  Results equal the host-counter entry points'. */
int ppf_step_counter_add(int64_t* counter, int64_t inc, ppf_stream_t s);
int ppf_warren_step_dev(const ppf_grid* g, const void* in, void* out, const void* bcvals,
                        const ppf_warren_params* p, uint64_t seed, const int64_t* step_base, int64_t step_offset,
                        ppf_stream_t s);
int ppf_ncomp_thermo_dev(const ppf_grid* g, const void* in, const void* bcvals, const ppf_temperature* t,
                         const ppf_grid* tg, void* transfer, const ppf_ncomp_params* p, const ppf_thermo* th,
                         uint64_t seed, const int64_t* step_base, int64_t step_offset, ppf_stream_t s);
int ppf_ncomp_step_dev(const ppf_grid* g, const void* in, void* out, const void* bcvals, const ppf_temperature* t,
                       const ppf_grid* tg, const void* transfer, const ppf_ncomp_params* p, uint64_t seed,
                       const int64_t* step_base, int64_t step_offset, ppf_stream_t s);
int ppf_ncomp_fused_step_dev(const ppf_grid* g, const void* in, void* out, const void* bcvals,
                             const ppf_temperature* t, const ppf_ncomp_params* p, const ppf_thermo* th,
                             uint64_t seed, const int64_t* step_base, int64_t step_offset, ppf_stream_t s);

/* Global FP64 reductions over the interior (conservation checks):
   kind 0: out[f] = sum of field f (mass), for every field of the stack;
   kind 1: out[0] = Cahn-Hilliard free energy sum[c^2(1-c)^2 + eps^2/2 |grad c|^2] dx^d
           (forward differences, halos folded), field 0 only.
   out: device pointer to FP64 results; work: device scratch of
   ppf_reduce_work_bytes(g) bytes. */
int64_t ppf_reduce_work_bytes(const ppf_grid* g);
int ppf_reduce(const ppf_grid* g, const void* in, const void* bcvals, int kind, double epsilon,
               double* out, void* work, ppf_stream_t s);

/* The same reductions without a workspace: no counterpart in the reference;
   for callers that cannot allocate device memory (code under HIP graph
   capture, tight serving loops, harnesses that forbid hipMalloc).  The
   partial sums live in a file-scope device array of the library
   (PPF_REDUCE_NOALLOC_MAX_FIELDS x 2048 doubles = 256 KiB per GPU, held until
   the process ends); kind 0 on more than PPF_REDUCE_NOALLOC_MAX_FIELDS fields
   is rejected.  Every call overwrites all the partial sums it reads, so
   nothing needs clearing; calls that overlap on one GPU share the array and
   race.  Allocates nothing (no hipMalloc, no library allocation). */
#define PPF_REDUCE_NOALLOC_MAX_FIELDS 16
int ppf_reduce_noalloc(const ppf_grid* g, const void* in, const void* bcvals, int kind, double epsilon,
                       double* out, ppf_stream_t s);

/* ---- AnisoDorr binary polycrystalline solidification (Engines/AnisoDorrGPU.py) ----
   2D; field stack (phi, q1, q4, c); transfer stack of 6 arrays with the layout
   convention of g (tg->nfields = 6).  The reference hard-codes these constants
   in both kernels (AnisoDorrGPU.py:63-105); derived ones (ebar, eqbar, W_A,
   W_B, M_A, M_B, dt) are computed by the caller as there. */
typedef struct ppf_aniso_params {
  double T, M_qmax, H, T_mA, T_mB, L_A, L_B, D_L, D_S, v_m, R, y_e;
  double beta;  /* clipping value of the squared quaternion gradients (1.5) */
  double alpha; /* noise amplitude (0.3) */
  double ebar, eqbar, W_A, W_B, M_A, M_B, dt;
} ppf_aniso_params;

/* Vertex and diffusion terms of every padded cell: replaces
   AnisoDorr_helper_kernel (AnisoDorrGPU.py:262-375). */
int ppf_aniso_helper(const ppf_grid* g, const void* in, const void* bcvals, const ppf_grid* tg, void* transfer,
                     const ppf_aniso_params* p, ppf_stream_t s);

/* One explicit step from the transfer stack ppf_aniso_helper wrote: replaces
   AnisoDorr_kernel (AnisoDorrGPU.py:54-245).  Noise: uniform [0,1) from the
   counter RNG keyed by (seed, step). */
int ppf_aniso_step(const ppf_grid* g, const void* in, void* out, const void* bcvals, const ppf_grid* tg,
                   const void* transfer, const ppf_aniso_params* p, uint64_t seed, int64_t step, ppf_stream_t s);

/* The same step as ppf_aniso_helper + ppf_aniso_step in one launch: the
   vertex and diffusion terms of each 64 x 8 tile stay in LDS (no transfer
   stack; identical arithmetic). */
int ppf_aniso_fused_step(const ppf_grid* g, const void* in, void* out, const void* bcvals,
                         const ppf_aniso_params* p, uint64_t seed, int64_t step, ppf_stream_t s);
/* the same with the step counter in device memory (see ppf_step_counter_add) */
int ppf_aniso_fused_step_dev(const ppf_grid* g, const void* in, void* out, const void* bcvals,
                             const ppf_aniso_params* p, uint64_t seed, const int64_t* step_base,
                             int64_t step_offset, ppf_stream_t s);

/* ---- implicit and Crank-Nicolson diffusion (Engines/Diffusion.py:18-633) ----
   Every matrix the reference builds (diffusion_matrix_1d/2d/3d,
   Diffusion.py:18-196) is periodic; these entry points require PERIODIC on
   every axis and at least 3 cells per axis.  They act on field 0 of the
   stack; `in` and `out` are distinct padded stacks with the layout of g and
   the halos of `out` are left untouched. */

/* Host only (no device work): coefficients of the cyclic tridiagonal system
     -a x[i-1] + (1+2a) x[i] - a x[i+1] = d[i]   (indices mod n)
   i.e. diffusion_matrix_1d(n, 1+2a, -a) (Diffusion.py:18-64), for the
   corner-free Thomas elimination plus a Sherman-Morrison correction.
   table must hold ppf_cyclic_table_len(n) doubles (-1 when n < 3). */
int64_t ppf_cyclic_table_len(int64_t n);
int ppf_cyclic_table(int64_t n, double a, double* table);

/* For every grid line along `axis` (0 = x, 1 = y, 2 = z):
     out = M^-1 rhs,  rhs = E(in),
   M = the cyclic matrix of `table` (a device copy of ppf_cyclic_table(n[axis], a)),
   E = identity (explicit_axis = -1) or the explicit half step along
   explicit_axis, (1-2b) c + b (c[-1] + c[+1]) with periodic neighbours.
   Replaces np.linalg.solve / the np.dot(inv, line) loops of the 1D and
   alternating-direction solvers (Diffusion.py:210-241, 309-343, 413-633). */
int ppf_diffusion_line_solve(const ppf_grid* g, const void* in, void* out, int axis, const double* table,
                             int explicit_axis, double b, ppf_stream_t s);

/* The same line solves with parallelism along each line: the matrix factors
   as (a/r)(I - r S)(I - r S^T), r = ((1+2a) - sqrt(1+4a))/(2a), S the cyclic
   shift, so each solve is a forward and a backward first-order recurrence
   closed around the ring, evaluated in chunks of m cells (one thread per
   line and chunk) with carry scans between chunks.  table: a device copy of
   ppf_scan_table(n[axis], a, m) (ppf_scan_table_len(m) doubles, host only);
   work: ppf_line_scan_work_bytes(g, axis, m) bytes of device scratch.
   Results equal ppf_diffusion_line_solve's to rounding. */
int64_t ppf_scan_table_len(int64_t m);
int ppf_scan_table(int64_t n, double a, int64_t m, double* table);
int64_t ppf_line_scan_work_bytes(const ppf_grid* g, int axis, int64_t m);
int ppf_diffusion_line_solve_scan(const ppf_grid* g, const void* in, void* out, int axis, const double* table,
                                  int64_t m, int explicit_axis, double b, void* work, ppf_stream_t s);

/* Full implicit / Crank-Nicolson systems of diffusion_matrix_2d/3d
   (Diffusion.py:243-308, 344-412):
     (1 + 2*ndim*a) x - a * sum_{6 or 4 neighbours} x = rhs,
     rhs = in (b = 0) or (1 - 2*ndim*b) in + b * sum_nb in,
   solved by matrix-free conjugate gradients from x0 = in.  work holds
   ppf_diffusion_cg_work_bytes(g) bytes of device memory, the first
   sizeof(ppf_cg_state) of them the solver state.  ppf_diffusion_cg_begin
   writes x0 into out and the initial residual; ppf_diffusion_cg_iterate
   enqueues `iterations` iterations numbered first_iteration.. (the caller
   numbers them consecutively from 0 after a begin); once r.r <= tol^2 rhs.rhs
   the state is marked converged and further iterations do nothing.  No host
   synchronisation: the caller reads the state (e.g. every few dozen
   iterations) to decide whether to continue. */
typedef struct ppf_cg_state {
  double bb;          /* rhs . rhs */
  double rr;          /* r . r at the latest convergence test */
  int32_t iterations; /* iterations taken once converged, else -1 */
  int32_t converged;
} ppf_cg_state;
int64_t ppf_diffusion_cg_work_bytes(const ppf_grid* g);
int ppf_diffusion_cg_begin(const ppf_grid* g, const void* in, void* out, double a, double b, void* work,
                           ppf_stream_t s);
int ppf_diffusion_cg_iterate(const ppf_grid* g, void* out, double a, double tol, int64_t first_iteration,
                             int iterations, void* work, ppf_stream_t s);

/* Allen-Cahn semi-implicit steps (Engines/CahnAllen.py:137-280), periodic
   boxes as for the diffusion solvers.  The double-well force
   16 W (4 phi^3 - 6 phi^2 + 2 phi) is linearised as 16 W s(phi_old) phi_new,
   s(phi) = 4 phi^2 - 6 phi + 2, which adds src_k * s to the diagonal
   (src_k = 16 W M dt for a full solve, 8 W M dt per 2D alternating-direction
   sweep, 16 W M dt / 3 per 3D sweep).

   Line solve along `axis`: rhs d = E(in) (explicit half step along
   explicit_axis with coefficient b, as in ppf_diffusion_line_solve), then
   with imex_k != 0 d <- d - imex_k (4 d^3 - 6 d^2 + 2 d); the cyclic system
   has off-diagonal -a and diagonal 1 + 2a + src_k s(d) per cell.  work holds
   ppf_allencahn_line_work_bytes(g) bytes of device scratch. */
int64_t ppf_allencahn_line_work_bytes(const ppf_grid* g);
int ppf_allencahn_line_solve(const ppf_grid* g, const void* in, void* out, int axis, double a, int explicit_axis,
                             double b, double src_k, double imex_k, void* work, ppf_stream_t s);

/* Full 2D/3D semi-implicit system (the GMRES engine, CahnAllen.py:165-177):
     (1 + 2*ndim*a + src_k s(in)) x - a * sum_nb x = in,
   conjugate gradients exactly as ppf_diffusion_cg_begin/_iterate (same work
   size and state) with the per-cell diagonal; `in` must stay unchanged until
   the iterations are done.  Symmetric positive definite while
   1 + src_k s(in) > 0 (src_k < 4); otherwise the state does not converge. */
int ppf_allencahn_cg_begin(const ppf_grid* g, const void* in, void* out, double a, double src_k, void* work,
                           ppf_stream_t s);
int ppf_allencahn_cg_iterate(const ppf_grid* g, const void* in, void* out, double a, double src_k, double tol,
                             int64_t first_iteration, int iterations, void* work, ppf_stream_t s);

/* Cahn-Hilliard semi-implicit 1D (engine_CahnHilliardImplicit1D and its
   GMRES and Crank-Nicolson variants, Engines/CahnHilliard.py:148-222, matrix
   of implicit_matrix_1d, :20-70): solves, on a periodic 1D grid of n >= 5
   cells, the pentadiagonal system whose row i is
     (1 + alpha (2 s_i + 6 e)) x_i - alpha (s_i + 4 e)(x_{i-1} + x_{i+1})
       + alpha e (x_{i-2} + x_{i+2}) = c_i,   s_i = 4 c_i^2 - 6 c_i + 2,
   c = field 0 of `in`, x = field 0 of `out` (interior cells only; indices
   mod n), alpha = M dt / dx^2, e = eps^2 / dx^2.  One work group: banded
   elimination of the non-periodic band, then a rank-4 Sherman-Morrison-
   Woodbury correction for the corner entries, in FP64 for both storage
   types.  Needs ppf_cahnhilliard_line_work_bytes(g) bytes of device scratch.
   Crank-Nicolson: ppf_cahnhilliard_step with dt/2, then this solve with
   alpha = M (dt/2) / dx^2. */
int64_t ppf_cahnhilliard_line_work_bytes(const ppf_grid* g);
int ppf_cahnhilliard_line_solve(const ppf_grid* g, const void* in, void* out, double alpha, double e, void* work,
                                ppf_stream_t s);

/* Plain thermodynamics at scattered points: for each of n points with
   temperature T[i] and component fractions c[i*(N-1)+l] (c_N implied),
   write for each phase (liquid then solid) G and dG/dc_l:
   out[i*2N + 0] = G_L, out[i*2N + 1..N-1] = dG_L/dc_l,
   out[i*2N + N] = G_S, out[i*2N + N+1..2N-1] = dG_S/dc_l.  FP64 throughout.
   Used to generate surrogate training data. */
int ppf_thermo_eval(const ppf_thermo* th, int64_t n, const double* T, const double* c,
                    double* out, ppf_stream_t s);

/* ---- surrogate thermodynamics (FP16 / BF16 MFMA, FP32 accumulate) -------
   No counterpart in the reference: a learned replacement for the per-cell
   Gibbs-energy evaluation of NComponent_helper_kernel (NCGPU_new.py:295-305),
   trained by the kernels below on samples of the plain evaluation. */

/* C[M x N] = alpha A B + beta C with A(i,k) = A[i*sai + k*sak],
   B(k,j) = B[k*sbk + j*sbj], C row-major (ldc); FP32 in memory, MFMA
   products with FP32 accumulation.  split_b selects the operand formats:
   0 = A and B rounded to BF16; 1 = A in BF16, B as the BF16 pair hi + lo
   (hi = bf16(b), lo = bf16(b - hi)); 2 = A in FP16, B as the FP16 pair
   hi + lo (the inference kernel's formats); other values are rejected.  ksplit > 1
   splits K across blocks with atomic accumulation (beta must be 0 and C
   zeroed by the caller). */
int ppf_mlp_gemm(int M, int N, int K, const float* A, int64_t sai, int64_t sak, const float* B,
                 int64_t sbk, int64_t sbj, float* C, int64_t ldc, float alpha, float beta, int ksplit,
                 int split_b, ppf_stream_t s);
/* H[i,j] = act(Z[i,j] + b[i]) on [M x N] row-major; act 0 = identity, 1 = ReLU; b may be NULL */
int ppf_mlp_bias_act(const float* Z, const float* b, float* H, int M, int N, int act, ppf_stream_t s);
/* dZ = dH * (H > 0) */
int ppf_mlp_relu_bwd(const float* dH, const float* H, float* dZ, int64_t n, ppf_stream_t s);
/* out[i] = sum_j X[i,j] on [M x N] row-major */
int ppf_mlp_rowsum(const float* X, float* out, int M, int N, ppf_stream_t s);
/* number of per-block partial sums ppf_mlp_mse writes */
int64_t ppf_mlp_mse_partials(int64_t n);
/* dY = 2 (Y - T) / n; partial[b] = sum of squared errors of block b */
int ppf_mlp_mse(const float* Y, const float* T, float* dY, float* partial, int64_t n, ppf_stream_t s);
/* Adam update of n parameters at step t (>= 1) */
int ppf_mlp_adam(float* p, const float* g, float* m, float* v, int64_t n, float lr, float b1, float b2,
                 float eps, int64_t t, ppf_stream_t s);

/* Two trained 3-64-64-3 ReLU networks (liquid, solid), inputs (T, c1, c2),
   FP16 weights, FP32 accumulation; inputs as FP16 hi + lo pairs, hidden
   activations as pairs or FP16 (flags).  The packed layers carry power-of-two
   scales that keep every hidden value of an accepted cell in [0, 1]
   (pyphasefield_amd.surrogate.pack_phase); b2 is stored divided by its scale.
   Outputs (normalised): liquid (G_ni, dG_ni/dc1, dG_ni/dc2); solid
   (G_S,ni - G_L,ni, dG_S,ni/dc1, dG_S,ni/dc2), where "ni" is G minus the ideal
   mixing term R T sum c ln c (evaluated analytically in the kernel). */
typedef struct ppf_mlp_net {
  const void* weights[2];  /* per phase: packed FP16 MFMA operands (see pyphasefield_amd.surrogate) */
  const float* bias[2];    /* per phase: b1[64], b2[64], b3[4] (FP32) */
  float in_lo[3], in_scale[3];          /* x_hat = (x - lo) * scale - 1 */
  float out_mu[2][3], out_sigma[2][3];  /* y = y_hat * sigma + mu */
  float T_lo, T_hi, c_lo;  /* range gate: cells outside use the plain FP64 evaluation */
  int32_t flags;           /* PPF_MLP_SPLIT2 / PPF_MLP_SPLIT3: the inputs of layer 2 / 3 enter as
                              FP16 hi + lo pairs (else FP16 only) */
  int32_t* marks;          /* optional device int32[2], zero-initialised once: counts of cells
                              handed to the FP64 fallback (lets it skip its scan when none are);
                              NULL: the fallback always scans */
  const ppf_thermo* thermo; /* optional device copy of the ppf_thermo passed with the call: when
                              set, cells outside the trained box are evaluated in the network
                              kernel itself (FP64 plain evaluation) and no second kernel runs */
} ppf_mlp_net;
#define PPF_MLP_SPLIT3 1
#define PPF_MLP_SPLIT2 2

/* N-component thermodynamic pass with the surrogate (FP32, N = 3): same
   outputs as ppf_ncomp_thermo. */
int ppf_ncomp_thermo_mlp(const ppf_grid* g, const void* in, const void* bcvals, const ppf_temperature* t,
                         const ppf_grid* tg, void* transfer, const ppf_ncomp_params* p, const ppf_thermo* th,
                         const ppf_mlp_net* net, uint64_t seed, int64_t step, ppf_stream_t s);
/* the same with the step counter in device memory (see ppf_step_counter_add) */
int ppf_ncomp_thermo_mlp_dev(const ppf_grid* g, const void* in, const void* bcvals, const ppf_temperature* t,
                             const ppf_grid* tg, void* transfer, const ppf_ncomp_params* p, const ppf_thermo* th,
                             const ppf_mlp_net* net, uint64_t seed, const int64_t* step_base, int64_t step_offset,
                             ppf_stream_t s);

/* ---- Batched ensemble (no reference counterpart: the reference runs one
   Simulation per instance, each stepping with engine_CahnAllenExplicit,
   Engines/CahnAllen.py:128-135, or engine_CahnHilliardExplicit,
   Engines/CahnHilliard.py:134-146).  Independent 2D FP32 instances, one
   persistent launch per step over a host-built tile work-list. ---- */
#define PPF_ENS_AC 0
#define PPF_ENS_CH 1
#define PPF_ENS_XCDS 8
typedef struct ppf_ens_instance {
  int32_t kind;            /* PPF_ENS_AC or PPF_ENS_CH */
  int32_t bc[2];           /* x, y: PPF_PERIODIC or PPF_NEUMANN (zero gradient) */
  int32_t pad_;
  int64_t n[2];            /* interior cells along x, y */
  int64_t pitch;           /* elements per padded row */
  int64_t offset;          /* element offset of padded cell (0, 0) in both arenas */
  double M, W, eps, dt, dx; /* Allen-Cahn: M, W, eps; Cahn-Hilliard: M, eps (W unused) */
} ppf_ens_instance;        /* 88 bytes */

/* number of tile configurations, and the tile shapes of one:
   shape = {Allen-Cahn TX, TY, Cahn-Hilliard TX, TY}.  Configurations with
   TX = {62, 60} (64 lanes less the halo) are wave tiles held in registers,
   the others block tiles staged through LDS; both give identical results. */
int ppf_ensemble_num_configs(void);
int ppf_ensemble_config(int config, int32_t* shape);
/* One explicit step of every instance: out = step(in).  instances: device
   array of n_instances descriptors; tiles: device array of n_tiles x 4 int32
   (instance, x0, y0, 0) with x0, y0 the padded coordinates of the tile's
   first output cell (1 + multiples of the configuration's tile shape for the
   instance's kind); segments: host array of PPF_ENS_XCDS + 1 tile indices,
   0 = s[0] <= s[1] <= ... <= s[8] = n_tiles (tiles of XCD k are
   [s[k], s[k+1])); blocks: persistent grid size, >= 1 (with fewer than
   PPF_ENS_XCDS blocks, block b walks segments b, b + blocks, ...).  Halo cells are never read
   (boundary rules folded into the loads); in and out must not alias. */
int ppf_ensemble_step(const ppf_ens_instance* instances, int32_t n_instances, const int32_t* tiles,
                      int32_t n_tiles, const int32_t* segments, int32_t config, int32_t blocks,
                      const float* in, float* out, ppf_stream_t s);

#ifdef __cplusplus
}
#endif
#endif /* PPF_HIP_H */
