"""CPU oracle for the pyphasefield engine hot path (FP64, numpy).

TEST INFRASTRUCTURE ONLY.  Nothing in the shipped package imports this
module: only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may use it, and only as the checker or
as the timed CPU baseline -- never as the thing measured or shipped.

The oracle restates, step by step, what the reference computes for one
explicit-Euler time step of each engine, with the defect resolutions listed
in the design notes (reference engines are not runnable as written).  Each
function cites the reference file:line it follows; paths are relative to
``pyphasefield/pyphasefield/`` of the reference tree.

Pinning (what checks each part, see tests/test_oracle_*.py):

* ``bc``        -- brute-force rules on tiny grids + invariants.
* ``rng``       -- reference-free (project-defined counter RNG): moments,
                   uniqueness, fixed known-answer values.
* ``diffusion`` -- exact discrete Fourier-mode eigenvalue, mass conservation.
* ``allencahn`` -- fixed points of the double well, discrete stationarity of
                   the closed-form tanh kink, Fourier-mode linearisation.
* ``cahnhilliard`` -- exact mass conservation, discrete dispersion relation,
                   free-energy decrease.
* ``warren``    -- reduces to Allen-Cahn-type relaxation, 4-fold symmetry,
                   solute conservation under zero-flux BC, isotropy at y_e=0,
                   term-by-term continuum convergence.
* ``ncomp``     -- scalar per-cell transcription vs vectorised restatement,
                   continuum convergence of the anisotropy, orientation and
                   flux operators to their analytic divergences, invariants
                   (|q|=1, phi clamp, solute conservation).
* ``thermo``    -- closed-form RK sums, analytic vs finite-difference
                   derivatives, symmetry L^k_ij = (-1)^k L^k_ji.

Parity unpinned: the reference's pycalphad/sympy path for real TDB files
(pycalphad is not installed) and numba's xoroshiro128p noise stream; neither
is used by the configurations measured here (synthetic Redlich-Kister
coefficients; noise from the counter RNG or amplitude 0).
"""

from . import bc, rng, engines, thermo  # noqa: F401
