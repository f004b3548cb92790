"""Counter-based noise generator shared by the oracle and the HIP kernels.

The reference draws noise from ``np.random`` (``Engines/Warren1995.py:85``)
and from numba's xoroshiro128p states (``Engines/NCGPU_new.py:188,226-227,
309``); neither stream can be reproduced on another device, and the numba
one races on shared states (``Engines/NCGPU.py:42``).  This project replaces
both with a stateless hash of (seed, step, stream, cell) so that any cell's
draw is a pure function of its coordinates -- bit-identical on CPU and GPU
and across checkpoint/resume.

    counter = (step * NSTREAMS + stream) * ncells + cell        (uint64)
    h       = splitmix64_finalize(seed + counter * GOLDEN)
    uniform_f64 = (h >> 11) * 2**-53          in [0, 1)
    uniform_f32 = (h >> 40) * 2**-24          in [0, 1)
    normal      = Marsaglia's polar method: attempt k = 0, 1, ... uses
                  h_0 = h, h_k = hash64(h_{k-1}, k); u = 2((h_k >> 32) + 1/2) 2**-32 - 1,
                  v = 2((h_k & 0xffffffff) + 1/2) 2**-32 - 1, s = u^2 + v^2; the first
                  attempt with 0 < s < 1 gives u * sqrt(-2 ln(s) / s), with ln the
                  shared FP64 logarithm of oracle.thermo.plog (no libm call whose
                  last bit could differ between host and device).

``cell`` is the linear index in the canonical padded layout
(k*(ny+2) + j)*(nx+2) + i, independent of the device memory pitch.
"""

import numpy as np

NSTREAMS = 16
GOLDEN = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)

# stream ids
S_WARREN_PHI = 0
S_NC_PHI = 1
S_NC_Q1 = 2
S_NC_Q4 = 3
S_NC_C0 = 4  # component l uses S_NC_C0 + l


def hash64(seed, counter):
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + np.asarray(counter, dtype=np.uint64) * GOLDEN
        z = (z ^ (z >> np.uint64(30))) * M1
        z = (z ^ (z >> np.uint64(27))) * M2
        z = z ^ (z >> np.uint64(31))
    return z


def counters(step, stream, ncells, cells):
    with np.errstate(over="ignore"):
        base = (np.uint64(step) * np.uint64(NSTREAMS) + np.uint64(stream)) * np.uint64(ncells)
        return base + np.asarray(cells, dtype=np.uint64)


def uniform_f64(seed, step, stream, ncells, cells):
    h = hash64(seed, counters(step, stream, ncells, cells))
    return (h >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def uniform_f32(seed, step, stream, ncells, cells):
    h = hash64(seed, counters(step, stream, ncells, cells))
    return ((h >> np.uint64(40)).astype(np.float32) * np.float32(2.0 ** -24)).astype(np.float32)


MAX_POLAR_ATTEMPTS = 64


def normal_f64(seed, step, stream, ncells, cells):
    from .thermo import plog
    h = hash64(seed, counters(step, stream, ncells, cells))
    out = np.zeros(h.shape, dtype=np.float64)
    todo = np.ones(h.shape, dtype=bool)
    for k in range(MAX_POLAR_ATTEMPTS):
        if k > 0:
            h = np.where(todo, hash64(h, np.uint64(k)), h)
        u = 2.0 * (((h >> np.uint64(32)).astype(np.float64) + 0.5) * (2.0 ** -32)) - 1.0
        v = 2.0 * (((h & np.uint64(0xFFFFFFFF)).astype(np.float64) + 0.5) * (2.0 ** -32)) - 1.0
        s = u * u + v * v
        ok = todo & (s < 1.0) & (s > 0.0)
        ss = np.where(ok, s, 0.5)
        val = u * np.sqrt(-2.0 * plog(ss) / ss)
        out = np.where(ok, val, out)
        todo = todo & ~ok
        if not todo.any():
            break
    return out


def padded_cells(shape_padded):
    """Canonical padded linear index of every cell of an array of this shape."""
    return np.arange(int(np.prod(shape_padded)), dtype=np.uint64).reshape(shape_padded)
