"""Gibbs energy of a substitutional phase and its derivatives (oracle).

The reference builds G for each phase from a TDB file with pycalphad's
``Model.redlich_kister_sum`` + ``ideal_mixing_energy`` (``simulation.py:
212-214``), lambdifies it (``simulation.py:243``) and evaluates it per cell in
``NComponent_helper_kernel`` (``Engines/NCGPU_new.py:295-296``) with the input
vector [c_1..c_{N-1}, c_N, T] (``Engines/NCGPU_new.py:286-292``).  For a single
sublattice that expression is

    G(c, T) = sum_i c_i G_i(T)
            + sum_{i<j} c_i c_j sum_k L^k_ij(T) (c_i - c_j)^k
            + R T sum_i c_i ln c_i

with G_i(T) = a + b T + c T lnT + d T^2 + e / T  (one SGTE range) and
L^k_ij(T) = A + B T.  The chemical-potential derivative is the reference's
forward difference with h = 1e-7 along e_l - e_N
(``Engines/NCGPU_new.py:301-313``), including its exact sequence of
perturb / restore operations on the input vector.

``c ln c`` is evaluated at max(c, CMIN): the reference does not guard
``log`` at c <= 0 (it returns NaN there); the clamp is applied identically in
the oracle and on the device.

Coefficient arrays (float64):
    g0 : [N, 5]         a, b, c, d, e of G_i(T)
    L  : [N(N-1)/2, 3, 2]  pair (i<j) in lexicographic order, k = 0..2, (A, B)

TDB phases (tables): binary orders 0..4, and three-species interaction
parameters as pycalphad's redlich_kister_sum builds them (Muggianu):
c_i c_j c_l sum_k v_k L^k_ijl, v_k = c_k + (1 - c_i - c_j - c_l)/3, where
a file giving only L^0 stands for L^0 = L^1 = L^2 (since sum_k v_k = 1 the
term is then c_i c_j c_l L^0).
"""

import numpy as np

CMIN = 1e-10

# Natural logarithm shared bit-for-bit with the device (FP64): frexp range
# reduction to m in [sqrt(1/2), sqrt(2)), then log m = 2 atanh(s) with
# s = (m-1)/(m+1), summed as the odd power series 2s(1 + z/3 + z^2/5 + ...),
# z = s^2, to z^11 (|s| <= 0.1716, so the truncation is below 1e-18); the
# exponent is added with ln 2 split into a 32-bit head and a tail so that
# e*LN2_HI is exact.  The reference's forward difference (h = 1e-7) turns a
# 1-ulp difference between two log implementations into a 1e-8 relative
# difference in the chemical potential; one shared definition keeps FP64
# device runs within 1e-10 of this oracle.
LN2_HI = float.fromhex("0x1.62e42fee00000p-1")
LN2_LO = float.fromhex("0x1.a39ef35793c76p-33")
SQRT_HALF = float.fromhex("0x1.6a09e667f3bcdp-1")
LOG_SERIES = [1.0 / (2 * k + 1) for k in range(12)]


def plog(x):
    x = np.asarray(x, dtype=np.float64)
    m, e = np.frexp(x)
    small = m < SQRT_HALF
    m = np.where(small, m * 2.0, m)
    e = np.where(small, e - 1, e).astype(np.float64)
    f = m - 1.0
    s = f / (2.0 + f)
    z = s * s
    p = np.full_like(z, LOG_SERIES[11])
    for k in range(10, -1, -1):
        p = p * z + LOG_SERIES[k]
    return e * LN2_HI + (e * LN2_LO + 2.0 * s * p)


FD_H = 0.0000001
FD_INV = 10000000.0


def pairs(n):
    return [(i, j) for i in range(n) for j in range(i + 1, n)]


def tpoly(tab, T, lnT):
    """Piecewise TDB temperature function in the device table format
    (T_hi[r], c[r][8], nranges): range r covers (T_hi[r-1], T_hi[r]], the
    first below and the last above; terms 1, T, T lnT, T^2, T^3, 1/T, T^7,
    T^-9 summed in that order (the device's ppf_tpoly order, so FP64 values
    agree bit for bit)."""
    T_hi, c, n = tab
    T = np.asarray(T, dtype=np.float64)
    if n <= 0:
        return np.zeros_like(T)
    r = np.zeros(T.shape, dtype=np.int64)
    for k in range(n - 1):
        r = np.where((r == k) & (T > T_hi[k]), k + 1, r)
    C = np.asarray(c)[r]  # [..., 8]
    u = 1.0 / T
    T2 = T * T
    T3 = T2 * T
    T7 = (T3 * T3) * T
    u3 = (u * u) * u
    u9 = (u3 * u3) * u3
    v = C[..., 0] + C[..., 1] * T
    v = v + C[..., 2] * (T * lnT)
    v = v + C[..., 3] * T2
    v = v + C[..., 4] * T3
    v = v + C[..., 5] * u
    v = v + C[..., 6] * T7
    v = v + C[..., 7] * u9
    return v


def _is_table(g0):
    return isinstance(g0, (list, tuple)) and len(g0) > 0 and isinstance(g0[0], tuple)


def _unary(g0, i, T, lnT):
    if _is_table(g0):
        return tpoly(g0[i], T, lnT)
    a, b, c, d, e = g0[i]
    return a + b * T + c * T * lnT + d * T * T + e / T


def triples(n):
    return [(i, j, l) for i in range(n) for j in range(i + 1, n) for l in range(j + 1, n)]


def _is_table_L(L):
    return isinstance(L, dict)


def _rk(L, p, T, lnT):
    """Binary orders of pair p: five (TDB tables: {"pairs": [[table] * 5],
    "triples": [[table] * 3]}) or three (coefficient array [P, 3, 2])."""
    if _is_table_L(L):
        return [tpoly(tab, T, lnT) for tab in L["pairs"][p]]
    return [L[p, k, 0] + L[p, k, 1] * T for k in range(3)]


def _rk_poly(Lk, dd):
    """sum_k L^k dd^k by Horner from the highest order (the device's order)."""
    if len(Lk) == 5:
        return Lk[0] + dd * (Lk[1] + dd * (Lk[2] + dd * (Lk[3] + dd * Lk[4])))
    return Lk[0] + dd * (Lk[1] + dd * Lk[2])


def _rk_dpoly(Lk, dd):
    if len(Lk) == 5:
        return Lk[1] + dd * (2.0 * Lk[2] + dd * (3.0 * Lk[3] + dd * (4.0 * Lk[4])))
    return Lk[1] + 2.0 * dd * Lk[2]


def _ternary(L, t, T, lnT):
    return [tpoly(tab, T, lnT) for tab in L["triples"][t]]


def ternary_weights(ci, cj, cl):
    """Muggianu weights v_k = c_k + (1 - c_i - c_j - c_l)/3 of a ternary
    term (pycalphad's redlich_kister_sum for a three-species interaction)."""
    w = (1.0 - ((ci + cj) + cl)) / 3.0
    return ci + w, cj + w, cl + w


def gibbs(cs, T, g0, L, R):
    """G for component fractions ``cs`` (list of N arrays, c_N last) and T.
    g0 / L: the Redlich-Kister coefficient arrays, or TDB tables (g0: list of
    (T_hi, c, nranges), see :func:`tpoly`; L: {"pairs", "triples"}).  TDB
    phases add binary orders 3..4 and ternary terms
    c_i c_j c_l sum_k v_k L^k_ijl (:func:`ternary_weights`)."""
    n = len(cs)
    lnT = plog(T)
    G = np.zeros(np.broadcast(cs[0], T).shape)
    for i in range(n):
        gi = _unary(g0, i, T, lnT)
        G = G + cs[i] * gi
    for p, (i, j) in enumerate(pairs(n)):
        dd = cs[i] - cs[j]
        P = _rk_poly(_rk(L, p, T, lnT), dd)
        G = G + cs[i] * cs[j] * P
    if _is_table_L(L):
        for t, (i, j, l) in enumerate(triples(n)):
            V0, V1, V2 = _ternary(L, t, T, lnT)
            v0, v1, v2 = ternary_weights(cs[i], cs[j], cs[l])
            Wt = (v0 * V0 + v1 * V1) + v2 * V2
            G = G + ((cs[i] * cs[j]) * cs[l]) * Wt
    ideal = np.zeros_like(G)
    for i in range(n):
        ci = np.maximum(cs[i], CMIN)
        ideal = ideal + ci * plog(ci)
    return G + R * T * ideal


def fd_derivatives(cs, T, g0, L, R, G=None):
    """Forward differences of Engines/NCGPU_new.py:301-313, same op order.

    Returns (G, [dG/dc_l for l < N-1]).
    """
    n = len(cs)
    arr = [np.array(c, dtype=np.float64, copy=True) for c in cs]
    if G is None:
        G = gibbs(arr, T, g0, L, R)
    arr[n - 1] = arr[n - 1] - FD_H
    out = []
    for l in range(n - 1):
        arr[l] = arr[l] + FD_H
        out.append(FD_INV * (gibbs(arr, T, g0, L, R) - G))
        arr[l] = arr[l] - FD_H
    return G, out


def analytic_derivatives(cs, T, g0, L, R):
    """dG/dc_l along e_l - e_N (exact), used by the FP32 device path."""
    n = len(cs)
    lnT = plog(T)
    dG = []
    part = []
    for i in range(n):
        gi = _unary(g0, i, T, lnT)
        ci = np.maximum(cs[i], CMIN)
        clamp = (cs[i] > CMIN)
        part.append(gi + R * T * np.where(clamp, plog(ci) + 1.0, 0.0))
    for p, (i, j) in enumerate(pairs(n)):
        Lk = _rk(L, p, T, lnT)
        dd = cs[i] - cs[j]
        P = _rk_poly(Lk, dd)
        dP = _rk_dpoly(Lk, dd)
        part[i] = part[i] + cs[j] * P + cs[i] * cs[j] * dP
        part[j] = part[j] + cs[i] * P - cs[i] * cs[j] * dP
    if _is_table_L(L):
        for t, (i, j, l) in enumerate(triples(n)):
            V0, V1, V2 = _ternary(L, t, T, lnT)
            v0, v1, v2 = ternary_weights(cs[i], cs[j], cs[l])
            Wt = (v0 * V0 + v1 * V1) + v2 * V2
            Vm = ((V0 + V1) + V2) / 3.0
            prod = (cs[i] * cs[j]) * cs[l]
            part[i] = part[i] + (cs[j] * cs[l]) * Wt + prod * (V0 - Vm)
            part[j] = part[j] + (cs[i] * cs[l]) * Wt + prod * (V1 - Vm)
            part[l] = part[l] + (cs[i] * cs[j]) * Wt + prod * (V2 - Vm)
    for l in range(n - 1):
        dG.append(part[l] - part[n - 1])
    return dG
