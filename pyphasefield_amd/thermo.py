"""Thermodynamic models for the N-component engine.

The reference builds its Gibbs energies from TDB files with pycalphad
(``simulation.py:179-243``) and reads the engine parameters from TDB
``FUNCTION`` entries (``Engines/NCGPU_new.py:366-383``).  pycalphad is not
part of this engine path; instead a :class:`ThermoModel` holds, per phase,
the coefficients of a single-sublattice Redlich-Kister + ideal-mixing Gibbs
energy

    G(c, T) = sum_i c_i (a + bT + cT lnT + dT^2 + e/T)
            + sum_{i<j} c_i c_j sum_{k=0..2} (A_k + B_k T)(c_i - c_j)^k
            + R T sum_i c_i ln c_i

which the device evaluates per cell, plus the engine parameters the TDB
FUNCTIONs carry.  Components and phases are kept sorted alphabetically, as
the reference sorts them (``simulation.py:201-202``), so the solid (FCC_A1)
is phase 0 and the liquid phase 1 (``Engines/NCGPU_new.py:353-354``).
"""

import numpy as np

from . import _lib

R_GAS = 8.314462618  # ideal-mixing gas constant of the synthetic models

# Engine parameters of tests/Ni-Cu_Ideal.tdb:36-70 (units cm, J, K, s).
NI_CU_IDEAL_FUNCTIONS = {
    "AL_L": 1080.0, "CU_L": 1728.0, "NI_L": 2350.0,        # latent heat, J/cm^3
    "AL_TM": 933.5, "CU_TM": 1358.0, "NI_TM": 1728.0,      # melting point, K
    "AL_S": 0.00002, "CU_S": 0.000029, "NI_S": 0.000037,   # surface energy, J/cm^2
    "AL_B": 0.36, "CU_B": 0.39, "NI_B": 0.33,              # kinetic coefficient, cm/(K s)
    "V_M": 7.42,                                           # molar volume, cm^3/mol
    "D_L": 0.00001, "D_S": 0.000000001,                    # diffusivities, cm^2/s
    "M_Q": 80000000.0,                                     # orientational mobility
    "H": 0.00000000001,                                    # orientational energy, J/(K cm)
    "Y_E": 0.12,                                           # anisotropy
}


def pairs(n):
    return [(i, j) for i in range(n) for j in range(i + 1, n)]


class RKPhase:
    """Coefficients of one phase: g0 [N, 5] and L [N(N-1)/2, 3, 2]."""

    def __init__(self, name, g0, L, R=R_GAS):
        self.name = name
        self.g0 = np.asarray(g0, dtype=np.float64)
        self.L = np.asarray(L, dtype=np.float64)
        self.R = float(R)
        n = self.g0.shape[0]
        if self.g0.shape != (n, 5) or self.L.shape != (n * (n - 1) // 2, 3, 2):
            raise ValueError("bad coefficient shapes for phase %s" % name)

    @property
    def ncomp(self):
        return self.g0.shape[0]

    def to_ctypes(self):
        p = _lib.RKPhase()
        p.ncomp = self.ncomp
        p.R = self.R
        for i in range(self.ncomp):
            for k in range(5):
                p.g0[i][k] = float(self.g0[i, k])
        for q in range(self.L.shape[0]):
            for k in range(3):
                for m in range(2):
                    p.L[q][k][m] = float(self.L[q, k, m])
        return p


# gas constant pycalphad's ideal_mixing_energy uses for TDB phases
R_PYCALPHAD = 8.3145


def triples(n):
    """(i, j, l), i < j < l, in lexicographic order (ppf_tdb_phase.T3)."""
    return [(i, j, l) for i in range(n) for j in range(i + 1, n) for l in range(j + 1, n)]


class TablePhase:
    """A TDB phase of one sublattice: unary G (Piecewise per component),
    binary Redlich-Kister terms L^k (Piecewise per component pair i < j and
    order k = 0..4) and ternary terms (Piecewise per triple i < j < l and
    index k = 0..2), evaluated on the device from ppf_tdb_tables."""

    def __init__(self, name, g0, L, R=R_PYCALPHAD, T3=None):
        self.name = name
        self.g0 = list(g0)
        self.L = dict(L)
        self.T3 = dict(T3 or {})
        self.R = float(R)

    @property
    def ncomp(self):
        return len(self.g0)

    def tables(self):
        """(g0, L) in the oracle's table format: [(T_hi, c, n)] per component
        and {"pairs": [[(T_hi, c, n)] * 5] per pair, "triples": [[...] * 3] per
        triple}"""
        n = self.ncomp
        zero = (np.full(_lib.PPF_TPOLY_RANGES, np.inf), np.zeros((_lib.PPF_TPOLY_RANGES, _lib.PPF_TPOLY_TERMS)), 0)

        def tab(f):
            return f.device_table(_lib.PPF_TPOLY_RANGES) if f is not None else zero
        g0 = [tab(f) for f in self.g0]
        L = {"pairs": [[tab(self.L.get((p, k))) for k in range(_lib.PPF_RK_ORDERS)]
                       for p in range(n * (n - 1) // 2)],
             "triples": [[tab(self.T3.get((t, k))) for k in range(3)] for t in range(len(triples(n)))]}
        return g0, L

    def fill(self, st):
        """write this phase into a _lib.TdbPhase"""
        g0, L = self.tables()

        def put(dst, tab):
            T_hi, c, n = tab
            dst.nranges = int(n)
            for r in range(_lib.PPF_TPOLY_RANGES):
                dst.T_hi[r] = float(T_hi[r])
                for k in range(_lib.PPF_TPOLY_TERMS):
                    dst.c[r][k] = float(c[r, k])
        for i, tab in enumerate(g0):
            put(st.g0[i], tab)
        for p, row in enumerate(L["pairs"]):
            for k, tab in enumerate(row):
                put(st.L[p][k], tab)
        for t, row in enumerate(L["triples"]):
            for k, tab in enumerate(row):
                put(st.T3[t][k], tab)

    def to_ctypes(self):
        p = _lib.RKPhase()
        p.ncomp = self.ncomp
        p.R = self.R
        return p


class ThermoModel:
    """Phases (sorted) + components (sorted) + TDB-FUNCTION-style parameters.

    Plays the role of the reference's TDBContainer (ppf_utils.py:125-186) for
    the N-component engine: pass it as ``tdb_container``.
    """

    def __init__(self, components, phases, functions):
        self.components = sorted(components)
        if list(components) != self.components:
            raise ValueError("components must be given in alphabetical order")
        self.phases = dict(sorted(phases.items()))
        self.functions = dict(functions)
        for ph in self.phases.values():
            if ph.ncomp != len(self.components):
                raise ValueError("phase %s has the wrong number of components" % ph.name)

    @property
    def solid(self):
        return list(self.phases.values())[0]

    @property
    def liquid(self):
        return list(self.phases.values())[1]

    def to_ctypes(self, dg_mode=_lib.PPF_DG_FORWARD_DIFF):
        t = _lib.Thermo()
        t.liquid = self.liquid.to_ctypes()
        t.solid = self.solid.to_ctypes()
        t.dg_mode = dg_mode
        t.tables = self.device_tables()
        return t

    def device_tables(self):
        """Device copy of the TDB tables (uploaded once), or None for
        Redlich-Kister coefficient phases."""
        kinds = {isinstance(self.liquid, TablePhase), isinstance(self.solid, TablePhase)}
        if kinds == {False}:
            return None
        if kinds != {True}:
            raise ValueError("both phases must come from the same kind of model")
        if getattr(self, "_tables_dev", None) is None:
            import torch
            tab = _lib.TdbTables()
            self.liquid.fill(tab.liquid)
            self.solid.fill(tab.solid)
            raw = np.frombuffer(bytes(tab), dtype=np.uint8).copy()
            self._tables_dev = torch.from_numpy(raw).to("cuda")
        return self._tables_dev.data_ptr()

    def oracle_phase(self, which):
        """(g0, L, R) of the liquid or solid phase as the oracle takes them."""
        ph = self.liquid if which == "liquid" else self.solid
        if isinstance(ph, TablePhase):
            g0, L = ph.tables()
            return g0, L, ph.R
        return ph.g0, ph.L, ph.R

    def value(self, name):
        return float(self.functions[name])


def _hessian_min_eig(phase, T, n_grid=24):
    """Smallest eigenvalue of d2G/dc2 (independent c_1..c_{N-1}) over the
    simplex interior, by central differences on the closed form."""
    n = phase.ncomp
    lo = 0.02
    pts = []
    g = np.linspace(lo, 1 - lo, n_grid)
    if n == 2:
        pts = [[a] for a in g]
    elif n == 3:
        pts = [[a, b] for a in g for b in g if a + b < 1 - lo]
    else:
        pts = [[a, b, c] for a in g[::3] for b in g[::3] for c in g[::3] if a + b + c < 1 - lo]
    pts = np.array(pts)
    h = 1e-4

    def G(cs):
        full = list(cs.T) + [1 - cs.sum(axis=1)]
        lnT = np.log(T)
        out = np.zeros(len(cs))
        for i in range(n):
            a, b, c, d, e = phase.g0[i]
            out += full[i] * (a + b * T + c * T * lnT + d * T * T + e / T)
        for p, (i, j) in enumerate(pairs(n)):
            Lk = phase.L[p, :, 0] + phase.L[p, :, 1] * T
            dd = full[i] - full[j]
            out += full[i] * full[j] * (Lk[0] + dd * (Lk[1] + dd * Lk[2]))
        for i in range(n):
            out += phase.R * T * full[i] * np.log(np.maximum(full[i], 1e-10))
        return out

    m = n - 1
    H = np.zeros((len(pts), m, m))
    for a in range(m):
        for b in range(m):
            ea = np.zeros(m)
            eb = np.zeros(m)
            ea[a] = h
            eb[b] = h
            H[:, a, b] = (G(pts + ea + eb) - G(pts + ea - eb) - G(pts - ea + eb) + G(pts - ea - eb)) / (4 * h * h)
    return float(np.min(np.linalg.eigvalsh(H)))


def synthetic_rk_model(components=("AL", "CU", "NI"), seed=0, T_range=(1000.0, 3047.0), scale=1e4,
                       functions=None, max_tries=1000):
    """Seeded synthetic ternary (or N-ary) Redlich-Kister liquid + FCC model.

    Liquid: G_i = a_i + b_i T with a_i ~ N(0, scale), b_i ~ N(0, scale)/1000
    (J/mol and J/(mol K)); L^k_ij ~ N(0, scale) J/mol, independent of T.
    FCC:    G_i = liquid G_i - L_i v_m (T_M,i - T)/T_M,i (latent heat from the
    engine parameters, so each pure element melts at its T_M) and its own
    L^k_ij drawn like the liquid's.

    Draws are rejected (seed, seed+1, ...) until both phases are convex in
    composition over ``T_range`` (no spinodal: the N-component engine has no
    gradient energy in c, so a concave G would make the explicit scheme
    ill-posed).  Returns (model, accepted_seed).
    """
    comps = sorted(components)
    n = len(comps)
    fn = dict(NI_CU_IDEAL_FUNCTIONS if functions is None else functions)
    v_m = fn["V_M"]
    npair = n * (n - 1) // 2
    for s in range(seed, seed + max_tries):
        rng = np.random.default_rng(s)
        g0L = np.zeros((n, 5))
        g0L[:, 0] = rng.normal(0.0, scale, n)
        g0L[:, 1] = rng.normal(0.0, scale, n) / 1000.0
        LL = np.zeros((npair, 3, 2))
        LL[:, :, 0] = rng.normal(0.0, scale, (npair, 3))
        LS = np.zeros((npair, 3, 2))
        LS[:, :, 0] = rng.normal(0.0, scale, (npair, 3))
        g0S = g0L.copy()
        for i, cname in enumerate(comps):
            Lh = fn[cname + "_L"] * v_m   # J/mol
            Tm = fn[cname + "_TM"]
            # -Lh (Tm - T)/Tm = -Lh + (Lh/Tm) T
            g0S[i, 0] -= Lh
            g0S[i, 1] += Lh / Tm
        liquid = RKPhase("LIQUID", g0L, LL)
        solid = RKPhase("FCC_A1", g0S, LS)
        ok = True
        for T in np.linspace(T_range[0], T_range[1], 5):
            if _hessian_min_eig(liquid, T) <= 0 or _hessian_min_eig(solid, T) <= 0:
                ok = False
                break
        if ok:
            model = ThermoModel(comps, {"FCC_A1": solid, "LIQUID": liquid}, fn)
            return model, s
    raise RuntimeError("no convex synthetic model found")


def load_tdb_model(path, components=None, phases=None):
    """ThermoModel from a TDB file (one-sublattice phases: unary G, binary
    Redlich-Kister L^0..L^4 and ternary L^0..L^2), replacing pycalphad's
    Database + Model.redlich_kister_sum + ideal_mixing_energy
    (simulation.py:196-214).  Components default to the file's elements and
    phases to its phases, both sorted alphabetically as the reference sorts
    them (simulation.py:201-202): the first phase is the solid (FCC_A1), the
    second the liquid (NCGPU_new.py:353-354).  Only parameters whose
    constituents are all among the components enter (pycalphad's Model does
    the same); among those, what the model cannot represent (several
    sublattices, four-species interactions, binary orders above 4,
    expressions outside the temperature basis) is refused with its name.
    A ternary parameter's order k names its k-th constituent as written;
    when only order 0 of a triple is given, orders 1 and 2 take the same
    function (pycalphad's rule for symmetric ternary parameters).  Engine
    parameters are the file's constant FUNCTIONs (NCGPU_new.py:358-383)."""
    from . import tdb
    db = tdb.read(path)
    comps = sorted(c.upper() for c in (components or db.elements))
    names = sorted(p.upper() for p in (phases or list(db.phases) + list(db.multi_sublattice)))
    if len(names) != 2:
        raise ValueError("the N-component engine needs two phases, %s has %s" % (path, names))
    built = {}
    cset = set(comps)
    index = {c: i for i, c in enumerate(comps)}
    for name in names:
        if name in db.multi_sublattice:
            raise ValueError("phase %s of %s has %d sublattices; one is supported"
                             % (name, path, db.multi_sublattice[name]))
        cons = db.phases.get(name)
        if cons is None:
            raise ValueError("phase %s not in %s" % (name, path))
        missing = [c for c in comps if c not in cons]
        if missing:
            raise ValueError("phase %s has no constituent %s" % (name, missing))
        for (ph, cons_text, order), why in db.unsupported.items():
            species = set(x.strip() for x in cons_text.replace(":", ",").split(","))
            if ph == name and species <= cset | {"VA", "*"}:
                raise ValueError("parameter (%s, %s; %d) of %s: %s" % (name, cons_text, order, path, why))
        g0 = []
        for c in comps:
            f = db.parameters.get(("G", name, (c,), 0))
            if f is None:
                raise ValueError("no G(%s,%s;0) in %s" % (name, c, path))
            g0.append(f)
        L, T3, given3 = {}, {}, {}
        pair_index = {pr: p for p, pr in enumerate(pairs(len(comps)))}
        trip_index = {tr: t for t, tr in enumerate(triples(len(comps)))}
        for (kind, ph, sp, k), f in db.parameters.items():
            if kind != "L" or ph != name or not set(sp) <= cset:
                continue
            if len(set(sp)) != len(sp):
                raise ValueError("parameter L(%s,%s;%d) repeats a species" % (name, ",".join(sp), k))
            idx = [index[x] for x in sp]
            if len(sp) == 2:
                if k > 4:
                    raise ValueError("Redlich-Kister order %d of %s is beyond the supported 0..4" % (k, name))
                key = (pair_index[tuple(sorted(idx))], k)
                if key in L:
                    raise ValueError("L(%s,%s;%d) given twice" % (name, ",".join(sp), k))
                # L^k_ba = (-1)^k L^k_ab
                L[key] = f if idx[0] < idx[1] or k % 2 == 0 else -f
            elif len(sp) == 3:
                if k > 2:
                    raise ValueError("ternary parameter L(%s,%s;%d): orders 0..2 name a constituent"
                                     % (name, ",".join(sp), k))
                t = trip_index[tuple(sorted(idx))]
                slot = sorted(idx).index(idx[k])  # order k names constituent sp[k]
                if (t, slot) in T3:
                    raise ValueError("ternary L(%s,%s) given twice for %s" % (name, ",".join(sp), sp[k]))
                T3[(t, slot)] = f
                given3.setdefault(t, set()).add(k)
            else:
                raise ValueError("parameter L(%s,%s;%d): interactions of %d species are not supported"
                                 % (name, ",".join(sp), k, len(sp)))
        for t, ks in given3.items():
            if ks == {0}:
                f0 = [T3[(t, sl)] for sl in range(3) if (t, sl) in T3][0]
                for sl in range(3):
                    T3[(t, sl)] = f0
        built[name] = TablePhase(name, g0, L, T3=T3)
    functions = {}
    for fname in db.functions:
        f = db.functions[fname]
        if f is not None and f.is_const():
            functions[fname] = f.value_const()
    model = ThermoModel(comps, built, functions)
    model.path = str(path)
    return model