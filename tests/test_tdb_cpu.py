"""TDB reader (one-sublattice subset): the reference's own databases
(tests/golden/tdb/, copied from the reference's tests/ folder as data) are
read and pinned by what the thermodynamics fixes -- pure-element melting
points stated in the same files, continuity at range breakpoints, the
pair-order invariance of the Redlich-Kister excess -- and the oracle's table
evaluation is checked against the reader's own evaluator."""

import os

import numpy as np
import pytest

from oracle import thermo as ot
from pyphasefield_amd import tdb, thermo

GOLD = os.path.join(os.path.dirname(__file__), "golden", "tdb")


def _db(name):
    return tdb.read(os.path.join(GOLD, name))


@pytest.mark.parametrize("name,elements", [("Ni-Cu_Ideal.tdb", ["CU", "NI"]),
                                           ("Ni-Cu-Al_Ideal.tdb", ["AL", "CU", "NI"]),
                                           ("Ni-Nb_Simplified.tdb", ["NB", "NI"])])
def test_reads_reference_databases(name, elements):
    db = _db(name)
    assert sorted(db.elements) == elements
    assert sorted(db.phases) == ["FCC_A1", "LIQUID"]
    for ph, cons in db.phases.items():
        assert sorted(cons) == elements
    assert db.constant("V_M") == 7.42


@pytest.mark.parametrize("el", ["AL", "CU", "NI"])
def test_liquid_and_fcc_cross_at_the_stated_melting_point(el):
    """G_liquid - G_fcc of a pure element changes sign at its melting point
    (the file's <EL>_TM) with a negative slope of about the entropy of fusion."""
    db = _db("Ni-Cu-Al_Ideal.tdb")
    tm = db.constant(el + "_TM")
    d = db.parameters[("G", "LIQUID", (el,), 0)] - db.parameters[("G", "FCC_A1", (el,), 0)]
    v = d.evaluate(np.array([tm - 2.0, tm, tm + 2.0]))
    assert v[0] > 0 > v[2]
    assert abs(v[1]) < 5.0
    slope = (v[2] - v[0]) / 4.0
    assert -15.0 < slope < -5.0


def test_niobium_liquid_meets_its_reference_state_at_its_melting_point():
    """G(LIQUID,NB) is referenced to GHSERNB (bcc): their difference vanishes
    at NB_TM; fcc Nb is metastable by G(FCC_A1,NB) - GHSERNB > 0 there."""
    db = _db("Ni-Nb_Simplified.tdb")
    tm = db.constant("NB_TM")
    liq = db.parameters[("G", "LIQUID", ("NB",), 0)] - db.functions["GHSERNB"]
    fcc = db.parameters[("G", "FCC_A1", ("NB",), 0)] - db.functions["GHSERNB"]
    assert abs(float(liq.evaluate(tm))) < 2.0
    assert float(fcc.evaluate(tm)) > 1e4


@pytest.mark.parametrize("name", ["Ni-Cu_Ideal.tdb", "Ni-Cu-Al_Ideal.tdb", "Ni-Nb_Simplified.tdb"])
def test_unary_functions_are_continuous_at_breakpoints(name):
    db = _db(name)
    for key, f in db.parameters.items():
        for h in f.breakpoints[:-1]:
            lo, hi = f.evaluate(np.array([h - 1e-7, h + 1e-7]))
            # the files carry rounded coefficients (Cu fcc jumps by 0.66 J/mol at
            # 1357.77 K); a dropped or mis-signed term would jump by 1e2-1e4 J/mol
            assert abs(lo - hi) < 1e-4 * abs(lo) + 0.1, (key, h, lo, hi)


def test_excess_energy_independent_of_pair_order():
    """The file gives L(LIQUID,NI,NB;k) for the pair in the order (NI, NB); the
    model stores pairs in sorted order (NB, NI) with L^k -> (-1)^k L^k.  The
    excess x_Ni x_Nb sum_k L^k (x_Ni - x_Nb)^k must not change."""
    db = _db("Ni-Nb_Simplified.tdb")
    model = thermo.load_tdb_model(os.path.join(GOLD, "Ni-Nb_Simplified.tdb"))
    assert model.components == ["NB", "NI"]
    T = np.array([1500.0, 1900.0, 2300.0])
    x_nb = np.array([0.2, 0.45, 0.7])
    x_ni = 1 - x_nb
    want = np.zeros(3)
    for k in range(3):
        Lk = db.parameters[("L", "LIQUID", ("NI", "NB"), k)].evaluate(T)
        want += x_ni * x_nb * Lk * (x_ni - x_nb) ** k
    g0, L, R = model.oracle_phase("liquid")
    cs = [x_nb, x_ni]
    G = ot.gibbs(cs, T, g0, L, R)
    unary = sum(c * ot.tpoly(g, T, ot.plog(T)) for c, g in zip(cs, g0))
    ideal = R * T * sum(c * ot.plog(c) for c in cs)
    np.testing.assert_allclose(G - unary - ideal, want, rtol=1e-12, atol=1e-8)


def test_oracle_table_evaluation_matches_reader():
    model = thermo.load_tdb_model(os.path.join(GOLD, "Ni-Cu-Al_Ideal.tdb"))
    T = np.random.default_rng(0).uniform(300, 3000, 2000)
    for ph in (model.solid, model.liquid):
        g0, _ = ph.tables()
        for f, tab in zip(ph.g0, g0):
            np.testing.assert_allclose(ot.tpoly(tab, T, ot.plog(T)), f.evaluate(T), rtol=1e-13, atol=1e-8)


def test_engine_parameters_and_defaults():
    model = thermo.load_tdb_model(os.path.join(GOLD, "Ni-Cu-Al_Ideal.tdb"))
    assert model.solid.name == "FCC_A1" and model.liquid.name == "LIQUID"
    assert abs(model.value("M_Q") - 80000000 / 1574.0) < 1e-9
    assert model.value("NI_TM") == 1728.0
    assert model.solid.R == thermo.R_PYCALPHAD


def test_unsupported_features_are_recorded_then_refused_when_used(tmp_path):
    """A large database is read leniently; what the model cannot represent is
    refused when a model for components that touch it is built."""
    db = tdb.Database("PHASE X % 2 1 1 ! CONSTITUENT X :A:B: ! PARAMETER G(X,A:B;0) 298.15 1; 6000 N ! "
                      "PARAMETER TC(X,A:B;0) 298.15 1; 6000 N ! FUNCTION F 298.15 1+G; 6000 N !")
    assert db.multi_sublattice == {"X": 2} and "X" not in db.phases
    assert db.ignored == {"TC": 1}
    with pytest.raises(tdb.TDBError):
        db._function("F")
    db = tdb.Database("FUNCTION F 298.15 1+T**4; 6000 N !")
    with pytest.raises(tdb.TDBError):
        db.functions["F"].device_table(6)

    def model(params, comps=("A", "B", "C"), phases=None, extra=""):
        text = ("ELEMENT A X 1 1 1 ! ELEMENT B X 1 1 1 ! ELEMENT C X 1 1 1 ! ELEMENT D X 1 1 1 ! "
                "PHASE LIQUID % 1 1 ! CONSTITUENT LIQUID :A,B,C,D: ! PHASE FCC_A1 % 1 1 ! "
                "CONSTITUENT FCC_A1 :A,B,C,D: ! " + extra)
        for ph in ("LIQUID", "FCC_A1"):
            for el in "ABCD":
                text += "PARAMETER G(%s,%s;0) 298.15 -1000*T; 6000 N ! " % (ph, el)
        path = tmp_path / "m.tdb"
        path.write_text(text + params)
        return thermo.load_tdb_model(str(path), components=list(comps), phases=phases)

    quaternary = "PARAMETER L(LIQUID,A,B,C,D;0) 298.15 5; 6000 N !"
    with pytest.raises(ValueError, match="4 species"):
        model(quaternary, comps="ABCD")
    model(quaternary)  # D not selected: the parameter does not enter
    with pytest.raises(ValueError, match="0..4"):
        model("PARAMETER L(LIQUID,A,B;5) 298.15 5; 6000 N !")
    with pytest.raises(ValueError, match="undefined"):
        model("PARAMETER L(LIQUID,A,B;0) 298.15 5+UNDEF#; 6000 N !")
    model("PARAMETER L(LIQUID,A,D;0) 298.15 5+UNDEF#; 6000 N !")  # D not selected
    with pytest.raises(ValueError, match="sublattices"):
        model("", phases=["LIQUID", "X"], extra="PHASE X % 2 1 1 ! CONSTITUENT X :A,B,C,D:VA: ! ")


def _excess(model, which, cs, T):
    g0, L, R = model.oracle_phase(which)
    lnT = ot.plog(T)
    unary = sum(c * ot.tpoly(g, T, lnT) for c, g in zip(cs, g0))
    ideal = R * T * sum(np.maximum(c, ot.CMIN) * ot.plog(np.maximum(c, ot.CMIN)) for c in cs)
    return ot.gibbs(cs, T, g0, L, R) - unary - ideal


def _ab_model(tmp_path, params):
    text = ("ELEMENT A X 1 1 1 ! ELEMENT B X 1 1 1 ! ELEMENT C X 1 1 1 ! "
            "PHASE LIQUID % 1 1 ! CONSTITUENT LIQUID :A,B,C: ! PHASE FCC_A1 % 1 1 ! CONSTITUENT FCC_A1 :A,B,C: ! ")
    for ph in ("LIQUID", "FCC_A1"):
        for el in "ABC":
            text += "PARAMETER G(%s,%s;0) 298.15 -1000*T; 6000 N ! " % (ph, el)
    path = tmp_path / "m.tdb"
    path.write_text(text + params)
    return thermo.load_tdb_model(str(path))


def _comps(seed=0, n=400):
    c = np.random.default_rng(seed).dirichlet([1.5, 1.5, 1.5], n)
    return [c[:, 0], c[:, 1], c[:, 2]], np.random.default_rng(seed + 1).uniform(400, 2500, n)


def test_binary_orders_up_to_four_with_reversed_pair(tmp_path):
    """L(LIQUID,B,A;k) for k = 0..4 written with the pair reversed: the excess
    is c_A c_B sum_k L^k_BA (c_B - c_A)^k, computed here with explicit powers
    (the oracle and device use Horner in (c_A - c_B) with (-1)^k L^k)."""
    Lk = [-12000.0, 3100.0, -2200.0, 1700.0, -900.0]
    params = "".join("PARAMETER L(LIQUID,B,A;%d) 298.15 %r+%r*T; 6000 N ! " % (k, v, 0.5 * k)
                     for k, v in enumerate(Lk))
    m = _ab_model(tmp_path, params)
    cs, T = _comps()
    want = sum(cs[0] * cs[1] * (v + 0.5 * k * T) * np.power(cs[1] - cs[0], k) for k, v in enumerate(Lk))
    np.testing.assert_allclose(_excess(m, "liquid", cs, T), want, rtol=1e-11, atol=1e-7)
    np.testing.assert_allclose(_excess(m, "solid", cs, T), 0.0, atol=1e-7)


def test_ternary_terms(tmp_path):
    """Three-species interactions (pycalphad's Muggianu form):
    * only L^0 given: the ternary excess is c_A c_B c_C L^0 (the three
      weights sum to one);
    * L^0, L^1, L^2 distinct and written in the order (C, A, B): order k
      names the k-th written constituent, so the model equals the one
      written (A, B, C) with the values permuted; at c_A = c_B = c_C every
      weight is 1/3;
    * the analytic derivatives along e_l - e_N agree with central
      differences of G (orders 3..4 and ternary terms included)."""
    cs, T = _comps(3)
    m = _ab_model(tmp_path, "PARAMETER L(LIQUID,A,B,C;0) 298.15 7000-2*T; 6000 N ! ")
    np.testing.assert_allclose(_excess(m, "liquid", cs, T), cs[0] * cs[1] * cs[2] * (7000 - 2 * T),
                               rtol=1e-12, atol=1e-9)
    V = {"A": 4000.0, "B": -2500.0, "C": 1200.0}
    written = "".join("PARAMETER L(LIQUID,C,A,B;%d) 298.15 %r; 6000 N ! " % (k, V[x]) for k, x in enumerate("CAB"))
    canon = "".join("PARAMETER L(LIQUID,A,B,C;%d) 298.15 %r; 6000 N ! " % (k, V[x]) for k, x in enumerate("ABC"))
    m1, m2 = _ab_model(tmp_path, written), _ab_model(tmp_path, canon)
    np.testing.assert_array_equal(_excess(m1, "liquid", cs, T), _excess(m2, "liquid", cs, T))
    third = [np.full(3, 1 / 3.0)] * 3
    Tt = np.full(3, 1000.0)
    np.testing.assert_allclose(_excess(m2, "liquid", third, Tt), (1 / 27.0) * sum(V.values()) / 3.0, rtol=1e-12,
                               atol=1e-9)  # G holds -1e6 J/mol of unary terms
    # derivatives: binary orders 3..4 and a ternary together
    extra = ("PARAMETER L(LIQUID,A,B;3) 298.15 1500; 6000 N ! PARAMETER L(LIQUID,B,C;4) 298.15 -800; 6000 N ! ")
    m3 = _ab_model(tmp_path, canon + extra)
    g0, L, R = m3.oracle_phase("liquid")
    d = ot.analytic_derivatives(cs, T, g0, L, R)
    h = 1e-6
    for l in range(2):
        up = [c.copy() for c in cs]
        dn = [c.copy() for c in cs]
        up[l] = up[l] + h
        up[2] = up[2] - h
        dn[l] = dn[l] - h
        dn[2] = dn[2] + h
        fd = (ot.gibbs(up, T, g0, L, R) - ot.gibbs(dn, T, g0, L, R)) / (2 * h)
        np.testing.assert_allclose(d[l], fd, rtol=1e-6, atol=1e-3)


def test_reference_multicomponent_database():
    """mc_simplified.tdb (the reference's tests): 22 elements, LIQUID and a
    magnetic FCC_A1 whose TC / BMAGN parameters the reference's model leaves
    out (it sums G and L only, simulation.py:206-214).  For Al-Cr-Ni: the
    liquid's Al-Ni orders reach 4, and the one ternary L(LIQUID,AL,CR,NI;0) =
    16000 stands for all three orders."""
    db = _db("mc_simplified.tdb")
    assert db.ignored == {"TC": 30, "BMAGN": 20} and not db.multi_sublattice
    m = thermo.load_tdb_model(os.path.join(GOLD, "mc_simplified.tdb"), components=["AL", "CR", "NI"])
    assert m.solid.name == "FCC_A1" and m.liquid.name == "LIQUID"
    assert (1, 4) in m.liquid.L  # pair (AL, NI), order 4
    T = np.array([1000.0, 1700.0])
    for k in range(3):
        np.testing.assert_array_equal(m.liquid.T3[(0, k)].evaluate(T), [16000.0, 16000.0])
    assert m.value("NI_TM") > 1700
