"""The oracle's EutecticModel pinned against the reference.

* Golden vectors written by the reference's own src/phase (compiled unmodified;
  tests/golden/make_phase_golden.py): bit-for-bit equality.
* The known answers and identities of src/phase/test/test_eutectic.cpp:
  scales (:326-332), regions by forward construction (:122-183), boundary
  ownership and nextafter continuity (:185-235), conservation identities on
  random states (:237-254), finite-difference derivatives (:256-300), invalid
  inputs.
* The hypereutectic mirror against the reference evaluated on the mirrored
  diagram.
"""
import math
import os

import numpy as np
import pytest

import oracle as O

COLS = ["phi1", "phi2", "phil", "cl", "cs", "region", "TDp", "Tep", "dTD_dCD", "dTD_dHD"]


def material(row):
    names = [n for n, _ in O.Material._fields_]
    return O.Material.defaults(**dict(zip(names, row)))


def test_oracle_matches_reference_golden_bitwise(golden_phase):
    g = golden_phase
    models = [O.PhaseModel(material(r)) for r in g["params"]]
    checked = 0
    for k in range(g["H"].size):
        m = models[g["set"][k]]
        status = g["status"][k]
        if status != 0:
            with pytest.raises(O.OracleError):
                m.evaluate(g["H"][k], g["C"][k], g["P"][k])
            continue
        s = m.evaluate(g["H"][k], g["C"][k], g["P"][k])
        for j, name in enumerate(COLS):
            a, b = s[name], g["out"][k, j]
            assert a == b or (math.isnan(a) and math.isnan(b)), (name, k, a, b)
        flags = (s["has_solid"] and 1) + (s["has_liquid"] and 2) + \
            (s["finite_temperature_derivatives"] and 4)
        assert flags == g["out"][k, 10]
        checked += 1
    assert checked > 5000


def test_derived_scales_match_reference(golden_phase):
    g = golden_phase
    for row, d in zip(g["params"], g["derived"]):
        m = O.PhaseModel(material(row))
        got = [m.derived[n] for n in ("dT", "h0", "gamma", "LD", "TDm0", "TDe0", "invk0", "l0",
                                      "u0", "p0", "t0")]
        assert got == list(d)


def test_default_scales_closed_form():
    d = O.PhaseModel().derived
    assert d["dT"] == pytest.approx(520.0, rel=3e-12)
    assert d["h0"] == pytest.approx(624000.0, rel=3e-12)
    assert d["TDm0"] == pytest.approx(50.0 / 13.0, rel=3e-12)
    assert d["TDe0"] == pytest.approx(37.0 / 13.0, rel=3e-12)
    assert d["LD"] == pytest.approx(125.0 / 156.0, rel=3e-12)


def _sets():
    alt = dict(Tm0=2100.0, Te0=1500.0, cp=1350.0, L=6.0e5, Xe=0.3, nu=8.0e6, rho=3100.0,
               rhor=450.0, mus=3.0e19, mul=2.0, k0=2.0e-8, g=9.81, alpha0=4.0e-5)
    return [O.Material.defaults(), O.Material.defaults(**alt),
            O.Material.defaults(**dict(alt, nu=-8.0e6))]


PRESSURES = (0.0, 1.0e5, 1.0e9, -1.0e7)


def near(a, b, rtol=3e-12, atol=3e-13):
    assert abs(a - b) <= atol + rtol * max(abs(a), abs(b)), (a, b)


def check_identities(m, s, H, C, P):
    for v in (s["phi1"], s["phi2"], s["phil"], s["cs"], s["cl"]):
        assert 0.0 <= v <= 1.0
    near(s["phi1"] + s["phi2"] + s["phil"], 1.0)
    near(s["TDp"] + m.derived["LD"] * s["phil"], H)
    near(s["phi2"] + s["phil"] * s["cl"], C, 3e-12, 0.0)
    assert s["region"] == m.classify(H, C, P)


@pytest.mark.parametrize("mi", range(3))
def test_regions_forward_construction(mi):
    mat = _sets()[mi]
    m = O.PhaseModel(mat)
    LD, Xe = m.derived["LD"], mat.Xe
    for P in PRESSURES:
        C = 0.4 * Xe
        b = m.boundaries(C, P)
        s = m.evaluate(b["Tep"] - 0.25, C, P)
        assert s["region"] == O.REGION["TwoSolids"]
        for liquid in (0.1, 0.2, 0.3):
            s = m.evaluate(b["Tep"] + LD * liquid, C, P)
            assert s["region"] == O.REGION["Eutectic"]
            near(s["phil"], liquid)
            near(s["phi2"], C - liquid * Xe)
        for off in (0.15, 0.5, 0.85):
            for liquid in (0.05, 0.3, 0.8, 0.99):
                T = b["Tep"] + off
                cl = Xe * (1.0 - off)
                H, Cc = T + LD * liquid, liquid * cl
                s = m.evaluate(H, Cc, P)
                assert s["region"] == O.REGION["SolidLiquid"]
                near(s["TDp"], T)
                near(s["phil"], liquid)
                near(s["cl"], cl)
                check_identities(m, s, H, Cc, P)
        s = m.evaluate(b["H_liquidus"] + 0.3, C, P)
        assert s["region"] == O.REGION["Liquid"] and s["phil"] == 1.0
        pure = m.boundaries(0.0, P)
        assert m.evaluate(pure["Tmp"] - 0.3, 0.0, P)["region"] == O.REGION["PureSolid"]
        for liquid in (0.1, 0.5, 0.9):
            s = m.evaluate(pure["Tmp"] + LD * liquid, 0.0, P)
            assert s["region"] == O.REGION["PureMelting"]
            near(s["phil"], liquid)


@pytest.mark.parametrize("mi", range(3))
def test_boundary_ownership_and_continuity(mi):
    mat = _sets()[mi]
    m = O.PhaseModel(mat)
    for P in PRESSURES:
        for C in (0.0, 1e-300, 1e-18, 0.4 * mat.Xe, np.nextafter(mat.Xe, 0.0), mat.Xe):
            b = m.boundaries(C, P)
            for H in (b["Tep"], b["Tmp"], b["H_eutectic_end"], b["H_liquidus"]):
                lo = m.evaluate(np.nextafter(H, -np.inf), C, P)
                at = m.evaluate(H, C, P)
                hi = m.evaluate(np.nextafter(H, np.inf), C, P)
                near(lo["TDp"], at["TDp"])
                near(hi["TDp"], at["TDp"])
                near(lo["phil"], at["phil"])
                near(hi["phil"], at["phil"])
            assert m.classify(b["H_liquidus"], C, P) == O.REGION["Liquid"]
        C = 0.4 * mat.Xe
        b = m.boundaries(C, P)
        assert m.classify(b["Tep"], C, P) == O.REGION["TwoSolids"]
        assert m.classify(np.nextafter(b["Tep"], np.inf), C, P) == O.REGION["Eutectic"]
        assert m.classify(b["H_eutectic_end"], C, P) == O.REGION["Eutectic"]
        assert m.classify(np.nextafter(b["H_eutectic_end"], np.inf), C, P) == O.REGION["SolidLiquid"]
        assert m.classify(np.nextafter(b["H_liquidus"], -np.inf), C, P) == O.REGION["SolidLiquid"]
        pure = m.boundaries(0.0, P)
        corner = m.evaluate(pure["Tmp"], 0.0, P)
        assert corner["region"] == O.REGION["PureSolid"]
        assert math.isinf(corner["dTD_dCD"]) and corner["dTD_dCD"] < 0
        e = m.boundaries(mat.Xe, P)
        assert e["H_eutectic_end"] == e["H_liquidus"]


def test_conservation_identities_random_states():
    rng = np.random.default_rng(13579)
    for mat in _sets():
        m = O.PhaseModel(mat)
        LD = m.derived["LD"]
        for P in PRESSURES:
            b = m.boundaries(0.4 * mat.Xe, P)
            for i in range(600):
                C = mat.Xe * rng.random() if i % 2 == 0 else mat.Xe * 10.0 ** (-300 * rng.random())
                H = b["Tep"] - 0.5 + (2.0 + LD) * rng.random()
                check_identities(m, m.evaluate(H, C, P), H, C, P)


def test_derivatives_by_finite_differences():
    for mat in _sets():
        m = O.PhaseModel(mat)
        for P in PRESSURES:
            C = 0.4 * mat.Xe
            b = m.boundaries(C, P)
            for H in (b["Tep"] - 0.2, (b["Tep"] + b["H_eutectic_end"]) / 2,
                      (b["H_eutectic_end"] + b["H_liquidus"]) / 2, b["H_liquidus"] + 0.2):
                s = m.evaluate(H, C, P)
                st = 1e-6
                fdH = (m.evaluate(H + st, C, P)["TDp"] - m.evaluate(H - st, C, P)["TDp"]) / (2 * st)
                fdC = (m.evaluate(H, C + st, P)["TDp"] - m.evaluate(H, C - st, P)["TDp"]) / (2 * st)
                near(s["dTD_dHD"], fdH, 2e-8, 2e-8)
                near(s["dTD_dCD"], fdC, 2e-8, 2e-8)


def test_invalid_inputs_raise():
    m = O.PhaseModel()
    with pytest.raises(O.OracleError):
        m.evaluate(1.0, -1e-12, 0.0)
    with pytest.raises(O.OracleError):
        m.evaluate(1.0, np.nextafter(m.material.Xe, 1.0), 0.0)
    with pytest.raises(O.OracleError):
        m.evaluate(float("nan"), 0.1, 0.0)
    with pytest.raises(O.OracleError):
        O.PhaseModel(O.Material.defaults(Tm0=1000.0, Te0=1500.0))
    with pytest.raises(O.OracleError):
        O.PhaseModel(O.Material.defaults(Xe=1.0))


def test_hypereutectic_mirror_matches_reference_on_mirrored_diagram(golden_phase):
    """For C > Xe the extension evaluates the reference model at 1-C on the diagram
    with eutectic 1-Xe and maps phases back; the golden set 4 is that diagram."""
    g = golden_phase
    cfg = O.PhaseModel(material(g["params"][3]))
    sel = np.nonzero((g["set"] == 4) & (g["status"] == 0))[0]
    assert sel.size > 500
    for k in sel:
        H, Cm, P = g["H"][k], g["C"][k], g["P"][k]
        C = 1.0 - Cm
        if not (C > cfg.material.Xe) or 1.0 - C != Cm:
            continue   # only inputs whose mirror is exact
        s = cfg.evaluate(H, C, P, hyper=True)
        ref = dict(zip(COLS, g["out"][k]))
        assert s["TDp"] == ref["TDp"] and s["phil"] == ref["phil"]
        assert s["phi1"] == ref["phi2"] and s["phi2"] == ref["phi1"]
        assert s["region"] == int(ref["region"])
        if s["has_liquid"]:
            assert s["cl"] == 1.0 - ref["cl"]
        check_identities_mirror(cfg, s, H, C)


def check_identities_mirror(m, s, H, C):
    near(s["phi1"] + s["phi2"] + s["phil"], 1.0)
    near(s["TDp"] + m.derived["LD"] * s["phil"], H)
    near(s["phi2"] + s["phil"] * s["cl"], C, 3e-12, 1e-15)


def test_hypereutectic_continuity_at_eutectic():
    m = O.PhaseModel(O.Material.defaults(Tm0=1600.0, Te0=600.0, cp=1200.0, L=1.2e6, Xe=0.4))
    Xe = 0.4
    for H in np.linspace(0.3, 2.8, 41):
        a = m.evaluate(H, Xe, 0.0, hyper=True)
        b = m.evaluate(H, np.nextafter(Xe, 1.0), 0.0, hyper=True)
        near(a["TDp"], b["TDp"], 1e-12, 1e-12)
        near(a["phil"], b["phil"], 1e-12, 1e-12)


def test_batch_equals_scalar():
    m = O.PhaseModel(O.Material.defaults(Tm0=1600.0, Te0=600.0, cp=1200.0, L=1.2e6, Xe=0.4))
    rng = np.random.default_rng(3)
    H = 0.2 + 2.5 * rng.random(500)
    C = rng.random(500)
    T, phi, cl, cs, reg = m.batch(H, C, 0.0, hyper=True)
    for k in range(0, 500, 7):
        s = m.evaluate(H[k], C[k], 0.0, hyper=True)
        assert (T[k], phi[k], cl[k], cs[k], reg[k]) == (s["TDp"], s["phil"], s["cl"], s["cs"],
                                                        s["region"])


@pytest.mark.skipif(not O.ref_phase_available(), reason="reference phase build absent")
def test_oracle_matches_reference_build_on_fresh_vectors():
    rng = np.random.default_rng(99)
    for mat in _sets():
        m = O.PhaseModel(mat)
        n = 4000
        C = mat.Xe * rng.random(n)
        H = m.boundaries(0.0, 0.0)["Tep"] - 0.5 + 3 * rng.random(n)
        out, st = O.ref_phase_evaluate(mat, H, C, np.zeros(n))
        for k in range(0, n, 3):
            s = m.evaluate(H[k], C[k], 0.0)
            assert [s[c] for c in COLS] == list(out[k, :10])


# --------------------------------------------------------- static pressure
def test_static_pressure_matches_reference_golden_bitwise():
    """EutecticModel::staticPressure (eutectic_rescaled.cpp:151-161) against the
    reference's own values (tests/golden/make_static_pressure_golden.py)."""
    g = dict(np.load(os.path.join(os.path.dirname(__file__), "golden",
                                  "static_pressure_golden.npz"), allow_pickle=False))
    models = [O.PhaseModel(material(r)) for r in g["params"]]
    ok = 0
    for k in range(g["zD"].size):
        m = models[g["set"][k]]
        length = float(g["length"][k]) if g["scaled"][k] else None
        if g["status"][k] != 0:
            with pytest.raises(O.OracleError):
                m.static_pressure(g["zD"][k], length)
            continue
        assert m.static_pressure(g["zD"][k], length) == g["P"][k], k
        ok += 1
    assert ok > 1000 and (g["status"] == 1).sum() > 100 and (g["status"] == 2).sum() > 0


def test_static_pressure_known_answers():
    """src/phase/test/test_eutectic.cpp:337-341, 455-463."""
    for mat in (O.Material.defaults(), O.Material.defaults(rho=3100.0, g=9.81, mus=3e19, k0=2e-8,
                                                           mul=2.0)):
        m = O.PhaseModel(mat)
        assert m.static_pressure(0.0) == 0.0
        assert math.isclose(m.static_pressure(0.5), mat.rho * mat.g * 0.5 * m.derived["l0"],
                            rel_tol=1e-15)
        assert math.isclose(m.static_pressure(0.5, 1000.0), mat.rho * mat.g * 500.0, rel_tol=1e-15)
        for bad in (math.nan, math.inf, -math.inf):
            with pytest.raises(O.OracleError, match="zD"):
                m.static_pressure(bad)
        for length in (0.0, -1.0, math.inf, math.nan):
            with pytest.raises(O.OracleError, match="length_scale"):
                m.static_pressure(0.5, length)
    # linear in depth: P(a) + P(b) = P(a + b) to rounding
    m = O.PhaseModel()
    assert math.isclose(m.static_pressure(0.25) + m.static_pressure(0.5), m.static_pressure(0.75),
                        rel_tol=4e-16)
