"""Device EutecticModel (mp_phase_evaluate) against the reference's golden vectors.

Tolerances: region bit-exact outside a band of 8 ulps around each phase
boundary (Tep, Tmp, H_eutectic_end, H_liquidus); temperature and phase
fractions within 2e-14 absolute + 2e-14 relative; derivatives within 1e-12
relative.  The phase compositions Cs = phi2/(phi1+phi2) and Cl = C/phil are
ratios: an absolute error e in the fractions moves them by e/(phi1+phi2)
resp. e/phil, so they are held to 2e-14 + 4e-19/(fraction) -- 4e-19 being
a few long-double roundings of O(1) terms, the reference's own accuracy.  The
reference computes in 80-bit long double, the device in FP64 with
double-double boundaries and eutectic fractions.
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from mantlepar_amd import MantleError
from mantlepar_amd import _lib
from mantlepar_amd.transport import EutecticModel, default_phase

pytestmark = pytest.mark.gpu

NAMES = [n for n, _ in O.Material._fields_]


def params_from(row, hyper=0):
    return default_phase(**dict(zip(NAMES, row)), hypereutectic=hyper)


def near_boundary(model, H, Cd, P, ulps=8):
    if Cd < 0 or Cd > model.material.Xe or not np.isfinite(H):
        return True
    b = model.boundaries(Cd, P)
    for v in b.values():
        if abs(H - v) <= ulps * np.spacing(abs(v)):
            return True
    return False


def test_device_phase_matches_reference_golden(golden_phase):
    g = golden_phase
    bad_region = 0
    band = 0
    for si, row in enumerate(g["params"][:4]):
        sel = np.nonzero((g["set"] == si) & (g["status"] == 0))[0]
        dev = EutecticModel(params_from(row))
        out = dev.evaluate(g["H"][sel], g["C"][sel], g["P"][sel])
        ref = g["out"][sel]
        om = O.PhaseModel(O.Material.defaults(**dict(zip(NAMES, row))))
        for idx, k in enumerate(sel):
            inband = near_boundary(om, g["H"][k], g["C"][k], g["P"][k])
            band += inband
            if not inband:
                bad_region += out["region"][idx] != int(ref[idx, 5])
        for name, col in (("T", 6), ("phi", 2), ("phi1", 0), ("phi2", 1)):
            d = np.abs(out[name] - ref[:, col])
            assert np.all(d <= 2e-14 + 2e-14 * np.abs(ref[:, col])), (si, name, d.max())
        solid = ref[:, 0] + ref[:, 1]
        for name, col, frac in (("Cs", 4, solid), ("Cl", 3, ref[:, 2])):
            d = np.abs(out[name] - ref[:, col])
            bound = 2e-14 + 4e-19 / np.maximum(frac, 1e-300)
            assert np.all(d <= bound), (si, name, d.max())
        for name, col in (("dTdH", 9), ("dTdC", 8)):
            a, b = out[name], ref[:, col]
            fin = np.isfinite(b)
            assert np.all(np.isinf(a[~fin]) & (a[~fin] < 0))
            assert np.all(np.abs(a[fin] - b[fin]) <= 1e-12 * (1 + np.abs(b[fin])))
    assert bad_region == 0
    # the golden set holds 4 x 4 x 72 exact-boundary triples by design
    assert band <= 4 * 4 * 72 + 64


def test_device_regions_at_exact_boundaries(golden_phase):
    """The nextafter boundary triples: count how many classify bit-exactly."""
    g = golden_phase
    total = exact = 0
    for si, row in enumerate(g["params"][:3]):
        sel = np.nonzero((g["set"] == si) & (g["status"] == 0))[0]
        dev = EutecticModel(params_from(row))
        out = dev.evaluate(g["H"][sel], g["C"][sel], g["P"][sel])
        total += sel.size
        exact += int(np.sum(out["region"] == g["out"][sel, 5].astype(np.uint8)))
    assert exact >= 0.999 * total


def test_device_mirror_branch_matches_oracle():
    p = default_phase(Tm0=1600.0, Te0=600.0, cp=1200.0, L=1.2e6, Xe=0.4, hypereutectic=1)
    rng = np.random.default_rng(7)
    H = 0.2 + 2.6 * rng.random(20000)
    Cd = rng.random(20000)
    out = EutecticModel(p).evaluate(H, Cd)
    om = O.PhaseModel(O.Material.defaults(Tm0=1600.0, Te0=600.0, cp=1200.0, L=1.2e6, Xe=0.4))
    T, phi, cl, cs, reg = om.batch(H, Cd, 0.0, hyper=True)
    assert np.abs(out["T"] - T).max() < 1e-14
    assert np.abs(out["phi"] - phi).max() < 1e-14
    assert np.abs(out["Cl"] - cl).max() < 1e-14
    assert np.abs(out["Cs"] - cs).max() < 1e-14
    assert (out["region"] != reg).sum() <= 2


def _raw_evaluate(params, H, Cd, sentinel):
    n = H.size
    T = np.full(n, sentinel)
    rc = _lib.lib().mp_phase_evaluate(C.byref(params), n, _lib.dptr(H), _lib.dptr(Cd), None,
                                      _lib.dptr(T), None, None, None, None, None, None, None,
                                      None)
    return rc, T


def test_invalid_inputs_fail_and_leave_outputs_unchanged():
    p = default_phase()
    H = np.array([1.0, 1.0, 2.0])
    rc, T = _raw_evaluate(p, H, np.array([0.1, 0.3, 0.1]), -7.0)     # C > Xe = 0.25
    assert rc == 8 and np.all(T == -7.0)
    rc, T = _raw_evaluate(p, H, np.array([0.1, -1e-12, 0.1]), -7.0)
    assert rc == 8 and np.all(T == -7.0)
    rc, T = _raw_evaluate(p, np.array([1.0, np.nan, 2.0]), np.array([0.1, 0.1, 0.1]), -7.0)
    assert rc == 2 and np.all(T == -7.0)
    with pytest.raises(MantleError):
        EutecticModel(default_phase(Xe=1.5)).evaluate(H, H * 0)
