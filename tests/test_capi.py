"""The C-ABI library loads and exports every entry point include/*.h declares
(no compute calls: these run without a GPU)."""
import ctypes as C
import glob
import os
import re

from mantlepar_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for header in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(header).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(mp_\w+)\s*\(", text, re.M):
            names.add(m.group(1))
    return names


def test_header_declares_the_entry_points():
    names = declared_functions()
    for required in ("mp_transport_create", "mp_transport_step", "mp_transport_phase",
                     "mp_phase_evaluate", "mp_last_error"):
        assert required in names


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(_lib.LIB_PATH)
    missing = [n for n in sorted(declared_functions()) if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    bound = {name for name, _, _ in _lib.SIGNATURES}
    assert declared_functions() <= bound | set(), declared_functions() - bound


def test_defaults_mirror_the_reference():
    from mantlepar_amd.transport import default_phase, default_weno
    w = default_weno()
    assert (w.epsilon, w.s, w.large_weight, list(w.small_weight), w.use_constant) == \
        (1e-2, 1, 1.0, [1.0] * 4, 0)
    assert w.constant_linear_weight == 1e-5
    p = default_phase()
    assert (p.Tm0, p.Te0, p.nu, p.L, p.cp, p.Xe) == (2000.0, 1480.0, 6.5e6, 5.0e5, 1200.0, 0.25)
    assert (p.rho, p.rhor, p.mus, p.mul, p.k0, p.g, p.alpha0) == \
        (3000.0, 500.0, 1.0e19, 1.0, 1.0e-8, 10.0, 3.0e-5)


def test_phase_boundaries_entry_point_matches_oracle():
    """mp_phase_boundaries is host arithmetic (long double, like the reference)."""
    import oracle as O
    from mantlepar_amd.transport import EutecticModel
    m = EutecticModel()
    o = O.PhaseModel()
    for P in (0.0, 1e5, 1e9, -1e7):
        for c in (0.0, 1e-300, 0.1, 0.25):
            assert m.boundaries(c, P) == o.boundaries(c, P)


def test_missing_library_fails_loudly(tmp_path, monkeypatch):
    # monkeypatch restores the path and the loaded handle afterwards; the
    # module is not reloaded, so its exception class stays the one other tests
    # imported
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "absent.so"))
    monkeypatch.setattr(_lib, "_lib", None)
    try:
        _lib.lib()
    except ImportError as e:
        assert "no CPU fallback" in str(e)
    else:
        raise AssertionError("expected ImportError")


def _golden(name):
    import numpy as np
    return dict(np.load(os.path.join(ROOT, "tests", "golden", name), allow_pickle=False))


def _params(row):
    from mantlepar_amd.transport import default_phase
    import oracle as O
    names = [n for n, _ in O.Material._fields_]
    return default_phase(**dict(zip(names, row)))


def test_derived_scales_match_reference_golden():
    """mp_phase_derive (host) is bitwise the reference's DerivedParameters."""
    import numpy as np
    g = _golden("phase_golden.npz")
    for row, ref in zip(g["params"], g["derived"]):
        out = np.zeros(11)
        assert _lib.lib().mp_phase_derive(C.byref(_params(row)), _lib.dptr(out)) == 0
        assert np.array_equal(out, ref)
    bad = _params(g["params"][0])
    bad.Xe = 1.0
    out = np.zeros(11)
    assert _lib.lib().mp_phase_derive(C.byref(bad), _lib.dptr(out)) == 1   # invalid_argument


def test_static_pressure_matches_reference_golden():
    """mp_phase_static_pressure (host) against the reference's staticPressure:
    values bitwise, and its failures in the reference's categories
    (1 invalid_argument -> MP_ERR_ARG, 2 overflow_error -> MP_ERR_FP)."""
    g = _golden("static_pressure_golden.npz")
    out = C.c_double()
    for k in range(g["zD"].size):
        p = _params(g["params"][int(g["set"][k])])
        rc = _lib.lib().mp_phase_static_pressure(C.byref(p), float(g["zD"][k]), float(g["length"][k]),
                                                 int(g["scaled"][k]), C.byref(out))
        st = int(g["status"][k])
        assert rc == {0: 0, 1: 1, 2: 2}[st], (k, rc, st)
        if st == 0:
            assert out.value == g["P"][k] or (out.value == 0 and g["P"][k] == 0), k


def test_cpp_facade_program_links():
    """The compiled C++ caller of include/mantlepar_amd.hpp loads the library
    (usage message, no device call)."""
    import subprocess
    exe = os.path.join(ROOT, "tests", "cpp", "facade_check")
    assert os.path.exists(exe), "build() compiles tests/cpp/facade_check"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr
