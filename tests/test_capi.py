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
