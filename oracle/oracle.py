"""ctypes bindings for the C oracle (test infrastructure only)."""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# MPO_ORACLE_LIB selects another build of the same source (the sanitizer build)
LIB_PATH = os.environ.get("MPO_ORACLE_LIB") or os.path.join(HERE, "liboracle_mp.so")
REF_PATH = os.path.join(HERE, "_ref", "libref_phase.so")

__all__ = [
    "lib", "OracleError", "build", "Material", "PhaseModel", "Transport", "gauss_rule",
    "ref_phase_available", "ref_phase_evaluate", "ref_phase_derived", "ref_static_pressure", "WenoOptions",
    "StencilPoly", "Recon", "llf_point", "integrate_edge_flux", "REGION",
    "lagrange_weights", "Sampling", "diffusive_sampling", "diffusive_flux", "LAGRANGE",
    "SIDE", "DIFF_BC", "MAX_SAMPLES", "CoupledParams", "transport_velocities", "edge_count",
]

LAGRANGE = {"Middle": 0, "EndLo": 1, "EndHi": 2}
SIDE = {"Left": 0, "Right": 1, "Boundary": 2}
DIFF_BC = {"PrescribedState": 0, "PrescribedQuantity": 1, "PrescribedFlux": 2, "ZeroFlux": 3}
MAX_SAMPLES = 8

REGION = {"PureSolid": 1, "TwoSolids": 2, "Eutectic": 3, "SolidLiquid": 4, "Liquid": 5,
          "PureMelting": 6}


class OracleError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"oracle error {code}: {message}")
        self.code = code


def build(ref=False):
    """Compile the oracle (and the reference phase library when ``ref``)."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
    if ref:
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


class Point(C.Structure):
    _fields_ = [("p", C.c_double * 2)]


class Material(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("Tm0", "Te0", "nu", "L", "cp", "Xe", "rho", "rhor", "mus", "mul", "k0", "g",
                 "alpha0")]

    @classmethod
    def defaults(cls, **kw):
        m = cls()
        lib().mpo_material_defaults(C.byref(m))
        for k, v in kw.items():
            setattr(m, k, v)
        return m

    def as_array(self):
        return np.array([getattr(self, n) for n, _ in self._fields_], dtype=np.float64)


class Derived(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("dT", "h0", "gamma", "LD", "TDm0", "TDe0", "invk0", "l0", "u0", "p0", "t0")]


class _PhaseModel(C.Structure):
    _fields_ = [("p", Material), ("d", Derived)]


class Bounds(C.Structure):
    _fields_ = [("Tep", C.c_double), ("Tmp", C.c_double), ("H_eutectic_end", C.c_double),
                ("H_liquidus", C.c_double)]


class PhaseState(C.Structure):
    _fields_ = [("phi1", C.c_double), ("phi2", C.c_double), ("phil", C.c_double),
                ("cl", C.c_double), ("cs", C.c_double), ("region", C.c_int),
                ("TDp", C.c_double), ("Tep", C.c_double), ("dTD_dCD", C.c_double),
                ("dTD_dHD", C.c_double), ("has_solid", C.c_int), ("has_liquid", C.c_int),
                ("finite_temperature_derivatives", C.c_int)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


MAX_STENCIL = 12
MAX_CAND = 8


class Stencil(C.Structure):
    _fields_ = [("start_i", C.c_int), ("start_j", C.c_int), ("size_i", C.c_int),
                ("size_j", C.c_int), ("count", C.c_int), ("order", C.c_int),
                ("center", Point), ("h", C.c_double),
                ("coef", C.c_double * (MAX_STENCIL * MAX_STENCIL)),
                ("sigma", C.c_double * (MAX_STENCIL * MAX_STENCIL)),
                ("target_index", C.c_int),
                ("target_sigma", C.c_double * (MAX_STENCIL * MAX_STENCIL)),
                ("corners", (Point * 4) * MAX_STENCIL)]


class Family(C.Structure):
    _fields_ = [("size_i", C.c_int), ("size_j", C.c_int), ("order", C.c_int),
                ("noffsets", C.c_int), ("off_i", C.c_int * MAX_CAND),
                ("off_j", C.c_int * MAX_CAND), ("linear_weight", C.c_double * MAX_CAND),
                ("has_weights", C.c_int), ("target_smoothness", C.c_int)]


class WenoOptions(C.Structure):
    _fields_ = [("large", Family), ("small", Family), ("use_constant", C.c_int),
                ("constant_linear_weight", C.c_double), ("epsilon", C.c_double),
                ("s", C.c_int)]

    @classmethod
    def production(cls):
        o = cls()
        lib().mpo_production_options(C.byref(o))
        return o


class Candidate(C.Structure):
    _fields_ = [("family", C.c_int), ("configured_index", C.c_int), ("linear_index", C.c_int),
                ("linear_weight", C.c_double), ("stencil", C.POINTER(Stencil)),
                ("start_i", C.c_int), ("start_j", C.c_int), ("center", Point), ("h", C.c_double),
                ("centered", C.c_double * MAX_STENCIL), ("smoothness", C.c_double),
                ("weight", C.c_double)]


class _Recon(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("ti", C.c_int), ("tj", C.c_int),
                ("req_bi", C.c_int), ("req_bj", C.c_int), ("req_ei", C.c_int),
                ("req_ej", C.c_int), ("opt", WenoOptions), ("ncand", C.c_int),
                ("cand", Candidate * MAX_CAND), ("has_weights", C.c_int),
                ("target_average", C.c_double), ("constant_weight", C.c_double),
                ("area", C.c_double)]


class _Mesh(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("coords", C.POINTER(C.c_double))]


_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


class Sampling(C.Structure):
    """Normal-line diffusion sampling of one edge (mpo_sampling)."""
    _fields_ = [("nq", C.c_int), ("count", C.c_int), ("boundary", C.c_int),
                ("length", C.c_double), ("normal", Point), ("points", Point * 16),
                ("spacing", C.c_double * 16), ("weights", C.c_double * 16),
                ("coefficients", C.c_double * (MAX_SAMPLES + 1)),
                ("samples", Point * (16 * (MAX_SAMPLES + 1))),
                ("sides", C.c_int * (16 * (MAX_SAMPLES + 1)))]

    def sample_points(self):
        n = self.nq * self.count
        return np.array([[self.samples[a].p[0], self.samples[a].p[1]] for a in range(n)])

    def sample_sides(self):
        return np.array([self.sides[a] for a in range(self.nq * self.count)])

    def quadrature_points(self):
        return np.array([[self.points[q].p[0], self.points[q].p[1]] for q in range(self.nq)])


class CoupledParams(C.Structure):
    """Parameters of the coupled RHS (mpo_coupled_params)."""
    _fields_ = [("LD", C.c_double), ("theta", C.c_double), ("cooling", C.c_double),
                ("kappa", C.c_double), ("lf_mode", C.c_int), ("global_speed", C.c_double),
                ("cell_points", C.c_int), ("samples", C.c_int), ("extent", C.c_double),
                ("pressure", C.c_double), ("lithostatic", C.c_int), ("surface_y", C.c_double),
                ("hypereutectic", C.c_int), ("admissible_c_max", C.c_double)]

    @classmethod
    def make(cls, **kw):
        d = dict(LD=1.0, theta=1.0, cooling=0.0, kappa=0.0, lf_mode=0, global_speed=0.0,
                 cell_points=5, samples=4, extent=0.9, pressure=0.0, lithostatic=0,
                 surface_y=0.0, hypereutectic=0, admissible_c_max=0.4)
        d.update(kw)
        return cls(**d)


class _Flow(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("edge_stokes", "edge_darcy", "edge_phi", "cell_stokes",
                                          "cell_darcy", "cell_phi")]


class _CoupledRates(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("boundaryH", "boundaryC", "sourceH", "maximumRate",
                                          "divergence")]


class _CoupledTime(C.Structure):
    _fields_ = [("cfl", C.c_double), ("initial", C.c_double), ("maximum", C.c_double),
                ("minimum", C.c_double), ("end", C.c_double), ("fixed", C.c_int)]


class _CoupledStep(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("dt", "courant", "balanceH", "balanceC", "boundaryH",
                                          "boundaryC", "sourceH", "maximumRate")] + \
               [("retries", C.c_int)]


def edge_count(nx, ny):
    return nx * (ny + 1) + (nx + 1) * ny


def _flow(flow):
    """flow: dict of arrays edge_stokes [E][q][2], edge_darcy, edge_phi [E][q],
    cell_stokes [N][n*n][2], cell_darcy, cell_phi [N][n*n]."""
    keep = {k: np.ascontiguousarray(flow[k], dtype=np.float64) for k in
            ("edge_stokes", "edge_darcy", "edge_phi", "cell_stokes", "cell_darcy", "cell_phi")}
    f = _Flow(**{k: v.ctypes.data_as(C.c_void_p) for k, v in keep.items()})
    return f, keep


def transport_velocities(solid, relative, phi, theta, a, b):
    """PhaseTransportVelocities (src/couple/evolve.cpp:36-52); a, b = (cl, cs)."""
    m = (C.c_double * 2)()
    e = (C.c_double * 2)()
    lib().mpo_phase_transport_velocities((C.c_double * 2)(*solid), (C.c_double * 2)(*relative),
                                         phi, theta, a[0], a[1], b[0], b[1], m, e)
    return np.array(m[:]), np.array(e[:])


def _declare(L):
    L.mpo_last_error.restype = C.c_char_p
    L.mpo_gauss_rule.argtypes = [C.c_int, _dp, _dp]
    L.mpo_material_defaults.argtypes = [C.POINTER(Material)]
    L.mpo_phase_set_parameters.argtypes = [C.POINTER(_PhaseModel), C.POINTER(Material)]
    L.mpo_phase_boundaries.argtypes = [C.POINTER(_PhaseModel), C.c_double, C.c_double,
                                       C.POINTER(Bounds)]
    L.mpo_phase_evaluate.argtypes = [C.POINTER(_PhaseModel), C.c_double, C.c_double,
                                     C.c_double, C.POINTER(PhaseState)]
    L.mpo_phase_evaluate_ext.argtypes = L.mpo_phase_evaluate.argtypes
    L.mpo_phase_classify.argtypes = [C.POINTER(_PhaseModel), C.c_double, C.c_double,
                                     C.c_double, C.POINTER(C.c_int)]
    L.mpo_static_pressure.argtypes = [C.POINTER(_PhaseModel), C.c_double, C.POINTER(C.c_double)]
    L.mpo_static_pressure_scaled.argtypes = [C.POINTER(_PhaseModel), C.c_double, C.c_double,
                                             C.POINTER(C.c_double)]
    L.mpo_phase_batch.argtypes = [C.POINTER(_PhaseModel), C.c_int64, _dp, _dp, C.c_double,
                                  C.c_int, _dp, _dp, _dp, _dp, _u8p]
    L.mpo_production_options.argtypes = [C.POINTER(WenoOptions)]
    L.mpo_stencil_init.argtypes = [C.POINTER(Stencil), C.c_void_p, C.c_int, C.c_int, C.c_int,
                                   C.c_int, C.c_int, C.c_double]
    L.mpo_stencil_set_target.argtypes = [C.POINTER(Stencil), C.c_int, C.c_int]
    L.mpo_stencil_smoothness.argtypes = [C.POINTER(Stencil), _dp, C.c_int,
                                         C.POINTER(C.c_double)]
    L.mpo_stencil_eval.argtypes = [C.POINTER(Stencil), _dp, Point, C.c_int, C.c_int,
                                   C.POINTER(C.c_double)]
    L.mpo_recon_storage_count.argtypes = [C.POINTER(WenoOptions)]
    L.mpo_recon_init.argtypes = [C.POINTER(_Recon), C.POINTER(_Mesh), C.c_int, C.c_int,
                                 C.POINTER(WenoOptions), C.c_double, C.c_void_p, C.c_void_p]
    L.mpo_recon_update.argtypes = [C.POINTER(_Recon), _dp, C.c_int, C.c_int, C.c_int, C.c_int,
                                   C.c_double]
    L.mpo_recon_eval.argtypes = [C.POINTER(_Recon), Point, C.c_int, C.c_int,
                                 C.POINTER(C.c_double)]
    L.mpo_llf_point.argtypes = [C.c_double, C.c_double, Point, Point, C.c_int, C.c_double,
                                C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.mpo_integrate_edge_flux.argtypes = [C.POINTER(Point), C.c_int, _dp, _dp,
                                          C.POINTER(Point), _dp, _dp, C.c_int, C.c_double,
                                          C.POINTER(C.c_double)]
    L.mpo_transport_create.argtypes = [C.POINTER(C.c_void_p), C.POINTER(_Mesh), C.c_int,
                                       C.POINTER(WenoOptions), _dp, _dp, C.c_void_p,
                                       C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.c_void_p, C.c_int, C.c_int]
    L.mpo_transport_destroy.argtypes = [C.c_void_p]
    L.mpo_transport_rates.argtypes = [C.c_void_p, _dp, C.POINTER(C.c_double)]
    L.mpo_transport_rhs.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, C.POINTER(C.c_double),
                                    C.POINTER(C.c_double)]
    L.mpo_transport_reconstruct.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
    L.mpo_transport_advance.argtypes = [C.c_void_p, _dp, _dp, C.c_int, C.c_double, C.c_int,
                                        C.c_void_p]
    L.mpo_transport_integral.argtypes = [C.c_void_p, _dp]
    L.mpo_transport_integral.restype = C.c_double
    L.mpo_lagrange_weights.argtypes = [C.c_int, C.c_int, _dp]
    L.mpo_diffusive_sampling.argtypes = [C.POINTER(Point), C.POINTER(Point), C.c_void_p, C.c_int,
                                         _dp, _dp, C.c_int, C.c_double, C.POINTER(Sampling)]
    L.mpo_diffusive_flux.argtypes = [C.POINTER(Sampling), _dp, _dp, C.c_void_p, C.c_void_p,
                                     C.POINTER(C.c_double)]
    L.mpo_transport_set_limiter.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double,
                                            C.c_double]
    L.mpo_transport_set_lf.argtypes = [C.c_void_p, C.c_int, C.c_double]
    L.mpo_transport_limits.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
    L.mpo_transport_traces.argtypes = [C.c_void_p, C.c_int, _dp, _dp]
    L.mpo_transport_set_diffusion.argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_double,
                                              C.c_void_p, C.c_void_p]
    L.mpo_phase_transport_velocities.argtypes = [C.c_double * 2, C.c_double * 2, C.c_double,
                                                 C.c_double, C.c_double, C.c_double,
                                                 C.c_double, C.c_double, C.c_double * 2,
                                                 C.c_double * 2]
    L.mpo_phase_transport_velocities.restype = None
    L.mpo_transport_set_composition_boundary.argtypes = [C.c_void_p] + [C.c_void_p] * 4
    L.mpo_coupled_rhs.argtypes = [C.c_void_p, C.POINTER(_PhaseModel), C.POINTER(CoupledParams),
                                  C.POINTER(_Flow), _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                  C.POINTER(_CoupledRates)]
    L.mpo_coupled_run.argtypes = [C.c_void_p, C.POINTER(_PhaseModel), C.POINTER(CoupledParams),
                                  C.POINTER(_Flow), C.POINTER(_CoupledTime), _dp, _dp,
                                  C.POINTER(C.c_double), C.c_int, C.POINTER(_CoupledStep)]
    L.mpo_mix64.argtypes = [C.c_uint64]
    L.mpo_mix64.restype = C.c_uint64


def _check(code):
    if code:
        raise OracleError(code, lib().mpo_last_error().decode())


def gauss_rule(n):
    x = np.zeros(n)
    w = np.zeros(n)
    _check(lib().mpo_gauss_rule(n, x, w))
    return x, w


class PhaseModel:
    """EutecticModel restatement (src/phase/eutectic_rescaled.h:103-136)."""

    def __init__(self, material=None, **kw):
        self.material = material if material is not None else Material.defaults(**kw)
        self._m = _PhaseModel()
        _check(lib().mpo_phase_set_parameters(C.byref(self._m), C.byref(self.material)))

    @property
    def derived(self):
        return {n: getattr(self._m.d, n) for n, _ in Derived._fields_}

    def boundaries(self, Cd, P=0.0):
        b = Bounds()
        _check(lib().mpo_phase_boundaries(C.byref(self._m), Cd, P, C.byref(b)))
        return {n: getattr(b, n) for n, _ in Bounds._fields_}

    def evaluate(self, H, Cd, P=0.0, hyper=False):
        s = PhaseState()
        f = lib().mpo_phase_evaluate_ext if hyper else lib().mpo_phase_evaluate
        _check(f(C.byref(self._m), H, Cd, P, C.byref(s)))
        return s.as_dict()

    def static_pressure(self, zD, length=None):
        """staticPressure(zD[, length_scale]) (eutectic_rescaled.cpp:151-161)."""
        out = C.c_double()
        if length is None:
            _check(lib().mpo_static_pressure(C.byref(self._m), zD, C.byref(out)))
        else:
            _check(lib().mpo_static_pressure_scaled(C.byref(self._m), zD, length, C.byref(out)))
        return out.value

    def classify(self, H, Cd, P=0.0):
        r = C.c_int()
        _check(lib().mpo_phase_classify(C.byref(self._m), H, Cd, P, C.byref(r)))
        return r.value

    def batch(self, H, Cd, P=0.0, hyper=True):
        H = np.ascontiguousarray(H, dtype=np.float64).ravel()
        Cd = np.ascontiguousarray(Cd, dtype=np.float64).ravel()
        n = H.size
        T, phi, cl, cs = (np.zeros(n) for _ in range(4))
        reg = np.zeros(n, dtype=np.uint8)
        _check(lib().mpo_phase_batch(C.byref(self._m), n, H, Cd, P, int(hyper), T, phi, cl, cs,
                                     reg))
        return T, phi, cl, cs, reg


_ref = None


def ref_phase_available():
    return os.path.exists(REF_PATH)


def _reflib():
    global _ref
    if _ref is None:
        _ref = C.CDLL(REF_PATH)
        _ref.ref_phase_evaluate.argtypes = [_dp, C.c_long, _dp, _dp, _dp, _dp,
                                            np.ctypeslib.ndpointer(dtype=np.int32)]
        _ref.ref_phase_derived.argtypes = [_dp, _dp]
        _ref.ref_static_pressure.argtypes = [_dp, C.c_long, _dp, _dp, C.c_int, _dp,
                                             np.ctypeslib.ndpointer(dtype=np.int32)]
    return _ref


def ref_phase_evaluate(material, H, Cd, P):
    """The reference's own EutecticModel::evaluate on arrays (oracle/_ref)."""
    H = np.ascontiguousarray(H, dtype=np.float64)
    Cd = np.ascontiguousarray(Cd, dtype=np.float64)
    P = np.ascontiguousarray(np.broadcast_to(P, H.shape), dtype=np.float64)
    out = np.zeros((H.size, 11))
    status = np.zeros(H.size, dtype=np.int32)
    rc = _reflib().ref_phase_evaluate(material.as_array(), H.size, H, Cd, P, out, status)
    if rc:
        raise OracleError(rc, "reference model construction failed")
    return out, status


def ref_phase_derived(material):
    d = np.zeros(11)
    rc = _reflib().ref_phase_derived(material.as_array(), d)
    return rc, d


def ref_static_pressure(material, zD, length=None):
    """The reference's own EutecticModel::staticPressure on arrays (oracle/_ref)."""
    zD = np.ascontiguousarray(zD, dtype=np.float64).ravel()
    ln = np.ascontiguousarray(np.broadcast_to(0.0 if length is None else length, zD.shape),
                              dtype=np.float64)
    out = np.zeros(zD.size)
    status = np.zeros(zD.size, dtype=np.int32)
    rc = _reflib().ref_static_pressure(material.as_array(), zD.size, zD, ln,
                                       0 if length is None else 1, out, status)
    if rc:
        raise OracleError(rc, "reference model construction failed")
    return out, status



def _corners_array(corners):
    arr = (Point * 4 * MAX_STENCIL)()
    for c, quad in enumerate(corners):
        for k in range(4):
            arr[c][k].p[0] = quad[k][0]
            arr[c][k].p[1] = quad[k][1]
    return arr


class StencilPoly:
    """TensorStencilPoly restatement (geometry-only Initialize)."""

    def __init__(self, corners, start, size, order, h):
        self._s = Stencil()
        self._corners = _corners_array(corners)
        _check(lib().mpo_stencil_init(C.byref(self._s), C.addressof(self._corners), start[0],
                                      start[1], size[0], size[1], order, h))

    def set_target(self, li, lj):
        _check(lib().mpo_stencil_set_target(C.byref(self._s), li, lj))

    def smoothness(self, averages, target=False):
        v = C.c_double()
        _check(lib().mpo_stencil_smoothness(C.byref(self._s),
                                            np.ascontiguousarray(averages, dtype=np.float64),
                                            int(target), C.byref(v)))
        return v.value

    def evaluate(self, averages, point, dx=0, dy=0):
        v = C.c_double()
        pt = Point()
        pt.p[0], pt.p[1] = point
        _check(lib().mpo_stencil_eval(C.byref(self._s),
                                      np.ascontiguousarray(averages, dtype=np.float64), pt, dx,
                                      dy, C.byref(v)))
        return v.value

    @property
    def center(self):
        return (self._s.center.p[0], self._s.center.p[1])

    @property
    def coefficients(self):
        n = self._s.count
        return np.array(self._s.coef[:n * n]).reshape(n, n)


class Recon:
    """Reconstruction restatement: Initialize/Update/Evaluate on a vertex mesh."""

    def __init__(self, coords, nx, ny, target, options=None, h=None):
        self._coords = np.ascontiguousarray(coords, dtype=np.float64)
        self._mesh = _Mesh(nx, ny, self._coords.ctypes.data_as(C.POINTER(C.c_double)))
        self.options = options if options is not None else WenoOptions.production()
        n = lib().mpo_recon_storage_count(C.byref(self.options))
        self._storage = (Stencil * max(n, 1))()
        self._r = _Recon()
        if h is None:
            h = 1.0
        _check(lib().mpo_recon_init(C.byref(self._r), C.byref(self._mesh), target[0], target[1],
                                    C.byref(self.options), h, C.addressof(self._storage), None))

    @property
    def candidate_count(self):
        return self._r.ncand

    def candidate(self, k):
        c = self._r.cand[k]
        s = c.stencil.contents
        return {"family": c.family, "configured_index": c.configured_index,
                "linear_weight": c.linear_weight, "smoothness": c.smoothness,
                "weight": c.weight, "start": (c.start_i, c.start_j),
                "size": (s.size_i, s.size_j), "order": s.order}

    @property
    def required(self):
        r = self._r
        return (r.req_bi, r.req_bj, r.req_ei, r.req_ej)

    @property
    def constant_weight(self):
        return self._r.constant_weight

    def update(self, values, rng, area):
        _check(lib().mpo_recon_update(C.byref(self._r),
                                      np.ascontiguousarray(values, dtype=np.float64).ravel(),
                                      rng[0], rng[1], rng[2], rng[3], area))

    def evaluate(self, point, dx=0, dy=0):
        v = C.c_double()
        pt = Point()
        pt.p[0], pt.p[1] = point
        _check(lib().mpo_recon_eval(C.byref(self._r), pt, dx, dy, C.byref(v)))
        return v.value


def llf_point(left, right, velocity, normal, mode=0, global_speed=0.0):
    v = Point(); n = Point()
    v.p[0], v.p[1] = velocity
    n.p[0], n.p[1] = normal
    f = C.c_double()
    s = C.c_double()
    _check(lib().mpo_llf_point(left, right, v, n, mode, global_speed, C.byref(f), C.byref(s)))
    return f.value


def integrate_edge_flux(edge, velocity, left, right, nq=5, mode=0, global_speed=0.0):
    x, w = gauss_rule(nq)
    e = (Point * 2)()
    for k in range(2):
        e[k].p[0], e[k].p[1] = edge[k]
    vel = (Point * nq)()
    for q in range(nq):
        vel[q].p[0], vel[q].p[1] = velocity[q]
    f = C.c_double()
    _check(lib().mpo_integrate_edge_flux(e, nq, x, w, vel,
                                         np.ascontiguousarray(left, dtype=np.float64),
                                         np.ascontiguousarray(right, dtype=np.float64), mode,
                                         global_speed, C.byref(f)))
    return f.value


def lagrange_weights(npts, location="Middle"):
    """First-derivative weights on nodes 0..npts-1 (src/mlweno/diffusiveflux.cpp:203-254)."""
    w = np.zeros(npts)
    _check(lib().mpo_lagrange_weights(npts, LAGRANGE[location], w))
    return w


def _points(pts):
    arr = (Point * len(pts))()
    for k, p in enumerate(pts):
        arr[k].p[0], arr[k].p[1] = float(p[0]), float(p[1])
    return arr


def diffusive_sampling(edge, left, right=None, rule=5, n=4, extent=0.9):
    """DiffusiveSampling::Build (src/mlweno/diffusiveflux.cpp:256-358); ``rule`` is a
    Gauss point count or a (points, weights) pair; right=None builds the boundary form."""
    qp, qw = gauss_rule(rule) if np.isscalar(rule) else (np.asarray(rule[0], float),
                                                          np.asarray(rule[1], float))
    s = Sampling()
    rarr = _points(right) if right is not None else None
    _check(lib().mpo_diffusive_sampling(_points(edge), _points(left),
                                        C.cast(rarr, C.c_void_p) if rarr is not None else None,
                                        len(qp), np.ascontiguousarray(qp),
                                        np.ascontiguousarray(qw), n, extent, C.byref(s)))
    return s


def diffusive_flux(s, samples, kappa, bc_type=None, bc_value=None):
    """IntegrateDiffusiveFlux / IntegrateDiffusiveBoundaryFlux with the identity quantity."""
    out = C.c_double()
    samples = np.ascontiguousarray(samples, dtype=np.float64)
    kappa = np.ascontiguousarray(kappa, dtype=np.float64)
    bt = bv = None
    if bc_type is not None:
        bt = np.ascontiguousarray(bc_type, dtype=np.int32)
        bv = np.ascontiguousarray(bc_value, dtype=np.float64)
    _check(lib().mpo_diffusive_flux(C.byref(s), samples, kappa,
                                    bt.ctypes.data_as(C.c_void_p) if bt is not None else None,
                                    bv.ctypes.data_as(C.c_void_p) if bv is not None else None,
                                    C.byref(out)))
    return out.value


def _u8_or_none(a):
    if a is None:
        return None, None
    arr = np.ascontiguousarray(a, dtype=np.uint8)
    return arr, arr.ctypes.data_as(C.c_void_p)


class Transport:
    """Reference-faithful H/C transport operator on a structured quad mesh.

    vnx: [q][ny][nx+1] normal velocity along +x at vertical-face Gauss points,
    vny: [q][ny+1][nx] normal velocity along +y at horizontal-face Gauss points,
    both in ascending Gauss order along +y resp. +x.  bc: dict side->uint8 array
    of per-face policies (0 zero flux, 1 outflow only, 2 inflow/outflow).
    """

    def __init__(self, coords, nx, ny, vnx, vny, q=5, options=None, bc=None, inflow_H=None,
                 inflow_C=None, table_mode=0, threads=0):
        self.nx, self.ny, self.q = nx, ny, q
        self._coords = np.ascontiguousarray(coords, dtype=np.float64).ravel()
        self._mesh = _Mesh(nx, ny, self._coords.ctypes.data_as(C.POINTER(C.c_double)))
        self.options = options if options is not None else WenoOptions.production()
        bc = bc or {}
        keep = []
        ptrs = []
        for side in ("bottom", "right", "top", "left"):
            arr, ptr = _u8_or_none(bc.get(side))
            keep.append(arr)
            ptrs.append(ptr)
        infl = []
        for a in (inflow_H, inflow_C):
            if a is None:
                infl.append(None)
            else:
                arr = np.ascontiguousarray(a, dtype=np.float64).ravel()
                keep.append(arr)
                infl.append(arr.ctypes.data_as(C.c_void_p))
        self._keep = keep
        h = C.c_void_p()
        _check(lib().mpo_transport_create(
            C.byref(h), C.byref(self._mesh), q, C.byref(self.options),
            np.ascontiguousarray(vnx, dtype=np.float64).ravel(),
            np.ascontiguousarray(vny, dtype=np.float64).ravel(), *ptrs, infl[0], infl[1],
            table_mode, threads))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().mpo_transport_destroy(h)
            self._h = None

    @property
    def N(self):
        return self.nx * self.ny

    def rates(self):
        r = np.zeros(self.N)
        m = C.c_double()
        _check(lib().mpo_transport_rates(self._h, r, C.byref(m)))
        return r.reshape(self.ny, self.nx), m.value

    def rhs(self, H, Cc):
        H = np.ascontiguousarray(H, dtype=np.float64).ravel()
        Cc = np.ascontiguousarray(Cc, dtype=np.float64).ravel()
        LH = np.zeros(self.N)
        LC = np.zeros(self.N)
        bh = C.c_double()
        bc = C.c_double()
        _check(lib().mpo_transport_rhs(self._h, H, Cc, LH, LC, C.byref(bh), C.byref(bc)))
        return LH.reshape(self.ny, self.nx), LC.reshape(self.ny, self.nx), bh.value, bc.value

    def reconstruct(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64).ravel()
        sig = np.zeros(self.N * 5)
        om = np.zeros(self.N * 5)
        tr = np.zeros(self.N * 4 * self.q)
        _check(lib().mpo_transport_reconstruct(self._h, u, sig, om, tr))
        return (sig.reshape(self.ny, self.nx, 5), om.reshape(self.ny, self.nx, 5),
                tr.reshape(self.ny, self.nx, 4, self.q))

    def advance(self, H, Cc, nsteps, dt, scheme=3):
        H = np.array(H, dtype=np.float64).ravel()
        Cc = np.array(Cc, dtype=np.float64).ravel()
        totals = np.zeros(4 * max(nsteps, 1))
        _check(lib().mpo_transport_advance(self._h, H, Cc, nsteps, dt, scheme,
                                           totals.ctypes.data_as(C.c_void_p)))
        return (H.reshape(self.ny, self.nx), Cc.reshape(self.ny, self.nx),
                totals.reshape(-1, 4)[:nsteps])

    def set_lf(self, mode, global_speed=0.0):
        """LF speed of the interior faces (src/mlweno/advectiveflux.cpp:129-164):
        0 local |v.n|, 1 the global speed."""
        _check(lib().mpo_transport_set_lf(self._h, int(mode), float(global_speed)))

    def set_limiter(self, enable=True, cell_points=5, samples=4, extent=0.9, c_upper=0.4):
        """Bound limiter of src/couple/state_reconstruction.cpp:136-192 on every trace."""
        _check(lib().mpo_transport_set_limiter(self._h, int(enable), cell_points, samples,
                                               extent, c_upper))

    def set_diffusion(self, kappa, samples=4, extent=0.9, types=None, values=None):
        """Legacy preheat thermal diffusion of H (src/couple/preheat.cpp:176-190).
        types: dict side -> uint8 per boundary face, or per face and point [n][q]
        (0 zero flux, 1 temperature, 2 outward flux); values: [sum of side
        faces][q] in side order bottom, right, top, left."""
        t = v = None
        if types is not None:
            parts = []
            for sd, n in (("bottom", self.nx), ("right", self.ny), ("top", self.nx),
                          ("left", self.ny)):
                a = np.asarray(types.get(sd, np.zeros(n)), dtype=np.uint8)
                parts.append(np.broadcast_to(a.reshape(n, -1), (n, self.q)).ravel())
            t = np.ascontiguousarray(np.concatenate(parts))
        if values is not None:
            v = np.ascontiguousarray(values, dtype=np.float64).ravel()
        self._dif_keep = (t, v)
        _check(lib().mpo_transport_set_diffusion(
            self._h, kappa, samples, extent,
            t.ctypes.data_as(C.c_void_p) if t is not None else None,
            v.ctypes.data_as(C.c_void_p) if v is not None else None))

    def set_composition_boundary(self, bc):
        keep, ptrs = [], []
        for side in ("bottom", "right", "top", "left"):
            arr, ptr = _u8_or_none(bc.get(side))
            keep.append(arr)
            ptrs.append(ptr)
        self._bcC_keep = keep
        _check(lib().mpo_transport_set_composition_boundary(self._h, *ptrs))

    def coupled_rhs(self, phase, params, flow, H, Cc):
        """One coupled RHS (src/couple/evolve.cpp:279-384); phase: PhaseModel."""
        f, keep = _flow(flow)
        out = {k: np.zeros(self.N) for k in ("rhsH", "rhsC", "rate", "div", "T", "phi")}
        r = _CoupledRates()
        _check(lib().mpo_coupled_rhs(self._h, C.byref(phase._m), C.byref(params), C.byref(f),
                                     np.ascontiguousarray(H, dtype=np.float64).ravel(),
                                     np.ascontiguousarray(Cc, dtype=np.float64).ravel(),
                                     out["rhsH"], out["rhsC"], out["rate"], out["div"],
                                     out["T"], out["phi"], C.byref(r)))
        del keep
        res = {k: v.reshape(self.ny, self.nx) for k, v in out.items()}
        res.update({n: getattr(r, n) for n, _ in _CoupledRates._fields_})
        return res

    def coupled_run(self, phase, params, flow, H, Cc, nsteps, cfl=0.5, initial=1e-2,
                    maximum=1e-2, minimum=1e-12, end=1e30, fixed=False, time=0.0):
        """SSPRK2 steps with retries (src/couple/evolve.cpp:428-511), flow fixed."""
        f, keep = _flow(flow)
        tc = _CoupledTime(cfl, initial, maximum, minimum, end, int(fixed))
        H = np.array(H, dtype=np.float64).ravel()
        Cc = np.array(Cc, dtype=np.float64).ravel()
        t = C.c_double(time)
        steps = (_CoupledStep * max(nsteps, 1))()
        _check(lib().mpo_coupled_run(self._h, C.byref(phase._m), C.byref(params), C.byref(f),
                                     C.byref(tc), H, Cc, C.byref(t), nsteps, steps))
        del keep
        rec = [{n: getattr(st, n) for n, _ in _CoupledStep._fields_} for st in steps[:nsteps]]
        return H.reshape(self.ny, self.nx), Cc.reshape(self.ny, self.nx), t.value, rec

    def traces(self, u, field=0):
        tr = np.zeros(self.N * 4 * self.q)
        _check(lib().mpo_transport_traces(self._h, field,
                                          np.ascontiguousarray(u, dtype=np.float64).ravel(), tr))
        return tr.reshape(self.ny, self.nx, 4, self.q)

    def limits(self, H, Cc):
        tH = np.zeros(self.N)
        tC = np.zeros(self.N)
        _check(lib().mpo_transport_limits(self._h, np.ascontiguousarray(H, dtype=np.float64).ravel(),
                                          np.ascontiguousarray(Cc, dtype=np.float64).ravel(),
                                          tH, tC))
        return tH.reshape(self.ny, self.nx), tC.reshape(self.ny, self.nx)

    def integral(self, u):
        return lib().mpo_transport_integral(self._h,
                                            np.ascontiguousarray(u, dtype=np.float64).ravel())
