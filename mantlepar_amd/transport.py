"""Python mirror of the transport path's interface over the C ABI.

``EutecticModel`` mirrors ``mantle::phase::EutecticModel``
(src/phase/eutectic_rescaled.h:103-136) with batched device evaluation;
``Transport`` mirrors one rank's StateReconstruction + RHS + stage loop
(src/couple/state_reconstruction.h:11-22, src/couple/preheat.cpp:160-284).
Errors raise ``MantleError`` carrying the library's status and message.
"""
import ctypes as C

import numpy as np

from ._lib import (BC_INFLOW_OUTFLOW, BC_OUTFLOW, BC_ZERO_FLUX, PREC_FP64, PREC_MIXED,
                   LAUNCH_AUTO, LAUNCH_PER_STAGE, LAUNCH_PER_STEP, SCHEME_FE, SCHEME_SSPRK2,
                   SCHEME_SSPRK3, MantleError, MeshDesc, PhaseParams,
                   StepReport, WenoOptions, check, dptr, lib, u8ptr)

__all__ = ["EutecticModel", "Transport", "MantleError", "MeshDesc", "PhaseParams",
           "WenoOptions", "PREC_FP64", "PREC_MIXED", "SCHEME_FE", "SCHEME_SSPRK2",
           "SCHEME_SSPRK3", "BC_ZERO_FLUX", "BC_OUTFLOW", "BC_INFLOW_OUTFLOW",
           "default_weno", "default_phase"]

REGION = {1: "PureSolid", 2: "TwoSolids", 3: "Eutectic", 4: "SolidLiquid", 5: "Liquid",
          6: "PureMelting"}


def default_weno(**kw):
    o = WenoOptions()
    lib().mp_default_weno_options(C.byref(o))
    for k, v in kw.items():
        if k == "small_weight":
            for i, w in enumerate(v):
                o.small_weight[i] = w
        else:
            setattr(o, k, v)
    return o


def default_phase(**kw):
    p = PhaseParams()
    lib().mp_default_phase_params(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class EutecticModel:
    """Batched EutecticModel on the device (MaterialParameters + pressure)."""

    def __init__(self, params=None, **kw):
        self.params = params if params is not None else default_phase(**kw)

    def evaluate(self, H, Cd, P=None):
        """Arrays in, dict of arrays out: T, phi (liquid), Cl, Cs, region, phi1,
        phi2, dTdC, dTdH.  Raises MantleError like the reference throws."""
        H = _f64(H).ravel()
        Cd = _f64(Cd).ravel()
        n = H.size
        if Cd.size != n:
            raise ValueError("H and C sizes differ")
        Pa = None if P is None else _f64(np.broadcast_to(P, (n,))).copy()
        out = {k: np.empty(n) for k in ("T", "phi", "Cl", "Cs", "phi1", "phi2", "dTdC", "dTdH")}
        reg = np.empty(n, dtype=np.uint8)
        check(lib().mp_phase_evaluate(C.byref(self.params), n, dptr(H), dptr(Cd), dptr(Pa),
                                      dptr(out["T"]), dptr(out["phi"]), dptr(out["Cl"]),
                                      dptr(out["Cs"]), u8ptr(reg), dptr(out["phi1"]),
                                      dptr(out["phi2"]), dptr(out["dTdC"]), dptr(out["dTdH"])))
        out["region"] = reg
        return out

    def boundaries(self, Cd, P=0.0):
        v = [C.c_double() for _ in range(4)]
        check(lib().mp_phase_boundaries(C.byref(self.params), Cd, P, *[C.byref(x) for x in v]))
        return dict(zip(("Tep", "Tmp", "H_eutectic_end", "H_liquidus"), (x.value for x in v)))


class Transport:
    """One mesh's H/C transport state resident on the GPU."""

    def __init__(self, nx, ny, x0=0.0, x1=0.5, y0=0.0, y1=1.0, q=5, weno=None, phase=None,
                 precision=PREC_FP64, stream=None):
        self.nx, self.ny, self.q = nx, ny, q
        self.mesh = MeshDesc(nx, ny, x0, x1, y0, y1, q)
        self.weno = weno if weno is not None else default_weno()
        self.phase = phase if phase is not None else default_phase()
        self.precision = precision
        h = C.c_void_p()
        check(lib().mp_transport_create(C.byref(self.mesh), C.byref(self.weno),
                                        C.byref(self.phase), precision, stream, C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().mp_transport_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    @property
    def N(self):
        return self.nx * self.ny

    def upload(self, H, Cd):
        H = _f64(H).reshape(-1)
        Cd = _f64(Cd).reshape(-1)
        assert H.size == self.N and Cd.size == self.N
        check(lib().mp_transport_upload_state(self._h, dptr(H), dptr(Cd)))

    def download(self):
        H = np.empty(self.N)
        Cd = np.empty(self.N)
        check(lib().mp_transport_download_state(self._h, dptr(H), dptr(Cd)))
        return H.reshape(self.ny, self.nx), Cd.reshape(self.ny, self.nx)

    def set_face_velocity(self, vnx, vny):
        vnx = _f64(vnx).reshape(-1)
        vny = _f64(vny).reshape(-1)
        assert vnx.size == self.q * self.ny * (self.nx + 1)
        assert vny.size == self.q * (self.ny + 1) * self.nx
        check(lib().mp_transport_set_face_velocity(self._h, dptr(vnx), dptr(vny)))

    def set_boundary(self, bottom=None, right=None, top=None, left=None):
        arrs = []
        for a, n in ((bottom, self.nx), (right, self.ny), (top, self.nx), (left, self.ny)):
            if a is None:
                arrs.append(None)
            else:
                b = np.ascontiguousarray(np.broadcast_to(a, (n,)), dtype=np.uint8).copy()
                arrs.append(b)
        self._bc = arrs
        check(lib().mp_transport_set_boundary(self._h, *[u8ptr(a) for a in arrs]))

    def set_inflow(self, inflow_H=None, inflow_C=None):
        a = None if inflow_H is None else _f64(inflow_H).reshape(-1)
        b = None if inflow_C is None else _f64(inflow_C).reshape(-1)
        check(lib().mp_transport_set_inflow(self._h, dptr(a), dptr(b)))

    def stable_dt(self, cfl):
        dt = C.c_double()
        check(lib().mp_transport_stable_dt(self._h, cfl, C.byref(dt)))
        return dt.value

    def step(self, nsteps, dt, scheme=SCHEME_SSPRK3):
        rep = StepReport()
        check(lib().mp_transport_step(self._h, nsteps, scheme, dt, C.byref(rep)))
        return rep.as_dict()

    def step_noalloc(self, nsteps, dt, scheme=SCHEME_SSPRK3):
        """mp_transport_step with its work memory in the library's static scratch."""
        rep = StepReport()
        check(lib().mp_transport_step_noalloc(self._h, nsteps, scheme, dt, C.byref(rep)))
        return rep.as_dict()

    def phase_state(self):
        out = {k: np.empty(self.N) for k in ("T", "phi", "Cl", "Cs")}
        reg = np.empty(self.N, dtype=np.uint8)
        check(lib().mp_transport_phase(self._h, dptr(out["T"]), dptr(out["phi"]),
                                       dptr(out["Cl"]), dptr(out["Cs"]), u8ptr(reg)))
        out["region"] = reg
        return {k: v.reshape(self.ny, self.nx) for k, v in out.items()}

    def rhs(self):
        rh = np.empty(self.N)
        rc = np.empty(self.N)
        b = np.empty(2)
        check(lib().mp_transport_rhs(self._h, dptr(rh), dptr(rc), dptr(b)))
        return rh.reshape(self.ny, self.nx), rc.reshape(self.ny, self.nx), b[0], b[1]

    def evaluate(self, cells, x, y):
        """Reconstructed (limited) H and C of cells[k] (flat j*nx+i) at points (x[k], y[k])."""
        cells = np.ascontiguousarray(cells, dtype=np.int32).ravel()
        x, y = _f64(x).ravel(), _f64(y).ravel()
        n = cells.size
        assert x.size == n and y.size == n
        h, c = np.empty(n), np.empty(n)
        check(lib().mp_transport_evaluate(self._h, n, cells.ctypes.data, dptr(x), dptr(y),
                                          dptr(h), dptr(c)))
        return h, c

    def reconstruct(self, field):
        sig = np.empty(self.N * 5)
        om = np.empty(self.N * 5)
        tr = np.empty(self.N * 4 * self.q)
        check(lib().mp_transport_reconstruct(self._h, field, dptr(sig), dptr(om), dptr(tr)))
        return (sig.reshape(self.ny, self.nx, 5), om.reshape(self.ny, self.nx, 5),
                tr.reshape(self.ny, self.nx, 4, self.q))

    def set_launch(self, per_step=True):
        """One launch per step (True), one per stage (False) or chosen by mesh
        size (None, the default); results are bitwise equal."""
        mode = LAUNCH_AUTO if per_step is None else (LAUNCH_PER_STEP if per_step else LAUNCH_PER_STAGE)
        check(lib().mp_transport_set_launch(self._h, mode))

    def set_balance_tolerance(self, tol=1e-10):
        """Relative tolerance of every step's conservation balance check."""
        check(lib().mp_transport_set_balance_tolerance(self._h, tol))

    def layout(self):
        """rows per item, items per stage, first tail row, one launch per SSPRK3 step."""
        v = [C.c_int32() for _ in range(4)]
        check(lib().mp_transport_layout(self._h, *[C.byref(x) for x in v]))
        return {"rows": v[0].value, "items": v[1].value, "tail_split": v[2].value,
                "one_launch_steps": bool(v[3].value)}

    def set_tile(self, rows):
        check(lib().mp_transport_set_tile(self._h, rows))

    def set_limiter(self, enable=True, cell_points_per_axis=5, diffusion_samples=4,
                    extent_fraction=0.9):
        """StateReconstruction::Limit (src/couple/state_reconstruction.cpp:136-192) on
        every reconstruction of the stage kernels (either precision)."""
        check(lib().mp_transport_set_limiter(self._h, int(enable), cell_points_per_axis,
                                             diffusion_samples, extent_fraction))

    def set_diffusion(self, kappa, diffusion_samples=4, extent_fraction=0.9):
        """Legacy preheat thermal diffusion of H (src/couple/preheat.cpp:176-190)."""
        check(lib().mp_transport_set_diffusion(self._h, kappa, diffusion_samples,
                                               extent_fraction))

    def set_thermal_boundary(self, types=None, values=None):
        """types/values: arrays [2*(nx+ny)*q] (sides bottom, right, top, left, each
        [face][q]); types 0 zero flux, 1 prescribed temperature, 2 outward flux."""
        t = None if types is None else np.ascontiguousarray(types, dtype=np.uint8).ravel()
        v = None if values is None else _f64(values).ravel()
        check(lib().mp_transport_set_thermal_boundary(self._h, u8ptr(t), dptr(v)))

    def set_lithostatic(self, surface_y=None):
        """Pressure model "lithostatic" (src/couple/initialization.cpp:155-162): the
        phase of each cell at P = staticPressure(surface_y - y_centre); None
        returns to the constant phase pressure."""
        check(lib().mp_transport_set_lithostatic(self._h, int(surface_y is not None),
                                                 0.0 if surface_y is None else surface_y))

    def limits(self):
        """Limiter factors (theta_H, theta_C) of the current state."""
        tH = np.empty(self.N)
        tC = np.empty(self.N)
        check(lib().mp_transport_limits(self._h, dptr(tH), dptr(tC)))
        return tH.reshape(self.ny, self.nx), tC.reshape(self.ny, self.nx)

    def profile(self, nsteps, dt, scheme=SCHEME_SSPRK3):
        vals = [C.c_double() for _ in range(3)]
        check(lib().mp_transport_profile(self._h, nsteps, scheme, dt,
                                         *[C.byref(v) for v in vals]))
        return dict(zip(("ms_total", "ms_stage", "ms_phase_stage"), (v.value for v in vals)))

    def synchronize(self):
        check(lib().mp_transport_synchronize(self._h))
