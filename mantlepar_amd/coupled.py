"""The coupled transport law on the device (mp_coupled_* of the C ABI).

One rank's share of MantlePar's phase-coupled evolution
(src/couple/evolve.cpp:279-507) on a uniform mesh, FP64: limited traces,
temperature flux with the latent term, effective composition velocity,
normal-line diffusion of T, cooling source, and the SSPRK2 loop with
admissibility/CFL retries.  Flow samples (the Stokes-Darcy solve's output)
are inputs and stay fixed over a call.
"""
import ctypes as C

import numpy as np

from ._lib import (CoupledParams, CoupledRates, CoupledStep, CoupledTime, MeshDesc, check,
                   dptr, lib, u8ptr)
from .transport import _f64, default_phase


def default_coupled(**kw):
    p = CoupledParams()
    lib().mp_default_coupled_params(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


class Coupled:
    def __init__(self, nx, ny, x0=0.0, x1=0.5, y0=0.0, y1=1.0, q=5, phase=None, params=None,
                 stream=None):
        self.nx, self.ny, self.q = nx, ny, q
        self.mesh = MeshDesc(nx, ny, x0, x1, y0, y1, q)
        self.phase = phase if phase is not None else default_phase()
        self.params = params if params is not None else default_coupled()
        h = C.c_void_p()
        check(lib().mp_coupled_create(C.byref(self.mesh), C.byref(self.phase),
                                      C.byref(self.params), stream, C.byref(h)))
        self._h = h

    @property
    def N(self):
        return self.nx * self.ny

    def close(self):
        if getattr(self, "_h", None):
            lib().mp_coupled_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def set_flow(self, flow):
        """flow: dict edge_stokes [E][q][2], edge_darcy, edge_phi [E][q],
        cell_stokes [N][n*n][2], cell_darcy, cell_phi [N][n*n]."""
        a = [_f64(flow[k]).ravel() for k in ("edge_stokes", "edge_darcy", "edge_phi",
                                             "cell_stokes", "cell_darcy", "cell_phi")]
        check(lib().mp_coupled_set_flow(self._h, *[dptr(x) for x in a]))

    def set_boundary(self, field, bc):
        arrs = [None if bc.get(s) is None else np.ascontiguousarray(bc[s], dtype=np.uint8)
                for s in ("bottom", "right", "top", "left")]
        check(lib().mp_coupled_set_boundary(self._h, field, *[u8ptr(x) for x in arrs]))

    def set_inflow(self, inflow_H=None, inflow_C=None):
        h = None if inflow_H is None else _f64(inflow_H).ravel()
        c = None if inflow_C is None else _f64(inflow_C).ravel()
        check(lib().mp_coupled_set_inflow(self._h, dptr(h), dptr(c)))

    def set_thermal_boundary(self, types=None, values=None):
        t = None if types is None else np.ascontiguousarray(types, dtype=np.uint8).ravel()
        v = None if values is None else _f64(values).ravel()
        check(lib().mp_coupled_set_thermal_boundary(self._h, u8ptr(t), dptr(v)))

    def set_lithostatic(self, surface_y=None):
        """Pressure model "lithostatic" (src/couple/initialization.cpp:155-162): every
        phase evaluation at P = staticPressure(surface_y - y) of its own point;
        None returns to the constant phase pressure."""
        check(lib().mp_coupled_set_lithostatic(self._h, int(surface_y is not None),
                                               0.0 if surface_y is None else surface_y))

    def upload(self, H, Cd):
        check(lib().mp_coupled_upload_state(self._h, dptr(_f64(H).ravel()), dptr(_f64(Cd).ravel())))

    def download(self):
        H = np.empty(self.N)
        Cd = np.empty(self.N)
        check(lib().mp_coupled_download_state(self._h, dptr(H), dptr(Cd)))
        return H.reshape(self.ny, self.nx), Cd.reshape(self.ny, self.nx)

    def rhs(self):
        out = {k: np.empty(self.N) for k in ("rhsH", "rhsC", "rate", "div", "T", "phi")}
        r = CoupledRates()
        check(lib().mp_coupled_rhs(self._h, *[dptr(out[k]) for k in ("rhsH", "rhsC", "rate", "div",
                                                                       "T", "phi")], C.byref(r)))
        res = {k: v.reshape(self.ny, self.nx) for k, v in out.items()}
        res.update({n: getattr(r, n) for n, _ in CoupledRates._fields_})
        return res

    def run(self, nsteps, cfl=0.5, initial=1e-2, maximum=1e-2, minimum=1e-12, end=1e30,
            fixed=False, time=0.0):
        tc = CoupledTime(cfl, initial, maximum, minimum, end, int(fixed))
        t = np.array([time], dtype=np.float64)
        steps = (CoupledStep * max(nsteps, 1))()
        acc = C.c_int32()
        check(lib().mp_coupled_run(self._h, C.byref(tc), dptr(t), nsteps, steps, C.byref(acc)))
        rec = [{n: getattr(s, n) for n, _ in CoupledStep._fields_ if n != "pad"}
               for s in steps[:acc.value]]
        return float(t[0]), rec


def run_batch(handles, nsteps, controls, times=None):
    """Advance several Coupled handles together (mp_coupled_run_batch): one
    step attempt of every running handle per round, one launch per kernel over
    all of them.  controls: one dict of Coupled.run's time keywords (cfl,
    initial, maximum, minimum, end, fixed) per handle, or one dict for all.
    Returns a list of (status, time, step records) per handle; a handle that
    failed keeps its last accepted state.  Raises nothing for per-handle
    failures: check each status (0 is success)."""
    n = len(handles)
    if isinstance(controls, dict):
        controls = [controls] * n
    keys = ("cfl", "initial", "maximum", "minimum", "end", "fixed")
    dflt = {"cfl": 0.5, "initial": 1e-2, "maximum": 1e-2, "minimum": 1e-12, "end": 1e30, "fixed": False}
    tcs = (CoupledTime * n)()
    for k, c in enumerate(controls):
        v = {**dflt, **c}
        tcs[k] = CoupledTime(*(float(v[x]) for x in keys[:5]), int(v["fixed"]))
    t = np.zeros(n) if times is None else np.array(times, dtype=np.float64)
    hs = (C.c_void_p * n)(*(h._h for h in handles))
    steps = (CoupledStep * max(n * nsteps, 1))()
    acc = (C.c_int32 * n)()
    st = (C.c_int32 * n)()
    code = lib().mp_coupled_run_batch(hs, n, tcs, dptr(t), nsteps, steps, acc, st)
    if code != 0 and all(s == 0 for s in st):
        check(code)                     # the call itself failed
    out = []
    for k in range(n):
        rec = [{f: getattr(s, f) for f, _ in CoupledStep._fields_ if f != "pad"}
               for s in steps[k * nsteps:k * nsteps + acc[k]]]
        out.append((int(st[k]), float(t[k]), rec))
    return out
