"""Many meshes on one device: ``Batch`` (lock-step stepping, hipGraph replay,
schedule autotuner) and ``Server`` (continuous batching of transport runs).

Each mesh in either is one rank's share of the reference's transport loop
(src/couple/state_reconstruction.cpp:35-120, src/couple/preheat.cpp:234-287);
its results are bitwise those of a ``Transport`` handle on the same inputs.
"""
import ctypes as C

import numpy as np

from ._lib import (BATCH_EAGER, BATCH_GRAPH, LAUNCH_AUTO, LAUNCH_PER_STAGE, LAUNCH_PER_STEP,
                   MP_ERR_PENDING, PREC_FP64, SCHEME_SSPRK3,
                   MantleError, MeshDesc, Request, RequestResult, ServerStats, StepReport,
                   check, dptr, lib, u8ptr)
from .transport import default_phase, default_weno

__all__ = ["Batch", "Server", "BATCH_EAGER", "BATCH_GRAPH"]


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64).reshape(-1)


def _bc_arrays(nx, ny, bc):
    if bc is None:
        return [None] * 4
    out = []
    for a, n in zip(bc, (nx, ny, nx, ny)):
        out.append(None if a is None else
                   np.ascontiguousarray(np.broadcast_to(a, (n,)), dtype=np.uint8).copy())
    return out


class Batch:
    """Independent meshes advanced together, one persistent launch per stage."""

    def __init__(self, capacity, weno=None, phase=None, precision=PREC_FP64,
                 hbm_budget_bytes=0, stream=None):
        self.weno = weno if weno is not None else default_weno()
        self.phase = phase if phase is not None else default_phase()
        self.capacity = capacity
        self.meshes = {}
        h = C.c_void_p()
        check(lib().mp_batch_create(capacity, C.byref(self.weno), C.byref(self.phase), precision,
                                    int(hbm_budget_bytes), stream, C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().mp_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def add_mesh(self, nx, ny, x0=0.0, x1=0.5, y0=0.0, y1=1.0, q=5):
        d = MeshDesc(nx, ny, x0, x1, y0, y1, q)
        k = C.c_int32()
        check(lib().mp_batch_add_mesh(self._h, C.byref(d), C.byref(k)))
        self.meshes[k.value] = (nx, ny, q)
        return k.value

    def remove_mesh(self, k):
        check(lib().mp_batch_remove_mesh(self._h, k))
        self.meshes.pop(k, None)

    def upload(self, k, H, Cd):
        H, Cd = _f64(H), _f64(Cd)
        nx, ny, _ = self.meshes[k]
        assert H.size == nx * ny and Cd.size == nx * ny
        check(lib().mp_batch_upload_state(self._h, k, dptr(H), dptr(Cd)))

    def download(self, k):
        nx, ny, _ = self.meshes[k]
        H, Cd = np.empty(nx * ny), np.empty(nx * ny)
        check(lib().mp_batch_download_state(self._h, k, dptr(H), dptr(Cd)))
        return H.reshape(ny, nx), Cd.reshape(ny, nx)

    def set_face_velocity(self, k, vnx, vny):
        nx, ny, q = self.meshes[k]
        vnx, vny = _f64(vnx), _f64(vny)
        assert vnx.size == q * ny * (nx + 1) and vny.size == q * (ny + 1) * nx
        check(lib().mp_batch_set_face_velocity(self._h, k, dptr(vnx), dptr(vny)))

    def set_boundary(self, k, bottom=None, right=None, top=None, left=None):
        nx, ny, _ = self.meshes[k]
        arrs = _bc_arrays(nx, ny, (bottom, right, top, left))
        check(lib().mp_batch_set_boundary(self._h, k, *[u8ptr(a) for a in arrs]))

    def set_inflow(self, k, inflow_H=None, inflow_C=None):
        a = None if inflow_H is None else _f64(inflow_H)
        b = None if inflow_C is None else _f64(inflow_C)
        check(lib().mp_batch_set_inflow(self._h, k, dptr(a), dptr(b)))

    def set_lithostatic(self, k, surface_y=None):
        """Pressure model "lithostatic" for mesh k (src/couple/initialization.cpp:155-162);
        None returns to the constant phase pressure."""
        check(lib().mp_batch_set_lithostatic(self._h, k, int(surface_y is not None),
                                             0.0 if surface_y is None else surface_y))

    def set_limiter(self, k, enable=True, cell_points_per_axis=5, diffusion_samples=4,
                    extent_fraction=0.9):
        """Bound limiter of mesh k (StateReconstruction::Limit); enable=False: off."""
        check(lib().mp_batch_set_limiter(self._h, k, int(enable), cell_points_per_axis,
                                         diffusion_samples, extent_fraction))

    def set_diffusion(self, k, kappa, diffusion_samples=4, extent_fraction=0.9):
        """Legacy preheat thermal diffusion of H of mesh k (src/couple/preheat.cpp:176-190)."""
        check(lib().mp_batch_set_diffusion(self._h, k, kappa, diffusion_samples, extent_fraction))

    def set_thermal_boundary(self, k, types=None, values=None):
        """As Transport.set_thermal_boundary, for mesh k."""
        t = None if types is None else np.ascontiguousarray(types, dtype=np.uint8).ravel()
        v = None if values is None else _f64(values).ravel()
        check(lib().mp_batch_set_thermal_boundary(self._h, k, u8ptr(t), dptr(v)))

    def stable_dt(self, k, cfl):
        dt = C.c_double()
        check(lib().mp_batch_stable_dt(self._h, k, cfl, C.byref(dt)))
        return dt.value

    def set_dt(self, k, dt):
        check(lib().mp_batch_set_dt(self._h, k, dt))

    def step(self, nsteps, scheme=SCHEME_SSPRK3, graph=False):
        """Advance all meshes; returns {mesh id: report dict}."""
        reps = (StepReport * self.capacity)()
        check(lib().mp_batch_step(self._h, nsteps, scheme, BATCH_GRAPH if graph else BATCH_EAGER,
                                  reps))
        return {k: reps[k].as_dict() for k in self.meshes}

    def phase_state(self, k):
        nx, ny, _ = self.meshes[k]
        out = {n: np.empty(nx * ny) for n in ("T", "phi", "Cl", "Cs")}
        reg = np.empty(nx * ny, dtype=np.uint8)
        check(lib().mp_batch_phase(self._h, k, dptr(out["T"]), dptr(out["phi"]), dptr(out["Cl"]),
                                   dptr(out["Cs"]), u8ptr(reg)))
        out["region"] = reg
        return {n: v.reshape(ny, nx) for n, v in out.items()}

    def set_schedule(self, rows, blocks):
        check(lib().mp_batch_set_schedule(self._h, rows, blocks))

    def schedule(self):
        r, b = C.c_int32(), C.c_int32()
        check(lib().mp_batch_get_schedule(self._h, C.byref(r), C.byref(b)))
        return r.value, b.value

    def set_launch(self, per_step=True):
        """Mixed batches: one launch per step (True), one per stage (False), or
        the autotuner's choice (None, the default); results are bitwise equal."""
        mode = LAUNCH_AUTO if per_step is None else (LAUNCH_PER_STEP if per_step else LAUNCH_PER_STAGE)
        check(lib().mp_batch_set_launch(self._h, mode))

    def set_balance_tolerance(self, tol=1e-10):
        """Relative tolerance of every mesh's per-step heat-balance check."""
        check(lib().mp_batch_set_balance_tolerance(self._h, tol))

    def per_step_launch(self):
        """Whether steps currently run in one launch."""
        v = C.c_int32()
        check(lib().mp_batch_get_launch(self._h, C.byref(v)))
        return bool(v.value)

    def autotune(self, nsteps=2, scheme=SCHEME_SSPRK3, cache_path=None):
        r, b, ms = C.c_int32(), C.c_int32(), C.c_double()
        path = None if cache_path is None else str(cache_path).encode()
        check(lib().mp_batch_autotune(self._h, nsteps, scheme, path, C.byref(r), C.byref(b),
                                      C.byref(ms)))
        return {"rows": r.value, "blocks": b.value, "ms_per_step": ms.value}

    def profile(self, nsteps, scheme=SCHEME_SSPRK3):
        vals = [C.c_double() for _ in range(3)]
        check(lib().mp_batch_profile(self._h, nsteps, scheme, *[C.byref(v) for v in vals]))
        return dict(zip(("ms_total", "ms_stage", "ms_phase_stage"), (v.value for v in vals)))

    def info(self):
        m, c, it, b = C.c_int32(), C.c_int64(), C.c_int64(), C.c_uint64()
        check(lib().mp_batch_info(self._h, C.byref(m), C.byref(c), C.byref(it), C.byref(b)))
        return {"meshes": m.value, "cells": c.value, "items": it.value, "hbm_in_use": b.value}


class Server:
    """Continuous batching of transport runs (one server thread in the library)."""

    def __init__(self, max_slots=512, weno=None, phase=None, precision=PREC_FP64,
                 scheme=SCHEME_SSPRK3, hbm_budget_bytes=200 * 10**9, max_epoch_steps=8):
        self.weno = weno if weno is not None else default_weno()
        self.phase = phase if phase is not None else default_phase()
        h = C.c_void_p()
        check(lib().mp_server_create(max_slots, C.byref(self.weno), C.byref(self.phase), precision,
                                     scheme, int(hbm_budget_bytes), max_epoch_steps, C.byref(h)))
        self._h = h
        self._live = {}   # id -> arrays that must outlive the request

    def close(self):
        if getattr(self, "_h", None):
            lib().mp_server_destroy(self._h)
            self._h = None
            self._live.clear()

    def __del__(self):
        self.close()

    def set_tile(self, rows):
        check(lib().mp_server_set_tile(self._h, rows))

    def now(self):
        return lib().mp_server_now(self._h)

    def submit(self, nx, ny, H, Cd, vnx, vny, nsteps, cfl, x0=0.0, x1=0.5, y0=0.0, y1=1.0, q=5,
               bc=None, outputs=("H", "C", "T", "phi", "Cl", "Cs", "region"), surface_y=None,
               limiter=None, diffusion=None):
        """Queue one run; returns its id.  Inputs are kept referenced until done.
        surface_y: lithostatic phase pressure below that surface (None: constant).
        limiter: None (off) or (cell_points_per_axis, diffusion_samples,
        extent_fraction) of the bound limiter.  diffusion: None (off) or
        (kappa, thermal types or None, thermal values or None)."""
        ins = [_f64(H), _f64(Cd), _f64(vnx), _f64(vny)]
        n = nx * ny
        assert ins[0].size == n and ins[1].size == n
        assert ins[2].size == q * ny * (nx + 1) and ins[3].size == q * (ny + 1) * nx
        outs = {k: (np.empty(n, dtype=np.uint8) if k == "region" else np.empty(n))
                for k in outputs}
        bcs = _bc_arrays(nx, ny, bc)
        r = Request()
        r.mesh = MeshDesc(nx, ny, x0, x1, y0, y1, q)
        r.H, r.C, r.vnx, r.vny = (a.ctypes.data for a in ins)
        for i, a in enumerate(bcs):
            r.bc[i] = None if a is None else a.ctypes.data
        r.cfl = cfl
        r.nsteps = nsteps
        r.lithostatic = int(surface_y is not None)
        r.surface_y = 0.0 if surface_y is None else surface_y
        if limiter is not None:
            r.limiter = 1
            r.cell_points_per_axis, r.diffusion_samples, r.extent_fraction = limiter
        thermal = []
        if diffusion is not None:
            kappa, types, values = diffusion
            r.thermal_diffusivity = kappa
            if types is not None:
                thermal = [np.ascontiguousarray(types, dtype=np.uint8).ravel(),
                           None if values is None else _f64(values).ravel()]
                r.thermal_types = thermal[0].ctypes.data
                r.thermal_values = None if thermal[1] is None else thermal[1].ctypes.data
        names = {"H": "H_out", "C": "C_out", "T": "T", "phi": "phi", "Cl": "Cl", "Cs": "Cs",
                 "region": "region"}
        for k, a in outs.items():
            setattr(r, names[k], a.ctypes.data)
        rid = C.c_int64()
        check(lib().mp_server_submit(self._h, C.byref(r), C.byref(rid)))
        self._live[rid.value] = (ins + thermal, outs, bcs, (ny, nx))
        return rid.value

    def wait(self, rid, timeout=-1.0):
        """Block until run rid is done; returns (result dict, outputs dict).
        Raises MantleError for a failed run, and with code MP_ERR_PENDING on
        timeout (the run stays queued)."""
        res = RequestResult()
        code = lib().mp_server_wait(self._h, rid, float(timeout), C.byref(res))
        if code == MP_ERR_PENDING:
            raise MantleError(code, lib().mp_last_error().decode())
        ins, outs, bcs, shape = self._live.pop(rid)
        check(code)
        return res.as_dict(), {k: v.reshape(shape) for k, v in outs.items()}

    def drain(self):
        check(lib().mp_server_drain(self._h))

    def stats(self):
        s = ServerStats()
        check(lib().mp_server_stats_get(self._h, C.byref(s)))
        return s.as_dict()

    def progress(self, cap=65536):
        """(t, cumulative cell-steps) per completed epoch, oldest first."""
        t, w = np.empty(cap), np.empty(cap)
        n = C.c_int32()
        check(lib().mp_server_progress(self._h, dptr(t), dptr(w), cap, C.byref(n)))
        return t[:n.value].copy(), w[:n.value].copy()
