"""Formats either side of the transport step, over the C ABI (host only):

* the H/C handoff (cellH1.dat, cellC1.dat, transport_state.json) written
  after a preheat and read before an evolution
  (src/couple/initialization.cpp:642-686, 823-841; src/couple/evolve.cpp:54-125);
* face velocities from the flow solve's edge Gauss-point samples
  (FlowPointSample, src/couple/visualization.h:8-12; src/couple/preheat.cpp:83-90).
"""
import ctypes as C
import os

import numpy as np

from ._lib import EdgeSample, HandoffMeta, check, dptr, lib

__all__ = ["HandoffMeta", "handoff_meta", "write_handoff", "read_handoff", "edge_id",
           "face_velocity_from_samples"]


def handoff_meta(nx, ny, domain=(0.0, 0.5, 0.0, 1.0), family="rectangular", perturbation=0.0,
                 seed=0, enthalpy_Jkg=1.0, length_m=1.0, time_scale_s=1.0, time=0.0,
                 accepted_steps=0):
    m = HandoffMeta()
    m.family = family.encode()
    m.nx, m.ny = nx, ny
    for k, v in enumerate(domain):
        m.domain[k] = v
    m.perturbation, m.seed = perturbation, seed
    m.enthalpy_Jkg, m.length_m, m.time_scale_s = enthalpy_Jkg, length_m, time_scale_s
    m.time, m.accepted_steps = time, accepted_steps
    return m


def write_handoff(directory, meta, H, Cc):
    H = np.ascontiguousarray(H, dtype=np.float64).reshape(-1)
    Cc = np.ascontiguousarray(Cc, dtype=np.float64).reshape(-1)
    assert H.size == meta.nx * meta.ny and Cc.size == H.size
    check(lib().mp_handoff_write(os.fsencode(directory), C.byref(meta), dptr(H), dptr(Cc)))


def read_handoff(directory, expect, c_max):
    """Returns (H, C, found meta); H, C shaped (ny, nx)."""
    n = expect.nx * expect.ny
    H, Cc = np.empty(n), np.empty(n)
    found = HandoffMeta()
    check(lib().mp_handoff_read(os.fsencode(directory), C.byref(expect), c_max, dptr(H),
                                dptr(Cc), C.byref(found)))
    return H.reshape(expect.ny, expect.nx), Cc.reshape(expect.ny, expect.nx), found


def edge_id(nx, ny, vertical, i, j):
    """Reference edge numbering (src/core/mesh_info.h:79): horizontal edges
    first (j*nx + i), then vertical (nx*(ny+1) + j*(nx+1) + i)."""
    return nx * (ny + 1) + j * (nx + 1) + i if vertical else j * nx + i


def face_velocity_from_samples(nx, ny, q, edge_ids, qs, vx, vy):
    """vnx[q][ny][nx+1], vny[q][ny+1][nx] from per-sample arrays."""
    n = len(edge_ids)
    arr = (EdgeSample * max(n, 1))()
    for k in range(n):
        arr[k].edge_id, arr[k].q, arr[k].vx, arr[k].vy = int(edge_ids[k]), int(qs[k]), float(vx[k]), float(vy[k])
    vnx = np.empty(q * ny * (nx + 1))
    vny = np.empty(q * (ny + 1) * nx)
    check(lib().mp_face_velocity_from_samples(nx, ny, q, n, arr, dptr(vnx), dptr(vny)))
    return vnx.reshape(q, ny, nx + 1), vny.reshape(q, ny + 1, nx)
