"""The full-chip workload's ingredients at reduced size: face velocities
refreshed every few steps (psi amplitude drift), a through-flow entering on
the left and leaving through an outflow-only top segment.

Bars: FP64 against the oracle (1e-12 relative, 32x64 cells, 20 SSP-RK3
steps with a refresh every 10); the conservation balance
total(end) - total(start) + outward boundary flux = 0 to 1e-12 of the total in
both precisions; mixed within 1e-6 of FP64 in H and C, and within 1e-9 in the
global totals, after 100 steps on 512x512 cells.
"""
import numpy as np
import pytest

import mantlepar_amd as M
import oracle as O
from mantlepar_amd import workloads as W

pytestmark = pytest.mark.gpu


def velocities(spec, amp):
    bx, by = W.face_velocity(spec)
    tx, ty = W.throughflow_velocity(spec)
    return amp * bx + tx, amp * by + ty


def run(spec, prec, steps, every, cfl=0.4):
    H, Cc = W.enthalpy(spec), W.composition(spec)
    bc, inflow = W.throughflow_boundary(spec)
    t = M.Transport(spec.nx, spec.ny, spec.x0, spec.x1, spec.y0, spec.y1, q=spec.q,
                    phase=W.config_phase(), precision=prec)
    t.upload(H, Cc)
    t.set_boundary(bc["bottom"], bc["right"], bc["top"], bc["left"])
    t.set_inflow(*inflow)
    amps = W.refresh_amplitudes(steps // every, seed=spec.seed)
    area = spec.dx * spec.dy
    t0 = (H.sum() * area, Cc.sum() * area)
    bH = bC = 0.0
    dts = []
    rep = None
    for a in amps:
        t.set_face_velocity(*velocities(spec, a))
        dt = t.stable_dt(cfl)
        dts.append(dt)
        rep = t.step(every, dt)
        bH += rep["boundary_H"]
        bC += rep["boundary_C"]
    Hf, Cf = t.download()
    t.close()
    return dict(H=Hf, C=Cf, t0=t0, t1=(rep["total_H"], rep["total_C"]), b=(bH, bC), dts=dts,
                H0=H, C0=Cc, bc=bc, inflow=inflow, amps=amps)


def test_fp64_matches_oracle_with_refresh_and_outflow():
    spec = W.MeshSpec(32, 64, seed=5)
    g = run(spec, M.PREC_FP64, 20, 10)
    H, Cc = g["H0"], g["C0"]
    for a, dt in zip(g["amps"], g["dts"]):
        vx, vy = velocities(spec, a)
        o = O.Transport(W.vertex_coords(spec), spec.nx, spec.ny, vx, vy, q=spec.q, bc=g["bc"],
                        inflow_H=g["inflow"][0], inflow_C=g["inflow"][1], table_mode=0)
        H, Cc, _ = o.advance(H, Cc, 10, dt, 3)
    assert np.abs(g["H"] - H).max() <= 1e-12 * np.abs(H).max()
    assert np.abs(g["C"] - Cc).max() <= 1e-12 * np.abs(Cc).max()


def test_conservation_and_mixed_vs_fp64():
    spec = W.MeshSpec(512, 512, seed=5)
    out = {p: run(spec, p, 100, 50) for p in (M.PREC_FP64, M.PREC_MIXED)}
    for p, g in out.items():
        for f in (0, 1):
            defect = g["t1"][f] - g["t0"][f] + g["b"][f]
            assert abs(defect) <= 1e-12 * abs(g["t0"][f]), (p, f, defect)
        # the through-flow did carry mass out through the top segment
        assert g["b"][0] != 0.0
    a, b = out[M.PREC_MIXED], out[M.PREC_FP64]
    assert np.abs(a["H"] - b["H"]).max() <= 1e-6 * np.abs(b["H"]).max()
    assert np.abs(a["C"] - b["C"]).max() <= 1e-6 * np.abs(b["C"]).max()
    for f in (0, 1):
        assert abs(a["t1"][f] - b["t1"][f]) <= 1e-9 * abs(b["t1"][f])
