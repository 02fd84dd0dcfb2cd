"""The full-chip workload's ingredients at reduced size: face velocities
refreshed every few steps (psi amplitude drift), a through-flow entering on
the left and leaving through an outflow-only top segment.

Bars: FP64 against the oracle (1e-12 relative, 32x64 cells, 20 SSP-RK3
steps with a refresh every 10); the conservation balance
total(end) - total(start) + outward boundary flux = 0 to 1e-12 of the total in
both precisions; mixed within 1e-6 of FP64 in H and C, and within 1e-9 in the
global totals, after 100 steps on 512x512 cells; the same workload at
2048x2048 for 250 steps (see test_full_chip_workload_at_scale).
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


def test_full_chip_workload_at_scale():
    """The full-chip workload (domain, refresh drift, through-flow and
    outflow-only top segment) on 2048 x 2048 cells for 250 SSP-RK3 steps with
    a refresh every 50, in both precisions and in FP64 from a state perturbed
    by 2^-24 (FP32 resolution).  Bars: the conservation balance closes to
    1e-12 of the totals in every run; the global totals of mixed and FP64
    agree to 1e-9; pointwise, mixed is within 1e-6 of FP64 in H and C, or
    within 10x the FP64 spread where the FP64 method itself carries an
    FP32-sized initial difference further than that (over long runs the
    grid-scale random C field and the inflow front do)."""
    spec = W.MeshSpec(2048, 2048, seed=5)
    steps, every = 250, 50
    H, Cc = W.enthalpy(spec), W.composition(spec)
    bx, by = W.face_velocity(spec)
    tx, ty = W.throughflow_velocity(spec)
    bc, inflow = W.throughflow_boundary(spec)
    amps = W.refresh_amplitudes(steps // every, seed=spec.seed)
    area = spec.dx * spec.dy
    rng = np.random.default_rng(17)
    Hp = H * (1.0 + 2.0 ** -24 * rng.standard_normal(H.shape))
    Cp = Cc * (1.0 + 2.0 ** -24 * rng.standard_normal(Cc.shape))
    out = {}
    for tag, prec, H0, C0 in (("fp64", M.PREC_FP64, H, Cc), ("fp64p", M.PREC_FP64, Hp, Cp),
                              ("mixed", M.PREC_MIXED, H, Cc)):
        print(f"full chip at 2048x2048: {tag}", flush=True)
        t = M.Transport(spec.nx, spec.ny, spec.x0, spec.x1, spec.y0, spec.y1, q=spec.q,
                        phase=W.config_phase(), precision=prec)
        t.upload(H0, C0)
        t.set_boundary(bc["bottom"], bc["right"], bc["top"], bc["left"])
        t.set_inflow(*inflow)
        t0 = (H0.sum() * area, C0.sum() * area)
        b = [0.0, 0.0]
        rep = None
        for a in amps:
            t.set_face_velocity(a * bx + tx, a * by + ty)
            rep = t.step(every, t.stable_dt(0.4))
            b[0] += rep["boundary_H"]
            b[1] += rep["boundary_C"]
        Hf, Cf = t.download()
        t.close()
        t1 = (rep["total_H"], rep["total_C"])
        for f in (0, 1):
            defect = t1[f] - t0[f] + b[f]
            assert abs(defect) <= 1e-12 * abs(t0[f]), (tag, f, defect)
        assert b[0] != 0.0
        out[tag] = (Hf, Cf, t1)

    def rel(a, ref):
        return max(np.abs(a[0] - ref[0]).max() / np.abs(ref[0]).max(),
                   np.abs(a[1] - ref[1]).max() / np.abs(ref[1]).max())

    spread = rel(out["fp64p"], out["fp64"])
    err = rel(out["mixed"], out["fp64"])
    print(f"full chip at 2048x2048, 250 steps: mixed vs FP64 {err:.3e}, FP64 spread {spread:.3e}",
          flush=True)
    assert err <= max(1e-6, 10.0 * spread), (err, spread)
    for f in (0, 1):
        assert abs(out["mixed"][2][f] - out["fp64"][2][f]) <= 1e-9 * abs(out["fp64"][2][f])
