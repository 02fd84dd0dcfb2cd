"""HIP transport path against the CPU oracle, through the C ABI.

Tolerances (FP64 path): reconstruction traces and RHS within 1e-12 of the
oracle relative to the field's maximum; indicators within 1e-11 relative to
the largest indicator of the field; weights within 1e-11 absolute; states
after the config-1 run (10 SSP-RK3 steps) within 1e-12 relative (max norm).
Mixed path: states within 1e-6 relative of the FP64 oracle.  Phase regions
must match exactly; T and phi within 1e-13.
"""
import numpy as np
import pytest

import mantlepar_amd as M
import oracle as O
from mantlepar_amd import workloads as W

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def setup(spec, prec=M.PREC_FP64, vel=None, bc=None, inflow=None, phase=None):
    H, Cc = W.enthalpy(spec), W.composition(spec)
    vx, vy = vel if vel is not None else W.face_velocity(spec)
    t = M.Transport(spec.nx, spec.ny, spec.x0, spec.x1, spec.y0, spec.y1, q=spec.q,
                    phase=phase or W.config_phase(), precision=prec)
    t.upload(H, Cc)
    t.set_face_velocity(vx, vy)
    if bc is not None:
        t.set_boundary(bc.get("bottom"), bc.get("right"), bc.get("top"), bc.get("left"))
    if inflow is not None:
        t.set_inflow(*inflow)
    o = O.Transport(W.vertex_coords(spec), spec.nx, spec.ny, vx, vy, q=spec.q, bc=bc,
                    inflow_H=None if inflow is None else inflow[0],
                    inflow_C=None if inflow is None else inflow[1],
                    table_mode=0 if spec.cells <= 4096 else 1)
    return t, o, H, Cc


MESHES = [W.MeshSpec(20, 20, seed=1), W.MeshSpec(2, 2, seed=2), W.MeshSpec(3, 7, seed=3),
          W.MeshSpec(64, 5, seed=4), W.MeshSpec(130, 97, seed=5), W.MeshSpec(17, 9, q=3, seed=6),
          W.MeshSpec(9, 11, q=8, seed=7), W.MeshSpec(12, 12, q=1, seed=8)]


@pytest.mark.parametrize("spec", MESHES, ids=lambda s: f"{s.nx}x{s.ny}q{s.q}")
def test_reconstruction_matches_oracle(spec):
    t, o, H, Cc = setup(spec)
    for field, u in ((0, H), (1, Cc)):
        sg, om, tr = t.reconstruct(field)
        so, oo, to = o.reconstruct(u)
        assert np.abs(sg - so).max() <= 1e-11 * max(so.max(), 1e-300)
        assert np.abs(om - oo).max() <= 1e-11
        assert rel(tr, to) <= 1e-12


@pytest.mark.parametrize("spec", MESHES[:5], ids=lambda s: f"{s.nx}x{s.ny}")
def test_rhs_matches_oracle(spec):
    t, o, H, Cc = setup(spec)
    LH, LC, bh, bc = t.rhs()
    oLH, oLC, obh, obc = o.rhs(H, Cc)
    assert rel(LH, oLH) <= 1e-12 and rel(LC, oLC) <= 1e-12


def open_case(spec, seed=0):
    q = spec.q
    vx = np.full((q, spec.ny, spec.nx + 1), 0.3)
    vy = np.full((q, spec.ny + 1, spec.nx), -0.2)
    rng = np.random.default_rng(seed)
    n = 2 * (spec.nx + spec.ny) * q
    inflow = (rng.random(n), rng.random(n) * 0.4)
    bc = {"bottom": np.full(spec.nx, 2, np.uint8), "right": np.full(spec.ny, 1, np.uint8),
          "top": np.full(spec.nx, 2, np.uint8), "left": np.full(spec.ny, 2, np.uint8)}
    bc["bottom"][::3] = 0           # some zero-flux faces
    return (vx, vy), bc, inflow


@pytest.mark.parametrize("prec", [M.PREC_FP64, M.PREC_MIXED])
def test_open_boundaries_match_oracle(prec):
    spec = W.MeshSpec(23, 19, seed=9)
    vel, bc, inflow = open_case(spec)
    t, o, H, Cc = setup(spec, prec, vel, bc, inflow)
    LH, LC, bh, bcf = t.rhs()
    oLH, oLC, obh, obc = o.rhs(H, Cc)
    tol = 1e-12 if prec == M.PREC_FP64 else 1e-6
    assert rel(LH, oLH) <= tol and rel(LC, oLC) <= tol
    assert abs(bh - obh) <= tol * abs(obh) and abs(bcf - obc) <= tol * abs(obc)
    dt = t.stable_dt(0.4)
    rep = t.step(3, dt)
    Hg, Cg = t.download()
    Ho, Co, tot = o.advance(H, Cc, 3, dt, 3)
    assert rel(Hg, Ho) <= 10 * tol and rel(Cg, Co) <= 10 * tol
    # conservation: total(after) = total(before) - sum dt * weighted boundary flux
    area = spec.dx * spec.dy
    assert abs(rep["total_H"] + rep["boundary_H"] - area * H.sum()) < 1e-12 * area * H.sum()
    assert abs(rep["total_C"] + rep["boundary_C"] - area * Cc.sum()) < 1e-12 * area * Cc.sum()


@pytest.mark.parametrize("scheme", [1, 2, 3])
def test_config1_run_matches_oracle_fp64(scheme):
    spec, nsteps, cfl = W.config1()
    t, o, H, Cc = setup(spec)
    dt = t.stable_dt(cfl)
    _, mr = o.rates()
    assert abs(dt - cfl / mr) <= 1e-14 * dt
    rep = t.step(nsteps, dt, scheme)
    Hg, Cg = t.download()
    Ho, Co, tot = o.advance(H, Cc, nsteps, dt, scheme)
    assert rel(Hg, Ho) <= 1e-12 and rel(Cg, Co) <= 1e-12
    assert abs(rep["total_H"] - tot[-1, 0]) <= 1e-14 and abs(rep["total_C"] - tot[-1, 1]) <= 1e-14
    ph = t.phase_state()
    om = O.PhaseModel(O.Material.defaults(Tm0=1600.0, Te0=600.0, cp=1200.0, L=1.2e6, Xe=0.4))
    T, phi, cl, cs, reg = om.batch(Ho, Co, 0.0, hyper=True)
    assert np.array_equal(ph["region"].ravel(), reg)
    assert np.abs(ph["T"].ravel() - T).max() <= 1e-13
    assert np.abs(ph["phi"].ravel() - phi).max() <= 1e-13


def test_config1_mixed_precision_within_1e6_of_oracle():
    spec, nsteps, cfl = W.config1()
    t, o, H, Cc = setup(spec, M.PREC_MIXED)
    dt = t.stable_dt(cfl)
    t.step(nsteps, dt)
    Hg, Cg = t.download()
    Ho, Co, _ = o.advance(H, Cc, nsteps, dt, 3)
    assert rel(Hg, Ho) <= 1e-6 and rel(Cg, Co) <= 1e-6


def test_results_do_not_depend_on_tile_shape():
    spec = W.MeshSpec(150, 77, seed=12)
    out = []
    for rows in (4, 5, 64, 200):
        t, o, H, Cc = setup(spec)
        t.set_tile(rows)
        t.step(2, t.stable_dt(0.4))
        out.append(t.download())
    for h, c in out[1:]:
        assert np.array_equal(h, out[0][0]) and np.array_equal(c, out[0][1])


@pytest.mark.parametrize("prec", [M.PREC_FP64, M.PREC_MIXED], ids=["fp64", "mixed"])
def test_tail_items_leave_results_unchanged(prec):
    """On a mesh that fills the chip, 12-row items leave the top rows to 4-row
    tail items dispatched last; 4-row items have no tail.  The states are the
    same bit for bit."""
    spec = W.MeshSpec(2048, 3400, seed=19)
    H, Cc = W.enthalpy(spec), W.composition(spec)
    out = []
    for rows in (4, 12):
        t = M.Transport(spec.nx, spec.ny, spec.x0, spec.x1, spec.y0, spec.y1, q=spec.q,
                        phase=W.config_phase(), precision=prec)
        t.upload(H, Cc)
        t.set_face_velocity(*W.face_velocity(spec))
        t.set_tile(rows)
        lay = t.layout()
        # the tail exists only with items taller than 4 rows; the test is
        # meaningful only where the layout has one
        assert (lay["tail_split"] < spec.ny) == (rows > 4), lay
        t.step(2, t.stable_dt(0.4))
        out.append(t.download() + (t.phase_state()["T"],))
        t.close()
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)


def test_backflow_error_leaves_state_unchanged():
    spec = W.MeshSpec(16, 8, seed=3)
    vx = np.full((spec.q, spec.ny, spec.nx + 1), 0.3)     # inflow through the left side
    vy = np.zeros((spec.q, spec.ny + 1, spec.nx))
    t, o, H, Cc = setup(spec, vel=(vx, vy), bc={"left": np.full(spec.ny, 1, np.uint8)})
    before = t.download()
    with pytest.raises(M.MantleError) as e:
        t.step(2, t.stable_dt(0.4))
    assert e.value.code == 7
    after = t.download()
    assert np.array_equal(before[0], after[0]) and np.array_equal(before[1], after[1])


def test_composition_outside_diagram_is_rejected_without_mirror():
    spec = W.MeshSpec(10, 10, seed=4)
    t, o, H, Cc = setup(spec, phase=W.config_phase(hypereutectic=0))
    before = t.download()
    with pytest.raises(M.MantleError) as e:
        t.step(1, t.stable_dt(0.4))
    assert e.value.code == 8
    after = t.download()
    assert np.array_equal(before[1], after[1])


def test_step_before_velocity_is_a_state_error():
    t = M.Transport(8, 8)
    with pytest.raises(M.MantleError) as e:
        t.step(1, 1e-3)
    assert e.value.code == 3


def test_full_size_config2_properties():
    """2048x4096 at its real shape: conservation, and mixed vs FP64 within 1e-6."""
    spec, _, cfl = W.config2()
    res = {}
    for prec in (M.PREC_FP64, M.PREC_MIXED):
        t, _, H, Cc = (None, None, None, None)
        H, Cc = W.enthalpy(spec), W.composition(spec)
        vx, vy = W.face_velocity(spec)
        t = M.Transport(spec.nx, spec.ny, q=spec.q, phase=W.config_phase(), precision=prec)
        t.upload(H, Cc)
        t.set_face_velocity(vx, vy)
        rep = t.step(10, t.stable_dt(cfl))
        res[prec] = (t.download(), rep)
        area = spec.dx * spec.dy
        assert abs(rep["total_H"] - area * np.sum(H)) <= 1e-12 * area * np.sum(H)
        assert abs(rep["total_C"] - area * np.sum(Cc)) <= 1e-12 * area * np.sum(Cc)
        t.close()
    (h64, c64), _ = res[M.PREC_FP64]
    (hm, cm), _ = res[M.PREC_MIXED]
    assert rel(hm, h64) <= 1e-6 and rel(cm, c64) <= 1e-6


def row_pressure(om, spec, surface):
    """staticPressure(surface - y_c) per row, y_c = MapCellPoint((0, 0)) over the
    uniform vertices (src/core/integral.cpp:91-106)."""
    y = W.vertex_coords(spec)[:, 0, 1]
    lower, upper = 0.5 * y[:-1] + 0.5 * y[:-1], 0.5 * y[1:] + 0.5 * y[1:]
    yc = 0.5 * lower + 0.5 * upper
    return np.array([om.static_pressure(surface - v) for v in yc])


def oracle_phase_rows(om, H, Cc, P):
    out = [om.batch(H[j], Cc[j], P[j], hyper=True) for j in range(H.shape[0])]
    return [np.stack([o[k] for o in out]) for k in range(5)]


@pytest.mark.parametrize("prec", [M.PREC_FP64, M.PREC_MIXED], ids=["fp64", "mixed"])
def test_lithostatic_phase_per_row(prec):
    """Pressure model lithostatic: each cell's phase at the static pressure of
    its centre, before stepping (state phase kernel) and fused into the last
    stage of a step; regions exact, T and phi within 1e-13."""
    spec = W.MeshSpec(40, 36, seed=3)
    t, o, H, Cc = setup(spec, prec, phase=W.config_phase(rho=600.0))
    om = O.PhaseModel(O.Material.defaults(Tm0=1600.0, Te0=600.0, cp=1200.0, L=1.2e6, Xe=0.4,
                                          rho=600.0))
    P = row_pressure(om, spec, 1.2)
    flat = t.phase_state()
    t.set_lithostatic(1.2)
    for stepped in (False, True):
        if stepped:
            t.step(5, t.stable_dt(0.4))
        Hs, Cs = t.download()
        ph = t.phase_state()
        T, phi, cl, cs, reg = oracle_phase_rows(om, Hs, Cs, P)
        assert np.array_equal(ph["region"], reg), stepped
        assert np.abs(ph["T"] - T).max() <= 1e-13
        assert np.abs(ph["phi"] - phi).max() <= 1e-13
    # the pressure moves the eutectic: many cells change region
    assert (flat["region"] != oracle_phase_rows(om, H, Cc, P)[4]).sum() > 50
    with pytest.raises(M.MantleError, match="surface_y must be at or above the mesh top"):
        t.set_lithostatic(0.9)
    t.set_lithostatic(None)
    Hs, Cs = t.download()
    T0, _, _, _, reg0 = om.batch(Hs.ravel(), Cs.ravel(), 0.0, hyper=True)
    assert np.array_equal(t.phase_state()["region"].ravel(), reg0)
    t.close()


@pytest.mark.parametrize("prec", [M.PREC_FP64, M.PREC_MIXED], ids=["fp64", "mixed"])
def test_step_rejects_a_failed_heat_balance(prec):
    """Every step is checked against the preheat's conservation balance
    (src/couple/preheat.cpp:268-273).  Forced from the host: with tolerance 0
    only an exact balance passes, and on a through-flow the rounding residue is
    not zero, so the step is rejected with the reference's message, the state
    unchanged; at the reference's 1e-10 the same steps pass (as they do in
    every other test here)."""
    spec = W.MeshSpec(96, 128, seed=41)
    vel = W.throughflow_velocity(spec)
    bc, inflow = W.throughflow_boundary(spec)
    t, _, _, _ = setup(spec, prec, vel=vel, bc=bc, inflow=inflow)
    dt = t.stable_dt(0.4)
    H0, C0 = t.download()
    t.set_balance_tolerance(0.0)
    with pytest.raises(M.MantleError) as e:
        t.step(3, dt)
    assert e.value.code == 10 and "heat balance failed; candidate rejected" in str(e.value)
    H1, C1 = t.download()
    assert np.array_equal(H0, H1) and np.array_equal(C0, C1)
    t.set_balance_tolerance()
    rep = t.step(3, dt)
    assert rep["steps_done"] == 3
    with pytest.raises(M.MantleError):
        t.set_balance_tolerance(-1.0)
    t.close()
