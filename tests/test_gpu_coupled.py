"""Coupled transport law on the GPU against the oracle, through the C ABI.

Bars (FP64 on both sides): the RHS fields (rhs_H, rhs_C, rate, mixture
divergence, cell-averaged T and phi) within 1e-12 of the oracle relative to
each field's maximum, the totals within 1e-12; the SSPRK2 loop takes the same
number of retries on every step, the same dt to 1e-12, and ends within 1e-11
of the oracle's state; errors carry the reference's messages.
"""
import numpy as np
import pytest

import mantlepar_amd as M
import oracle as O
from mantlepar_amd import workloads as W

pytestmark = pytest.mark.gpu

XE = 0.4
THETA, COOLING, KAPPA = 1.0, 0.05, 2e-4


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def pair(spec, darcy=0.05, H_in=0.55, kappa=KAPPA, state=None, surface=None):
    """Device handle and oracle on the same workload; surface: lithostatic
    pressure below surface_y (rho = 600) instead of P = 0."""
    flow = W.coupled_flow(spec, darcy=darcy)
    bc, inflow, ttypes, tvals = W.coupled_boundary(spec, H_in=H_in)
    H, Cc = W.coupled_state(spec) if state is None else state
    mat = {} if surface is None else {"rho": 600.0}
    g = M.Coupled(spec.nx, spec.ny, spec.x0, spec.x1, spec.y0, spec.y1, q=spec.q,
                  phase=W.config_phase(hypereutectic=0, **mat),
                  params=M.default_coupled(theta=THETA, cooling=COOLING,
                                           thermal_diffusivity=kappa))
    g.set_flow(flow)
    g.set_boundary(0, bc)
    g.set_boundary(1, bc)
    g.set_inflow(*inflow)
    g.set_thermal_boundary(ttypes, tvals)
    g.upload(H, Cc)
    vx = np.zeros((spec.q, spec.ny, spec.nx + 1))
    vy = np.zeros((spec.q, spec.ny + 1, spec.nx))
    o = O.Transport(W.vertex_coords(spec), spec.nx, spec.ny, vx, vy, q=spec.q, bc=bc,
                    inflow_H=inflow[0], inflow_C=inflow[1])
    o.set_limiter(True, 5, 4, 0.9, XE)
    o.set_diffusion(kappa, 4, 0.9, types={"bottom": np.ones(spec.nx)}, values=tvals)
    pm = O.PhaseModel(O.Material.defaults(Tm0=1600.0, Te0=600.0, cp=1200.0, L=1.2e6, Xe=XE, **mat))
    cp = O.CoupledParams.make(LD=pm.derived["LD"], theta=THETA, cooling=COOLING, kappa=kappa,
                              admissible_c_max=XE, lithostatic=int(surface is not None),
                              surface_y=surface or 0.0)
    if surface is not None:
        g.set_lithostatic(surface)
    return g, o, pm, cp, flow, H, Cc


@pytest.mark.parametrize("shape", [(16, 14), (40, 36)])
def test_rhs_matches_oracle(shape):
    spec = W.MeshSpec(*shape, seed=2)
    g, o, pm, cp, flow, H, Cc = pair(spec)
    got = g.rhs()
    want = o.coupled_rhs(pm, cp, flow, H, Cc)
    for k in ("rhsH", "rhsC", "rate", "div", "T", "phi"):
        assert rel(got[k], want[k]) <= 1e-12, (k, rel(got[k], want[k]))
    for a, b in (("boundary_H", "boundaryH"), ("boundary_C", "boundaryC"), ("source_H", "sourceH"),
                 ("maximum_rate", "maximumRate"), ("divergence", "divergence")):
        assert abs(got[a] - want[b]) <= 1e-12 * abs(want[b]), (a, got[a], want[b])
    # the workload exercises every term
    assert want["boundaryH"] != 0 and want["sourceH"] != 0 and np.abs(want["div"]).max() > 0
    g.close()


def test_step_loop_matches_oracle_with_retries():
    spec = W.MeshSpec(16, 14, seed=2)
    _, Cc = W.coupled_state(spec)
    H = np.full_like(Cc, 0.599)
    g, o, pm, cp, flow, H, Cc = pair(spec, darcy=0.5, H_in=0.9, kappa=0.0, state=(H, Cc))
    tg, sg = g.run(6, cfl=0.5, initial=1.0, maximum=1.0)
    Ho, Co, to, so = o.coupled_run(pm, cp, flow, H, Cc, 6, cfl=0.5, initial=1.0, maximum=1.0)
    assert len(sg) == len(so) == 6
    assert [s["retries"] for s in sg] == [s["retries"] for s in so]
    assert sum(s["retries"] for s in sg) >= 3
    for a, b in zip(sg, so):
        assert abs(a["dt"] - b["dt"]) <= 1e-12 * b["dt"]
        assert abs(a["balance_H"]) < 1e-10 and abs(a["balance_C"]) < 1e-10
    assert abs(tg - to) <= 1e-12 * to
    Hg, Cg = g.download()
    assert rel(Hg, Ho) <= 1e-11 and rel(Cg, Co) <= 1e-11
    g.close()


def test_failed_fixed_step_rolls_back():
    """A fixed step at the CFL bound that the retry workload cannot take: both
    sides fail with the reference's message, and the device state is the
    state before the step, although the attempt ran both stages."""
    spec = W.MeshSpec(16, 14, seed=2)
    _, Cc = W.coupled_state(spec)
    H = np.full_like(Cc, 0.599)
    g, o, pm, cp, flow, H, Cc = pair(spec, darcy=0.5, H_in=0.9, kappa=0.0, state=(H, Cc))
    _, _, _, so = o.coupled_run(pm, cp, flow, H, Cc, 1, cfl=0.5, initial=1.0, maximum=1.0)
    assert so[0]["retries"] >= 1          # the step at the bound is rejected
    dt = 0.5 / g.rhs()["maximum_rate"]
    msg = "Fixed coupled step failed stage admissibility/CFL"
    with pytest.raises(O.OracleError, match=msg):
        o.coupled_run(pm, cp, flow, H, Cc, 1, cfl=0.5, initial=dt, maximum=dt, fixed=True)
    with pytest.raises(M.MantleError, match=msg):
        g.run(1, cfl=0.5, initial=dt, maximum=dt, fixed=True)
    Hg, Cg = g.download()
    assert np.array_equal(Hg, H) and np.array_equal(Cg, Cc)
    # the handle goes on from the kept state as a fresh one would
    tg, sg = g.run(3, cfl=0.5, initial=1.0, maximum=1.0)
    Ho, Co, to, so = o.coupled_run(pm, cp, flow, H, Cc, 3, cfl=0.5, initial=1.0, maximum=1.0)
    assert [s["retries"] for s in sg] == [s["retries"] for s in so]
    Hg, Cg = g.download()
    assert rel(Hg, Ho) <= 1e-11 and rel(Cg, Co) <= 1e-11
    g.close()


def test_general_run_matches_oracle():
    spec = W.MeshSpec(24, 20, seed=2)
    g, o, pm, cp, flow, H, Cc = pair(spec)
    tg, sg = g.run(5, cfl=0.5, initial=1e-3, maximum=1.0)
    Ho, Co, to, so = o.coupled_run(pm, cp, flow, H, Cc, 5, cfl=0.5, initial=1e-3, maximum=1.0)
    assert [s["retries"] for s in sg] == [s["retries"] for s in so]
    Hg, Cg = g.download()
    assert rel(Hg, Ho) <= 1e-11 and rel(Cg, Co) <= 1e-11
    g.close()


def test_errors_keep_the_state():
    spec = W.MeshSpec(16, 14, seed=2)
    H, Cc = W.coupled_state(spec)
    bad = Cc.copy()
    bad[3, 3] = XE + 0.01
    g, *_ = pair(spec, state=(H, bad))
    with pytest.raises(M.MantleError, match="Inadmissible cell average") as ei:
        g.rhs()
    assert ei.value.code == 1
    g.close()
    g, o, pm, cp, flow, H, Cc = pair(spec)
    r = g.rhs()
    big = 4.0 / r["maximum_rate"]
    with pytest.raises(M.MantleError, match="exceeds the phase-coupled CFL bound"):
        g.run(1, cfl=0.5, initial=big, maximum=big, fixed=True)
    Hg, Cg = g.download()
    assert np.array_equal(Hg, H) and np.array_equal(Cg, Cc)
    g.close()


# ------------------------------------------------------- lithostatic pressure
@pytest.mark.parametrize("shape", [(16, 14), (40, 36)])
def test_lithostatic_rhs_matches_oracle(shape):
    """Every phase evaluation at its own point's static pressure: cell points,
    both sides of each face point, inflow temperatures and diffusion samples."""
    spec = W.MeshSpec(*shape, seed=2)
    g, o, pm, cp, flow, H, Cc = pair(spec, surface=1.2)
    got = g.rhs()
    want = o.coupled_rhs(pm, cp, flow, H, Cc)
    for k in ("rhsH", "rhsC", "rate", "div", "T", "phi"):
        assert rel(got[k], want[k]) <= 1e-12, (k, rel(got[k], want[k]))
    for a, b in (("boundary_H", "boundaryH"), ("boundary_C", "boundaryC"), ("source_H", "sourceH"),
                 ("maximum_rate", "maximumRate")):
        assert abs(got[a] - want[b]) <= 1e-12 * abs(want[b]), (a, got[a], want[b])
    # the pressure matters here
    flat = o.coupled_rhs(pm, O.CoupledParams.make(LD=pm.derived["LD"], theta=THETA, cooling=COOLING,
                                                 kappa=KAPPA, admissible_c_max=XE), flow, H, Cc)
    assert rel(flat["T"], want["T"]) > 1e-3
    g.close()


def test_lithostatic_run_matches_oracle():
    spec = W.MeshSpec(24, 20, seed=2)
    g, o, pm, cp, flow, H, Cc = pair(spec, surface=1.2)
    tg, sg = g.run(5, cfl=0.5, initial=1e-3, maximum=1.0)
    Ho, Co, to, so = o.coupled_run(pm, cp, flow, H, Cc, 5, cfl=0.5, initial=1e-3, maximum=1.0)
    assert [s["retries"] for s in sg] == [s["retries"] for s in so]
    for a, b in zip(sg, so):
        assert abs(a["dt"] - b["dt"]) <= 1e-12 * b["dt"]
    Hg, Cg = g.download()
    assert rel(Hg, Ho) <= 1e-11 and rel(Cg, Co) <= 1e-11
    g.close()


def test_lithostatic_eutectic_layer_closed_form():
    """H = Tep(y) + 0.25, C = 0.2: T = Tep(P(y)) at every point, linear in
    depth with slope B.  Uniform upward flow v entering at the bottom with
    H_in = Tep(y0) + 0.25 and diffusion kappa with insulated sides give
    rhs_H = v*B everywhere plus kappa*B/dy in the top row and -kappa*B/dy in
    the bottom row; rhs_C = 0."""
    spec = W.MeshSpec(10, 12, seed=1)
    surface, v, kappa = 1.2, 0.05, 1e-2
    g, o, pm, cp, flow, _, _ = pair(spec, surface=surface, kappa=kappa)
    B = pm.static_pressure(1.0) / pm.material.nu / pm.derived["dT"]

    def tep(y):
        return pm.boundaries(0.2, pm.static_pressure(surface - y))["Tep"]

    yc = spec.y0 + (np.arange(spec.ny) + 0.5) * spec.dy
    H = np.repeat(np.array([tep(y) + 0.25 for y in yc])[:, None], spec.nx, axis=1)
    Cc = np.full_like(H, 0.2)
    for k in flow:
        flow[k][:] = 0.0
    flow["edge_stokes"][..., 1] = v
    flow["cell_stokes"][..., 1] = v
    g.set_flow(flow)
    n = 2 * (spec.nx + spec.ny) * spec.q
    g.set_inflow(np.full(n, tep(spec.y0) + 0.25), np.full(n, 0.2))
    g.set_thermal_boundary(None)
    g.upload(H, Cc)
    got = g.rhs()
    want = np.full_like(H, v * B)
    want[-1] += kappa * B / spec.dy
    want[0] -= kappa * B / spec.dy
    # the workload's cooling source -cooling*T*v_y, averaged over the cell's
    # points: T is linear in y, so its average is Tep at the centre
    cool = -COOLING * v * np.array([tep(y) for y in yc])[:, None]
    assert np.abs(got["rhsH"] - (want + cool)).max() <= 1e-11 * np.abs(want).max()
    assert np.abs(got["rhsC"]).max() <= 1e-13
    g.close()


# --------------------------------------------- options of the coupled RHS
def options_pair(spec, hyper=0, lf_mode=0, global_speed=0.0, bcC=None, state=None, inflow_C=0.2):
    """pair() with the hypereutectic mirror, the global LF speed and separate
    composition policies; the same settings on both sides."""
    flow = W.coupled_flow(spec)
    bc, inflow, ttypes, tvals = W.coupled_boundary(spec, C_in=inflow_C)
    H, Cc = W.coupled_state(spec) if state is None else state
    c_max = 1.0 if hyper else XE
    g = M.Coupled(spec.nx, spec.ny, spec.x0, spec.x1, spec.y0, spec.y1, q=spec.q,
                  phase=W.config_phase(hypereutectic=hyper),
                  params=M.default_coupled(theta=THETA, cooling=COOLING, thermal_diffusivity=KAPPA,
                                           lf_mode=lf_mode, global_speed=global_speed))
    g.set_flow(flow)
    g.set_boundary(0, bc)
    g.set_boundary(1, bc if bcC is None else bcC)
    g.set_inflow(*inflow)
    g.set_thermal_boundary(ttypes, tvals)
    g.upload(H, Cc)
    vx = np.zeros((spec.q, spec.ny, spec.nx + 1))
    vy = np.zeros((spec.q, spec.ny + 1, spec.nx))
    o = O.Transport(W.vertex_coords(spec), spec.nx, spec.ny, vx, vy, q=spec.q, bc=bc,
                    inflow_H=inflow[0], inflow_C=inflow[1])
    if bcC is not None:
        o.set_composition_boundary(bcC)
    o.set_limiter(True, 5, 4, 0.9, c_max)
    o.set_diffusion(KAPPA, 4, 0.9, types={"bottom": np.ones(spec.nx)}, values=tvals)
    pm = O.PhaseModel(O.Material.defaults(Tm0=1600.0, Te0=600.0, cp=1200.0, L=1.2e6, Xe=XE))
    cp = O.CoupledParams.make(LD=pm.derived["LD"], theta=THETA, cooling=COOLING, kappa=KAPPA,
                              lf_mode=lf_mode, global_speed=global_speed, hypereutectic=hyper,
                              admissible_c_max=c_max)
    return g, o, pm, cp, flow, H, Cc


def check_rhs(g, o, pm, cp, flow, H, Cc, spec):
    """Fields within 1e-12 of their maximum; totals within 1e-12 relative, or
    within 1e-15 of the field's sum |K| |rhs| where the total is a cancelling
    sum of terms of both signs (the mirror case's cooling source sums to 1e-9
    from terms near 1e-5)."""
    got = g.rhs()
    want = o.coupled_rhs(pm, cp, flow, H, Cc)
    for k in ("rhsH", "rhsC", "rate", "div", "T", "phi"):
        assert rel(got[k], want[k]) <= 1e-12, (k, rel(got[k], want[k]))
    area = spec.dx * spec.dy
    for a, b, f in (("boundary_H", "boundaryH", "rhsH"), ("boundary_C", "boundaryC", "rhsC"),
                    ("source_H", "sourceH", "rhsH"), ("maximum_rate", "maximumRate", "rate")):
        tol = 1e-12 * abs(want[b]) + 1e-15 * area * np.abs(want[f]).sum()
        assert abs(got[a] - want[b]) <= tol, (a, got[a], want[b])
    return got, want


def test_global_lf_and_separate_composition_policies():
    """Global Lax-Friedrichs speed (alpha fixed above every |v.n|) and a
    composition that may not enter through the bottom (zero-flux C policy there,
    inflow for H)."""
    spec = W.MeshSpec(24, 20, seed=2)
    bcC = {"bottom": np.zeros(spec.nx, np.uint8), "right": np.zeros(spec.ny, np.uint8),
           "top": np.ones(spec.nx, np.uint8), "left": np.zeros(spec.ny, np.uint8)}
    g, o, pm, cp, flow, H, Cc = options_pair(spec, lf_mode=1, global_speed=2.0, bcC=bcC)
    got, want = check_rhs(g, o, pm, cp, flow, H, Cc, spec)
    # the zero-flux top changes C's boundary flux against the shared policies
    g2, o2, pm2, cp2, flow2, _, _ = options_pair(spec, lf_mode=1, global_speed=2.0)
    assert abs(g2.rhs()["boundary_C"] - got["boundary_C"]) > 1e-6 * abs(got["boundary_C"])
    tg, sg = g.run(3, cfl=0.5, initial=1e-3, maximum=1.0)
    Ho, Co, to, so = o.coupled_run(pm, cp, flow, H, Cc, 3, cfl=0.5, initial=1e-3, maximum=1.0)
    Hg, Cg = g.download()
    assert rel(Hg, Ho) <= 1e-11 and rel(Cg, Co) <= 1e-11
    g.close()
    g2.close()
    # a global speed below a flow speed is rejected
    g3, *_ = options_pair(spec, lf_mode=1, global_speed=1e-4)
    with pytest.raises(M.MantleError):
        g3.rhs()
    g3.close()


def test_hypereutectic_mirror():
    """C on both sides of the eutectic (0.2-0.6 about Xe = 0.4), evaluated on the
    mirrored diagram above Xe; admissible up to C = 1."""
    spec = W.MeshSpec(20, 18, seed=3)
    H, _ = W.coupled_state(spec)
    xc = (np.arange(spec.nx) + 0.5) / spec.nx
    yc = (np.arange(spec.ny) + 0.5) / spec.ny
    Cc = 0.2 + 0.4 * (0.5 + 0.5 * np.outer(np.cos(np.pi * yc), np.sin(2 * np.pi * xc)))
    g, o, pm, cp, flow, H, Cc = options_pair(spec, hyper=1, state=(H, Cc), inflow_C=0.5)
    assert Cc.max() > XE and Cc.min() < XE
    check_rhs(g, o, pm, cp, flow, H, Cc, spec)
    tg, sg = g.run(3, cfl=0.5, initial=1e-3, maximum=1.0)
    Ho, Co, to, so = o.coupled_run(pm, cp, flow, H, Cc, 3, cfl=0.5, initial=1e-3, maximum=1.0)
    assert [s["retries"] for s in sg] == [s["retries"] for s in so]
    Hg, Cg = g.download()
    assert rel(Hg, Ho) <= 1e-11 and rel(Cg, Co) <= 1e-11
    g.close()
