"""Oracle pins for the coupled transport law (CPU only).

The reference's coupled RHS (src/couple/evolve.cpp:279-384) and step loop
(:428-511) need the Stokes-Darcy solve and PETSc, so they do not run here.
The oracle's restatement is pinned by:

* PhaseTransportVelocities: the reference's own cases
  (src/couple/test/test_evolution.cpp:5-27: wet, dry, pure and
  liquid-only limits at 1e-14);
* reduction to the transport operator: in a sub-solidus state (T = H) with
  dry flow (flow porosity 0: mixture = effective = Stokes velocity), no
  diffusion and no cooling, the C equation is the limited transport operator
  exactly, and the H equation is that operator plus LD times the discrete
  divergence of the Stokes velocity (the latent term -LD*(1-phi)*v_s.n);
* the cooling source: on a uniform state with flow only at the cell points,
  rhs_H = sum_ab w_a w_b/4 * (-cooling * T * v_mixture_y) per cell;
* diffusion of T with reconstructed samples: a field linear in x has zero
  rhs in the cells whose samples all come from unlimited reconstructions;
* conservation: sum |K| rhs_H = source_H - boundary_H and
  sum |K| rhs_C = -boundary_C on a general case (wet partial melt, inflow,
  outflow, Dirichlet temperature);
* the step loop: accepted steps pass the reference's balance and
  admissibility checks, dt halves on rejection, fixed steps above the CFL
  bound and inadmissible averages are rejected with the reference's messages.
"""
import numpy as np
import pytest

import oracle as O
from mantlepar_amd import workloads as W

XE = 0.4


def phase_model():
    return O.PhaseModel(O.Material.defaults(Tm0=1600.0, Te0=600.0, cp=1200.0, L=1.2e6, Xe=XE))


def edge_normal_velocity(spec, flow):
    """vnx/vny arrays of the transport oracle from the flow's Stokes samples."""
    nh = spec.nx * (spec.ny + 1)
    st = flow["edge_stokes"]
    vny = st[:nh, :, 1].reshape(spec.ny + 1, spec.nx, spec.q).transpose(2, 0, 1)
    vnx = st[nh:, :, 0].reshape(spec.ny, spec.nx + 1, spec.q).transpose(2, 0, 1)
    return np.ascontiguousarray(vnx), np.ascontiguousarray(vny)


def transport(spec, flow, bc=None, inflow=None):
    vx, vy = edge_normal_velocity(spec, flow)
    t = O.Transport(W.vertex_coords(spec), spec.nx, spec.ny, vx, vy, q=spec.q, bc=bc,
                    inflow_H=None if inflow is None else inflow[0],
                    inflow_C=None if inflow is None else inflow[1])
    t.set_limiter(True, 5, 4, 0.9, XE)
    return t


def test_phase_transport_velocities_reference_cases():
    solid, darcy = [.01, .02], [.3, -.4]
    m, e = O.transport_velocities(solid, darcy, .1, .5, (.2, .03), (.24, .05))
    q, bulk = .1 ** 1.5, .1 * .22 + .9 * .04
    assert np.allclose(m, np.add(solid, q * np.asarray(darcy)), rtol=0, atol=1e-14)
    assert np.allclose(e, np.add(solid, .22 / bulk * q * np.asarray(darcy)), rtol=0, atol=1e-14)
    m, e = O.transport_velocities(solid, darcy, 0.0, .5, (.2, .03), (.24, .05))
    assert np.allclose(m, solid, atol=1e-14) and np.allclose(e, solid, atol=1e-14)
    m, e = O.transport_velocities(solid, darcy, .1, 0.0, (0, 0), (0, 0))
    assert np.all(np.isfinite(e)) and np.allclose(e, solid, atol=1e-14)
    m, e = O.transport_velocities(solid, darcy, .1, 0.0, (.2, 0), (.2, 0))
    assert np.allclose(e, np.add(solid, darcy), atol=1e-14)


def test_dry_solid_state_reduces_to_limited_transport():
    spec = W.MeshSpec(14, 12, seed=3)
    H, Cc = W.coupled_state(spec)
    H = 0.3 + 0.25 * (H - H.min()) / (H.max() - H.min())       # below Tep = 0.6: T = H
    flow = W.coupled_flow(spec, phi_max=0.0)
    # a compressible part, so the latent term's divergence is not rounding-sized
    rng = np.random.default_rng(5)
    flow["edge_stokes"] += 0.02 * rng.standard_normal(flow["edge_stokes"].shape)
    nh = spec.nx * (spec.ny + 1)
    flow["edge_stokes"][:spec.nx, :, 1] = 0.0                    # bottom wall
    flow["edge_stokes"][nh - spec.nx:nh, :, 1] = 0.0             # top wall
    side = flow["edge_stokes"][nh:].reshape(spec.ny, spec.nx + 1, spec.q, 2)
    side[:, 0, :, 0] = 0.0                                       # left wall
    side[:, -1, :, 0] = 0.0                                      # right wall
    t = transport(spec, flow)
    pm = phase_model()
    LD = pm.derived["LD"]
    cp = O.CoupledParams.make(LD=LD)
    r = t.coupled_rhs(pm, cp, flow, H, Cc)
    LH, LC, _, _ = t.rhs(H, Cc)
    assert np.abs(r["rhsC"] - LC).max() <= 1e-13 * np.abs(LC).max()
    # discrete divergence of the Stokes samples, outward, per cell
    vx, vy = edge_normal_velocity(spec, flow)
    g, w = np.polynomial.legendre.leggauss(spec.q)
    Fx = np.einsum("q,qji->ji", w * spec.dy / 2, vx)
    Fy = np.einsum("q,qji->ji", w * spec.dx / 2, vy)
    div = (Fx[:, 1:] - Fx[:, :-1] + Fy[1:, :] - Fy[:-1, :]) / (spec.dx * spec.dy)
    assert np.abs(r["div"] - div).max() <= 1e-12 * np.abs(div).max()
    want = LH + LD * div
    assert np.abs(r["rhsH"] - want).max() <= 1e-12 * np.abs(want).max()
    assert r["sourceH"] == 0.0


def test_cooling_source_closed_form():
    spec = W.MeshSpec(6, 5, seed=1)
    H = np.full((spec.ny, spec.nx), 0.7)
    Cc = np.full((spec.ny, spec.nx), 0.25)
    flow = W.coupled_flow(spec)
    for k in ("edge_stokes", "edge_darcy"):
        flow[k][:] = 0.0
    t = transport(spec, flow)
    pm = phase_model()
    cp = O.CoupledParams.make(LD=pm.derived["LD"], theta=1.0, cooling=0.3)
    r = t.coupled_rhs(pm, cp, flow, H, Cc)
    T = pm.evaluate(0.7, 0.25)["TDp"]
    _, wg = np.polynomial.legendre.leggauss(5)
    wab = np.outer(wg, wg).ravel() / 4           # [b*n + a]
    factor = flow["cell_phi"] ** 2               # phi^(1 + theta)
    mixy = flow["cell_stokes"][..., 1] + factor * flow["cell_darcy"][..., 1]
    want = (wab[None, :] * (-0.3 * T * mixy)).sum(1).reshape(spec.ny, spec.nx)
    assert np.abs(r["rhsH"] - want).max() <= 1e-13 * np.abs(want).max()
    assert np.all(r["rhsC"] == 0)
    assert np.allclose(r["T"], T, rtol=0, atol=1e-15)
    assert abs(r["sourceH"] - (want * spec.dx * spec.dy).sum()) <= 1e-13 * abs(r["sourceH"])


def test_diffusion_of_a_linear_field():
    spec = W.MeshSpec(12, 10, seed=1)
    xc = spec.x0 + (np.arange(spec.nx) + 0.5) * spec.dx
    H = np.repeat((0.3 + 0.4 * (xc - spec.x0))[None, :], spec.ny, axis=0)   # T = H, solid
    Cc = np.full_like(H, 0.2)
    flow = W.coupled_flow(spec)
    for k in ("edge_stokes", "edge_darcy", "cell_stokes", "cell_darcy"):
        flow[k][:] = 0.0
    t = transport(spec, flow)
    pm = phase_model()
    cp = O.CoupledParams.make(LD=pm.derived["LD"], kappa=1e-2)
    r = t.coupled_rhs(pm, cp, flow, H, Cc)
    inner = r["rhsH"][:, 3:-3]
    assert np.abs(inner).max() <= 1e-12 * 1e-2 / spec.dx ** 2
    # and the interior flux itself is kappa*0.4 per unit length: a step in the
    # field's slope shows up in exactly the cells next to it
    H2 = H.copy()
    H2[:, 6:] += 0.1 * (xc[6:] - xc[5])[None, :]
    r2 = t.coupled_rhs(pm, cp, flow, H2, Cc)
    assert np.abs(r2["rhsH"][:, 3:-3]).max() > 1e-3


def general_case(spec, nsteps=None):
    H, Cc = W.coupled_state(spec)
    flow = W.coupled_flow(spec)
    bc, inflow, ttypes, tvals = W.coupled_boundary(spec)
    t = transport(spec, flow, bc, inflow)
    t.set_diffusion(2e-4, 4, 0.9, types={"bottom": np.ones(spec.nx)}, values=tvals)
    pm = phase_model()
    cp = O.CoupledParams.make(LD=pm.derived["LD"], theta=1.0, cooling=0.05, kappa=2e-4)
    return t, pm, cp, flow, H, Cc


def test_general_case_conserves():
    spec = W.MeshSpec(16, 14, seed=2)
    t, pm, cp, flow, H, Cc = general_case(spec)
    r = t.coupled_rhs(pm, cp, flow, H, Cc)
    area = spec.dx * spec.dy
    assert abs((r["rhsH"] * area).sum() - (r["sourceH"] - r["boundaryH"])) <= 1e-13 * np.abs(r["rhsH"] * area).sum()
    assert abs((r["rhsC"] * area).sum() + r["boundaryC"]) <= 1e-13 * np.abs(r["rhsC"] * area).sum()
    assert r["boundaryH"] != 0 and r["boundaryC"] != 0 and r["sourceH"] != 0
    assert r["maximumRate"] > 0 and np.isfinite(r["divergence"])


def retry_case(spec):
    """Hot inflow (H 0.9) into a state just below the eutectic (H 0.599): the
    first stage melts cells, the phase sensitivity 1 + |dT/dC| jumps, and the
    second stage's rate breaks the CFL bound of the first: dt is halved."""
    flow = W.coupled_flow(spec, darcy=0.5)
    bc, inflow, _, _ = W.coupled_boundary(spec, H_in=0.9)
    t = transport(spec, flow, bc, inflow)
    pm = phase_model()
    cp = O.CoupledParams.make(LD=pm.derived["LD"], theta=1.0, cooling=0.05)
    _, Cc = W.coupled_state(spec)
    return t, pm, cp, flow, np.full_like(Cc, 0.599), Cc


def test_step_loop_retries_and_balance():
    spec = W.MeshSpec(16, 14, seed=2)
    t, pm, cp, flow, H, Cc = retry_case(spec)
    H1, C1, time, steps = t.coupled_run(pm, cp, flow, H, Cc, 6, cfl=0.5, initial=1.0,
                                        maximum=1.0)
    assert sum(s["retries"] for s in steps) >= 3
    for s in steps:
        assert s["courant"] <= 0.5 * (1 + 1e-12)
        assert abs(s["balanceH"]) < 1e-10 and abs(s["balanceC"]) < 1e-10
    assert abs(time - sum(s["dt"] for s in steps)) <= 1e-14 * time
    assert H1.min() >= 0 and C1.min() >= 0 and C1.max() <= XE
    t, pm, cp, flow, H, Cc = general_case(spec)
    r = t.coupled_rhs(pm, cp, flow, H, Cc)
    big = 4.0 / r["maximumRate"]
    with pytest.raises(O.OracleError, match="exceeds the phase-coupled CFL bound"):
        t.coupled_run(pm, cp, flow, H, Cc, 1, cfl=0.5, initial=big, maximum=big, fixed=True)
    bad = Cc.copy()
    bad[3, 3] = XE + 0.01
    with pytest.raises(O.OracleError, match="Inadmissible cell average"):
        t.coupled_rhs(pm, cp, flow, H, bad)


# ------------------------------------------------------- lithostatic pressure
# P = staticPressure(surface_y - y) at each point where the phase is evaluated
# (src/couple/initialization.cpp:155-162, evolve.cpp:273 and :318).  With
# rho = 600 the eutectic temperature rises by B = rho*g*l0/(nu*dT) = 0.29 per
# unit depth; surface 1.2 over a mesh spanning y in [0, 1].
SURFACE = 1.2


def litho_model():
    return O.PhaseModel(O.Material.defaults(Tm0=1600.0, Te0=600.0, cp=1200.0, L=1.2e6, Xe=XE,
                                            rho=600.0))


def litho_slope(pm):
    return pm.static_pressure(1.0) / pm.material.nu / pm.derived["dT"]


def tep_at(pm, y, c=0.2):
    return pm.boundaries(c, pm.static_pressure(SURFACE - y))["Tep"]


def still_flow(spec):
    flow = W.coupled_flow(spec)
    for k in flow:
        flow[k][:] = 0.0
    return flow


def eutectic_layer(spec, pm):
    """H = Tep(y) + 0.25 at the cell centres, C = 0.2: the eutectic region at
    every depth (H_eutectic_end - Tep = LD*C/Xe = 0.5), so T = Tep(P(y)),
    linear in y."""
    yc = spec.y0 + (np.arange(spec.ny) + 0.5) * spec.dy
    H = np.repeat(np.array([tep_at(pm, y) + 0.25 for y in yc])[:, None], spec.nx, axis=1)
    return H, np.full_like(H, 0.2)


def test_lithostatic_diffusion_of_the_eutectic_layer():
    """No flow: the normal-line diffusion of T = Tep(y) carries kappa*B upward
    through every horizontal face, so interior rows have zero rhs and the
    insulated top and bottom rows gain and lose kappa*B/dy; the cell-averaged T
    is Tep at the cell centre."""
    spec = W.MeshSpec(10, 12, seed=1)
    pm = litho_model()
    B, kappa = litho_slope(pm), 1e-2
    H, Cc = eutectic_layer(spec, pm)
    flow = still_flow(spec)
    t = transport(spec, flow)
    t.set_diffusion(kappa, 4, 0.9)
    cp = O.CoupledParams.make(LD=pm.derived["LD"], kappa=kappa, lithostatic=1, surface_y=SURFACE)
    r = t.coupled_rhs(pm, cp, flow, H, Cc)
    scale = kappa * B / spec.dy
    want = np.zeros_like(H)
    want[-1] = scale
    want[0] = -scale
    assert np.abs(r["rhsH"] - want).max() <= 1e-11 * scale
    yc = spec.y0 + (np.arange(spec.ny) + 0.5) * spec.dy
    assert np.abs(r["T"] - np.array([tep_at(pm, y) for y in yc])[:, None]).max() <= 1e-14
    # eutectic melting: phi = (H - Tep)/LD at every point
    assert np.abs(r["phi"] - 0.25 / pm.derived["LD"]).max() <= 1e-14
    # at constant pressure the same state is no longer one layer
    r0 = t.coupled_rhs(pm, O.CoupledParams.make(LD=pm.derived["LD"], kappa=kappa), flow, H, Cc)
    assert np.abs(r0["rhsH"] - want).max() > 1e3 * 1e-11 * scale


def test_lithostatic_vertical_flow_through_the_eutectic_layer():
    """Uniform upward Stokes flow v through the layer, entering at the bottom
    with H_in = Tep(y0) + 0.25: every horizontal face carries v*Tep(y_face)
    minus the latent term, both sides of a face see the same point and
    pressure, and the inflow temperature is taken at the boundary point's
    pressure, so rhs_H = v*B in every cell and rhs_C = 0."""
    spec = W.MeshSpec(10, 12, seed=1)
    pm = litho_model()
    B, v = litho_slope(pm), 0.05
    H, Cc = eutectic_layer(spec, pm)
    flow = still_flow(spec)
    flow["edge_stokes"][..., 1] = v
    flow["cell_stokes"][..., 1] = v
    bc, _, _, _ = W.coupled_boundary(spec)
    n = 2 * (spec.nx + spec.ny) * spec.q
    inflow = (np.full(n, tep_at(pm, spec.y0) + 0.25), np.full(n, 0.2))
    t = transport(spec, flow, bc, inflow)
    cp = O.CoupledParams.make(LD=pm.derived["LD"], lithostatic=1, surface_y=SURFACE)
    r = t.coupled_rhs(pm, cp, flow, H, Cc)
    assert np.abs(r["rhsH"] - v * B).max() <= 1e-11 * v * B
    assert np.abs(r["rhsC"]).max() <= 1e-14


def test_lithostatic_inflow_through_vertical_faces():
    """Uniform H = 1.8, C = 0.2 in the solid-liquid region at every depth, where
    T depends on depth nonlinearly; uniform flow u in x enters on the left with
    H_in = 1.6.  Only the left column changes:
    rhs_H = -u/(dx dy) * sum_q w_q dy/2 (T(1.8, P(y_q)) - T(1.6, P(y_q))),
    the sum over the face's Gauss points at their own depths."""
    spec = W.MeshSpec(10, 12, seed=1)
    pm = litho_model()
    u = 0.05
    H = np.full((spec.ny, spec.nx), 1.8)
    Cc = np.full_like(H, 0.2)
    flow = still_flow(spec)
    flow["edge_stokes"][..., 0] = u
    flow["cell_stokes"][..., 0] = u
    bc = {"bottom": np.zeros(spec.nx, np.uint8), "right": np.ones(spec.ny, np.uint8),
          "top": np.zeros(spec.nx, np.uint8), "left": np.full(spec.ny, 2, np.uint8)}
    n = 2 * (spec.nx + spec.ny) * spec.q
    t = transport(spec, flow, bc, (np.full(n, 1.6), np.full(n, 0.2)))
    cp = O.CoupledParams.make(LD=pm.derived["LD"], lithostatic=1, surface_y=SURFACE)
    r = t.coupled_rhs(pm, cp, flow, H, Cc)
    g, w = np.polynomial.legendre.leggauss(spec.q)

    def T(h, y):
        return pm.evaluate(h, 0.2, pm.static_pressure(SURFACE - y))["TDp"]

    want = np.zeros_like(H)
    centre = np.zeros(spec.ny)
    for j in range(spec.ny):
        ys = spec.y0 + (j + 0.5 * (1 + g)) * spec.dy
        d = sum(wq * spec.dy / 2 * (T(1.8, y) - T(1.6, y)) for wq, y in zip(w, ys))
        want[j, 0] = -u * d / (spec.dx * spec.dy)
        yc = spec.y0 + (j + 0.5) * spec.dy
        centre[j] = -u * spec.dy * (T(1.8, yc) - T(1.6, yc)) / (spec.dx * spec.dy)
    scale = np.abs(want).max()
    assert np.abs(r["rhsH"] - want).max() <= 1e-11 * scale
    assert np.abs(r["rhsC"]).max() <= 1e-14
    # the Gauss points' own depths matter: the face-centre pressure misses by far more
    assert np.abs(centre - want[:, 0]).max() > 1e3 * 1e-11 * scale


def test_lithostatic_surface_below_the_mesh_top_is_rejected():
    spec = W.MeshSpec(10, 12, seed=1)
    pm = litho_model()
    H, Cc = eutectic_layer(spec, pm)
    flow = still_flow(spec)
    t = transport(spec, flow)
    cp = O.CoupledParams.make(LD=pm.derived["LD"], lithostatic=1, surface_y=0.99)
    with pytest.raises(O.OracleError, match="surface_y must be at or above the mesh top"):
        t.coupled_rhs(pm, cp, flow, H, Cc)
