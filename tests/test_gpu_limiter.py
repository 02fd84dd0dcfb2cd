"""Bound limiter on the GPU (FP64 and mixed sweeps) against the oracle, through the C ABI.

Bars: limiter factors within 1e-9 of the oracle's (they are ratios of
rounding-sized differences where the reconstruction barely leaves its bounds,
so their own last digits are not meaningful); limited traces within 1e-13 of
the field maximum; states after 10 SSP-RK3 steps with the limiter on every
stage within 1e-12 relative (max norm), every state inside H >= 0 and
0 <= C <= Xe; switching the limiter off restores the unlimited results
bitwise.  The mixed sweep's limiter: states and traces within 1e-6.
"""
import numpy as np
import pytest

import mantlepar_amd as M
import oracle as O
from mantlepar_amd import workloads as W

pytestmark = pytest.mark.gpu

XE = 0.4


def phase():
    return W.config_phase(hypereutectic=0)


def fields(spec, seed=7):
    """A step in H, grid-scale noise in C, and isolated cells exactly on the
    bounds 0 and Xe.  (Plateaus of equal averages are avoided here: where the
    reconstruction is mathematically equal to a bound that equals the mean, at
    some point of the set, theta is 0 or 1 depending on the sign of a
    rounding-sized residue, in the reference as on the device.)"""
    rng = np.random.default_rng(seed)
    H = W.enthalpy(spec)
    H[:, spec.nx // 3:] += 0.3                    # a step the reconstruction overshoots
    H[spec.ny // 2, 4] = 0.0
    Cc = XE * rng.random((spec.ny, spec.nx))      # grid-scale noise in [0, Xe]
    for k in range(6):
        Cc[5 + 5 * k, 3 + 7 * k] = XE
        Cc[3 + 6 * k, 40 - 6 * k] = 0.0
    return H, Cc


def interior_fields(spec, seed=7):
    """A step in H and grid-scale noise in C strictly inside (0, Xe).  A cell
    whose average sits exactly on 0 or Xe with non-constant neighbours can leave
    the admissible set in one stage at any dt (its limited traces may exceed the
    bound at the edges while averaging to it), which the reference's coupled
    step reports as a failed step; stepping tests therefore start inside."""
    rng = np.random.default_rng(seed)
    H = W.enthalpy(spec)
    H[:, spec.nx // 3:] += 0.3
    Cc = 0.05 + 0.3 * rng.random((spec.ny, spec.nx))
    return H, Cc


def make(spec, H, Cc, limiter=True, points=5, samples=4, extent=0.9, precision=M.PREC_FP64):
    vx, vy = W.face_velocity(spec)
    t = M.Transport(spec.nx, spec.ny, spec.x0, spec.x1, spec.y0, spec.y1, q=spec.q,
                    phase=phase(), precision=precision)
    t.upload(H, Cc)
    t.set_face_velocity(vx, vy)
    if limiter:
        t.set_limiter(True, points, samples, extent)
    o = O.Transport(W.vertex_coords(spec), spec.nx, spec.ny, vx, vy, q=spec.q)
    if limiter:
        o.set_limiter(True, points, samples, extent, XE)
    return t, o


def test_factors_and_traces_match_oracle():
    spec = W.MeshSpec(48, 40, seed=3)
    H, Cc = fields(spec)
    t, o = make(spec, H, Cc)
    gH, gC = t.limits()
    oH, oC = o.limits(H, Cc)
    assert (oH < 1).sum() > 50 and (oC < 1).sum() > 200 and (oC == 0).sum() > 0
    for g, o_, u in ((gH, oH, H), (gC, oC, Cc)):
        bad = np.argwhere(np.abs(g - o_) > 1e-9)
        assert len(bad) == 0, [(int(j), int(i), u[j, i], g[j, i], o_[j, i]) for j, i in bad[:6]]
    for field, u in ((0, H), (1, Cc)):
        _, _, tr = t.reconstruct(field)
        want = o.traces(u, field)
        assert np.abs(tr - want).max() <= 1e-13 * np.abs(u).max(), field
    t.close()


@pytest.mark.parametrize("q,points,samples,extent", [(5, 5, 4, 0.9), (3, 3, 6, 0.75)],
                         ids=["q5-5pts-4samples", "q3-3pts-6samples"])
def test_limited_steps_match_oracle(q, points, samples, extent):
    """Ten limited SSP-RK3 steps; the second case runs the sweep's generic edge-rule
    instance with other cell-point and normal-line sample tables."""
    spec = W.MeshSpec(48, 40, seed=3, q=q)
    H, Cc = interior_fields(spec)
    t, o = make(spec, H, Cc, points=points, samples=samples, extent=extent)
    dt = t.stable_dt(0.4)
    for chunk in range(2):
        t.step(5, dt)
        Hg, Cg = t.download()
        assert Hg.min() >= 0 and Cg.min() >= 0 and Cg.max() <= XE, chunk
    Ho, Co, _ = o.advance(H, Cc, 10, dt, 3)
    assert np.abs(Hg - Ho).max() <= 1e-12 * np.abs(Ho).max()
    assert np.abs(Cg - Co).max() <= 1e-12 * np.abs(Co).max()
    # the limiter was active: the unlimited run differs and reaches new extremes
    u, _ = make(spec, H, Cc, limiter=False)
    u.step(10, dt)
    Hu, Cu = u.download()
    assert np.abs(Hu - Hg).max() > 1e-6
    assert Cu.min() < Cg.min() or Cu.max() > Cg.max()
    t.close()
    u.close()


def test_off_again_is_bitwise_unlimited():
    spec = W.MeshSpec(48, 40, seed=3)
    H, Cc = fields(spec)
    a, _ = make(spec, H, Cc, limiter=False)
    b, _ = make(spec, H, Cc, limiter=True)
    b.set_limiter(False)
    dt = a.stable_dt(0.3)
    a.step(4, dt)
    b.step(4, dt)
    for x, y in zip(a.download(), b.download()):
        assert np.array_equal(x, y)
    with pytest.raises(M.MantleError) as ei:
        b.limits()
    assert ei.value.code == 3            # MP_ERR_STATE
    a.close()
    b.close()


def test_errors():
    spec = W.MeshSpec(48, 40, seed=3)
    H, Cc = fields(spec)
    bad = Cc.copy()
    bad[10, 10] = XE + 1e-3
    t, _ = make(spec, H, bad)
    with pytest.raises(M.MantleError, match="Inadmissible cell average") as ei:
        t.step(1, t.stable_dt(0.3))
    assert ei.value.code == 1            # MP_ERR_ARG
    for args in ((True, 0, 4, 0.9), (True, 5, 3, 0.9), (True, 5, 4, 0.0), (True, 5, 4, 1.5)):
        with pytest.raises(M.MantleError):
            t.set_limiter(*args)
    t.close()
    m, _ = make(spec, H, bad, precision=M.PREC_MIXED)
    with pytest.raises(M.MantleError, match="Inadmissible cell average") as ei:
        m.step(1, m.stable_dt(0.3))
    assert ei.value.code == 1
    m.close()


@pytest.mark.parametrize("q,points,samples,extent", [(5, 5, 4, 0.9), (3, 3, 6, 0.75)],
                         ids=["q5-5pts-4samples", "q3-3pts-6samples"])
def test_mixed_limited_steps_match_oracle(q, points, samples, extent):
    """The mixed sweep's limiter (FP32 extremes at the same points, bounds from
    the FP64 differences): ten limited SSP-RK3 steps within 1e-6 of the
    oracle's FP64 run relative to the field maximum, states inside the bounds,
    and its limiter factors within 1e-5 of the FP64 sweep's where they act."""
    spec = W.MeshSpec(48, 40, seed=3, q=q)
    H, Cc = interior_fields(spec)
    t, o = make(spec, H, Cc, points=points, samples=samples, extent=extent, precision=M.PREC_MIXED)
    dt = t.stable_dt(0.4)
    for chunk in range(2):
        t.step(5, dt)
        Hg, Cg = t.download()
        assert Hg.min() >= 0 and Cg.min() >= 0 and Cg.max() <= XE, chunk
    Ho, Co, _ = o.advance(H, Cc, 10, dt, 3)
    assert np.abs(Hg - Ho).max() <= 1e-6 * np.abs(Ho).max()
    assert np.abs(Cg - Co).max() <= 1e-6 * np.abs(Co).max()
    # the limiter acted: the unlimited mixed run is further from the oracle
    u, _ = make(spec, H, Cc, limiter=False, precision=M.PREC_MIXED)
    u.step(10, dt)
    Hu, _ = u.download()
    assert np.abs(Hu - Ho).max() > 1e2 * np.abs(Hg - Ho).max()
    t.close()
    u.close()


def test_mixed_limited_traces_and_evaluation():
    """Limited traces of the mixed handle (FP32 blend, limiter factor from it)
    within 1e-6 of the oracle's, on states with cells on the bounds."""
    spec = W.MeshSpec(48, 40, seed=3)
    H, Cc = fields(spec)
    t, o = make(spec, H, Cc, precision=M.PREC_MIXED)
    for field, u in ((0, H), (1, Cc)):
        _, _, tr = t.reconstruct(field)
        want = o.traces(u, field)
        assert np.abs(tr - want).max() <= 1e-6 * np.abs(u).max(), field
    t.close()
