"""One launch per step against one launch per stage (mixed precision), bitwise.

The per-step launch runs all stages of an SSP-RK step in one grid, each work
item waiting for the items whose output rows it reads; it must give states,
phase outputs and step totals bit for bit equal to one launch per stage, on
meshes with and without tail items, with open boundaries, diffusion and the
global LF speed.  Its bounded wait is exercised from the host: the stall check
waits for a completion that is never published, and the step must fail with
MP_ERR_DEVICE and leave the state as it was.
"""
import numpy as np
import pytest

import mantlepar_amd as M
from mantlepar_amd import workloads as W
from mantlepar_amd._lib import LAUNCH_STALL_CHECK, MantleError, check, lib

pytestmark = pytest.mark.gpu

REPORT_KEYS = ("dt", "total_H", "total_C", "boundary_H", "boundary_C", "error_bits", "steps_done")


def make(spec, per_step, vel=None, bc=None, inflow=None, weno=None, kappa=0.0, rows=None):
    vx, vy = vel if vel is not None else W.face_velocity(spec)
    t = M.Transport(spec.nx, spec.ny, spec.x0, spec.x1, spec.y0, spec.y1, q=spec.q, weno=weno,
                    phase=W.config_phase(), precision=M.PREC_MIXED)
    t.set_launch(per_step)
    if rows:
        t.set_tile(rows)
    t.upload(W.enthalpy(spec), W.composition(spec))
    t.set_face_velocity(vx, vy)
    if bc is not None:
        t.set_boundary(bc.get("bottom"), bc.get("right"), bc.get("top"), bc.get("left"))
    if inflow is not None:
        t.set_inflow(*inflow)
    if kappa:
        t.set_diffusion(kappa, 4, 0.9)
    return t


def compare(spec, nsteps=5, schemes=(M.SCHEME_SSPRK3, M.SCHEME_SSPRK2), **kw):
    print(f"step launch: {spec.nx}x{spec.ny} q{spec.q} {sorted(kw)}", flush=True)
    a = make(spec, False, **kw)
    b = make(spec, True, **kw)
    dt = a.stable_dt(0.4)
    for scheme in schemes:
        ra = a.step(nsteps, dt, scheme=scheme)
        rb = b.step(nsteps, dt, scheme=scheme)
        for k in REPORT_KEYS:
            assert ra[k] == rb[k], (k, ra[k], rb[k])
        Ha, Ca = a.download()
        Hb, Cb = b.download()
        assert np.array_equal(Ha, Hb) and np.array_equal(Ca, Cb)
        pa, pb = a.phase_state(), b.phase_state()
        for k in pa:
            assert np.array_equal(pa[k], pb[k]), k
    a.close()
    b.close()


@pytest.mark.parametrize("shape", [(300, 260, 5), (64, 5, 5), (2, 2, 5), (17, 9, 3), (130, 97, 8),
                                   (1000, 700, 5)], ids=lambda s: f"{s[0]}x{s[1]}q{s[2]}")
def test_per_step_launch_matches_per_stage(shape):
    nx, ny, q = shape
    compare(W.MeshSpec(nx, ny, q=q, seed=nx + ny))


@pytest.mark.parametrize("rows", [4, 5, 12, 16])
def test_per_step_launch_with_tail_items_and_short_bands(rows):
    # 2048 x 3400 gives every wave slot several items, so the layout ends on
    # 4-row tail items; 3400 is not a multiple of these heights, so the last
    # band below the tail can be one row tall
    compare(W.MeshSpec(2048, 3400, seed=21), nsteps=3, schemes=(M.SCHEME_SSPRK3,), rows=rows)


def test_per_step_launch_open_boundaries_diffusion_and_global_lf():
    spec = W.MeshSpec(256, 384, seed=22)
    vel = W.throughflow_velocity(spec)
    bc, inflow = W.throughflow_boundary(spec)
    compare(spec, vel=vel, bc=bc, inflow=inflow)
    compare(spec, vel=vel, bc=bc, inflow=inflow, kappa=1e-4)
    compare(W.MeshSpec(200, 150, seed=23), weno=M.default_weno(lf_mode=1, global_speed=1.0))


def test_stalled_wait_fails_the_step_and_keeps_the_state():
    spec = W.MeshSpec(600, 400, seed=24)
    t = make(spec, True)
    ref = make(spec, False)
    dt = t.stable_dt(0.4)
    H0, C0 = t.download()
    print("step launch: stall check on 600x400", flush=True)
    check(lib().mp_transport_set_launch(t._h, LAUNCH_STALL_CHECK))
    with pytest.raises(MantleError) as e:
        t.step(2, dt)
    assert e.value.code == 5 and "not completed in time" in str(e.value)
    H1, C1 = t.download()
    assert np.array_equal(H0, H1) and np.array_equal(C0, C1)
    # back to the normal launch, the handle steps as one that never stalled
    t.set_launch(True)
    ra, rb = t.step(3, dt), ref.step(3, dt)
    assert all(ra[k] == rb[k] for k in REPORT_KEYS)
    assert all(np.array_equal(x, y) for x, y in zip(t.download(), ref.download()))
    t.close()
    ref.close()
