"""mp_transport_step_noalloc against mp_transport_step, through the C ABI.

The allocation-free step keeps its work memory in the library's static device
scratch; its states, phase outputs and report must be bitwise those of the
allocating step on the same inputs, on consecutive calls that reuse the
scratch with a smaller mesh the second time, in both precisions.
"""
import numpy as np
import pytest

import mantlepar_amd as M
from mantlepar_amd import workloads as W
from mantlepar_amd._lib import MantleError

pytestmark = pytest.mark.gpu

REPORT_KEYS = ("dt", "courant", "total_H", "total_C", "boundary_H", "boundary_C", "error_bits",
               "steps_done")


def make(spec, prec, vel=None, bc=None, inflow=None):
    vx, vy = vel if vel is not None else W.face_velocity(spec)
    t = M.Transport(spec.nx, spec.ny, spec.x0, spec.x1, spec.y0, spec.y1, q=spec.q,
                    phase=W.config_phase(), precision=prec)
    t.upload(W.enthalpy(spec), W.composition(spec))
    t.set_face_velocity(vx, vy)
    if bc is not None:
        t.set_boundary(bc.get("bottom"), bc.get("right"), bc.get("top"), bc.get("left"))
    if inflow is not None:
        t.set_inflow(*inflow)
    return t


def same_run(spec, prec, nsteps, calls=2, vel=None, bc=None, inflow=None):
    a = make(spec, prec, vel, bc, inflow)
    b = make(spec, prec, vel, bc, inflow)
    dt = a.stable_dt(0.4)
    for _ in range(calls):
        ra = a.step(nsteps, dt)
        rb = b.step_noalloc(nsteps, dt)
        for k in REPORT_KEYS:
            assert ra[k] == rb[k], (spec, prec, k, ra[k], rb[k])
        Ha, Ca = a.download()
        Hb, Cb = b.download()
        assert np.array_equal(Ha, Hb) and np.array_equal(Ca, Cb)
        pa, pb = a.phase_state(), b.phase_state()
        for k in pa:
            assert np.array_equal(pa[k], pb[k]), k
    a.close()
    b.close()


@pytest.mark.parametrize("prec", [M.PREC_FP64, M.PREC_MIXED], ids=["fp64", "mixed"])
def test_noalloc_step_matches_step_then_smaller_mesh(prec):
    print(f"noalloc: 300x260 then 40x30, precision {prec}", flush=True)
    same_run(W.MeshSpec(300, 260, seed=11), prec, nsteps=6)
    # second use of the scratch: a smaller mesh, fewer items and steps
    same_run(W.MeshSpec(40, 30, seed=12), prec, nsteps=3)


def test_noalloc_open_boundaries_and_schemes():
    spec = W.MeshSpec(128, 256, seed=13)
    vel = W.throughflow_velocity(spec)
    bc, inflow = W.throughflow_boundary(spec)
    print("noalloc: 128x256 through-flow, SSPRK2 and FE", flush=True)
    for prec in (M.PREC_FP64, M.PREC_MIXED):
        a = make(spec, prec, vel, bc, inflow)
        b = make(spec, prec, vel, bc, inflow)
        dt = a.stable_dt(0.3)
        for scheme in (M.SCHEME_SSPRK2, M.SCHEME_FE):
            ra = a.step(4, dt, scheme=scheme)
            rb = b.step_noalloc(4, dt, scheme=scheme)
            assert all(ra[k] == rb[k] for k in REPORT_KEYS)
            assert all(np.array_equal(x, y) for x, y in zip(a.download(), b.download()))
        a.close()
        b.close()


def test_noalloc_rejects_what_exceeds_the_scratch():
    spec = W.MeshSpec(64, 64, seed=14)
    t = make(spec, M.PREC_FP64)
    H0, C0 = t.download()
    dt = t.stable_dt(0.4)
    with pytest.raises(MantleError) as e:
        t.step_noalloc(2_000_000, dt)      # 64 MB of per-step totals
    assert e.value.code == 6 and "scratch" in str(e.value)
    H1, C1 = t.download()
    assert np.array_equal(H0, H1) and np.array_equal(C0, C1)
    # the handle still steps, with and without allocation
    t.step_noalloc(2, dt)
    t.step(1, dt)
    t.close()
