"""Face flux records of the mixed sweep and the LF speed modes, against the oracle.

Under the local LF speed the mixed sweep reads each face's flux moments in a
16-byte one-sided form: where no point of the face changes sign, one side's
moments are exactly zero.  A face whose points change sign, and every face
under the global LF speed, is read from its full 32-byte record.  These tests
cover faces of both kinds side by side in one wave and the global speed on the
single-mesh and batched handles.

Bars (as in test_gpu_transport.py): FP64 within 1e-12 of the oracle relative
to each field's maximum (RHS and 3 SSP-RK3 steps); mixed within 1e-6 (RHS) and
1e-5 (3 steps); a batch under the global speed equals the single-mesh handle
bit for bit.
"""
import numpy as np
import pytest

import mantlepar_amd as M
import oracle as O
from mantlepar_amd import workloads as W

pytestmark = pytest.mark.gpu

PRECS = [M.PREC_FP64, M.PREC_MIXED]


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def sign_changing_flow(spec, amp=0.5, seed=4):
    """The workload flow plus per-point noise on interior faces, so that about
    half of the faces have points of both signs; walls stay closed."""
    vx, vy = W.face_velocity(spec)
    s = amp * max(np.abs(vx).max(), np.abs(vy).max())
    rng = np.random.default_rng(seed)
    vx = vx.copy()
    vy = vy.copy()
    vx[:, :, 1:-1] += s * rng.uniform(-1, 1, vx[:, :, 1:-1].shape)
    vy[:, 1:-1, :] += s * rng.uniform(-1, 1, vy[:, 1:-1, :].shape)
    return vx, vy


def changing_fraction(v):
    return float(((v > 0).any(0) & (v < 0).any(0)).mean())


def pair(spec, prec, vel, weno=None, lf=(0, 0.0)):
    H, Cc = W.enthalpy(spec), W.composition(spec)
    vx, vy = vel
    t = M.Transport(spec.nx, spec.ny, spec.x0, spec.x1, spec.y0, spec.y1, q=spec.q,
                    weno=weno, phase=W.config_phase(), precision=prec)
    t.upload(H, Cc)
    t.set_face_velocity(vx, vy)
    o = O.Transport(W.vertex_coords(spec), spec.nx, spec.ny, vx, vy, q=spec.q, table_mode=0)
    o.set_lf(*lf)
    return t, o, H, Cc


def check(t, o, H, Cc, prec):
    tol = 1e-12 if prec == M.PREC_FP64 else 1e-6
    LH, LC, _, _ = t.rhs()
    oLH, oLC, _, _ = o.rhs(H, Cc)
    assert rel(LH, oLH) <= tol and rel(LC, oLC) <= tol, (rel(LH, oLH), rel(LC, oLC))
    dt = t.stable_dt(0.4)
    t.step(3, dt)
    Hg, Cg = t.download()
    Ho, Co, _ = o.advance(H, Cc, 3, dt, 3)
    run_tol = 1e-12 if prec == M.PREC_FP64 else 1e-5
    assert rel(Hg, Ho) <= run_tol and rel(Cg, Co) <= run_tol, (rel(Hg, Ho), rel(Cg, Co))


@pytest.mark.parametrize("prec", PRECS, ids=["fp64", "mixed"])
def test_faces_whose_points_change_sign(prec):
    spec = W.MeshSpec(140, 40, seed=21)
    vel = sign_changing_flow(spec)
    for v in (vel[0][:, :, 1:-1], vel[1][:, 1:-1, :]):
        assert 0.1 < changing_fraction(v) < 0.9
    t, o, H, Cc = pair(spec, prec, vel)
    check(t, o, H, Cc, prec)
    t.close()


@pytest.mark.parametrize("prec", PRECS, ids=["fp64", "mixed"])
def test_global_lf_speed(prec):
    spec = W.MeshSpec(140, 40, seed=22)
    vel = W.face_velocity(spec)
    a = 1.3 * max(np.abs(vel[0]).max(), np.abs(vel[1]).max())
    weno = M.default_weno(lf_mode=1, global_speed=a)
    t, o, H, Cc = pair(spec, prec, vel, weno=weno, lf=(1, a))
    check(t, o, H, Cc, prec)
    t.close()
    slow = M.default_weno(lf_mode=1, global_speed=0.5 * a)
    t = M.Transport(spec.nx, spec.ny, spec.x0, spec.x1, spec.y0, spec.y1, q=spec.q,
                    weno=slow, phase=W.config_phase(), precision=prec)
    with pytest.raises(M.MantleError, match="Global LF speed is smaller than the flow normal speed"):
        t.set_face_velocity(*vel)
    t.close()


@pytest.mark.parametrize("prec", PRECS, ids=["fp64", "mixed"])
def test_batch_global_lf_matches_single_mesh_bitwise(prec):
    specs = [W.MeshSpec(20, 20, seed=31), W.MeshSpec(130, 97, seed=32), W.MeshSpec(3, 7, seed=33),
             W.MeshSpec(200, 33, seed=34, psi_amplitude=0.6)]
    vels = [W.face_velocity(s) for s in specs]
    a = 1.3 * max(max(np.abs(vx).max(), np.abs(vy).max()) for vx, vy in vels)
    weno = M.default_weno(lf_mode=1, global_speed=a)
    b = M.Batch(8, weno=weno, phase=W.config_phase(), precision=prec)
    ids = []
    for spec, (vx, vy) in zip(specs, vels):
        k = b.add_mesh(spec.nx, spec.ny, spec.x0, spec.x1, spec.y0, spec.y1, q=spec.q)
        b.upload(k, W.enthalpy(spec), W.composition(spec))
        b.set_face_velocity(k, vx, vy)
        b.set_dt(k, b.stable_dt(k, 0.4))
        ids.append(k)
    b.set_schedule(16, 0)
    b.step(4)
    for k, spec, vel in zip(ids, specs, vels):
        t, _, _, _ = pair(spec, prec, vel, weno=weno, lf=(1, a))
        t.set_tile(16)
        t.step(4, t.stable_dt(0.4))
        H1, C1 = t.download()
        ph = t.phase_state()
        t.close()
        H, Cc = b.download(k)
        assert np.array_equal(H, H1) and np.array_equal(Cc, C1)
        bp = b.phase_state(k)
        for n in ("T", "phi", "Cl", "Cs", "region"):
            assert np.array_equal(bp[n], ph[n]), n
    b.close()
