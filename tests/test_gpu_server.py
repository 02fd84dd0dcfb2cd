"""Serving loop (mp_server_*): continuous batching with HBM-budget admission.

Bar: a run's outputs (H, C, T, phi, Cl, Cs, region) are bitwise the same
whether it ran alone or arrived among others at random times, in FP64 and in
mixed precision, and bitwise those of a single-mesh handle on the same
inputs.  Admission never exceeds the slot count or the HBM budget; a failing
run fails alone.
"""
import time

import numpy as np
import pytest

import mantlepar_amd as M
from mantlepar_amd import workloads as W

pytestmark = pytest.mark.gpu


def make_requests(count, seed=21, max_steps=24):
    rng = np.random.default_rng(seed)
    specs = W.batch_specs(count, seed=seed)
    reqs = []
    for spec in specs:
        # keep the test quick: shrink the largest meshes
        spec.ny = min(spec.ny, 160)
        H, Cc = W.enthalpy(spec), W.composition(spec)
        vx, vy = W.face_velocity(spec)
        reqs.append(dict(spec=spec, H=H, C=Cc, vx=vx, vy=vy,
                         nsteps=int(rng.integers(1, max_steps + 1))))
    return reqs


def submit(srv, r, cfl=0.4):
    s = r["spec"]
    return srv.submit(s.nx, s.ny, r["H"], r["C"], r["vx"], r["vy"], r["nsteps"], cfl,
                      s.x0, s.x1, s.y0, s.y1, q=s.q, bc=r.get("bc"), surface_y=r.get("surface"),
                      limiter=r.get("limiter"))


OUT = ("H", "C", "T", "phi", "Cl", "Cs", "region")


@pytest.mark.parametrize("prec", [M.PREC_FP64, M.PREC_MIXED], ids=["fp64", "mixed"])
def test_alone_equals_mixed(prec):
    reqs = make_requests(24)
    alone = []
    srv = M.Server(max_slots=8, phase=W.config_phase(), precision=prec, max_epoch_steps=4)
    for r in reqs:
        rid = submit(srv, r)
        res, out = srv.wait(rid)
        assert res["status"] == 0 and res["steps_done"] == r["nsteps"]
        alone.append((res, out))
    srv.close()

    # Poisson arrivals while earlier runs are on the device
    srv = M.Server(max_slots=8, phase=W.config_phase(), precision=prec, max_epoch_steps=4)
    rng = np.random.default_rng(5)
    ids = []
    for r in reqs:
        time.sleep(float(rng.exponential(0.002)))
        ids.append(submit(srv, r))
    srv.drain()
    st = srv.stats()
    assert st["completed"] == len(reqs) and st["failed"] == 0
    assert st["max_resident"] > 1        # runs did share launches
    assert st["max_resident"] <= 8
    for rid, (ra, oa) in zip(ids, alone):
        res, out = srv.wait(rid)
        for n in OUT:
            assert np.array_equal(out[n], oa[n]), n
        for n in ("dt", "total_H", "total_C", "boundary_H", "boundary_C"):
            assert res[n] == ra[n], n
    srv.close()

    # and equal to the single-mesh handle
    for r, (ra, oa) in list(zip(reqs, alone))[:4]:
        s = r["spec"]
        t = M.Transport(s.nx, s.ny, s.x0, s.x1, s.y0, s.y1, q=s.q, phase=W.config_phase(),
                        precision=prec)
        t.set_tile(12)                 # the server's rows per work item
        t.upload(r["H"], r["C"])
        t.set_face_velocity(r["vx"], r["vy"])
        rep = t.step(r["nsteps"], t.stable_dt(0.4))
        H, Cc = t.download()
        ph = t.phase_state()
        assert np.array_equal(H, oa["H"]) and np.array_equal(Cc, oa["C"])
        for n in ("T", "phi", "Cl", "Cs", "region"):
            assert np.array_equal(ph[n], oa[n]), n
        assert rep["total_H"] == ra["total_H"] and rep["boundary_H"] == ra["boundary_H"]
        t.close()


@pytest.mark.parametrize("prec", [M.PREC_FP64, M.PREC_MIXED], ids=["fp64", "mixed"])
def test_lithostatic_requests_equal_single_mesh(prec):
    """Requests with the lithostatic pressure model, mixed with constant-pressure
    ones, give the single-mesh handle's outputs bit for bit; a surface below the
    mesh top fails that request alone."""
    reqs = make_requests(8, seed=41, max_steps=10)
    for i, r in enumerate(reqs):
        if i % 2 == 0:
            r["surface"] = r["spec"].y1 + 0.1 * (i + 1)
    bad = dict(reqs[1], surface=0.5 * reqs[1]["spec"].y1)
    srv = M.Server(max_slots=4, phase=W.config_phase(), precision=prec, max_epoch_steps=4)
    ids = [submit(srv, r) for r in reqs]
    bad_id = submit(srv, bad)
    srv.drain()
    with pytest.raises(M.MantleError, match="surface_y must be at or above the mesh top"):
        srv.wait(bad_id)
    for r, rid in zip(reqs, ids):
        res, out = srv.wait(rid)
        s = r["spec"]
        t = M.Transport(s.nx, s.ny, s.x0, s.x1, s.y0, s.y1, q=s.q, phase=W.config_phase(),
                        precision=prec)
        t.set_tile(12)
        t.upload(r["H"], r["C"])
        t.set_face_velocity(r["vx"], r["vy"])
        if "surface" in r:
            t.set_lithostatic(r["surface"])
        t.step(r["nsteps"], t.stable_dt(0.4))
        H, Cc = t.download()
        ph = t.phase_state()
        assert np.array_equal(H, out["H"]) and np.array_equal(Cc, out["C"])
        for n in ("T", "phi", "Cl", "Cs", "region"):
            assert np.array_equal(ph[n], out[n]), n
        t.close()
    srv.close()


def test_budget_limits_residency():
    reqs = make_requests(10, seed=33)
    from_bytes = []
    for r in reqs:
        s = r["spec"]
        # generous upper bound of one slot: 10 states, 2 phase sets, faces, moments
        from_bytes.append(8 * s.nx * s.ny * 40 + 8 * 12 * s.q * (s.nx + 1) * (s.ny + 1))
    budget = int(2.2 * max(from_bytes))
    srv = M.Server(max_slots=64, phase=W.config_phase(), precision=M.PREC_MIXED,
                   hbm_budget_bytes=budget)
    ids = [submit(srv, r) for r in reqs]
    srv.drain()
    st = srv.stats()
    assert st["completed"] == len(reqs)
    assert st["hbm_peak"] <= budget
    assert st["hbm_in_use"] == 0
    for rid in ids:
        res, _ = srv.wait(rid)
        assert res["t_admit"] >= res["t_submit"] and res["t_done"] >= res["t_admit"]
    srv.close()


def test_oversize_request_rejected():
    srv = M.Server(max_slots=4, phase=W.config_phase(), hbm_budget_bytes=1 << 20)
    r = make_requests(1, seed=40)[0]
    with pytest.raises(M.MantleError) as ei:
        submit(srv, r)
    assert ei.value.code == 6      # MP_ERR_MEMORY
    srv.close()


def test_failed_run_is_isolated():
    reqs = make_requests(6, seed=44)
    bad = dict(reqs[0])
    s = bad["spec"]
    bad["vx"] = np.full((s.q, s.ny, s.nx + 1), -0.2)
    bad["vy"] = np.zeros((s.q, s.ny + 1, s.nx))
    bad["bc"] = (None, np.full(s.ny, M.BC_OUTFLOW, np.uint8), None, None)
    srv = M.Server(max_slots=8, phase=W.config_phase(), precision=M.PREC_FP64)
    ids = [submit(srv, bad)] + [submit(srv, r) for r in reqs[1:]]
    srv.drain()
    with pytest.raises(M.MantleError) as ei:
        srv.wait(ids[0])
    assert ei.value.code == 7      # MP_ERR_BACKFLOW
    for rid in ids[1:]:
        res, _ = srv.wait(rid)
        assert res["status"] == 0
    st = srv.stats()
    assert st["failed"] == 1 and st["completed"] == len(reqs) - 1
    srv.close()


def test_zero_step_run_returns_inputs():
    r = make_requests(1, seed=50)[0]
    r["nsteps"] = 0
    srv = M.Server(max_slots=2, phase=W.config_phase())
    rid = submit(srv, r)
    res, out = srv.wait(rid)
    assert res["steps_done"] == 0
    assert np.array_equal(out["H"], r["H"]) and np.array_equal(out["C"], r["C"])
    ph = M.EutecticModel(W.config_phase()).evaluate(r["H"], r["C"])
    assert np.array_equal(out["region"].ravel(), ph["region"])
    srv.close()


def test_wait_timeout_reports_pending():
    reqs = make_requests(1, seed=60, max_steps=1)
    r = reqs[0]
    r["nsteps"] = 400
    srv = M.Server(max_slots=2, phase=W.config_phase())
    rid = submit(srv, r)
    with pytest.raises(M.MantleError) as ei:
        srv.wait(rid, timeout=0.0)
    assert ei.value.code == 9      # MP_ERR_PENDING
    res, _ = srv.wait(rid)
    assert res["steps_done"] == 400
    srv.close()


def test_nonfinite_inputs_rejected_alone():
    """The stagers check inputs as the single-mesh upload does: a NaN average
    or velocity fails its own request with the reference's message."""
    reqs = make_requests(5, seed=71)
    bad_h, bad_v = dict(reqs[0]), dict(reqs[1])
    bad_h["H"] = reqs[0]["H"].copy()
    bad_h["H"][1, 2] = np.nan
    bad_v["vy"] = np.array(reqs[1]["vy"], dtype=float, copy=True)
    bad_v["vy"].flat[7] = np.inf
    srv = M.Server(max_slots=4, phase=W.config_phase(), precision=M.PREC_MIXED)
    ids = [submit(srv, bad_h), submit(srv, bad_v)] + [submit(srv, r) for r in reqs[2:]]
    srv.drain()
    for rid, msg in ((ids[0], "Cell averages must be finite"), (ids[1], "Edge velocity must be finite")):
        with pytest.raises(M.MantleError) as ei:
            srv.wait(rid)
        assert ei.value.code == 1 and msg in str(ei.value)
    for rid in ids[2:]:
        res, _ = srv.wait(rid)
        assert res["status"] == 0
    srv.close()


def test_progress_curve_accounts_every_step():
    reqs = make_requests(16, seed=72, max_steps=30)
    srv = M.Server(max_slots=6, phase=W.config_phase(), precision=M.PREC_MIXED, max_epoch_steps=4)
    ids = [submit(srv, r) for r in reqs]
    srv.drain()
    st = srv.stats()
    t, w = srv.progress()
    for rid in ids:
        srv.wait(rid)
    srv.close()
    assert len(t) == st["epochs"]
    assert np.all(np.diff(t) >= 0) and np.all(np.diff(w) > 0)
    assert w[-1] == st["cell_steps"] == sum(r["spec"].cells * r["nsteps"] for r in reqs)


@pytest.mark.parametrize("prec", [M.PREC_FP64, M.PREC_MIXED], ids=["fp64", "mixed"])
def test_limited_requests_equal_single_mesh(prec):
    """Requests with the bound limiter (default and other tables) among requests
    without it: each result is bitwise a single-mesh handle's with the same
    limiter setting, alone and in the mix."""
    ph = W.config_phase(hypereutectic=0)
    reqs = make_requests(6, seed=81, max_steps=9)
    rng = np.random.default_rng(3)
    for k, r in enumerate(reqs):
        spec = r["spec"]
        r["H"] = r["H"].copy()
        r["H"][:, spec.nx // 3:] += 0.3
        r["C"] = 0.05 + 0.3 * rng.random((spec.ny, spec.nx))
        if k % 3 == 0:
            r["limiter"] = (5, 4, 0.9)
        elif k % 3 == 1 and spec.q == 5:
            r["limiter"] = (3, 6, 0.75)
    srv = M.Server(max_slots=8, phase=ph, precision=prec, max_epoch_steps=4)
    ids = [submit(srv, r) for r in reqs]
    srv.drain()
    outs = [srv.wait(i) for i in ids]
    srv.close()
    for r, (res, out) in zip(reqs, outs):
        s = r["spec"]
        t = M.Transport(s.nx, s.ny, s.x0, s.x1, s.y0, s.y1, q=s.q, phase=ph, precision=prec)
        t.set_tile(12)
        t.upload(r["H"], r["C"])
        t.set_face_velocity(r["vx"], r["vy"])
        if r.get("limiter"):
            t.set_limiter(True, *r["limiter"])
        rep = t.step(r["nsteps"], t.stable_dt(0.4))
        H, Cc = t.download()
        assert np.array_equal(H, out["H"]) and np.array_equal(Cc, out["C"]), r.get("limiter")
        ph_s = t.phase_state()
        for n in ("T", "phi", "region"):
            assert np.array_equal(ph_s[n], out[n]), n
        assert rep["total_H"] == res["total_H"] and rep["boundary_C"] == res["boundary_C"]
        t.close()


@pytest.mark.parametrize("prec", [M.PREC_FP64, M.PREC_MIXED], ids=["fp64", "mixed"])
def test_diffusion_requests_equal_single_mesh(prec):
    """Requests with thermal diffusion (one with thermal boundary data) among
    requests without it: each result is bitwise a single-mesh handle's."""
    reqs = make_requests(5, seed=91, max_steps=9)
    for k, r in enumerate(reqs):
        s = r["spec"]
        if k == 0:
            q = s.q
            types = np.concatenate([np.ones(s.nx * q), np.zeros(s.ny * q), np.zeros(s.nx * q),
                                    np.ones(s.ny * q)]).astype(np.uint8)
            vals = np.concatenate([np.full(s.nx * q, 0.9), np.zeros(s.ny * q), np.zeros(s.nx * q),
                                   np.full(s.ny * q, 0.7)])
            r["diffusion"] = (2e-4, types, vals)
        elif k == 2:
            r["diffusion"] = (1e-4, None, None)
    srv = M.Server(max_slots=8, phase=W.config_phase(), precision=prec, max_epoch_steps=4)
    ids = []
    for r in reqs:
        s = r["spec"]
        ids.append(srv.submit(s.nx, s.ny, r["H"], r["C"], r["vx"], r["vy"], r["nsteps"], 0.4,
                              s.x0, s.x1, s.y0, s.y1, q=s.q, diffusion=r.get("diffusion")))
    srv.drain()
    outs = [srv.wait(i) for i in ids]
    srv.close()
    for r, (res, out) in zip(reqs, outs):
        s = r["spec"]
        t = M.Transport(s.nx, s.ny, s.x0, s.x1, s.y0, s.y1, q=s.q, phase=W.config_phase(), precision=prec)
        t.set_tile(12)
        t.upload(r["H"], r["C"])
        t.set_face_velocity(r["vx"], r["vy"])
        if r.get("diffusion"):
            kappa, types, vals = r["diffusion"]
            t.set_diffusion(kappa, 4, 0.9)
            if types is not None:
                t.set_thermal_boundary(types, vals)
        dt = t.stable_dt(0.4)
        assert dt == res["dt"]
        rep = t.step(r["nsteps"], dt)
        H, Cc = t.download()
        assert np.array_equal(H, out["H"]) and np.array_equal(Cc, out["C"])
        assert rep["total_H"] == res["total_H"] and rep["boundary_H"] == res["boundary_H"]
        t.close()
