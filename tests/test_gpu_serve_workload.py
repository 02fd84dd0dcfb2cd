"""The serving workload itself: requests drawn from the batch size law (nx in
{40, 60, ..., 400}, ny log-uniform in [5, 800], full size), 20-200 SSP-RK3
steps each, Poisson arrivals, a 200 GB HBM budget, FP64 and mixed precision.

Bars:
* every request's outputs (H, C, T, phi, Cl, Cs, region) and step totals are
  bitwise the same run alone and run among the Poisson arrivals;
* requests small enough for the CPU oracle (cells x steps <= 3e5) match it:
  H and C within 1e-12 (FP64) and 1e-6 (mixed) relative to the field's
  maximum, except where the oracle's own FP64 run amplifies a difference of
  the size the path makes: there the bound is k times the spread between the
  oracle's runs from (H, C) and from (H, C)(1 + d*xi): d = 1e-15 and k = 10
  for FP64; d = 2^-24 (FP32 resolution) and k = 30 for mixed, which makes
  FP32-sized differences at every stage, not only in its initial state.
  Tall-cell meshes amplify such differences by orders of magnitude
  (test_oracle_sensitivity.py);
* the phase outputs are the oracle's EutecticModel of the returned state:
  regions equal, T and phi within 1e-13.
The per-run step loop is the reference's Run loop (src/couple/evolve.cpp:
449-507) with SSP-RK3 stages at a fixed dt = 0.4 / max rate.
"""
import time

import numpy as np
import pytest

import mantlepar_amd as M
import oracle as O
from mantlepar_amd import workloads as W

pytestmark = pytest.mark.gpu

COUNT = 40
OUT = ("H", "C", "T", "phi", "Cl", "Cs", "region")
ORACLE_WORK = 3e5


def requests(seed=4):
    rng = np.random.default_rng(seed)
    specs = W.batch_specs(COUNT, seed=seed)
    reqs = []
    for spec in specs:
        H, Cc = W.enthalpy(spec), W.composition(spec)
        vx, vy = W.face_velocity(spec)
        reqs.append(dict(spec=spec, H=H, C=Cc, vx=vx, vy=vy, nsteps=int(rng.integers(20, 201))))
    return reqs


REQS = requests()


def submit(srv, r):
    s = r["spec"]
    return srv.submit(s.nx, s.ny, r["H"], r["C"], r["vx"], r["vy"], r["nsteps"], 0.4,
                      s.x0, s.x1, s.y0, s.y1, q=s.q)


def oracle_run(r, dt, perturb=0.0):
    s = r["spec"]
    H, Cc = r["H"], r["C"]
    if perturb:
        rng = np.random.default_rng(s.seed)
        H = H * (1.0 + perturb * rng.standard_normal(H.shape))
        Cc = Cc * (1.0 + perturb * rng.standard_normal(Cc.shape))
    o = O.Transport(W.vertex_coords(s), s.nx, s.ny, r["vx"], r["vy"], q=s.q, table_mode=1,
                    threads=8)
    Ho, Co, _ = o.advance(H, Cc, r["nsteps"], dt, 3)
    return Ho.reshape(s.ny, s.nx), Co.reshape(s.ny, s.nx)


def rel(a, b):
    return max(np.abs(a[0] - b[0]).max() / np.abs(b[0]).max(),
               np.abs(a[1] - b[1]).max() / np.abs(b[1]).max())


_oracle = {}


def oracle_of(i, dt):
    """Oracle state of request i and its spreads under relative perturbations
    of 1e-15 and 2^-24 (computed once)."""
    if i not in _oracle:
        ref = oracle_run(REQS[i], dt)
        _oracle[i] = (ref, {d: rel(oracle_run(REQS[i], dt, perturb=d), ref) for d in (1e-15, 2.0 ** -24)})
    return _oracle[i]


def test_workload_shape():
    small = [r for r in REQS if r["spec"].cells * r["nsteps"] <= ORACLE_WORK]
    assert max(r["spec"].ny for r in REQS) > 400          # full-height meshes, not shrunk
    assert min(r["nsteps"] for r in REQS) >= 20 and max(r["nsteps"] for r in REQS) <= 200
    assert len(small) >= 5, len(small)


@pytest.mark.parametrize("prec", [M.PREC_FP64, M.PREC_MIXED], ids=["fp64", "mixed"])
def test_serving_workload_alone_equals_mixed_and_matches_oracle(prec):
    phase = W.config_phase()
    print(f"serve workload: {COUNT} requests alone, precision {prec}", flush=True)
    alone = []
    srv = M.Server(phase=phase, precision=prec, hbm_budget_bytes=200 * 10**9)
    for r in REQS:
        res, out = srv.wait(submit(srv, r))
        assert res["status"] == 0 and res["steps_done"] == r["nsteps"]
        alone.append((res, out))
    srv.close()

    print("serve workload: Poisson arrivals at 2000/s", flush=True)
    srv = M.Server(phase=phase, precision=prec, hbm_budget_bytes=200 * 10**9)
    rng = np.random.default_rng(9)
    ids = []
    for r in REQS:
        time.sleep(float(rng.exponential(1.0 / 2000.0)))
        ids.append(submit(srv, r))
    srv.drain()
    st = srv.stats()
    assert st["completed"] == COUNT and st["failed"] == 0
    assert st["max_resident"] > 1 and st["hbm_peak"] <= 200 * 10**9
    for rid, (ra, oa) in zip(ids, alone):
        res, out = srv.wait(rid)
        for n in OUT:
            assert np.array_equal(out[n], oa[n]), n
        for n in ("dt", "total_H", "total_C", "boundary_H", "boundary_C", "steps_done"):
            assert res[n] == ra[n], n
    srv.close()

    print("serve workload: oracle comparisons", flush=True)
    pm = O.PhaseModel(O.Material.defaults(Tm0=1600.0, Te0=600.0, cp=1200.0, L=1.2e6, Xe=0.4))
    tol, size, k = (1e-12, 1e-15, 10.0) if prec == M.PREC_FP64 else (1e-6, 2.0 ** -24, 30.0)
    checked = 0
    for i, (r, (ra, oa)) in enumerate(zip(REQS, alone)):
        # the phase outputs are the phase of the returned state
        T, phi, _, _, reg = pm.batch(oa["H"].ravel(), oa["C"].ravel(), 0.0, hyper=True)
        assert np.array_equal(oa["region"].ravel(), reg)
        assert np.abs(oa["T"].ravel() - T).max() <= 1e-13
        assert np.abs(oa["phi"].ravel() - phi).max() <= 1e-13
        if r["spec"].cells * r["nsteps"] > ORACLE_WORK:
            continue
        ref, spreads = oracle_of(i, ra["dt"])
        err = rel((oa["H"], oa["C"]), ref)
        spread = spreads[size]
        print(f"  request {i}: {r['spec'].nx}x{r['spec'].ny}, {r['nsteps']} steps: error {err:.2e}, "
              f"oracle spread {spread:.2e}", flush=True)
        assert err <= max(tol, k * spread), (i, r["spec"].nx, r["spec"].ny, err, spread)
        checked += 1
    assert checked >= 5
