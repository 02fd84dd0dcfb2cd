"""Batched coupled runs (mp_coupled_run_batch) against the same handles run
alone (mp_coupled_run), through the C ABI.

Bar: bitwise.  Every handle of a batch must end exactly as its own run ends --
status, time, number of accepted steps, every step record (dt, Courant number,
balances, boundary and source totals, retries) and the state -- because the
batched kernels run the single-mesh bodies and reductions on each handle's own
buffers, and the host takes the same decisions.  The batch mixes mesh sizes,
a workload whose steps need retries, the lithostatic pressure, the global LF
speed, a handle that stops at its end time and one whose fixed step fails and
rolls back while the others go on.  One handle is also checked against the
oracle (the single-handle loop's own bar, 1e-11).
"""
import numpy as np
import pytest

import mantlepar_amd as M
from mantlepar_amd import workloads as W

from test_gpu_coupled import options_pair, pair, rel

pytestmark = pytest.mark.gpu


def retry_case(spec):
    _, Cc = W.coupled_state(spec)
    H = np.full_like(Cc, 0.599)
    return pair(spec, darcy=0.5, H_in=0.9, kappa=0.0, state=(H, Cc))


# (name, builder, time controls); every case runs NSTEPS steps alone and in
# the batch, so whole runs are compared
NSTEPS = 5
CASES = [
    ("retries 16x14", lambda: retry_case(W.MeshSpec(16, 14, seed=2)), dict(cfl=0.5, initial=1.0, maximum=1.0)),
    ("general 24x20", lambda: pair(W.MeshSpec(24, 20, seed=2)), dict(cfl=0.5, initial=1e-3, maximum=1.0)),
    ("lithostatic 20x18", lambda: pair(W.MeshSpec(20, 18, seed=5), surface=1.0),
     dict(cfl=0.5, initial=1e-3, maximum=1.0)),
    ("global LF 24x20", lambda: options_pair(W.MeshSpec(24, 20, seed=2), lf_mode=1, global_speed=2.0),
     dict(cfl=0.5, initial=1e-3, maximum=1.0)),
    ("general 64x48", lambda: pair(W.MeshSpec(64, 48, seed=7)), dict(cfl=0.4, initial=1e-3, maximum=1.0)),
    ("end time 30x12", lambda: pair(W.MeshSpec(30, 12, seed=3)),
     dict(cfl=0.5, initial=1e-4, maximum=1e-4, end=2.5e-4)),
]


def alone(case, nsteps=NSTEPS):
    """Status, time, step records and final state of the case run alone."""
    g, *_ = case[1]()
    try:
        t, rec = g.run(nsteps, **case[2])
        st = 0
    except M.MantleError as e:
        t, rec, st = None, None, e.code
    H, Cc = g.download()
    g.close()
    return st, t, rec, H, Cc


def compare(got, want, name):
    st, t, rec, H, Cc = want
    gst, gt, grec, gH, gC = got
    assert gst == st, name
    if st == 0:
        assert gt == t, name
        assert grec == rec, name
    assert np.array_equal(gH, H) and np.array_equal(gC, Cc), name


def run_batched(cases, nsteps=NSTEPS):
    built = [c[1]() for c in cases]
    handles = [b[0] for b in built]
    res = M.coupled_run_batch(handles, nsteps, [c[2] for c in cases])
    out = []
    for k in range(len(cases)):
        H, Cc = handles[k].download()
        out.append((res[k][0], res[k][1], res[k][2], H, Cc))
    return out, built


def test_batch_equals_each_handle_alone():
    print("coupled batch: six handles, alone then batched", flush=True)
    want = [alone(c) for c in CASES]
    assert all(w[0] == 0 for w in want)
    assert sum(s["retries"] for s in want[0][2]) >= 1                  # the retry workload retried
    assert want[5][1] == 2.5e-4 and len(want[5][2]) < NSTEPS            # stopped at its end time
    got, built = run_batched(CASES)
    for case, g, w in zip(CASES, got, want):
        compare(g, w, case[0])
    # the 64x48 handle against the oracle
    o, pm, cp, flow, H0, C0 = built[4][1:]
    Ho, Co, to, so = o.coupled_run(pm, cp, flow, H0, C0, NSTEPS, cfl=0.4, initial=1e-3, maximum=1.0)
    assert [s["retries"] for s in got[4][2]] == [s["retries"] for s in so]
    assert abs(got[4][1] - to) <= 1e-12 * to
    assert rel(got[4][3], Ho) <= 1e-11 and rel(got[4][4], Co) <= 1e-11
    for b in built:
        b[0].close()


def test_failed_handle_rolls_back_alone():
    """A fixed step that the retry workload cannot take fails in the batch as
    it fails alone (same status, state kept); the other handles advance as
    they do alone."""
    print("coupled batch: one failing handle", flush=True)
    spec = W.MeshSpec(16, 14, seed=2)
    g, *_ = retry_case(spec)
    dt = 0.5 / g.rhs()["maximum_rate"]
    g.close()
    bad = ("fixed at the bound", lambda: retry_case(spec), dict(cfl=0.5, initial=dt, maximum=dt, fixed=True))
    cases = [CASES[1], bad, CASES[3]]
    want = [alone(c, 3) for c in cases]
    assert want[0][0] == 0 and want[1][0] != 0 and want[2][0] == 0
    got, built = run_batched(cases, 3)
    for case, g, w in zip(cases, got, want):
        compare(g, w, case[0])
    H0, C0 = built[1][5], built[1][6]
    assert np.array_equal(got[1][3], H0) and np.array_equal(got[1][4], C0)
    for b in built:
        b[0].close()


def test_batch_rejects_bad_arguments():
    spec = W.MeshSpec(16, 14, seed=2)
    g, *_ = pair(spec)
    with pytest.raises(M.MantleError):
        M.coupled_run_batch([g, g], 1, dict(cfl=0.5))
    res = M.coupled_run_batch([g], 1, dict(cfl=2.0))
    assert res[0][0] != 0
    g.close()
