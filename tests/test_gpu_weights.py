"""FP64 ML-WENO weights under non-default options, against the oracle.

The reference's weight exponents s*r + eta(r) are integers (7, 5 and 2 at the
default s = 1; src/mlweno/reconstruction.cpp:43-44, 225-233), so the FP64
reconstruction forms alpha_k = d_k / (sigma_k + eps*A)^p_k by a reciprocal and
repeated squaring wherever every active denominator lies in a range that keeps
the alphas normal and finite, and keeps the reference's log form elsewhere.
The cases run both forms:
* s = 2 (exponents 10, 7, 3) with the constant candidate on: the direct form,
  with the constant candidate's alpha formed on the host;
* epsilon = 1e-200 on a state with a constant block: the targets inside the
  block have sigma = 0, so eps*A lies far below the range and they take the
  log form, while the targets outside it take the direct form;
* s = 30 (exponent 94): beyond the direct form's range, the log form for every
  target.

Bar: as for the default options, within 1e-12 of the oracle relative to each
field's maximum, for the RHS and for three SSP-RK3 steps.
"""
import numpy as np
import pytest

import mantlepar_amd as M
import oracle as O
from mantlepar_amd import workloads as W

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def block_state(spec):
    """The workload state with H and C constant on a block of cells."""
    H, Cc = W.enthalpy(spec).copy(), W.composition(spec).copy()
    H[10:25, 30:70] = 0.6
    Cc[10:25, 30:70] = 0.3
    return H, Cc


CASES = {
    "s2 constant": (dict(s=2, use_constant=1), W.enthalpy, W.composition),
    "tiny epsilon": (dict(epsilon=1e-200), lambda s: block_state(s)[0], lambda s: block_state(s)[1]),
    "s30": (dict(s=30), W.enthalpy, W.composition),
}


@pytest.mark.parametrize("name", list(CASES))
def test_weights_match_oracle(name):
    opts, fH, fC = CASES[name]
    print(f"weights: {name}", flush=True)
    spec = W.MeshSpec(140, 40, seed=21)
    H, Cc = fH(spec), fC(spec)
    vx, vy = W.face_velocity(spec)
    t = M.Transport(spec.nx, spec.ny, spec.x0, spec.x1, spec.y0, spec.y1, q=spec.q,
                    weno=M.default_weno(**opts), phase=W.config_phase(), precision=M.PREC_FP64)
    oo = O.WenoOptions.production()
    for k, v in opts.items():
        setattr(oo, k, v)
    o = O.Transport(W.vertex_coords(spec), spec.nx, spec.ny, vx, vy, q=spec.q, options=oo, table_mode=0)
    try:
        t.upload(H, Cc)
        t.set_face_velocity(vx, vy)
        LH, LC, _, _ = t.rhs()
        oLH, oLC, _, _ = o.rhs(H, Cc)
        assert rel(LH, oLH) <= 1e-12 and rel(LC, oLC) <= 1e-12, (rel(LH, oLH), rel(LC, oLC))
        dt = t.stable_dt(0.4)
        t.step(3, dt)
        Hg, Cg = t.download()
        Ho, Co, _ = o.advance(H, Cc, 3, dt, 3)
        assert rel(Hg, Ho) <= 1e-12 and rel(Cg, Co) <= 1e-12, (rel(Hg, Ho), rel(Cg, Co))
    finally:
        t.close()
