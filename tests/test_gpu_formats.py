"""Handoff in, edge samples in, transport, handoff out: the formats either side
of the step around a device run, checked bitwise against the same run fed
directly."""
import numpy as np
import pytest

import mantlepar_amd as M
from mantlepar_amd import formats as F
from mantlepar_amd import workloads as W

pytestmark = pytest.mark.gpu


def test_handoff_and_edge_samples_around_a_run(tmp_path):
    spec, steps, cfl = W.config1()
    H, Cc = W.enthalpy(spec), W.composition(spec)
    vx, vy = W.face_velocity(spec)
    meta = F.handoff_meta(spec.nx, spec.ny, domain=(spec.x0, spec.x1, spec.y0, spec.y1))
    src = tmp_path / "src"
    src.mkdir()
    F.write_handoff(src, meta, H, Cc)
    H1, C1, _ = F.read_handoff(src, meta, 1.0)
    # edge samples as the flow solve emits them: every cell's four sides
    ids, qs, sx, sy = [], [], [], []
    for j in range(spec.ny):
        for i in range(spec.nx):
            for vertical, ii, jj in ((False, i, j), (True, i + 1, j), (False, i, j + 1), (True, i, j)):
                for p in range(spec.q):
                    ids.append(F.edge_id(spec.nx, spec.ny, vertical, ii, jj))
                    qs.append(p)
                    sx.append(vx[p, jj, ii] if vertical else 0.0)
                    sy.append(0.0 if vertical else vy[p, jj, ii])
    ax, ay = F.face_velocity_from_samples(spec.nx, spec.ny, spec.q, ids, qs, sx, sy)
    out = {}
    for tag, (h0, c0, u, v) in {"files": (H1, C1, ax, ay), "direct": (H, Cc, vx, vy)}.items():
        t = M.Transport(spec.nx, spec.ny, spec.x0, spec.x1, spec.y0, spec.y1, q=spec.q,
                        phase=W.config_phase(), precision=M.PREC_FP64)
        t.upload(h0, c0)
        t.set_face_velocity(u, v)
        t.step(steps, t.stable_dt(cfl))
        out[tag] = t.download()
        t.close()
    assert np.array_equal(out["files"][0], out["direct"][0])
    assert np.array_equal(out["files"][1], out["direct"][1])
    dst = tmp_path / "dst"
    dst.mkdir()
    meta.time, meta.accepted_steps = 0.5, steps
    F.write_handoff(dst, meta, *out["files"])
    H2, C2, found = F.read_handoff(dst, meta, 1.0)
    assert np.array_equal(H2, out["files"][0]) and np.array_equal(C2, out["files"][1])
    assert found.accepted_steps == steps
