"""Handoff files and edge-sample conversion (host code of the C ABI; no GPU).

Pins: round trips are bitwise (17 significant digits); a source directory
written the way the reference's own evolution test writes one
(src/couple/test/test_evolution.py:29-40: json.dumps manifest, numpy
savetxt '%.17g') is read; every check of LoadPreheatState
(src/couple/evolve.cpp:62-85) rejects with the reference's message; edge ids
follow the cell-edge closed form of src/core/test/test_mesh_info.cpp:125-126;
face velocities rebuilt from per-cell edge samples (each interior edge
sampled from both cells, src/couple/flow_sampling.cpp:101-115) equal the
source arrays exactly.
"""
import json

import numpy as np
import pytest

from mantlepar_amd import MantleError
from mantlepar_amd import formats as F

pytestmark = pytest.mark.filterwarnings("ignore")


def fields(nx, ny, seed=0):
    rng = np.random.default_rng(seed)
    H = 0.5 + rng.random((ny, nx))
    Cc = 0.4 * rng.random((ny, nx))
    H[0, 0] = 0.1 + 1e-17          # needs all 17 digits
    Cc[-1, -1] = 1.0 / 3.0
    return H, Cc


def meta(nx=7, ny=5, **kw):
    d = dict(domain=(0.0, 0.5, 0.0, 1.0), enthalpy_Jkg=1.2e6, length_m=3.2e4,
             time_scale_s=1.7e12, time=0.25, accepted_steps=42)
    d.update(kw)
    return F.handoff_meta(nx, ny, **d)


def test_round_trip_is_bitwise_and_reference_shaped(tmp_path):
    H, Cc = fields(7, 5)
    m = meta()
    F.write_handoff(tmp_path, m, H, Cc)
    H2, C2, found = F.read_handoff(tmp_path, m, 0.4)
    assert np.array_equal(H, H2) and np.array_equal(Cc, C2)
    assert found.time == 0.25 and found.accepted_steps == 42
    # the files read back with the tools the reference's tests use
    assert np.array_equal(np.loadtxt(tmp_path / "cellH1.dat"), H)
    assert np.array_equal(np.loadtxt(tmp_path / "cellC1.dat"), Cc)
    man = json.loads((tmp_path / "transport_state.json").read_text())
    assert list(man) == ["schema_version", "status", "representation", "ordering", "H", "C",
                         "time", "time_s", "accepted_steps", "mesh", "scales"]
    assert man["schema_version"] == 1 and man["status"] == "complete"
    assert man["representation"] == "cell averages" and man["ordering"] == "j*nx+i"
    assert man["H"] == {"file": "cellH1.dat", "units": "h/(cp*dT)"}
    assert man["C"] == {"file": "cellC1.dat", "units": "fraction"}
    assert man["time_s"] == 0.25 * 1.7e12
    assert man["mesh"] == {"family": "rectangular", "nx": 7, "ny": 5,
                           "domain": [0.0, 0.5, 0.0, 1.0], "perturbation": 0.0, "seed": 0}
    assert man["scales"] == {"enthalpy_Jkg": 1.2e6, "length_m": 3.2e4, "time_s": 1.7e12}
    # no temporary files left behind
    assert sorted(p.name for p in tmp_path.iterdir()) == ["cellC1.dat", "cellH1.dat",
                                                           "transport_state.json"]


def reference_style_source(path, nx, ny, H, Cc, family="rectangular", perturbation=0.0, seed=0):
    data = {"schema_version": 1, "status": "complete", "representation": "cell averages",
            "ordering": "j*nx+i",
            "H": {"file": "cellH1.dat", "units": "h/(cp*dT)"},
            "C": {"file": "cellC1.dat", "units": "fraction"}, "time": 15000,
            "mesh": {"family": family, "nx": nx, "ny": ny, "domain": [0.0, 0.5, 0.0, 1.0],
                     "perturbation": perturbation, "seed": seed},
            "scales": {"enthalpy_Jkg": 1.2e6, "length_m": 3.2e4, "time_s": 1.7e12}}
    (path / "transport_state.json").write_text(json.dumps(data))
    np.savetxt(path / "cellH1.dat", H, fmt="%.17g")
    np.savetxt(path / "cellC1.dat", Cc, fmt="%.17g")
    return data


def test_reads_a_source_written_like_the_reference_test(tmp_path):
    H, Cc = fields(6, 6, seed=3)
    reference_style_source(tmp_path, 6, 6, H, Cc)
    H2, C2, found = F.read_handoff(tmp_path, meta(6, 6), 0.4)
    assert np.array_equal(H, H2) and np.array_equal(Cc, C2)
    assert found.time == 15000.0 and found.accepted_steps == -1


def mutate(path, fn):
    d = json.loads((path / "transport_state.json").read_text())
    fn(d)
    (path / "transport_state.json").write_text(json.dumps(d))


CASES = [
    ("status", lambda d: d.update(status="incomplete"), "must be complete"),
    ("schema", lambda d: d.update(schema_version=2), "must be complete"),
    ("ordering", lambda d: d.update(ordering="i*ny+j"), "representation/order"),
    ("units", lambda d: d["H"].update(units="J/kg"), "units do not match"),
    ("filename", lambda d: d["C"].update(file="c.dat"), "filenames"),
    ("nx", lambda d: d["mesh"].update(nx=7), "mesh does not match"),
    ("family", lambda d: d["mesh"].update(family="perturbed_quadrilateral"), "mesh does not match"),
    ("domain", lambda d: d["mesh"].update(domain=[0.0, 0.5, 0.0, 1.0 + 1e-9]), "domain does not match"),
    ("scales", lambda d: d["scales"].update(length_m=3.3e4), "scales do not match"),
    ("time", lambda d: d.update(time=float("nan")), "Nonfinite source time"),
]


@pytest.mark.parametrize("name,fn,message", CASES, ids=[c[0] for c in CASES])
def test_manifest_rejections(tmp_path, name, fn, message):
    H, Cc = fields(6, 6)
    reference_style_source(tmp_path, 6, 6, H, Cc)
    mutate(tmp_path, fn)
    with pytest.raises(MantleError) as ei:
        F.read_handoff(tmp_path, meta(6, 6), 0.4)
    assert message in str(ei.value)


def test_domain_within_tolerance_is_accepted(tmp_path):
    H, Cc = fields(6, 6)
    reference_style_source(tmp_path, 6, 6, H, Cc)
    mutate(tmp_path, lambda d: d["mesh"].update(domain=[0.0, 0.5 * (1 + 5e-13), 0.0, 1.0]))
    F.read_handoff(tmp_path, meta(6, 6), 0.4)


def test_matrix_rejections(tmp_path):
    H, Cc = fields(6, 6)
    reference_style_source(tmp_path, 6, 6, H, Cc)
    good = (tmp_path / "cellH1.dat").read_text()
    cases = [
        (good + "0.5\n", "cellH1.dat", "Extra entries"),
        (good.rsplit(" ", 1)[0] + "\n", "cellH1.dat", "Incomplete"),
        (good.replace(good.split()[3], "-0.25", 1), "cellH1.dat", "Inadmissible"),
        (good.replace(good.split()[3], "nan", 1), "cellH1.dat", "Inadmissible"),
    ]
    for text, name, message in cases:
        (tmp_path / name).write_text(text)
        with pytest.raises(MantleError) as ei:
            F.read_handoff(tmp_path, meta(6, 6), 0.4)
        assert message in str(ei.value), (message, str(ei.value))
        (tmp_path / name).write_text(good)
    Cbad = Cc.copy()
    Cbad[2, 3] = 0.41
    np.savetxt(tmp_path / "cellC1.dat", Cbad, fmt="%.17g")
    with pytest.raises(MantleError) as ei:
        F.read_handoff(tmp_path, meta(6, 6), 0.4)
    assert "Inadmissible" in str(ei.value)
    F.read_handoff(tmp_path, meta(6, 6), 1.0)          # admissible with the mirror's bound
    (tmp_path / "cellC1.dat").unlink()
    with pytest.raises(MantleError) as ei:
        F.read_handoff(tmp_path, meta(6, 6), 0.4)
    assert "Cannot open" in str(ei.value)


def test_perturbed_family_checks_perturbation_and_seed(tmp_path):
    H, Cc = fields(6, 6)
    reference_style_source(tmp_path, 6, 6, H, Cc, family="perturbed_quadrilateral",
                           perturbation=0.2, seed=18446744073709551557)
    m = meta(6, 6, family="perturbed_quadrilateral", perturbation=0.2, seed=18446744073709551557)
    _, _, found = F.read_handoff(tmp_path, m, 0.4)
    assert found.seed == 18446744073709551557
    m2 = meta(6, 6, family="perturbed_quadrilateral", perturbation=0.2, seed=18446744073709551556)
    with pytest.raises(MantleError) as ei:
        F.read_handoff(tmp_path, m2, 0.4)
    assert "perturbation does not match" in str(ei.value)


def test_edge_ids_follow_the_reference_numbering():
    nx, ny = 5, 4
    cx, M, horizontal = nx, nx + 1, nx * (ny + 1)
    for j in range(ny):
        for i in range(nx):
            expected = [j * cx + i, horizontal + j * M + i + 1, (j + 1) * cx + i, horizontal + j * M + i]
            got = [F.edge_id(nx, ny, False, i, j), F.edge_id(nx, ny, True, i + 1, j),
                   F.edge_id(nx, ny, False, i, j + 1), F.edge_id(nx, ny, True, i, j)]
            assert got == expected


def cell_samples(nx, ny, q, vnx, vny, rng):
    """Edge samples as SampleInitialFlow emits them: every cell, its four
    sides, q in canonical order; interior edges appear twice."""
    ids, qs, vx, vy = [], [], [], []
    for j in range(ny):
        for i in range(nx):
            for vertical, ii, jj in ((False, i, j), (True, i + 1, j), (False, i, j + 1), (True, i, j)):
                for p in range(q):
                    ids.append(F.edge_id(nx, ny, vertical, ii, jj))
                    qs.append(p)
                    if vertical:
                        vx.append(vnx[p, jj, ii]); vy.append(rng.random())
                    else:
                        vx.append(rng.random()); vy.append(vny[p, jj, ii])
    order = rng.permutation(len(ids))
    return [np.asarray(a)[order] for a in (ids, qs, vx, vy)]


def test_face_velocity_from_samples():
    nx, ny, q = 6, 4, 5
    rng = np.random.default_rng(7)
    vnx = rng.standard_normal((q, ny, nx + 1))
    vny = rng.standard_normal((q, ny + 1, nx))
    ids, qs, vx, vy = cell_samples(nx, ny, q, vnx, vny, rng)
    ax, ay = F.face_velocity_from_samples(nx, ny, q, ids, qs, vx, vy)
    assert np.array_equal(ax, vnx) and np.array_equal(ay, vny)
    # the last sample of an (edge, q) wins
    e = F.edge_id(nx, ny, True, 3, 2)
    ids2 = np.append(ids, e); qs2 = np.append(qs, 1); vx2 = np.append(vx, 9.5); vy2 = np.append(vy, 0.0)
    bx, _ = F.face_velocity_from_samples(nx, ny, q, ids2, qs2, vx2, vy2)
    assert bx[1, 2, 3] == 9.5
    # a face point nobody sampled, an id outside the mesh, a non-finite value
    keep = ~((ids == e) & (qs == 4))
    with pytest.raises(MantleError, match="no velocity sample"):
        F.face_velocity_from_samples(nx, ny, q, ids[keep], qs[keep], vx[keep], vy[keep])
    with pytest.raises(MantleError, match="outside the mesh"):
        F.face_velocity_from_samples(nx, ny, q, np.append(ids, nx * (ny + 1) + (nx + 1) * ny),
                                     np.append(qs, 0), np.append(vx, 0.0), np.append(vy, 0.0))
    with pytest.raises(MantleError) as ei:
        F.face_velocity_from_samples(nx, ny, q, np.append(ids, 0), np.append(qs, 0),
                                     np.append(vx, 0.0), np.append(vy, np.inf))
    assert ei.value.code == 2
