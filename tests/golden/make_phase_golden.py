"""Write tests/golden/phase_golden.npz from the reference's own src/phase.

The reference EutecticModel is compiled unmodified from its sources by
``make -C oracle ref`` (oracle/_ref/libref_phase.so); this script evaluates it
on a fixed set of inputs and stores inputs, outputs and exception categories.
Run in a container where the reference sources are mounted:

    make -C oracle ref && python tests/golden/make_phase_golden.py

Cases follow the reference's own test groups (src/phase/test/test_eutectic.cpp):
the three parameter sets and four pressures, ordinary and trace compositions,
exact boundaries with their nextafter neighbours, forward-constructed
solid-liquid states and invalid inputs; plus the configs' parameter set and the
mirrored evaluations (1-C on a diagram with eutectic 1-Xe) that pin the
hypereutectic extension.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402

PRESSURES = (0.0, 1.0e5, 1.0e9, -1.0e7)


def parameter_sets():
    alt = dict(Tm0=2100.0, Te0=1500.0, cp=1350.0, L=6.0e5, Xe=0.3, nu=8.0e6, rho=3100.0,
               rhor=450.0, mus=3.0e19, mul=2.0, k0=2.0e-8, g=9.81, alpha0=4.0e-5)
    neg = dict(alt, nu=-8.0e6)
    cfg = dict(Tm0=1600.0, Te0=600.0, cp=1200.0, L=1.2e6, Xe=0.4)
    mirror = dict(cfg, Xe=1.0 - 0.4)
    return [O.Material.defaults(), O.Material.defaults(**alt), O.Material.defaults(**neg),
            O.Material.defaults(**cfg), O.Material.defaults(**mirror)]


def cases_for(model, m, rng, P):
    Xe, LD = m.Xe, model.derived["LD"]
    H, C = [], []
    b = model.boundaries(0.4 * Xe, P)
    n = 200
    u = rng.random(n)
    C += list(np.where(np.arange(n) % 2 == 0, Xe * u, Xe * 10.0 ** (-300.0 * rng.random(n))))
    H += list(b["Tep"] - 0.5 + (2.0 + LD) * rng.random(n))
    for c in (0.0, 1e-300, 1e-18, 0.4 * Xe, np.nextafter(Xe, 0.0), Xe):
        bb = model.boundaries(c, P)
        for h in (bb["Tep"], bb["Tmp"], bb["H_eutectic_end"], bb["H_liquidus"]):
            for hv in (np.nextafter(h, -np.inf), h, np.nextafter(h, np.inf)):
                H.append(hv)
                C.append(c)
    for off in (0.15, 0.5, 0.85):
        for liquid in (0.05, 0.3, 0.8, 0.99):
            T = b["Tep"] + off
            cl = Xe * (1.0 - off)
            H.append(T + LD * liquid)
            C.append(liquid * cl)
    # invalid inputs: C outside [0, Xe], non-finite H
    H += [b["Tep"], b["Tep"], np.nan, b["Tep"]]
    C += [-1e-12, np.nextafter(Xe, 1.0), 0.1 * Xe, np.inf]
    return np.array(H), np.array(C)


def main():
    if not O.ref_phase_available():
        sys.exit("oracle/_ref/libref_phase.so is missing: run `make -C oracle ref` first")
    rng = np.random.default_rng(20240613)
    mats = parameter_sets()
    rows_set, rows_H, rows_C, rows_P, rows_out, rows_status = [], [], [], [], [], []
    for si, m in enumerate(mats):
        model = O.PhaseModel(m)
        for P in PRESSURES:
            H, C = cases_for(model, m, rng, P)
            out, status = O.ref_phase_evaluate(m, H, C, np.full(H.size, P))
            rows_set.append(np.full(H.size, si, dtype=np.int32))
            rows_H.append(H)
            rows_C.append(C)
            rows_P.append(np.full(H.size, P))
            rows_out.append(out)
            rows_status.append(status)
    derived = []
    for m in mats:
        rc, d = O.ref_phase_derived(m)
        assert rc == 0
        derived.append(d)
    np.savez_compressed(
        os.path.join(HERE, "phase_golden.npz"),
        params=np.stack([m.as_array() for m in mats]),
        derived=np.stack(derived),
        set=np.concatenate(rows_set), H=np.concatenate(rows_H), C=np.concatenate(rows_C),
        P=np.concatenate(rows_P), out=np.concatenate(rows_out),
        status=np.concatenate(rows_status),
        columns=np.array(["phi1", "phi2", "phil", "cl", "cs", "region", "TDp", "Tep",
                          "dTD_dCD", "dTD_dHD", "flags"]))
    print("wrote", sum(h.size for h in rows_H), "cases")


if __name__ == "__main__":
    main()
