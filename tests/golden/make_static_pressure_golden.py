"""Write tests/golden/static_pressure_golden.npz from the reference's own src/phase.

The reference EutecticModel is compiled unmodified from its sources by
``make -C oracle ref`` (oracle/_ref/libref_phase.so); this script evaluates
EutecticModel::staticPressure on fixed inputs and stores inputs, outputs and
exception categories.  Run in a container where the reference sources are
mounted:

    make -C oracle ref && python tests/golden/make_static_pressure_golden.py

Cases: the parameter sets of make_phase_golden.py; depths at the surface, at
the mesh rows of a few lithostatic columns (surface_y - y for uniform rows),
random and negative depths; both the one-argument form (length scale l0) and
explicit length scales; non-finite depths and invalid length scales
(src/phase/test/test_eutectic.cpp:337-341, 455-463); an overflowing product.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

import oracle as O  # noqa: E402
from make_phase_golden import parameter_sets  # noqa: E402


def depths(rng):
    rows = []
    for ny, surface in ((14, 1.0), (36, 1.25), (64, 3.0)):
        y = np.arange(ny + 1) * (1.0 / ny)
        rows.append(surface - y)
    z = np.concatenate(rows + [np.array([0.0, -0.0, 0.5, 1.0, 1e-300, 5e-324, -0.25]),
                               rng.uniform(0, 4, 64), rng.uniform(-1, 0, 8),
                               np.exp(rng.uniform(-30, 30, 16))])
    return np.concatenate([z, [np.nan, np.inf, -np.inf, 1e300]])


def main():
    rng = np.random.default_rng(20260)
    z = depths(rng)
    lengths = np.array([1000.0, 0.5, 1e-3, 0.0, -1.0, np.inf, np.nan])
    mats = parameter_sets()
    sets, zz, ll, scaled, out, status = [], [], [], [], [], []
    for si, m in enumerate(mats):
        p, st = O.ref_static_pressure(m, z)
        sets.append(np.full(z.size, si)); zz.append(z); ll.append(np.zeros(z.size))
        scaled.append(np.zeros(z.size, dtype=np.int8)); out.append(p); status.append(st)
        zl = np.repeat(z[:24], lengths.size)
        lv = np.tile(lengths, 24)
        p, st = O.ref_static_pressure(m, zl, lv)
        sets.append(np.full(zl.size, si)); zz.append(zl); ll.append(lv)
        scaled.append(np.ones(zl.size, dtype=np.int8)); out.append(p); status.append(st)
    np.savez_compressed(os.path.join(HERE, "static_pressure_golden.npz"),
                        params=np.array([m.as_array() for m in mats]),
                        set=np.concatenate(sets), zD=np.concatenate(zz), length=np.concatenate(ll),
                        scaled=np.concatenate(scaled), P=np.concatenate(out),
                        status=np.concatenate(status))


if __name__ == "__main__":
    main()
