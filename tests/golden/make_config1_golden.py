"""Write tests/golden/config1_oracle.npz: the 20x20 workload (seeded H, C and
divergence-free face velocities on [0,0.5]x[0,1]) after 10 SSP-RK3 steps at
CFL 0.4, computed by the oracle (the CPU restatement of the reference's
ML-WENO transport, pinned by the tests under tests/test_oracle_*.py).  The
inputs are regenerated from mantlepar_amd.workloads; the fixture holds the
step size and the final cell averages.

Usage: python tests/golden/make_config1_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
from mantlepar_amd import workloads as W  # noqa: E402


def main():
    spec, steps, cfl = W.config1()
    H, Cc = W.enthalpy(spec), W.composition(spec)
    vx, vy = W.face_velocity(spec)
    o = O.Transport(W.vertex_coords(spec), spec.nx, spec.ny, vx, vy, q=spec.q)
    _, rate = o.rates()
    dt = cfl / rate
    H1, C1, _ = o.advance(H, Cc, steps, dt, 3)
    np.savez_compressed(os.path.join(HERE, "config1_oracle.npz"), dt=np.float64(dt),
                        steps=np.int32(steps), H=np.asarray(H1, np.float64).reshape(spec.ny, spec.nx),
                        C=np.asarray(C1, np.float64).reshape(spec.ny, spec.nx))
    print(f"config1_oracle.npz: dt {dt!r}, {steps} steps")


if __name__ == "__main__":
    main()
