"""How far the FP64 method itself carries a rounding-sized difference, on the
oracle alone (no GPU).

The batch workload draws meshes with 5-20 rows of cells 20-100 times taller
than wide.  On some of them two FP64 runs of the oracle whose initial states
differ by a relative 1e-15 end 50 SSP-RK3 steps more than 1e-6 apart, so no
implementation, FP64 or mixed, can be held to 1e-6 of another there.  On a
well-resolved mesh of the same draw the two runs stay within 1e-12.  The
amplifier is the nonlinear ML-WENO weights, alpha ~ (sigma + eps*A)^-7
(src/mlweno/reconstruction.cpp:212-265): with epsilon = 1e30 the weights are
frozen at their linear values, the scheme is linear, and the same perturbation
stays at rounding level on the same mesh.
"""
import numpy as np
import pytest

import oracle as O
from mantlepar_amd import workloads as W

SPECS = {s.seed: s for s in W.batch_specs(256, seed=3)}


def run(spec, H, C, steps=50, epsilon=None):
    vx, vy = W.face_velocity(spec)
    opts = O.WenoOptions.production()
    if epsilon is not None:
        opts.epsilon = epsilon
    o = O.Transport(W.vertex_coords(spec), spec.nx, spec.ny, vx, vy, q=spec.q, options=opts,
                    table_mode=1, threads=4)
    _, rate = o.rates()
    H1, C1, _ = o.advance(H, C, steps, 0.4 / rate, 3)
    return H1, C1


def spread(spec, epsilon=None):
    H, C = W.enthalpy(spec), W.composition(spec)
    rng = np.random.default_rng(spec.seed)
    a = run(spec, H, C, epsilon=epsilon)
    b = run(spec, H * (1.0 + 1e-15 * rng.standard_normal(H.shape)),
            C * (1.0 + 1e-15 * rng.standard_normal(C.shape)), epsilon=epsilon)
    return max(np.abs(a[0] - b[0]).max() / np.abs(a[0]).max(),
               np.abs(a[1] - b[1]).max() / np.abs(a[1]).max())


@pytest.mark.parametrize("seed", [300123, 300128], ids=["140x5", "180x7"])
def test_tall_cell_meshes_amplify_rounding_past_1e6(seed):
    spec = SPECS[seed]
    assert spec.ny <= 8 and spec.dy / spec.dx > 50
    assert spread(spec) > 1e-6


def test_well_resolved_mesh_does_not():
    spec = SPECS[300074]          # 40 x 65, cells 1.2 times taller than wide
    assert spread(spec) < 1e-12


def test_frozen_weights_do_not():
    """The same tall-cell mesh with linear weights (epsilon = 1e30)."""
    assert spread(SPECS[300123], epsilon=1e30) < 1e-12
