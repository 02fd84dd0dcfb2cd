"""The oracle's TensorStencilPoly / Reconstruction pinned by the reference's
known answers (src/mlweno/test/test_tensorstencilpoly.cpp,
src/mlweno/test/test_reconstruction.cpp) and by closed forms.

The reference's mlweno sources need PETSc and LAPACKE, which this image lacks,
so they cannot be compiled; these checks pin the restatement instead:
Gauss exactness, Q1/Q2 reproduction on hash-perturbed quads, analytic
smoothness indicators, the direct-formula weights, extreme weights at
epsilon*A = 1e-600, exact constants, boundary candidate counts and the
interior convergence order.
"""
import numpy as np
import pytest

import oracle as O
from mantlepar_amd import workloads as W


def perturbed_coords(nx, ny, L=1.0, Hh=1.0, perturbation=0.15, seed=1):
    """LogicRectMesh vertices (src/core/mesh.cpp:86-114): interior vertices
    moved by perturbation*h*SignedNoise(seed + 2*index (+1))."""
    hx, hy = L / nx, Hh / ny
    j, i = np.meshgrid(np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    x = i * hx
    y = j * hy
    index = (j * (nx + 1) + i).astype(np.uint64)
    nxs = (W.mix64(np.uint64(seed) + np.uint64(2) * index) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    nys = (W.mix64(np.uint64(seed) + np.uint64(2) * index + np.uint64(1)) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    interior = (i > 0) & (i < nx) & (j > 0) & (j < ny)
    x = np.where(interior, x + perturbation * hx * (2 * nxs - 1), x)
    y = np.where(interior, y + perturbation * hy * (2 * nys - 1), y)
    x[:, 0], x[:, -1], y[0, :], y[-1, :] = 0.0, L, 0.0, Hh
    return np.stack([x, y], axis=-1)


def corners(coords, i, j):
    return [coords[j, i], coords[j, i + 1], coords[j + 1, i + 1], coords[j + 1, i]]


def map_cell(c, r, s):
    a, b = 0.5 * (1 - r), 0.5 * (1 + r)
    lo = a * c[0] + b * c[1]
    hi = a * c[3] + b * c[2]
    return 0.5 * (1 - s) * lo + 0.5 * (1 + s) * hi


def jac(c, r, s):
    sx, t = 0.5 * (1 + r), 0.5 * (1 + s)
    dxi = 0.5 * ((1 - t) * (c[1] - c[0]) + t * (c[2] - c[3]))
    deta = 0.5 * ((1 - sx) * (c[3] - c[0]) + sx * (c[2] - c[1]))
    return dxi[0] * deta[1] - dxi[1] * deta[0]


def cell_average(c, f, n=8):
    g, w = np.polynomial.legendre.leggauss(n)
    tot = area = 0.0
    for a, wa in zip(g, w):
        for b, wb in zip(g, w):
            p = map_cell(c, a, b)
            jj = jac(c, a, b) * wa * wb
            tot += jj * f(p[0], p[1])
            area += jj
    return tot / area


def test_gauss_rule_exactness():
    for n in range(1, 9):
        x, w = O.gauss_rule(n)
        assert np.all(np.diff(x) > 0)
        for k in range(2 * n):
            exact = 0.0 if k % 2 else 2.0 / (k + 1)
            assert abs(np.dot(w, x ** k) - exact) < 1e-14


@pytest.mark.parametrize("size,order", [((3, 3), 2), ((2, 2), 1), ((4, 3), 3)])
def test_stencil_reproduces_its_polynomial_space(size, order):
    nx, ny = 8, 7
    coords = perturbed_coords(nx, ny, seed=11)
    rng = np.random.default_rng(1)
    coef = rng.normal(size=(size[1], size[0]))

    def poly(x, y):
        return sum(coef[b, a] * x ** a * y ** b for a in range(size[0]) for b in range(size[1]))

    start = (2, 3)
    cs = [corners(coords, start[0] + i, start[1] + j) for j in range(size[1]) for i in range(size[0])]
    avg = np.array([cell_average(c, poly) for c in cs])
    h = np.sqrt(1.0 / (nx * ny))
    s = O.StencilPoly(cs, start, size, order, h)
    for _ in range(20):
        c = cs[rng.integers(len(cs))]
        p = map_cell(c, *(2 * rng.random(2) - 1))
        assert abs(s.evaluate(avg, p) - poly(*p)) < 5e-11
    # the basis has unit cell averages: reconstruction of the averages themselves
    for k, c in enumerate(cs):
        got = cell_average(c, lambda x, y: s.evaluate(avg, (x, y)))
        assert abs(got - avg[k]) < 5e-11


def analytic_sigma(deriv_fn, region, h, order, maxd):
    """sum_{1<=dx+dy<=order} h^(2(dx+dy)) * mean_region (d^dx_x d^dy_y p)^2, 6-point Gauss."""
    g, w = np.polynomial.legendre.leggauss(6)
    total = 0.0
    for dy in range(maxd[1]):
        for dx in range(maxd[0]):
            k = dx + dy
            if k == 0 or k > order:
                continue
            acc = area = 0.0
            for a, wa in zip(g, w):
                for b, wb in zip(g, w):
                    p = map_cell(region, a, b)
                    jj = jac(region, a, b) * wa * wb
                    acc += jj * deriv_fn(p[0], p[1], dx, dy) ** 2
                    area += jj
            total += h ** (2 * k) * acc / area
    return total


def test_smoothness_indicators_match_analytic_integrals():
    nx, ny = 9, 9
    coords = perturbed_coords(nx, ny, seed=5)
    rng = np.random.default_rng(2)
    c2 = rng.normal(size=(3, 3))

    def d2(x, y, dx, dy):
        tot = 0.0
        for b in range(3):
            for a in range(3):
                if a >= dx and b >= dy:
                    fa = np.prod([a - k for k in range(dx)]) if dx else 1.0
                    fb = np.prod([b - k for k in range(dy)]) if dy else 1.0
                    tot += c2[b, a] * fa * fb * x ** (a - dx) * y ** (b - dy)
        return tot

    h = np.sqrt(1.0 / (nx * ny))
    start = (3, 2)
    cs = [corners(coords, start[0] + i, start[1] + j) for j in range(3) for i in range(3)]
    avg = np.array([cell_average(c, lambda x, y: d2(x, y, 0, 0)) for c in cs])
    s = O.StencilPoly(cs, start, (3, 3), 2, h)
    cx, cy = s.center
    square = [np.array(v) for v in ((cx - h / 2, cy - h / 2), (cx + h / 2, cy - h / 2),
                                    (cx + h / 2, cy + h / 2), (cx - h / 2, cy + h / 2))]
    ref = analytic_sigma(d2, square, h, 2, (3, 3))
    assert abs(s.smoothness(avg) - ref) <= 2e-11 * max(1.0, abs(ref))
    # TargetCell indicator of a Q1 candidate over a physical cell
    c1 = rng.normal(size=(2, 2))

    def d1(x, y, dx, dy):
        tot = 0.0
        for b in range(2):
            for a in range(2):
                if a >= dx and b >= dy:
                    tot += c1[b, a] * x ** (a - dx) * y ** (b - dy)
        return tot

    cs1 = [corners(coords, 4 + i, 4 + j) for j in range(2) for i in range(2)]
    avg1 = np.array([cell_average(c, lambda x, y: d1(x, y, 0, 0)) for c in cs1])
    s1 = O.StencilPoly(cs1, (4, 4), (2, 2), 1, h)
    s1.set_target(1, 0)
    ref1 = analytic_sigma(d1, cs1[1], h, 1, (2, 2))
    assert abs(s1.smoothness(avg1, target=True) - ref1) <= 2e-11 * max(1.0, abs(ref1))


def test_constant_data_is_exact_and_has_zero_indicator():
    nx, ny = 6, 6
    coords = perturbed_coords(nx, ny, seed=3)
    r = O.Recon(coords.reshape(-1), nx, ny, (3, 3), h=np.sqrt(1.0 / 36))
    vals = np.full(nx * ny, 1e8)
    r.update(vals, (0, 0, nx, ny), 0.01)
    assert r.evaluate(tuple(coords[3, 3])) == 1e8
    assert r.evaluate(tuple(coords[3, 3]), 1, 0) == 0.0
    for k in range(r.candidate_count):
        assert r.candidate(k)["smoothness"] == 0.0


def _options(small_offsets, small_weights=None, large_offsets=((-1, -1),), large_weights=None,
             large_order=2, target_small=True, use_constant=False, eps=1e-2, s=1, clw=1e-5):
    o = O.WenoOptions()
    o.large.size_i = o.large.size_j = 3
    o.large.order = large_order
    o.large.noffsets = len(large_offsets)
    for k, (a, b) in enumerate(large_offsets):
        o.large.off_i[k], o.large.off_j[k] = a, b
    if large_weights is not None:
        o.large.has_weights = 1
        for k, v in enumerate(large_weights):
            o.large.linear_weight[k] = v
    o.small.size_i = o.small.size_j = 2
    o.small.order = 1
    o.small.noffsets = len(small_offsets)
    for k, (a, b) in enumerate(small_offsets):
        o.small.off_i[k], o.small.off_j[k] = a, b
    if small_weights is not None:
        o.small.has_weights = 1
        for k, v in enumerate(small_weights):
            o.small.linear_weight[k] = v
    o.small.target_smoothness = int(target_small)
    o.use_constant = int(use_constant)
    o.constant_linear_weight = clw
    o.epsilon = eps
    o.s = s
    return o


TEST_SMALL = ((0, 0), (-1, 0), (0, -1), (-1, -1))


def test_boundary_candidate_counts_and_mapping():
    nx, ny = 12, 10
    coords = perturbed_coords(nx, ny, seed=7)
    opt = _options(TEST_SMALL, small_weights=(1, 2, 3, 4))
    for target, count in (((0, 0), 1), ((0, 5), 2), ((5, 5), 5), ((nx - 1, ny - 1), 1)):
        r = O.Recon(coords.reshape(-1), nx, ny, target, options=opt, h=0.1)
        assert r.candidate_count == count
    r = O.Recon(coords.reshape(-1), nx, ny, (0, 0), options=opt, h=0.1)
    c = r.candidate(0)
    assert c["family"] == 1 and c["configured_index"] == 0 and c["linear_weight"] == 1
    r = O.Recon(coords.reshape(-1), nx, ny, (nx - 1, ny - 1), options=opt, h=0.1)
    c = r.candidate(0)
    assert c["family"] == 1 and c["configured_index"] == 3 and c["linear_weight"] == 4


def test_weights_follow_the_direct_formula():
    nx, ny = 12, 10
    coords = perturbed_coords(nx, ny, seed=9)
    rng = np.random.default_rng(4)
    for constant in (False, True):
        opt = _options(TEST_SMALL, small_weights=(0.2, 0.5, 0.8, 1.1), large_weights=(0.7,),
                       use_constant=constant, eps=0.4, s=2, clw=2.0)
        r = O.Recon(coords.reshape(-1), nx, ny, (5, 5), options=opt, h=0.1)
        vals = rng.normal(size=nx * ny) * 0.3
        area = 0.6
        r.update(vals, (0, 0, nx, ny), area)
        alphas = []
        for k in range(r.candidate_count):
            c = r.candidate(k)
            rr = (2 if c["family"] == 0 else 1) + 1
            eta = 1 if rr == 1 else (3 if rr == 2 else 4)
            alphas.append(c["linear_weight"] / (c["smoothness"] + 0.4 * area) ** (2 * rr + eta))
        a0 = 2.0 / (0.4 * area) ** (2 + 1) if constant else 0.0
        total = sum(alphas) + a0
        for k in range(r.candidate_count):
            assert abs(r.candidate(k)["weight"] - alphas[k] / total) < 2e-11
        assert abs(r.constant_weight - a0 / total) < 2e-11


def test_extreme_weights_in_log_space():
    """epsilon*A = 1e-600 with exponent 105: direct arithmetic would underflow."""
    nx, ny = 12, 10
    coords = perturbed_coords(nx, ny, seed=2)
    opt = _options(TEST_SMALL, large_offsets=((-1, -1), (0, 0)), large_weights=(1, 3),
                   large_order=100, use_constant=True, eps=1e-300)
    r = O.Recon(coords.reshape(-1), nx, ny, (5, 5), options=opt, h=0.1)
    r.update(np.full(nx * ny, 1e8), (0, 0, nx, ny), 1e-300)
    assert abs(r.candidate(0)["weight"] - 0.25) < 3e-13
    assert abs(r.candidate(1)["weight"] - 0.75) < 3e-13


def test_weights_sum_to_one_and_update_errors_preserve_state():
    nx, ny = 12, 10
    coords = perturbed_coords(nx, ny, seed=4)
    r = O.Recon(coords.reshape(-1), nx, ny, (5, 5), h=0.1)
    vals = np.random.default_rng(8).random(nx * ny)
    r.update(vals, (0, 0, nx, ny), 0.01)
    tot = sum(r.candidate(k)["weight"] for k in range(r.candidate_count)) + r.constant_weight
    assert abs(tot - 1.0) < 5e-14
    before = r.evaluate(tuple(coords[5, 5]))
    with pytest.raises(O.OracleError):
        r.update(vals, (0, 0, nx, ny), 0.0)
    bad = vals.copy()
    bad[5 * nx + 5] = np.nan
    with pytest.raises(O.OracleError):
        r.update(bad, (0, 0, nx, ny), 0.01)
    assert r.evaluate(tuple(coords[5, 5])) == before


def test_interior_convergence_order():
    """Value error at interior face points converges at >= 3 - 0.4
    (src/mlweno/test/reconstruction_tests.cmake:29-52)."""
    def f(x, y):
        return np.sin(2.1 * x + 0.3) * np.cos(1.7 * y - 0.2)

    errs = []
    for n in (16, 32, 64):
        coords = W.vertex_coords(W.MeshSpec(n, n, 0.0, 1.0, 0.0, 1.0))
        vals = np.array([cell_average(corners(coords, i, j), f, n=5)
                         for j in range(n) for i in range(n)])
        err = 0.0
        for (i, j) in ((n // 2, n // 2), (n // 3, n // 2 + 1), (n // 2 - 2, n // 4)):
            r = O.Recon(coords.reshape(-1), n, n, (i, j), h=1.0 / n)
            r.update(vals, (0, 0, n, n), 1.0 / n ** 2)
            for p in (coords[j, i] * 0.5 + coords[j, i + 1] * 0.5, coords[j + 1, i + 1]):
                err = max(err, abs(r.evaluate(tuple(p)) - f(*p)))
        errs.append(err)
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert np.all(rates >= 3 - 0.4), (errs, rates)
