"""Synthetic transport workloads (seeded, counter-based, reproducible anywhere).

* H: 0.5 + 0.3*(1-y) plus three seeded Gaussian bumps (amplitude 0.1,
  sigma 0.05), cell averages by 5x5 Gauss (the reference's
  InitializeCellAverages, src/core/field_initialization.h:45; its "affine" and
  "gaussian" profiles, src/couple/initialization.cpp:64-85).
* C: i.i.d. uniform in [0.2, 0.6] per cell, keyed by SplitMix64 of
  (seed, cell index) -- the reference's stateless Mix64 (src/core/mesh.cpp:12-18).
* Face velocities from psi = A*sin(pi*(x-x0)/Lx)*sin(pi*(y-y0)/Ly) (on
  [0,0.5]x[0,1] this is 0.1*sin(2*pi*x)*sin(pi*y)), sampled at the q edge
  Gauss points, with a per-face constant correction so the Gauss-integrated
  face flux equals the psi difference across the face: every cell is
  discretely divergence-free.  psi vanishes on the whole boundary, so the
  boundary normal velocities are set to exactly zero.
All generators are separable in x and y, so an 8192x8192 mesh needs only the
output arrays.
"""
from dataclasses import dataclass

import numpy as np

from .transport import default_phase

_MASK = np.uint64(0xFFFFFFFFFFFFFFFF)


def mix64(v):
    """SplitMix64 finaliser on uint64 arrays (wrapping arithmetic)."""
    v = np.asarray(v, dtype=np.uint64)
    with np.errstate(over="ignore"):
        v = v + np.uint64(0x9E3779B97F4A7C15)
        v = (v ^ (v >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        v = (v ^ (v >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return v ^ (v >> np.uint64(31))


def unit_uniform(seed, index):
    """Uniform [0,1) double from (seed, index) with 53 random bits."""
    key = mix64(np.uint64(seed)) + np.asarray(index, dtype=np.uint64)
    return (mix64(key) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


def gauss_legendre(n):
    x, w = np.polynomial.legendre.leggauss(n)
    return x, w


@dataclass
class MeshSpec:
    nx: int
    ny: int
    x0: float = 0.0
    x1: float = 0.5
    y0: float = 0.0
    y1: float = 1.0
    q: int = 5
    seed: int = 1
    h_amplitude: float = 1.0     # scales the 0.1 bump amplitude
    psi_amplitude: float = 1.0   # scales the 0.1 streamfunction amplitude
    c_lo: float = 0.2
    c_hi: float = 0.6

    @property
    def dx(self):
        return (self.x1 - self.x0) / self.nx

    @property
    def dy(self):
        return (self.y1 - self.y0) / self.ny

    @property
    def cells(self):
        return self.nx * self.ny


def vertex_coords(spec):
    """(ny+1, nx+1, 2) uniform vertices with snapped endpoints (src/core/mesh.cpp:86-114)."""
    i = np.arange(spec.nx + 1, dtype=np.float64)
    j = np.arange(spec.ny + 1, dtype=np.float64)
    x = spec.x0 + i * ((spec.x1 - spec.x0) / spec.nx)
    y = spec.y0 + j * ((spec.y1 - spec.y0) / spec.ny)
    x[0], x[-1] = spec.x0, spec.x1
    y[0], y[-1] = spec.y0, spec.y1
    out = np.empty((spec.ny + 1, spec.nx + 1, 2))
    out[:, :, 0] = x[None, :]
    out[:, :, 1] = y[:, None]
    return out


def bump_centres(spec):
    u = unit_uniform(spec.seed + 7919, np.arange(6))
    xs = spec.x0 + (spec.x1 - spec.x0) * u[0::2]
    ys = spec.y0 + (spec.y1 - spec.y0) * u[1::2]
    return list(zip(xs, ys))


def _cell_gauss(lo, width, n, npts=5):
    g, w = gauss_legendre(npts)
    centres = lo + (np.arange(n) + 0.5) * width
    return centres[:, None] + 0.5 * width * g[None, :], 0.5 * w


def enthalpy(spec, npts=5):
    """Cell averages of H, shape (ny, nx)."""
    xs, wx = _cell_gauss(spec.x0, spec.dx, spec.nx, npts)
    ys, wy = _cell_gauss(spec.y0, spec.dy, spec.ny, npts)
    base_y = (0.5 + 0.3 * (1.0 - ys)) @ wy          # (ny,)
    H = np.repeat(base_y[:, None], spec.nx, axis=1)
    amp = 0.1 * spec.h_amplitude
    s2 = 2.0 * 0.05 ** 2
    for xc, yc in bump_centres(spec):
        gx = np.exp(-(xs - xc) ** 2 / s2) @ wx      # (nx,)
        gy = np.exp(-(ys - yc) ** 2 / s2) @ wy      # (ny,)
        H += amp * gy[:, None] * gx[None, :]
    return H


def composition(spec):
    """i.i.d. uniform [c_lo, c_hi] per cell, shape (ny, nx)."""
    u = unit_uniform(spec.seed, np.arange(spec.cells, dtype=np.uint64))
    return (spec.c_lo + (spec.c_hi - spec.c_lo) * u).reshape(spec.ny, spec.nx)


def face_velocity(spec):
    """vnx[q][ny][nx+1] (+x normal) and vny[q][ny+1][nx] (+y normal)."""
    g, w = gauss_legendre(spec.q)
    A = 0.1 * spec.psi_amplitude
    Lx, Ly = spec.x1 - spec.x0, spec.y1 - spec.y0
    i = np.arange(spec.nx + 1)
    j = np.arange(spec.ny + 1)
    sx = np.sin(np.pi * i / spec.nx)
    sy = np.sin(np.pi * j / spec.ny)
    sx[0] = sx[-1] = 0.0
    sy[0] = sy[-1] = 0.0
    # vertical faces: u = dpsi/dy = A*(pi/Ly)*sin(pi x~)*cos(pi y~)
    yq = (np.arange(spec.ny)[None, :] + 0.5 * (1.0 + g[:, None])) / spec.ny     # (q, ny) in [0,1]
    cy = np.cos(np.pi * yq)
    ku = A * np.pi / Ly
    S = ku * (0.5 * spec.dy) * (w @ cy)                                        # (ny,) per unit sin
    T = A * (sy[1:] - sy[:-1])                                                 # (ny,) per unit sin
    vnx = ku * cy[:, :, None] * sx[None, None, :]
    vnx += ((T - S) / spec.dy)[None, :, None] * sx[None, None, :]
    vnx[:, :, 0] = 0.0
    vnx[:, :, -1] = 0.0
    # horizontal faces: v = -dpsi/dx = -A*(pi/Lx)*cos(pi x~)*sin(pi y~)
    xq = (np.arange(spec.nx)[None, :] + 0.5 * (1.0 + g[:, None])) / spec.nx    # (q, nx)
    cx = np.cos(np.pi * xq)
    kv = -A * np.pi / Lx
    Sx = kv * (0.5 * spec.dx) * (w @ cx)                                       # (nx,)
    Tx = -A * (sx[1:] - sx[:-1])                                               # (nx,)
    vny = kv * cx[:, None, :] * sy[None, :, None]
    vny += ((Tx - Sx) / spec.dx)[None, None, :] * sy[None, :, None]
    vny[:, 0, :] = 0.0
    vny[:, -1, :] = 0.0
    return np.ascontiguousarray(vnx), np.ascontiguousarray(vny)


def _smoothstep(t):
    t = np.clip(t, 0.0, 1.0)
    return t * t * (3.0 - 2.0 * t)


def _dsmoothstep(t):
    return np.where((t > 0.0) & (t < 1.0), 6.0 * t * (1.0 - t), 0.0)


def throughflow_velocity(spec, Q=0.02, y_in=(0.9, 1.0), x_out=(0.0, 0.025)):
    """Face velocities of psi2 = Q*rho(y)*sigma(x), rho rising 0 -> 1 over y_in
    and sigma falling 1 -> 0 over x_out (smoothstep ramps).  Flow enters
    through the left side for y in y_in and leaves through the top for x in
    x_out; psi2 vanishes on the bottom and right sides.  Sampled at the edge
    Gauss points with the same per-face correction as face_velocity, so every
    cell stays discretely divergence-free.  Returns (vnx, vny) like
    face_velocity; the arrays are zero outside the corner block.

    The ramp ends are snapped to face edges: inside a ramp the smoothstep
    derivative is a cubic, which the edge Gauss rule integrates exactly, so
    the per-face corrections vanish to rounding and no Gauss point of an
    outflow face turns inward."""
    g, w = gauss_legendre(spec.q)
    ya, yb, xa, xb = _snapped_ramps(spec, y_in, x_out)
    xs = spec.x0 + np.arange(spec.nx + 1) * spec.dx          # vertical face abscissae
    ys = spec.y0 + np.arange(spec.ny + 1) * spec.dy          # horizontal face ordinates
    rho = _smoothstep((ys - ya) / (yb - ya))
    sig = 1.0 - _smoothstep((xs - xa) / (xb - xa))
    # vertical faces: u = dpsi2/dy = Q sigma(x) rho'(y)
    yq = ys[:-1][None, :] + 0.5 * (1.0 + g[:, None]) * spec.dy              # (q, ny)
    drho = _dsmoothstep((yq - ya) / (yb - ya)) / (yb - ya)
    S = (0.5 * spec.dy) * (w @ drho)
    fy = drho + ((rho[1:] - rho[:-1]) - S)[None, :] / spec.dy                # (q, ny)
    vnx = Q * fy[:, :, None] * sig[None, None, :]
    # horizontal faces: v = -dpsi2/dx = -Q sigma'(x) rho(y)
    xq = xs[:-1][None, :] + 0.5 * (1.0 + g[:, None]) * spec.dx              # (q, nx)
    dsig = -_dsmoothstep((xq - xa) / (xb - xa)) / (xb - xa)
    Sx = (0.5 * spec.dx) * (w @ dsig)
    fx = dsig + ((sig[1:] - sig[:-1]) - Sx)[None, :] / spec.dx               # (q, nx)
    vny = -Q * rho[None, :, None] * fx[:, None, :]
    return np.ascontiguousarray(vnx), np.ascontiguousarray(vny)


def _snapped_ramps(spec, y_in, x_out):
    def snap(v, lo, h):
        return lo + np.round((v - lo) / h) * h
    return (snap(y_in[0], spec.y0, spec.dy), snap(y_in[1], spec.y0, spec.dy),
            snap(x_out[0], spec.x0, spec.dx), snap(x_out[1], spec.x0, spec.dx))


def throughflow_boundary(spec, y_in=(0.9, 1.0), x_out=(0.0, 0.025), H_in=0.9, C_in=0.3):
    """Policies and inflow states for throughflow_velocity: inflow/outflow on
    the left faces that overlap y_in, outflow-only on the top faces that
    overlap x_out, zero flux elsewhere.  Returns (bc dict, (inflow_H,
    inflow_C)) in the layouts of Transport.set_boundary / set_inflow."""
    ya, _, _, xb = _snapped_ramps(spec, y_in, x_out)
    ys = spec.y0 + np.arange(spec.ny + 1) * spec.dy
    xs = spec.x0 + np.arange(spec.nx + 1) * spec.dx
    left = np.where(ys[1:] > ya + 0.5 * spec.dy, 2, 0).astype(np.uint8)
    top = np.where(xs[:-1] < xb - 0.5 * spec.dx, 1, 0).astype(np.uint8)
    bc = {"bottom": np.zeros(spec.nx, np.uint8), "right": np.zeros(spec.ny, np.uint8),
          "top": top, "left": left}
    n = 2 * (spec.nx + spec.ny) * spec.q
    return bc, (np.full(n, H_in), np.full(n, C_in))


def _flow_fields(spec, x, y, stokes, darcy, phi_max):
    """Analytic flow at points: Stokes velocity of psi = stokes*sin(pi X)sin(pi Y)
    (X, Y the unit coordinates), an upward Darcy flux whose x-component
    vanishes on the side walls, and a flow porosity in [0, phi_max]."""
    Lx, Ly = spec.x1 - spec.x0, spec.y1 - spec.y0
    X, Y = (x - spec.x0) / Lx, (y - spec.y0) / Ly
    # the walls carry exactly no Stokes flow (sin(pi) is not 0 in floating point)
    sX = np.where(np.isclose(X, 0.0, atol=1e-12) | np.isclose(X, 1.0, atol=1e-12), 0.0,
                  np.sin(np.pi * X))
    sY = np.where(np.isclose(Y, 0.0, atol=1e-12) | np.isclose(Y, 1.0, atol=1e-12), 0.0,
                  np.sin(np.pi * Y))
    su = stokes * np.pi / Ly * sX * np.cos(np.pi * Y)
    sv = -stokes * np.pi / Lx * np.cos(np.pi * X) * sY
    du = 0.2 * darcy * np.sin(2 * np.pi * X) * np.cos(np.pi * Y)
    dv = darcy * (1.0 + 0.5 * np.cos(np.pi * X))
    phi = phi_max * (0.5 + 0.5 * np.sin(2 * np.pi * X) * np.cos(1.5 * np.pi * Y)) ** 2
    return np.stack([su, sv], -1), np.stack([du, dv], -1), phi


def coupled_flow(spec, stokes=0.05, darcy=0.05, phi_max=0.2, cell_points=5):
    """Flow samples for the coupled law in the reference's FlowPointSample layout:
    edges by id (horizontal first, j*nx+i; then vertical nx*(ny+1)+j*(nx+1)+i)
    and canonical q (increasing x or y), cells by id and b*n+a."""
    g, _ = gauss_legendre(spec.q)
    t = (1.0 + g) / 2.0
    nx, ny = spec.nx, spec.ny
    jh, ih = np.meshgrid(np.arange(ny + 1), np.arange(nx), indexing="ij")
    xh = spec.x0 + (ih[..., None] + t) * spec.dx
    yh = np.broadcast_to((spec.y0 + jh * spec.dy)[..., None], xh.shape)
    jv, iv = np.meshgrid(np.arange(ny), np.arange(nx + 1), indexing="ij")
    yv = spec.y0 + (jv[..., None] + t) * spec.dy
    xv = np.broadcast_to((spec.x0 + iv * spec.dx)[..., None], yv.shape)
    x = np.concatenate([xh.reshape(-1, spec.q), xv.reshape(-1, spec.q)])
    y = np.concatenate([yh.reshape(-1, spec.q), yv.reshape(-1, spec.q)])
    es, ed, ep = _flow_fields(spec, x, y, stokes, darcy, phi_max)
    gc, _ = gauss_legendre(cell_points)
    tc = (1.0 + gc) / 2.0
    jc, ic = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    xc = spec.x0 + (ic.reshape(-1, 1, 1) + tc[None, None, :]) * spec.dx
    yc = spec.y0 + (jc.reshape(-1, 1, 1) + tc[None, :, None]) * spec.dy
    xc, yc = np.broadcast_arrays(xc, yc)
    cs, cd, cp = _flow_fields(spec, xc.reshape(nx * ny, -1), yc.reshape(nx * ny, -1), stokes,
                              darcy, phi_max)
    return dict(edge_stokes=es, edge_darcy=ed, edge_phi=ep, cell_stokes=cs, cell_darcy=cd,
                cell_phi=cp)


def coupled_boundary(spec, H_in=0.55, C_in=0.2, T_bottom=0.62):
    """Boundary data of the coupled workload: the Darcy flux enters through the
    bottom (inflow/outflow policy, prescribed H and C) and leaves through the
    top (outflow only); the side walls carry no normal flow (zero flux).  The
    bottom holds a prescribed temperature, the other sides are insulated.
    Returns (policies, (inflow_H, inflow_C), thermal types, thermal values) in
    the [side][face][q] layouts."""
    nx, ny, q = spec.nx, spec.ny, spec.q
    bc = {"bottom": np.full(nx, 2, np.uint8), "right": np.zeros(ny, np.uint8),
          "top": np.ones(nx, np.uint8), "left": np.zeros(ny, np.uint8)}
    n = 2 * (nx + ny) * q
    ttypes = np.zeros(n, np.uint8)
    tvals = np.zeros(n)
    ttypes[:nx * q] = 1
    tvals[:nx * q] = T_bottom
    return bc, (np.full(n, H_in), np.full(n, C_in)), ttypes, tvals


def coupled_state(spec, c_lo=0.05, c_hi=0.35):
    """H of the synthetic configs (partly above the eutectic temperature 0.6)
    and C seeded uniform in [c_lo, c_hi], inside (0, Xe) for Xe = 0.4."""
    u = unit_uniform(spec.seed + 77, np.arange(spec.cells, dtype=np.uint64))
    return enthalpy(spec), (c_lo + (c_hi - c_lo) * u).reshape(spec.ny, spec.nx)


def refresh_amplitudes(count, seed=5, drift=0.1):
    """psi amplitude factor of each velocity refresh: 1 + drift*(2u - 1)."""
    return 1.0 + drift * (2.0 * unit_uniform(seed + 31337, np.arange(count)) - 1.0)


def config_phase(**kw):
    """Phase parameters of the synthetic configs: latent heat LD = 1,
    eutectic at Xe = 0.4, Tep = 0.6 (Tm0 = 1600 K, Te0 = 600 K), P = 0, and
    the hypereutectic mirror for the half of C above Xe."""
    p = dict(Tm0=1600.0, Te0=600.0, cp=1200.0, L=1.2e6, Xe=0.4, pressure=0.0, hypereutectic=1)
    p.update(kw)
    return default_phase(**p)


def config1():
    """20x20 cells on [0,0.5]x[0,1], 10 SSP-RK3 steps at CFL 0.4."""
    return MeshSpec(20, 20, seed=1), 10, 0.4


def config2():
    """2048x4096 cells, 100 SSP-RK3 steps at CFL 0.4."""
    return MeshSpec(2048, 4096, seed=2), 100, 0.4


def config5():
    """Full-chip run: 8192x8192 cells, 1000 SSP-RK3 steps at CFL 0.4, face
    velocities refreshed every 50 steps (psi amplitude drift +-10%), outflow
    through the top segment x in [0, 0.025]."""
    return MeshSpec(8192, 8192, seed=5), 1000, 0.4, 50


def batch_specs(count=256, seed=3):
    """The batch workload: nx in {40,...,400} step 20, ny log-uniform in [5,800],
    amplitudes x U[0.5,1.5], each mesh seeded separately."""
    u = unit_uniform(seed, np.arange(4 * count)).reshape(count, 4)
    nxs = 40 + 20 * np.minimum((u[:, 0] * 19).astype(int), 18)
    nys = np.clip(np.rint(np.exp(np.log(5) + u[:, 1] * (np.log(800) - np.log(5)))), 5, 800).astype(int)
    specs = []
    for k in range(count):
        specs.append(MeshSpec(int(nxs[k]), int(nys[k]), seed=seed * 100003 + k,
                              h_amplitude=0.5 + u[k, 2], psi_amplitude=0.5 + u[k, 3]))
    return specs
