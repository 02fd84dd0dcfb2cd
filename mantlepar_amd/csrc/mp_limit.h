// Mean-preserving bound limiter of the coupled reconstruction
// (StateReconstruction::Limit, src/couple/state_reconstruction.cpp:136-192)
// on a uniform rectangular mesh.
//
// The reference evaluates each cell's reconstruction at a point set (centre,
// cellPointsPerAxis^2 Gauss points, the edge Gauss points and the in-cell
// normal-line diffusion samples of every edge: 86 points at the defaults),
// and takes theta = min over points of (bound - mean)/(R(p) - mean) where
// R(p) leaves [lo, hi].  For bound > mean that ratio falls as R(p) moves away,
// so the minimum is reached at the largest and smallest R(p): the device keeps
// only the two extremes of the centred blend P = R - mean and divides once.
//
// Every point of the set lies on a line x = const or y = const of the cell,
// so P is evaluated as a quadratic along that line after one reduction of the
// 3x3 coefficient block: two FMAs per point.
#pragma once

#include <hip/hip_runtime.h>

#include "mp_common.h"

namespace mp {

// Extremes of P(x, y) = sum B[b][a] x^a y^b along x = const over ys[0..n).
template <int N>
__device__ __forceinline__ void column_extremes(const double (&B)[3][3], double x, const double* ys,
                                                int n, double& lo, double& hi)
{
    const double q0 = B[0][0] + x * (B[0][1] + x * B[0][2]);
    const double q1 = B[1][0] + x * (B[1][1] + x * B[1][2]);
    const double q2 = B[2][0] + x * (B[2][1] + x * B[2][2]);
#pragma unroll
    for (int k = 0; k < (N > 0 ? N : kMaxQ); ++k) {
        if (N == 0 && k >= n) break;
        const double y = ys[k];
        const double v = q0 + y * (q1 + y * q2);
        lo = fmin(lo, v);
        hi = fmax(hi, v);
    }
}

// Extremes along y = const over xs[0..n).
template <int N>
__device__ __forceinline__ void row_extremes(const double (&B)[3][3], double y, const double* xs,
                                             int n, double& lo, double& hi)
{
    const double r0 = B[0][0] + y * (B[1][0] + y * B[2][0]);
    const double r1 = B[0][1] + y * (B[1][1] + y * B[2][1]);
    const double r2 = B[0][2] + y * (B[1][2] + y * B[2][2]);
#pragma unroll
    for (int k = 0; k < (N > 0 ? N : kMaxQ); ++k) {
        if (N == 0 && k >= n) break;
        const double x = xs[k];
        const double v = r0 + x * (r1 + x * r2);
        lo = fmin(lo, v);
        hi = fmax(hi, v);
    }
}

// Limiter factor of one field of one cell.  w: the 3x3 cell averages around
// the cell, with clamped (in-mesh) duplicates where the window leaves the
// mesh, so min/max over w are the reference's in-mesh neighbourhood bounds.
// B: the centred blend.  sides: bit 0 bottom, 1 right, 2 top, 3 left edge on
// the mesh boundary (its diffusion sampling has n in-cell samples, not n/2).
// upper: +inf for H, Xe for C.  NQ > 0 fixes the edge point count.
template <int NQ>
__device__ __forceinline__ double limit_factor(const MeshDev& m, const double (&w)[3][3],
                                               const double (&B)[3][3], unsigned sides,
                                               double upper, uint32_t& err)
{
    const double t = w[1][1];
    if (!(t >= 0.0 && t <= upper)) err |= kErrInadmissible;
    double lo = w[0][0], hi = w[0][0];
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            lo = fmin(lo, w[j][i]);
            hi = fmax(hi, w[j][i]);
        }
    lo = fmax(0.0, lo);
    hi = fmin(upper, hi);
    bool finite = true;
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int i = 0; i < 3; ++i) finite &= isfinite(B[j][i]);
    if (!finite) err |= kErrNonFinite;

    const int nq = NQ > 0 ? NQ : m.q;
    double pmin = B[0][0], pmax = B[0][0];                // cell centre
    for (int a = 0; a < m.lim_ncp; ++a)                  // cell Gauss points
        column_extremes<0>(B, m.lim_cx[a], m.lim_cy, m.lim_ncp, pmin, pmax);
    column_extremes<NQ>(B, m.hx, m.sv, nq, pmin, pmax);  // right and left edge points
    column_extremes<NQ>(B, -m.hx, m.sv, nq, pmin, pmax);
    row_extremes<NQ>(B, m.hy, m.sh, nq, pmin, pmax);     // top and bottom edge points
    row_extremes<NQ>(B, -m.hy, m.sh, nq, pmin, pmax);
    {   // right face samples; the left face's mirror them
        const double* xs = (sides & 2u) ? m.lim_xb : m.lim_xi;
        const int n = (sides & 2u) ? m.lim_nb : m.lim_ni;
        for (int k = 0; k < n; ++k) column_extremes<NQ>(B, xs[k], m.sv, nq, pmin, pmax);
        const double* xl = (sides & 8u) ? m.lim_xb : m.lim_xi;
        const int nl = (sides & 8u) ? m.lim_nb : m.lim_ni;
        for (int k = 0; k < nl; ++k) column_extremes<NQ>(B, -xl[k], m.sv, nq, pmin, pmax);
    }
    {   // top face samples; the bottom face's mirror them
        const double* ys = (sides & 4u) ? m.lim_yb : m.lim_yi;
        const int n = (sides & 4u) ? m.lim_nb : m.lim_ni;
        for (int k = 0; k < n; ++k) row_extremes<NQ>(B, ys[k], m.sh, nq, pmin, pmax);
        const double* yl = (sides & 1u) ? m.lim_yb : m.lim_yi;
        const int nl = (sides & 1u) ? m.lim_nb : m.lim_ni;
        for (int k = 0; k < nl; ++k) row_extremes<NQ>(B, -yl[k], m.sh, nq, pmin, pmax);
    }
    double theta = 1.0;
    const double dhi = hi - t, dlo = lo - t;
    if (pmax > dhi) theta = fmin(theta, dhi / pmax);
    if (pmin < dlo) theta = fmin(theta, dlo / pmin);
    // rounded inward so admissible 0 / Xe bounds survive the evaluation
    return theta < 1.0 ? fmax(0.0, theta * (1.0 - 1e-13)) : 1.0;
}

}  // namespace mp
