// Device ML-WENO reconstruction of one target cell on a uniform rectangular
// mesh, with the production candidate families (src/couple/state_reconstruction.cpp:64-68):
//   large: 3x3 Q2 candidate centred on the target, ReferenceSquare indicator;
//   small: four 2x2 Q1 candidates with offsets (0,-1),(0,0),(-1,-1),(-1,0),
//          TargetCell indicator.
//
// The reference builds each candidate from a moment matrix solved with dgesv
// and integrates its indicator with Gauss quadrature
// (src/mlweno/tensorstencilpoly.cpp:146-231, 318-398).  On rectangles both are
// exact closed forms: the moment matrix is a tensor product of 1-D 3x3 (or
// 2x2) systems, and the indicator is a sum of squares of the polynomial's
// coefficients.  Coordinates are normalised by h = sqrt(area) about the
// target centre, as the reference's polynomials are.  Candidates whose
// stencil leaves the mesh are discarded, not shifted
// (src/mlweno/reconstruction.cpp:150-153).
//
// The reconstruction is returned in centred form R(x) = t + P(x), P being the
// blend of the candidates' centred polynomials (src/mlweno/reconstruction.cpp:489-496),
// reduced to the four face lines: P on a face is a quadratic in the face's
// tangential coordinate s, P = c0 + s*(c1 + s*c2).
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "mp_common.h"

namespace mp {

template <typename R>
struct FaceLines {
    R c[4][3];   // faces 0 bottom (s = x), 1 right (s = y), 2 top (s = x), 3 left (s = y)
};

// Candidate validity bits for target (i, j): bit 0 large, bits 1..4 small k.
__device__ __forceinline__ unsigned candidate_mask(int i, int j, int nx, int ny)
{
    const bool il = i >= 1, ir = i <= nx - 2, jb = j >= 1, jt = j <= ny - 2;
    unsigned m = 0;
    if (il && ir && jb && jt) m |= 1u;
    if (ir && jb) m |= 2u;    // (0,-1)
    if (ir && jt) m |= 4u;    // (0,0)
    if (il && jb) m |= 8u;    // (-1,-1)
    if (il && jt) m |= 16u;   // (-1,0)
    return m;
}

// 1/s to within an ulp for a normal s: the hardware reciprocal refined by two
// Newton steps, each of which doubles the number of correct bits.
__device__ __forceinline__ double rcp64(double s)
{
    double r = __builtin_amdgcn_rcp(s);
    double e = fma(-s, r, 1.0);
    r = fma(r, e, r);
    e = fma(-s, r, 1.0);
    return fma(r, e, r);
}

// u^p (p >= 1) by repeated squaring; p is the same for every lane.
__device__ __forceinline__ double ipow(double u, int p)
{
    double r = (p & 1) ? u : 1.0;
    for (p >>= 1; p; p >>= 1) {
        u *= u;
        if (p & 1) r *= u;
    }
    return r;
}

template <typename R> __device__ __forceinline__ R wlog(R x);
template <typename R> __device__ __forceinline__ R wexp(R x);
template <> __device__ __forceinline__ double wlog<double>(double x) { return log(x); }
template <> __device__ __forceinline__ double wexp<double>(double x) { return exp(x); }
// FP32 weights work in base 2 with the hardware transcendental unit.
template <> __device__ __forceinline__ float wlog<float>(float x) { return __builtin_amdgcn_logf(x); }
template <> __device__ __forceinline__ float wexp<float>(float x) { return __builtin_amdgcn_exp2f(x); }
template <typename R> __device__ __forceinline__ R wscale();   // log(2)-base factor
template <> __device__ __forceinline__ double wscale<double>() { return 1.0; }
template <> __device__ __forceinline__ float wscale<float>() { return 1.4426950408889634f; }

// Centred blend of one target: B[b][a] is the coefficient of x^a y^b.
// w[dj+1][di+1] holds the cell averages around the target; entries of
// discarded candidates may be any finite value.  sigma_out/omega_out
// (optional) receive the five candidates' indicators and weights.
template <typename R>
__device__ __forceinline__ void blend(const MeshDev& m, const double (&w)[3][3], unsigned mask,
                                      R (&B)[3][3], uint32_t& err, double* sigma_out = nullptr,
                                      double* omega_out = nullptr)
{
    const double t = w[1][1];
    R c[3][3];
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int i = 0; i < 3; ++i) c[j][i] = static_cast<R>(w[j][i] - t);
    c[1][1] = R(0);

    // Large candidate: tensor-product 1-D quadratics, x then y.
    const R kx0 = R(m.kx0), kx1 = R(m.kx1), kx2 = R(m.kx2);
    const R ky0 = R(m.ky0), ky1 = R(m.ky1), ky2 = R(m.ky2);
    R g[3][3];   // g[row][a]
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const R l = c[j][0], mm = c[j][1], r = c[j][2];
        const R a2 = (r + l - R(2) * mm) * kx2;
        g[j][2] = a2;
        g[j][1] = (r - l) * kx1;
        g[j][0] = mm - a2 * kx0;
    }
    R P[3][3];   // P[b][a]: coefficient of x^a y^b
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const R l = g[0][a], mm = g[1][a], r = g[2][a];
        const R b2 = (r + l - R(2) * mm) * ky2;
        P[2][a] = b2;
        P[1][a] = (r - l) * ky1;
        P[0][a] = mm - b2 * ky0;
    }
    // ReferenceSquare indicator of the Q2 candidate: the mean over the unit
    // square of the squared first and second derivatives, as a sum of squares.
    R sigL;
    {
        const R p10 = P[0][1], p01 = P[1][0], p20 = P[0][2], p02 = P[2][0], p11 = P[1][1];
        const R p21 = P[1][2], p12 = P[2][1], p22 = P[2][2];
        const R e1 = p10 + p12 * R(1.0 / 12.0);
        const R e2 = p01 + p21 * R(1.0 / 12.0);
        const R x20 = R(2) * (p20 + p22 * R(1.0 / 12.0));
        const R y02 = R(2) * (p02 + p22 * R(1.0 / 12.0));
        sigL = e1 * e1 + e2 * e2 + R(13.0 / 12.0) * (x20 * x20 + y02 * y02) +
               R(7.0 / 6.0) * p11 * p11 + R(7.0 / 10.0) * (p12 * p12 + p21 * p21) +
               R(43.0 / 270.0) * p22 * p22;
    }
    // Small candidates: bilinear through the centres of a 2x2 block that holds
    // the target; TargetCell indicator = q10^2 + q01^2 + q11^2*(aw^2+bw^2)/12.
    const R iaw = R(m.inv_aw), ibw = R(m.inv_bw), iawbw = R(m.inv_awbw), q11w = R(m.q11w);
    R q10[4], q01[4], q11[4], sig[5];
    sig[0] = sigL;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int sx = (k < 2) ? 1 : -1;
        const int sy = (k == 0 || k == 2) ? -1 : 1;
        const R cx = c[1][1 + sx], cy = c[1 + sy][1], cxy = c[1 + sy][1 + sx];
        q10[k] = (sx > 0 ? cx : -cx) * iaw;
        q01[k] = (sy > 0 ? cy : -cy) * ibw;
        const R d = cxy - cx - cy;
        q11[k] = (sx * sy > 0 ? d : -d) * iawbw;
        sig[k + 1] = q10[k] * q10[k] + q01[k] * q01[k] + q11[k] * q11[k] * q11w;
    }
    R om[6];
    bool direct = false;
    if constexpr (std::is_same<R, double>::value) {
        // alpha_k = d_k / (sigma_k + epsA)^p_k with the reference's integer
        // exponents, normalised: the same weights as the log form below to
        // rounding, without FP64 log and exp.  Targets with a denominator out
        // of [wlo, whi] (or not finite) take the log form.
        if (m.wdirect) {
            bool ok = true;
            double sum = 0.0;
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                const bool on = k == 0 ? (mask & 1u) : (mask & (2u << (k - 1)));
                const double s = sig[k] + m.epsA;
                ok = ok && (!on || (s >= m.wlo && s <= m.whi));
                om[k] = on ? (k == 0 ? m.dL : m.dS[k - 1]) * ipow(rcp64(s), k == 0 ? m.ipL : m.ipS) : 0.0;
                sum += om[k];
            }
            om[5] = m.aC;
            sum += om[5];
            if (ok && sum > 0.0) {
                const double inv = 1.0 / sum;
#pragma unroll
                for (int k = 0; k < 6; ++k) om[k] *= inv;
                direct = true;
            }
        }
    }
    if (!direct) {
        // Normalised weights in log space (src/mlweno/reconstruction.cpp:212-265).
        const R sc = wscale<R>();
        const R epsA = R(m.epsA);
        R la[6];
        la[0] = (mask & 1u) ? R(m.ldL) * sc - R(m.pL) * wlog<R>(sig[0] + epsA) : R(-INFINITY);
#pragma unroll
        for (int k = 0; k < 4; ++k)
            la[k + 1] = (mask & (2u << k)) ? R(m.ldS[k]) * sc - R(m.pS) * wlog<R>(sig[k + 1] + epsA)
                                            : R(-INFINITY);
        la[5] = m.use_constant ? R(m.ldC) * sc - R(m.pC) * wlog<R>(epsA) : R(-INFINITY);
        R mx = la[0];
#pragma unroll
        for (int k = 1; k < 6; ++k) mx = la[k] > mx ? la[k] : mx;
        if (!(mx > R(-INFINITY)) || !isfinite(mx)) {
            err |= kErrWeights;
            mx = R(0);
        }
        R sum = R(0);
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            om[k] = la[k] > R(-INFINITY) ? wexp<R>(la[k] - mx) : R(0);
            sum += om[k];
        }
        const R inv = R(1) / sum;
#pragma unroll
        for (int k = 0; k < 6; ++k) om[k] *= inv;
    }
    if (sigma_out) {
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            const bool on = k == 0 ? (mask & 1u) : (mask & (2u << (k - 1)));
            sigma_out[k] = on ? double(sig[k]) : 0.0;
            omega_out[k] = double(om[k]);
        }
    }
    // Blend (Q1 candidates embed in the Q2 monomials; the constant candidate's
    // centred polynomial is zero).
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
        for (int a = 0; a < 3; ++a) B[b][a] = om[0] * P[b][a];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        B[0][1] += om[k + 1] * q10[k];
        B[1][0] += om[k + 1] * q01[k];
        B[1][1] += om[k + 1] * q11[k];
    }
}

// The blend on the four face lines.
template <typename R>
__device__ __forceinline__ void face_lines(const MeshDev& m, const R (&B)[3][3], FaceLines<R>& out)
{
    const R hx = R(m.hx), hy = R(m.hy);
#pragma unroll
    for (int b = 0; b < 3; ++b) {
        out.c[1][b] = B[b][0] + hx * (B[b][1] + hx * B[b][2]);   // right, x = +hx
        out.c[3][b] = B[b][0] - hx * (B[b][1] - hx * B[b][2]);   // left,  x = -hx
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        out.c[2][a] = B[0][a] + hy * (B[1][a] + hy * B[2][a]);   // top,    y = +hy
        out.c[0][a] = B[0][a] - hy * (B[1][a] - hy * B[2][a]);   // bottom, y = -hy
    }
}

// Reconstruct one target onto its face lines (see blend).
template <typename R>
__device__ __forceinline__ void reconstruct(const MeshDev& m, const double (&w)[3][3],
                                            unsigned mask, FaceLines<R>& out, uint32_t& err,
                                            double* sigma_out = nullptr,
                                            double* omega_out = nullptr)
{
    R B[3][3];
    blend<R>(m, w, mask, B, err, sigma_out, omega_out);
    face_lines<R>(m, B, out);
}

template <typename R>
__device__ __forceinline__ R line_at(const R (&c)[3], R s)
{
    return c[0] + s * (c[1] + s * c[2]);
}

}  // namespace mp
