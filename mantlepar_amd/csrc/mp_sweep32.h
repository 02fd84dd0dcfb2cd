// Mixed-precision stage sweep (FP32 reconstruction and flux corrections, FP64
// state, first-order flux part and update), written for gfx950.
//
// Layout as the FP64 sweep: one wavefront owns a strip of 62 columns plus a
// halo lane on each side and walks up a band of rows.  What differs is how
// little work each cell repeats:
//  * every lane loads only its own column of each new row; the left and right
//    neighbours arrive through DPP wavefront shifts (full-rate VALU moves, no
//    LDS), and only the two halo lanes load one extra column;
//  * the FP64 differences a reconstruction needs are formed once per row, not
//    once per cell: each row keeps u, the neighbour differences dl, dr and its
//    x-direction quadratic (a1, a2), and consecutive rows share the vertical
//    difference e.  The three cells that use a row (rows r-1, r, r+1 around
//    it) read these instead of recomputing eight differences each;
//  * face flux moments arrive as one 16-byte record per face (the one-sided
//    form of the local LF speed), 32 bytes where both sides carry weight.
//
// The arithmetic is the production ML-WENO of the reference
// (src/mlweno/reconstruction.cpp:212-265, 489-496) in the closed forms of
// mp_recon.h, regrouped around the per-row data: the small candidates' mixed
// term c_xy - c_x - c_y is the difference of two rows' neighbour
// differences, and their x and y slopes take two values each.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "mp_common.h"
#include "mp_device.h"
#include "mp_launch.h"
#include "mp_phase_items.h"
#include "mp_recon.h"

namespace mp {
namespace {

// Wavefront shifts (DPP wave_shr:1 and wave_shl:1).  Lane L receives lane
// L-1 (shr) or L+1 (shl); the end lane, which has no source, receives 0
// (bound_ctrl), so one DPP move does the shift with no copy of the operand
// first.  Every use discards the end lane's value: lane 0 (shr) and lane 63
// (shl) are the halo lanes, which take their outer neighbour from memory and
// own no cell.  All lanes must be active, so callers use them only in
// wave-uniform control flow.
__device__ __forceinline__ int shr1_b32(int v) { return __builtin_amdgcn_mov_dpp(v, 0x138, 0xf, 0xf, true); }
__device__ __forceinline__ int shl1_b32(int v) { return __builtin_amdgcn_mov_dpp(v, 0x130, 0xf, 0xf, true); }

__device__ __forceinline__ double shr1(double v)
{
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = unsigned(shr1_b32(int(unsigned(b)))), hi = unsigned(shr1_b32(int(unsigned(b >> 32))));
    return __builtin_bit_cast(double, (static_cast<unsigned long long>(hi) << 32) | lo);
}
__device__ __forceinline__ double shl1(double v)
{
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = unsigned(shl1_b32(int(unsigned(b)))), hi = unsigned(shl1_b32(int(unsigned(b >> 32))));
    return __builtin_bit_cast(double, (static_cast<unsigned long long>(hi) << 32) | lo);
}
__device__ __forceinline__ float shl1(float v) { return __builtin_bit_cast(float, shl1_b32(__builtin_bit_cast(int, v))); }

// Per-mesh constants in FP32, converted once per work item.
struct K32 {
    float kx0, kx1, kx2, ky0, ky1, ky2;
    float iaw, ibw, hx, hy;
    float q11s, q11u;   // scale of the mixed term's row differences in the indicator, and its inverse
    float epsA, pL, pS, ldL, ldS[4], laC;
    bool use_constant;
};

__device__ __forceinline__ K32 make_k32(const MeshDev& m)
{
    K32 k;
    k.kx0 = float(m.kx0); k.kx1 = float(m.kx1); k.kx2 = float(m.kx2);
    k.ky0 = float(m.ky0); k.ky1 = float(m.ky1); k.ky2 = float(m.ky2);
    k.iaw = float(m.inv_aw); k.ibw = float(m.inv_bw);
    k.hx = float(m.hx); k.hy = float(m.hy);
    // q11 * q11 * q11w == (q11 * sqrt(q11w))^2: the indicator squares the
    // scaled difference, the blend scales it back
    k.q11s = float(m.q11s);
    k.q11u = float(m.q11u);
    const float sc = wscale<float>();
    k.epsA = float(m.epsA);
    k.pL = float(m.pL);
    k.pS = float(m.pS);
    k.ldL = float(m.ldL) * sc;
#pragma unroll
    for (int s = 0; s < 4; ++s) k.ldS[s] = float(m.ldS[s]) * sc;
    k.use_constant = m.use_constant != 0;
    k.laC = k.use_constant ? float(m.ldC) * sc - float(m.pC) * wlog<float>(k.epsA) : -INFINITY;
    return k;
}

// One row of one field at this lane's column.
struct Row32 {
    double u;        // cell average
    float dl, dr;    // left / right neighbour minus u (FP64 difference, rounded once)
    float a1, a2;    // x quadratic through the three averages: slope and curvature terms
    float e;         // u minus the average of the row below
};

__device__ __forceinline__ Row32 make_row(double u, double outer, double below, bool first, bool last,
                                          const K32& k)
{
    const double l = shr1(u), r = shl1(u);
    Row32 R;
    R.u = u;
    R.e = float(u - below);
    R.dl = float((first ? outer : l) - u);
    R.dr = float((last ? outer : r) - u);
    R.a1 = (R.dr - R.dl) * k.kx1;
    R.a2 = (R.dr + R.dl) * k.kx2;
    return R;
}

// Centred blend B[b][a] (coefficient of x^a y^b) of the cell at row r from
// the rows below (R0), at (R1) and above (R2) it.  EDGE = false is the
// interior form: every candidate lies in the mesh and there is no constant
// candidate, so no weight needs masking.
template <bool EDGE>
__device__ __forceinline__ void blend32(const K32& k, const Row32& R0, const Row32& R1,
                                        const Row32& R2, unsigned mask, float (&B)[3][3],
                                        uint32_t& err)
{
    const float elo = R1.e, ehi = R2.e;   // u(r) - u(r-1), u(r+1) - u(r)
    // Large candidate: the rows' x quadratics, centred on the target, then y.
    const float g00 = -elo - R0.a2 * k.kx0, g10 = -R1.a2 * k.kx0, g20 = ehi - R2.a2 * k.kx0;
    float P[3][3];   // P[b][a]: coefficient of x^a y^b
    {
        const float l[3] = {g00, R0.a1, R0.a2}, mm[3] = {g10, R1.a1, R1.a2}, r[3] = {g20, R2.a1, R2.a2};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const float b2 = (r[a] + l[a] - 2.0f * mm[a]) * k.ky2;
            P[2][a] = b2;
            P[1][a] = (r[a] - l[a]) * k.ky1;
            P[0][a] = mm[a] - b2 * k.ky0;
        }
    }
    float sig[5];
    {
        const float p10 = P[0][1], p01 = P[1][0], p20 = P[0][2], p02 = P[2][0], p11 = P[1][1];
        const float p21 = P[1][2], p12 = P[2][1], p22 = P[2][2];
        // (13/12) (2 x)^2 = (13/3) x^2 for the second-derivative terms
        const float e1 = p10 + p12 * (1.0f / 12.0f);
        const float e2 = p01 + p21 * (1.0f / 12.0f);
        const float x20 = p20 + p22 * (1.0f / 12.0f);
        const float y02 = p02 + p22 * (1.0f / 12.0f);
        sig[0] = e1 * e1 + e2 * e2 + (13.0f / 3.0f) * (x20 * x20 + y02 * y02) +
                 (7.0f / 6.0f) * p11 * p11 + (7.0f / 10.0f) * (p12 * p12 + p21 * p21) +
                 (43.0f / 270.0f) * p22 * p22;
    }
    // Small candidates, offsets (0,-1),(0,0),(-1,-1),(-1,0): (sx, sy) =
    // (+,-), (+,+), (-,-), (-,+).  Slopes take two values per direction; the
    // mixed term is a difference of neighbour differences of two rows.
    const float q10p = R1.dr * k.iaw, q10m = -R1.dl * k.iaw;
    const float q01p = ehi * k.ibw, q01m = elo * k.ibw;
    float q11[4];   // mixed terms, scaled by sqrt(q11w)
    q11[0] = (R1.dr - R0.dr) * k.q11s;
    q11[1] = (R2.dr - R1.dr) * k.q11s;
    q11[2] = (R0.dl - R1.dl) * k.q11s;
    q11[3] = (R1.dl - R2.dl) * k.q11s;
    {
        const float xp = q10p * q10p, xm = q10m * q10m, yp = q01p * q01p, ym = q01m * q01m;
        sig[1] = xp + ym + q11[0] * q11[0];
        sig[2] = xp + yp + q11[1] * q11[1];
        sig[3] = xm + ym + q11[2] * q11[2];
        sig[4] = xm + yp + q11[3] * q11[3];
    }
    // Normalised weights in log space, base 2 (src/mlweno/reconstruction.cpp:212-265).
    float la[6], om[6], sum = 0.0f;
    if constexpr (EDGE) {
        la[0] = (mask & 1u) ? k.ldL - k.pL * wlog<float>(sig[0] + k.epsA) : -INFINITY;
#pragma unroll
        for (int s = 0; s < 4; ++s)
            la[s + 1] = (mask & (2u << s)) ? k.ldS[s] - k.pS * wlog<float>(sig[s + 1] + k.epsA) : -INFINITY;
        la[5] = k.laC;
        float mx = la[0];
#pragma unroll
        for (int s = 1; s < 6; ++s) mx = la[s] > mx ? la[s] : mx;
        if (!(mx > -INFINITY) || !isfinite(mx)) {
            err |= kErrWeights;
            mx = 0.0f;
        }
#pragma unroll
        for (int s = 0; s < 6; ++s) {
            om[s] = la[s] > -INFINITY ? wexp<float>(la[s] - mx) : 0.0f;
            sum += om[s];
        }
    } else {
        la[0] = k.ldL - k.pL * wlog<float>(sig[0] + k.epsA);
#pragma unroll
        for (int s = 0; s < 4; ++s) la[s + 1] = k.ldS[s] - k.pS * wlog<float>(sig[s + 1] + k.epsA);
        float mx = fmaxf(fmaxf(fmaxf(la[0], la[1]), fmaxf(la[2], la[3])), la[4]);
        if (!isfinite(mx)) {
            err |= kErrWeights;
            mx = 0.0f;
        }
#pragma unroll
        for (int s = 0; s < 5; ++s) {
            om[s] = wexp<float>(la[s] - mx);
            sum += om[s];
        }
        om[5] = 0.0f;
    }
    // sum lies in [1, 6] (the largest term is exp2(0) = 1): the hardware
    // reciprocal (1 ulp) needs none of a division's range fix-ups
    const float inv = __builtin_amdgcn_rcpf(sum);
#pragma unroll
    for (int s = 0; s < 5; ++s) om[s] *= inv;
    // Blend (Q1 candidates embed in the Q2 monomials; the constant candidate's
    // centred polynomial is zero).
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
        for (int a = 0; a < 3; ++a) B[b][a] = om[0] * P[b][a];
    B[0][1] += (om[1] + om[2]) * q10p + (om[3] + om[4]) * q10m;
    B[1][0] += (om[2] + om[4]) * q01p + (om[1] + om[3]) * q01m;
    B[1][1] += (om[1] * q11[0] + om[2] * q11[1] + om[3] * q11[2] + om[4] * q11[3]) * k.q11u;
}

// The blend's four face lines c0 + s*c1 + s^2*c2.
__device__ __forceinline__ void lines32(const K32& k, const float (&B)[3][3], FaceLines<float>& out)
{
    const float hx = k.hx, hy = k.hy;
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

template <bool EDGE>
__device__ __forceinline__ void recon32(const K32& k, const Row32& R0, const Row32& R1,
                                        const Row32& R2, unsigned mask, FaceLines<float>& out,
                                        uint32_t& err)
{
    float B[3][3];
    blend32<EDGE>(k, R0, R1, R2, mask, B, err);
    lines32(k, B, out);
}

// Extremes of P along x = const over ys[0..n) (FP32; the offsets are
// wave-uniform, read from the mesh record).  N > 0 fixes n at compile time.
template <int N>
__device__ __forceinline__ void col_ext32(const float (&B)[3][3], float x, const float* ys, int n,
                                          float& lo, float& hi)
{
    const float q0 = B[0][0] + x * (B[0][1] + x * B[0][2]);
    const float q1 = B[1][0] + x * (B[1][1] + x * B[1][2]);
    const float q2 = B[2][0] + x * (B[2][1] + x * B[2][2]);
#pragma unroll
    for (int k = 0; k < (N > 0 ? N : kMaxQ); ++k) {
        if (N == 0 && k >= n) break;
        const float y = ys[k];
        const float v = q0 + y * (q1 + y * q2);
        lo = fminf(lo, v);
        hi = fmaxf(hi, v);
    }
}

// Extremes along y = const over xs[0..n).
template <int N>
__device__ __forceinline__ void row_ext32(const float (&B)[3][3], float y, const float* xs, int n,
                                          float& lo, float& hi)
{
    const float r0 = B[0][0] + y * (B[1][0] + y * B[2][0]);
    const float r1 = B[0][1] + y * (B[1][1] + y * B[2][1]);
    const float r2 = B[0][2] + y * (B[1][2] + y * B[2][2]);
#pragma unroll
    for (int k = 0; k < (N > 0 ? N : kMaxQ); ++k) {
        if (N == 0 && k >= n) break;
        const float x = xs[k];
        const float v = r0 + x * (r1 + x * r2);
        lo = fminf(lo, v);
        hi = fmaxf(hi, v);
    }
}

// Limiter tables fixed at compile time (edge points NQ, cell points per
// axis NCP, normal-line samples NS) or read from the mesh record (all 0).
template <int NQ, int NCP, int NS> struct LimShape {
    static constexpr int nq = NQ, ncp = NCP, ns = NS;
};
using LimAny = LimShape<0, 0, 0>;
using LimDefault = LimShape<5, 5, 4>;     // q = 5, cellPointsPerAxis 5, 4 samples

// The bound limiter (StateReconstruction::Limit, src/couple/state_reconstruction.cpp:
// 136-192) on the mixed sweep's data: the limit_factor of mp_limit.h with the
// extremes of the centred blend at the same 86 points evaluated in FP32.  The
// neighbourhood bounds come relative to the cell mean from the rows' FP64
// differences (rounded once), so no FP64 average is re-read: dhi = max(0, the
// eight differences) capped at upper - t, dlo likewise floored at -t.  theta is
// rounded inward by 2^-22 so the limited FP32 values stay inside the bounds.
template <class S>
__device__ __forceinline__ float limit32(const MeshDev& m, const K32& k, const Row32& R0, const Row32& R1,
                                         const Row32& R2, const float (&B)[3][3], unsigned sides,
                                         double upper, uint32_t& err)
{
    constexpr int NQ = S::nq, NCP = S::ncp, NS = S::ns;
    const double t = R1.u;
    if (!(t >= 0.0 && t <= upper)) err |= kErrInadmissible;
    const float e0 = -R1.e, e2 = R2.e;        // centre rows below and above, minus t
    const float c00 = e0 + R0.dl, c01 = e0 + R0.dr, c20 = e2 + R2.dl, c21 = e2 + R2.dr;
    float dhi = fmaxf(fmaxf(fmaxf(0.0f, R1.dl), fmaxf(R1.dr, e0)), fmaxf(fmaxf(e2, c00), fmaxf(c01, fmaxf(c20, c21))));
    float dlo = fminf(fminf(fminf(0.0f, R1.dl), fminf(R1.dr, e0)), fminf(fminf(e2, c00), fminf(c01, fminf(c20, c21))));
    dhi = fminf(dhi, float(upper - t));
    dlo = fmaxf(dlo, float(-t));
    bool finite = true;
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int i = 0; i < 3; ++i) finite &= isfinite(B[j][i]);
    if (!finite) err |= kErrNonFinite;

    const int nq = NQ > 0 ? NQ : m.q;
    const int ncp = NCP > 0 ? NCP : m.lim_ncp;
    const int ni = NS > 0 ? NS / 2 : m.lim_ni, nb = NS > 0 ? NS : m.lim_nb;
    float pmin = B[0][0], pmax = B[0][0];                              // cell centre
#pragma unroll
    for (int a = 0; a < (NCP > 0 ? NCP : kMaxQ); ++a) {                // cell Gauss points
        if (NCP == 0 && a >= ncp) break;
        col_ext32<NCP>(B, m.lim32_cx[a], m.lim32_cy, ncp, pmin, pmax);
    }
    col_ext32<NQ>(B, k.hx, m.lim32_sv, nq, pmin, pmax);                 // right and left edge points
    col_ext32<NQ>(B, -k.hx, m.lim32_sv, nq, pmin, pmax);
    row_ext32<NQ>(B, k.hy, m.lim32_sh, nq, pmin, pmax);                 // top and bottom edge points
    row_ext32<NQ>(B, -k.hy, m.lim32_sh, nq, pmin, pmax);
    // normal-line samples: n/2 in-cell samples of an interior face, n of a
    // boundary face; the left and bottom faces mirror the right and top ones
    auto cols = [&](float sign, const float* xs, int n, auto nn) {
        constexpr int N = decltype(nn)::value;
#pragma unroll
        for (int s = 0; s < (N > 0 ? N : kMaxLimS); ++s) {
            if (N == 0 && s >= n) break;
            col_ext32<NQ>(B, sign * xs[s], m.lim32_sv, nq, pmin, pmax);
        }
    };
    auto rows = [&](float sign, const float* ys, int n, auto nn) {
        constexpr int N = decltype(nn)::value;
#pragma unroll
        for (int s = 0; s < (N > 0 ? N : kMaxLimS); ++s) {
            if (N == 0 && s >= n) break;
            row_ext32<NQ>(B, sign * ys[s], m.lim32_sh, nq, pmin, pmax);
        }
    };
    using Ni = std::integral_constant<int, NS / 2>;
    using Nb = std::integral_constant<int, NS>;
    if constexpr (NS > 0) {
        // Which table a face reads depends on the lane (a face on the mesh
        // boundary takes the boundary samples); read through a per-lane choice
        // of pointer, every offset would be a vector load waited for in turn.
        // With the table sizes fixed, each sample offset is a per-lane choice
        // between two uniform values, and only boundary faces take the samples
        // beyond the interior count.  The samples are visited in the same order.
        auto cols_fixed = [&](float sign, bool bnd) {
#pragma unroll
            for (int q = 0; q < NS; ++q) {
                if (q >= NS / 2 && !bnd) break;
                const float x = (q < NS / 2 && !bnd) ? m.lim32_xi[q] : m.lim32_xb[q];
                col_ext32<NQ>(B, sign * x, m.lim32_sv, nq, pmin, pmax);
            }
        };
        auto rows_fixed = [&](float sign, bool bnd) {
#pragma unroll
            for (int q = 0; q < NS; ++q) {
                if (q >= NS / 2 && !bnd) break;
                const float y = (q < NS / 2 && !bnd) ? m.lim32_yi[q] : m.lim32_yb[q];
                row_ext32<NQ>(B, sign * y, m.lim32_sh, nq, pmin, pmax);
            }
        };
        cols_fixed(1.0f, sides & 2u);
        cols_fixed(-1.0f, sides & 8u);
        rows_fixed(1.0f, sides & 4u);
        rows_fixed(-1.0f, sides & 1u);
    } else {
        if (sides & 2u) cols(1.0f, m.lim32_xb, nb, Nb{}); else cols(1.0f, m.lim32_xi, ni, Ni{});
        if (sides & 8u) cols(-1.0f, m.lim32_xb, nb, Nb{}); else cols(-1.0f, m.lim32_xi, ni, Ni{});
        if (sides & 4u) rows(1.0f, m.lim32_yb, nb, Nb{}); else rows(1.0f, m.lim32_yi, ni, Ni{});
        if (sides & 1u) rows(-1.0f, m.lim32_yb, nb, Nb{}); else rows(-1.0f, m.lim32_yi, ni, Ni{});
    }
    float theta = 1.0f;
    if (pmax > dhi) theta = fminf(theta, dhi / pmax);
    if (pmin < dlo) theta = fminf(theta, dlo / pmin);
    return theta < 1.0f ? fmaxf(0.0f, theta * (1.0f - 0x1p-22f)) : 1.0f;
}

// Limited reconstruction of one field: blend, limit factor, scaled face lines.
template <bool EDGE, class S>
__device__ __forceinline__ void recon32_limited(const MeshDev& m, const K32& k, const Row32& R0,
                                                const Row32& R1, const Row32& R2, unsigned mask,
                                                unsigned sides, double upper, FaceLines<float>& out,
                                                uint32_t& err)
{
    float B[3][3];
    blend32<EDGE>(k, R0, R1, R2, mask, B, err);
    const float th = limit32<S>(m, k, R0, R1, R2, B, sides, upper, err);
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
        for (int a = 0; a < 3; ++a) B[b][a] *= th;
    lines32(k, B, out);
}

// The record a sweep fetches per face: the one-sided form under the local LF
// speed (FULL = false), the full record under the global speed.
template <bool FULL> struct FaceFetch { using type = FaceRec; };
template <> struct FaceFetch<true> { using type = FaceMom; };

// Both fields through one interior face from its moment record: first-order
// part in FP64, high-order correction in FP32.
// D: thermal diffusion coefficient of H on this face (DIFF = false: none).
template <bool DIFF>
__device__ __forceinline__ void face_flux(const FaceMom& f, double D, double tLH,
                                          const float (&cLH)[3], double tRH,
                                          const float (&cRH)[3], double tLC,
                                          const float (&cLC)[3], double tRC,
                                          const float (&cRC)[3], double& FH, double& FC)
{
    // The sums are written as explicit fma chains: left to the compiler, a
    // sum of products may be contracted differently in different kernel
    // instances, and every instance must give the same bits.
    const float vp = float(f.vp), vm = float(f.vm);
    auto dot6 = [&](const float (&l)[3], const float (&r)[3]) {
        float c = l[0] * vp;
        c = fmaf(l[1], f.p1, c);
        c = fmaf(l[2], f.p2, c);
        c = fmaf(r[0], vm, c);
        c = fmaf(r[1], f.m1, c);
        return fmaf(r[2], f.m2, c);
    };
    const float ch = dot6(cLH, cRH), cc = dot6(cLC, cRC);
    const double adv = fma(tLH, f.vp, tRH * f.vm) + double(ch);
    FH = DIFF ? fma(D, tLH - tRH, adv) : adv;
    FC = fma(tLC, f.vp, tRC * f.vm) + double(cc);
}

// The same flux from the face's one-sided record (the local LF speed): the
// side that v's sign names is the only one with weight, so its centre and
// line are selected and dotted with the record's three moments, instead of
// expanding the record and multiplying the other side by zeros.  A flagged
// face (b NaN: weight on both sides) takes the full record.
template <bool DIFF>
__device__ __forceinline__ void face_flux_rec(const FaceRec& h, const FaceMom* full, size_t face, double D,
                                              double tLH, const float (&cLH)[3], double tRH,
                                              const float (&cRH)[3], double tLC,
                                              const float (&cLC)[3], double tRC,
                                              const float (&cRC)[3], double& FH, double& FC)
{
    if (__builtin_isnan(h.b)) {
        face_flux<DIFF>(gload(full + face), D, tLH, cLH, tRH, cRH, tLC, cLC, tRC, cRC, FH, FC);
        return;
    }
    const bool pos = h.v >= 0.0;
    const double tH = pos ? tLH : tRH, tC = pos ? tLC : tRC;
    const float v = float(h.v);
    const float ch = fmaf(pos ? cLH[2] : cRH[2], h.b, fmaf(pos ? cLH[1] : cRH[1], h.a, (pos ? cLH[0] : cRH[0]) * v));
    const float cc = fmaf(pos ? cLC[2] : cRC[2], h.b, fmaf(pos ? cLC[1] : cRC[1], h.a, (pos ? cLC[0] : cRC[0]) * v));
    const double adv = fma(tH, h.v, double(ch));
    FH = DIFF ? fma(D, tLH - tRH, adv) : adv;
    FC = fma(tC, h.v, double(cc));
}

// One work item of a stage.  The stage's phase and state sums come from
// phase_items_kernel after the last stage, so this sweep only writes the
// stage's boundary-flux partials (state-sum slots zero).  DIFF: the mesh has
// thermal diffusion or thermal boundary data; without it the diffusion terms
// are compiled out, which keeps the sweep within its register budget.
//
// SHARED (the one-launch step, where other items of the same launch, on any
// XCD, read what this item writes): the stage's input and output rows are
// read and written with relaxed agent-scope atomics, device-coherent accesses
// that bypass the caches that are not coherent across XCDs.  (Plain accesses
// in the first stage's loads and the last stage's stores, which only meet
// other launches, measured slower: 0.657 against 0.624 ms per 2048 x 4096
// step launch.)
// LIM: 0 no limiter, 1 the limiter with the mesh record's tables, 2 with the
// default tables fixed at compile time (LimDefault), 3 either, chosen per mesh
// as the single-mesh launch chooses (the batch's mixed lists), so that every
// mesh runs the arithmetic of its own single-mesh sweep.
// UNROLL: the row loop unrolled three times (the one-launch steps, which fit
// it in their register budget; the others would spill).  The arithmetic is
// the same either way: the sums whose contraction the compiler could choose
// differently per instance are written as explicit fma chains.
template <bool DIFF, bool FULL, bool SHARED = false, int LIM = 0, bool UNROLL = SHARED>
__device__ __forceinline__ void sweep32_item(const MeshDev& m, const StagePtrs& sp, int item, int strip, int j0,
                                             int j1)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int nx = m.nx, ny = m.ny;
    const int i = strip * kStripCells - 1 + lane;
    const bool col_ok = i >= 0 && i < nx;
    const bool owned = col_ok && lane >= 1 && lane <= kStripCells;
    const bool first = lane == 0, last = lane == kWave - 1;
    const int ic = clampi(i, 0, nx - 1);
    const int io = clampi(first ? i - 1 : i + 1, 0, nx - 1);   // halo lanes' outer column
    const size_t vxs = size_t(ny) * (nx + 1);
    const size_t vys = size_t(ny + 1) * nx;
    const K32 k = make_k32(m);
    uint32_t err = 0;
    // every lane's 3x3 neighbourhood inside the mesh (wave-uniform)
    const bool inner_strip = strip >= 1 && strip * kStripCells + kWave <= nx && !k.use_constant;

    auto load = [&](const double* base, int row, double& u, double& o) {
        const auto* p = gmem(base) + size_t(clampi(row, 0, ny - 1)) * nx;
        if constexpr (SHARED) {
            u = __hip_atomic_load(p + ic, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            o = (first || last) ? __hip_atomic_load(p + io, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0.0;
        } else {
            u = p[ic];
            o = (first || last) ? p[io] : 0.0;
        }
    };

    int r = j0 > 0 ? j0 - 1 : 0;
    Row32 H0, H1, H2, C0, C1, C2;
    {
        double u[3], o[3], v[3], w[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            load(sp.Hk, r - 1 + d, u[d], o[d]);
            load(sp.Ck, r - 1 + d, v[d], w[d]);
        }
        H0 = make_row(u[0], o[0], u[0], first, last, k);
        H1 = make_row(u[1], o[1], u[0], first, last, k);
        H2 = make_row(u[2], o[2], u[1], first, last, k);
        C0 = make_row(v[0], w[0], v[0], first, last, k);
        C1 = make_row(v[1], w[1], v[0], first, last, k);
        C2 = make_row(v[2], w[2], v[1], first, last, k);
    }

    // Face records and u_n are fetched one row ahead, like the averages, so
    // no load is consumed in the iteration that issues it.  Under the local LF
    // speed a lane fetches the 16-byte one-sided record and expands it where
    // it is used; under the global speed it fetches the full record.
    using Rec = typename FaceFetch<FULL>::type;
    const int ix = clampi(i + 1, 0, nx);           // vertical face x_{i+1}
    auto fy_index = [&](int row) { return size_t(row) * nx + ic; };
    auto fx_index = [&](int row) { return size_t(clampi(row, 0, ny - 1)) * (nx + 1) + ix; };
    auto load_fy = [&](int row) -> Rec {
        if constexpr (FULL) return gload(m.fy + fy_index(row));
        else return gload(m.ry + fy_index(row));
    };
    auto load_fx = [&](int row) -> Rec {
        if constexpr (FULL) return gload(m.fx + fx_index(row));
        else return gload(m.rx + fx_index(row));
    };
    const bool need_n = sp.a != 0.0;
    Rec fyc = load_fy(r), fxc = load_fx(r);
    double hn = 0.0, cn = 0.0;                     // u_n of the row being finished

    float topH[3] = {0.f, 0.f, 0.f}, topC[3] = {0.f, 0.f, 0.f};
    // (left - right + bottom) face fluxes of the row waiting for its top face,
    // summed in that order as soon as they are known.
    double accH = 0, accC = 0;
    double bH = 0, bC = 0;

    // One row of the sweep.  H0/C0, H1/C1, H2/C2 are rows r-1, r, r+1.
    // Unrolled, the row r+2 that the next call needs is built into H0/C0,
    // whose row is no longer used, and the caller passes the three rows
    // rotated, so the rotation is a renaming of registers, not moves; in the
    // plain loop the rows shift down and r+2 is built into H2/C2.
    auto row_step = [&](Row32& H0, Row32& H1, Row32& H2, Row32& C0, Row32& C1,
                        Row32& C2) __attribute__((always_inline)) {
        const bool have = r < ny;
        const bool preroll = r == j0 - 1;
        const bool shift = r + 1 <= j1;
        double nH = 0, oH = 0, nC = 0, oC = 0;
        if (shift) {
            load(sp.Hk, r + 2, nH, oH);
            load(sp.Ck, r + 2, nC, oC);
        }
        FaceLines<float> LH, LC;
        if constexpr (LIM != 0) {
            if (have) {
                // one field at a time: its blend, limit factor and face lines
                const unsigned sides = (r == 0 ? 1u : 0u) | (i == nx - 1 ? 2u : 0u) |
                                       (r == ny - 1 ? 4u : 0u) | (i == 0 ? 8u : 0u);
                uint32_t e2 = 0;
                auto limited = [&](auto shape) {
                    using S = decltype(shape);
                    if (inner_strip && r >= 1 && r <= ny - 2) {
                        recon32_limited<false, S>(m, k, H0, H1, H2, 31u, sides, INFINITY, LH, e2);
                        recon32_limited<false, S>(m, k, C0, C1, C2, 31u, sides, m.lim_cupper, LC, e2);
                    } else {
                        const unsigned mask = col_ok ? candidate_mask(i, r, nx, ny) : 2u;
                        recon32_limited<true, S>(m, k, H0, H1, H2, mask, sides, INFINITY, LH, e2);
                        recon32_limited<true, S>(m, k, C0, C1, C2, mask, sides, m.lim_cupper, LC, e2);
                    }
                };
                if constexpr (LIM == 1) limited(LimAny{});
                else if constexpr (LIM == 2) limited(LimDefault{});
                else if (limit_mode32(m) == 2) limited(LimDefault{});
                else limited(LimAny{});
                if (col_ok) err |= e2;
            } else {
#pragma unroll
                for (int f = 0; f < 4; ++f)
#pragma unroll
                    for (int c = 0; c < 3; ++c) { LH.c[f][c] = 0.f; LC.c[f][c] = 0.f; }
            }
        } else if (have && inner_strip && r >= 1 && r <= ny - 2) {
            recon32<false>(k, H0, H1, H2, 31u, LH, err);
            recon32<false>(k, C0, C1, C2, 31u, LC, err);
        } else if (have) {
            const unsigned mask = col_ok ? candidate_mask(i, r, nx, ny) : 2u;
            uint32_t e2 = 0;
            recon32<true>(k, H0, H1, H2, mask, LH, e2);
            // with full face records, the two reconstructions interleaved
            // spill past the register budget
            if constexpr (FULL) __builtin_amdgcn_sched_barrier(0);
            recon32<true>(k, C0, C1, C2, mask, LC, e2);
            if (col_ok) err |= e2;
        } else {
#pragma unroll
            for (int f = 0; f < 4; ++f)
#pragma unroll
                for (int c = 0; c < 3; ++c) { LH.c[f][c] = 0.f; LC.c[f][c] = 0.f; }
        }
        const double tH = H1.u, tC = C1.u;

        // Horizontal face y_r: top of row r-1 (centre H0.u), bottom of row r.
        double FvH = 0.0, FvC = 0.0;
        if (!preroll && col_ok) {
            if (r == 0 || r == ny) {
                const int side = r == 0 ? 0 : 2;
                const double sign = r == 0 ? -1.0 : 1.0;
                const double* vn = m.vny + size_t(r) * nx + ic;
                const double* inH = m.inflowH[side] ? m.inflowH[side] + size_t(ic) * m.q : nullptr;
                const double* inC = m.inflowC[side] ? m.inflowC[side] + size_t(ic) * m.q : nullptr;
                const uint8_t pol = m.bc[side] ? gmem(m.bc[side])[ic] : 0;
                const double oHf = (r == 0 ? boundary_flux<float>(m, vn, vys, sign, tH, LH.c[0], m.sh, m.wh, pol, inH, err)
                                           : boundary_flux<float>(m, vn, vys, sign, H0.u, topH, m.sh, m.wh, pol, inH, err)) +
                                   (DIFF ? thermal_flux(m, side, ic, r == 0 ? tH : H0.u, m.wh, m.dif_by) : 0.0);
                const double oCf = r == 0 ? boundary_flux<float>(m, vn, vys, sign, tC, LC.c[0], m.sh, m.wh, pol, inC, err)
                                          : boundary_flux<float>(m, vn, vys, sign, C0.u, topC, m.sh, m.wh, pol, inC, err);
                FvH = sign * oHf;
                FvC = sign * oCf;
                if (owned) { bH += oHf; bC += oCf; }
            } else {
                if constexpr (FULL)
                    face_flux<DIFF>(fyc, m.dif_y, H0.u, topH, tH, LH.c[0], C0.u, topC, tC, LC.c[0], FvH, FvC);
                else
                    face_flux_rec<DIFF>(fyc, m.fy, fy_index(r), m.dif_y, H0.u, topH, tH, LH.c[0], C0.u, topC,
                                        tC, LC.c[0], FvH, FvC);
            }
        }
        if (shift) fyc = load_fy(r + 1);   // used next iteration

        // Finish row r-1 now that its top face is known.
        if (!preroll && r > j0 && owned) {
            const size_t kk = size_t(r - 1) * nx + i;
            const double rhsH = (accH - FvH) * m.inv_area;
            const double rhsC = (accC - FvC) * m.inv_area;
            double outH = sp.b * (H0.u + sp.dt * rhsH), outC = sp.b * (C0.u + sp.dt * rhsC);
            if (need_n) {
                outH = sp.a * hn + outH;
                outC = sp.a * cn + outC;
            }
            if (!isfinite(outH) || !isfinite(outC)) err |= kErrNonFinite;
            if constexpr (SHARED) {
                __hip_atomic_store(gmem(sp.Ho) + kk, outH, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(gmem(sp.Co) + kk, outC, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            } else {
                gmem(sp.Ho)[kk] = outH;
                gmem(sp.Co)[kk] = outC;
            }
        }

        if (need_n && have) {               // u_n of row r, finished next iteration
            const size_t kn = size_t(r) * nx + ic;
            hn = gmem(sp.Hn)[kn];
            cn = gmem(sp.Cn)[kn];
        }

        // Vertical faces x_{i+1} of row r; lane i+1 supplies its centre and left line.
        if (!preroll && r < j1) {
            const double nbtH = shl1(tH), nbtC = shl1(tC);
            float nbH[3], nbC[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                nbH[c] = shl1(LH.c[3][c]);
                nbC[c] = shl1(LC.c[3][c]);
            }
            double FrH = 0.0, FrC = 0.0;
            if (i >= -1 && i <= nx - 1 && have) {
                if (i == -1 || i == nx - 1) {
                    const bool left = i == -1;
                    const int side = left ? 3 : 1;
                    const double sign = left ? -1.0 : 1.0;
                    const double* vn = m.vnx + size_t(r) * (nx + 1) + (i + 1);
                    const double* inH = m.inflowH[side] ? m.inflowH[side] + size_t(r) * m.q : nullptr;
                    const double* inC = m.inflowC[side] ? m.inflowC[side] + size_t(r) * m.q : nullptr;
                    const uint8_t pol = m.bc[side] ? gmem(m.bc[side])[r] : 0;
                    const double oHf = (left ? boundary_flux<float>(m, vn, vxs, sign, nbtH, nbH, m.sv, m.wv, pol, inH, err)
                                             : boundary_flux<float>(m, vn, vxs, sign, tH, LH.c[1], m.sv, m.wv, pol, inH, err)) +
                                       (DIFF ? thermal_flux(m, side, r, left ? nbtH : tH, m.wv, m.dif_bx) : 0.0);
                    const double oCf = left ? boundary_flux<float>(m, vn, vxs, sign, nbtC, nbC, m.sv, m.wv, pol, inC, err)
                                            : boundary_flux<float>(m, vn, vxs, sign, tC, LC.c[1], m.sv, m.wv, pol, inC, err);
                    FrH = sign * oHf;
                    FrC = sign * oCf;
                    if (left || owned) { bH += oHf; bC += oCf; }
                } else {
                    if constexpr (FULL)
                        face_flux<DIFF>(fxc, m.dif_x, tH, LH.c[1], nbtH, nbH, tC, LC.c[1], nbtC, nbC, FrH, FrC);
                    else
                        face_flux_rec<DIFF>(fxc, m.fx, fx_index(r), m.dif_x, tH, LH.c[1], nbtH, nbH, tC, LC.c[1],
                                            nbtC, nbC, FrH, FrC);
                }
            }
            accH = shr1(FrH) - FrH + FvH;
            accC = shr1(FrC) - FrC + FvC;
        }

#pragma unroll
        for (int c = 0; c < 3; ++c) { topH[c] = LH.c[2][c]; topC[c] = LC.c[2][c]; }
        if (shift) fxc = load_fx(r + 1);   // used next iteration
        if constexpr (UNROLL) {
            if (shift) {
                H0 = make_row(nH, oH, H2.u, first, last, k);
                C0 = make_row(nC, oC, C2.u, first, last, k);
            }
        } else if (shift) {
            H0 = H1; H1 = H2;
            C0 = C1; C1 = C2;
            H2 = make_row(nH, oH, H1.u, first, last, k);
            C2 = make_row(nC, oC, C1.u, first, last, k);
        }
    };
    if constexpr (UNROLL) {
        for (;;) {
            if (r > j1) break;
            row_step(H0, H1, H2, C0, C1, C2);
            ++r;
            if (r > j1) break;
            row_step(H1, H2, H0, C1, C2, C0);
            ++r;
            if (r > j1) break;
            row_step(H2, H0, H1, C2, C0, C1);
            ++r;
        }
    } else {
        for (; r <= j1; ++r) row_step(H0, H1, H2, C0, C1, C2);
    }
    item_partials(sp, item, 0.0, 0.0, 0.0, 0.0, bH, bC, err);
}

}  // namespace
}  // namespace mp
