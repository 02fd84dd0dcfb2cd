// Device binary-eutectic equilibrium: EutecticModel::evaluate restated for
// FP64 on gfx950 (src/phase/eutectic_rescaled.cpp:95-282).
//
// The reference evaluates in 80-bit long double and rounds each public value
// to double.  The device has no long double, so the phase boundaries, whose
// rounding decides the region, are formed in double-double (divisions by Xe,
// 1 - Xe and LD become products with host-formed double-double reciprocals)
// and rounded once;
// they then agree with the reference's long double -> double results except
// where the exact boundary lies within a few ulps of a rounding tie.  The
// eutectic-branch fractions (tiny differences of O(1) terms near the end of
// eutectic melting) are also formed in double-double.  All other quantities
// are plain FP64 and agree to a few ulps.
#pragma once

#include <hip/hip_runtime.h>

#include "mp_common.h"

namespace mp {

struct PhaseOut {
    double T, phil, cl, cs, phi1, phi2, dTdC, dTdH;
    int region;     // PhaseRegion: 1 PureSolid .. 6 PureMelting
};

struct DD { double hi, lo; };

// Error-free transformations must not be contracted: hipcc fuses a product
// feeding a sum into an FMA by default, which silently drops the rounding
// error these helpers exist to capture.  Each helper turns contraction off.

__device__ __forceinline__ DD dd_two_sum(double a, double b)
{
#pragma clang fp contract(off)
    const double s = a + b;
    const double bb = s - a;
    const double e = (a - (s - bb)) + (b - bb);
    return {s, e};
}

__device__ __forceinline__ DD dd_add(DD a, DD b)
{
#pragma clang fp contract(off)
    DD s = dd_two_sum(a.hi, b.hi);
    s.lo += a.lo + b.lo;
    return dd_two_sum(s.hi, s.lo);
}

// c / x as a double-double (quotient plus its fma residual).
__device__ __forceinline__ DD dd_div(double c, double x)
{
#pragma clang fp contract(off)
    const double q = c / x;
    const double r = fma(-q, x, c);
    return {q, r / x};
}

// (a.hi + a.lo) / b as a double-double.
__device__ __forceinline__ DD dd_div_d(DD a, double b)
{
#pragma clang fp contract(off)
    const double q1 = a.hi / b;
    const double r = fma(-q1, b, a.hi) + a.lo;
    return dd_two_sum(q1, r / b);
}

__device__ __forceinline__ DD dd_neg(DD a) { return {-a.hi, -a.lo}; }

__device__ __forceinline__ bool dd_less(DD a, DD b)
{
    return a.hi < b.hi || (a.hi == b.hi && a.lo < b.lo);
}

__device__ __forceinline__ DD dd_mul_d(DD a, double b)
{
#pragma clang fp contract(off)
    const double p = a.hi * b;
    const double e = fma(a.hi, b, -p) + a.lo * b;
    return dd_two_sum(p, e);
}

__device__ __forceinline__ DD dd_mul(DD a, DD b)
{
#pragma clang fp contract(off)
    const double p = a.hi * b.hi;
    const double e = fma(a.hi, b.hi, -p) + (a.hi * b.lo + a.lo * b.hi);
    return dd_two_sum(p, e);
}

// The phase parameters at entry k of a lithostatic table (tab null: p).
__device__ __forceinline__ PhaseDev phase_at(const PhaseDev& p, const TepPair* tab, size_t k)
{
    PhaseDev q = p;
    if (tab) {
        const TepPair t = tab[k];
        q.Tep = t.tep;
        q.Tmp = t.tmp;
    }
    return q;
}

// unitFraction (src/phase/eutectic_rescaled.cpp:108-114)
__device__ __forceinline__ double unit_fraction(double v, uint32_t& err)
{
    const double tol = 64.0 * 2.220446049250313e-16;
    if (!(v >= -tol && v <= 1.0 + tol)) err |= kErrFraction;
    return fmin(fmax(v, 0.0), 1.0);
}

// One evaluation on the side 0 <= C <= Xe of a diagram with eutectic Xe;
// (ixh, ixl) is 1/Xe as a double-double.  DERIV = false skips dT/dC and dT/dH
// (the stage kernels do not store them; analytically dT/dH lies in [0, 1]:
// root >= |a| and root^2 >= 4 LD r, so its unitFraction check cannot fire).
template <bool DERIV = true>
__device__ __forceinline__ PhaseOut phase_core(const PhaseDev& p, double Xe, double ixh, double ixl,
                                               double H, double C, uint32_t& err)
{
    PhaseOut s;
    s.T = 0; s.phil = 0; s.cl = 0; s.cs = 0; s.phi1 = 0; s.phi2 = 0; s.dTdC = 0; s.dTdH = 0;
    const double LD = p.LD, Tep = p.Tep, Tmp = p.Tmp;
    // Boundaries (:163-185) in plain FP64 first: C/Xe as C times the rounded 1/Xe,
    // each boundary within 3 eps (|Tep| + 1 + LD) of its correctly rounded value.
    // Only when H lies within a 16 eps margin of one of them are they redone in
    // double-double, which reproduces the reference's long double -> double
    // rounding, so the classification below is the reference's in every case.
    const double rf = C * ixh;
    double Heut = fma(rf, LD, Tep);
    double Hliq = C == 0.0 ? Tmp + LD : (Tep + (1.0 - rf)) + LD;
    const double tol = 16.0 * 2.220446049250313e-16 * (fabs(Tep) + 1.0 + LD);
    const bool near = fabs(H - Heut) <= tol || fabs(H - Hliq) <= tol;
    if (near) {
        const DD ratio = dd_mul_d({ixh, ixl}, C);
        Heut = dd_add({Tep, 0.0}, dd_mul_d(ratio, LD)).hi;
        if (C != 0.0) {
            const DD one_minus = dd_add({1.0, 0.0}, {-ratio.hi, -ratio.lo});
            Hliq = dd_add(dd_add({Tep, 0.0}, one_minus), {LD, 0.0}).hi;
        }
    }
    if (Tmp <= Tep || Hliq <= Tep || (C == 0.0 && Hliq <= Tmp)) err |= kErrPhaseRange;
    // selectRegion (:95-105)
    int region;
    if (C == 0.0) region = H <= Tmp ? 1 : (H < Hliq ? 6 : 5);
    else if (H <= Tep) region = 2;
    else if (H >= Hliq) region = 5;
    else if (H <= Heut) region = 3;
    else region = 4;
    s.region = region;
    switch (region) {
    case 1:
        s.phi1 = 1.0; s.T = H; s.dTdH = 1.0;
        break;
    case 6:
        s.phil = unit_fraction((H - Tmp) * p.ilh, err);
        s.phi1 = 1.0 - s.phil;
        s.T = Tmp;
        break;
    case 2:
        s.phi1 = 1.0 - C; s.phi2 = C; s.T = H; s.dTdH = 1.0;
        break;
    case 3: {
        // Fractions in FP64 while the solid fraction is at least 1/16 (then
        // Cs = phi2/(phi1+phi2) carries at most ~5e-15); near the boundaries and
        // near the end of eutectic melting, where both solid fractions are tiny
        // differences of O(1) terms, in double-double.
        const double e = (H - Tep) * p.ilh;
        double liquid = e < rf ? e : rf;
        double solid2 = fmax(fma(-Xe, liquid, C), 0.0);
        double solid1 = (1.0 - liquid) - solid2;
        if (near || !(solid1 + solid2 >= 0.0625)) {
            const DD maxl = dd_mul_d({ixh, ixl}, C);
            DD liq = maxl;
            if (H != Heut) {
                const DD ed = dd_mul(dd_two_sum(H, -Tep), {p.ilh, p.ill});
                if (dd_less(ed, maxl)) liq = ed;
            }
            DD s2 = dd_add({C, 0.0}, dd_mul_d(liq, -Xe));
            if (s2.hi < 0.0) s2 = {0.0, 0.0};
            const DD s1 = dd_add(dd_add({1.0, 0.0}, dd_neg(liq)), dd_neg(s2));
            liquid = liq.hi + liq.lo;
            solid2 = s2.hi + s2.lo;
            solid1 = s1.hi + s1.lo;
        }
        s.phil = unit_fraction(liquid, err);
        s.phi2 = unit_fraction(solid2, err);
        s.phi1 = unit_fraction(solid1, err);
        s.cl = Xe;
        s.T = Tep;
        break;
    }
    case 4: {
        // LD*phi^2 - (H-Tmp)*phi - C/Xe = 0 with the cancellation-free root
        // (:237-257); root = hypot(a, 2 sqrt(LD r)) with one square root.
        const double a = H - Tmp;
        const double r = rf;
        const double root = sqrt(fma(a, a, 4.0 * LD * r));
        const double liquid = a >= 0.0 ? (a + root) * (0.5 * p.ilh) : 2.0 * r / (root - a);
        s.phil = unit_fraction(liquid, err);
        s.phi1 = unit_fraction(1.0 - liquid, err);
        const double cl = C / liquid;
        s.cl = unit_fraction(cl, err);
        s.T = a >= 0.0 ? Tmp - cl * ixh : H - LD * liquid;    // r/liquid = cl/Xe
        if constexpr (DERIV) {
            s.dTdH = unit_fraction(a >= 0.0 ? 2.0 * (LD / root) * (r / (root + a))
                                            : 0.5 * (1.0 - a / root), err);
            s.dTdC = -((LD / root) / Xe);
        }
        break;
    }
    default:
        s.phil = 1.0; s.cl = C; s.T = H - LD; s.dTdH = 1.0;
        break;
    }
    if (DERIV && C == 0.0 && H > Tep && region != 5) {
        const double gap = fabs(H - Tmp);
        s.dTdC = gap == 0.0 ? -__builtin_inf() : -((LD / gap) / Xe);
    }
    if (!(s.phil > 0.0)) s.cl = 0.0;
    return s;   // cs is formed by the caller, after any mirroring
}

// EutecticModel::evaluate with optional hypereutectic mirror (C > Xe evaluates
// C' = 1-C on the diagram with eutectic 1-Xe and maps the phases back).  The
// mirror only selects the core's inputs, so a wave whose lanes lie on both
// sides of Xe runs the core once.
template <bool DERIV = true>
__device__ __forceinline__ PhaseOut phase_eval(const PhaseDev& p, double H, double C,
                                               uint32_t& err)
{
    if (!isfinite(H) || !isfinite(C)) {
        err |= kErrNonFinite;
        H = isfinite(H) ? H : 0.0;
        C = isfinite(C) ? C : 0.0;
    }
    const double Xe = p.Xe;
    const double cmax = p.hyper ? 1.0 : Xe;
    if (C < 0.0 || C > cmax) {
        err |= kErrComposition;
        C = fmin(fmax(C, 0.0), cmax);
    }
    const bool mirror = C > Xe;
    const double xs = mirror ? 1.0 - Xe : Xe;
    const double ih = mirror ? p.imh : p.ixh, il = mirror ? p.iml : p.ixl;
    const double cs = mirror ? 1.0 - C : C;
    PhaseOut s = phase_core<DERIV>(p, xs, ih, il, H, cs, err);
    if (mirror) {
        const double a = s.phi1;
        s.phi1 = s.phi2;
        s.phi2 = a;
        s.cl = s.phil > 0.0 ? 1.0 - s.cl : 0.0;
        s.dTdC = -s.dTdC;
    }
    const double solid = s.phi1 + s.phi2;
    s.cs = s.phi2 > 0.0 && solid > 0.0 ? s.phi2 / solid : 0.0;
    return s;
}

}  // namespace mp
