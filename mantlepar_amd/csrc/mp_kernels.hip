// Stage, diagnostic and utility kernels of the MI355X transport library.
//
// Stage kernel layout ("row sweep"): one wavefront owns a strip of 62 cell
// columns plus one halo column on each side, and walks up a band of rows.
// Each lane holds a 3x3 window of H and C cell averages in registers, loads
// one new row per step with coalesced 512-byte wave loads, reconstructs its
// cell, exchanges face traces with its neighbour lanes through cross-lane
// shuffles, and finishes the row below once that row's top face is known.
// There is no LDS tile and no halo recomputation in y; the only redundant
// work is the two halo lanes and one extra row per band.
//
// The flux scatter of the reference (rhs[left] -= F/|K|, rhs[right] += F/|K|,
// src/couple/preheat.cpp:193-196) becomes a gather: every cell sums its four
// face fluxes, so there are no atomics on state and results do not depend on
// launch order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "mp_common.h"
#include "mp_device.h"
#include "mp_launch.h"
#include "mp_limit.h"
#include "mp_phase.h"
#include "mp_phase_items.h"
#include "mp_recon.h"

namespace mp {

template <int MODE>
using RealT = typename std::conditional<MODE == 0, double, float>::type;

// Interior face flux (left state behind the +x/+y normal) for one field, per
// point, in the reference's LLF form F = (F_L + F_R - alpha*(uR - uL))/2
// (src/mlweno/advectiveflux.cpp:144-152).  Used by the diagnostics kernels.
template <int MODE>
__device__ __forceinline__ double interior_flux(const MeshDev& m, const double* vn, size_t vstride,
                                                double tL, const RealT<MODE> (&cL)[3], double tR,
                                                const RealT<MODE> (&cR)[3], const double* s,
                                                const double* W)
{
    using R = RealT<MODE>;
    double F = 0.0;
    for (int q = 0; q < m.q; ++q) {
        const double v = vn[q * vstride];
        const double alpha = m.lf_mode ? m.global_speed : fabs(v);
        const double uL = tL + double(line_at<R>(cL, R(s[q])));
        const double uR = tR + double(line_at<R>(cR, R(s[q])));
        F += W[q] * (0.5 * (v * uL + v * uR - alpha * (uR - uL)));
    }
    return F;
}

// One Gauss point of both fields' LLF fluxes, accumulated into fh, fc.  The
// compiler may fuse v*uL + v*uR either way, and may choose differently in
// different kernel instances of the same source; contraction is therefore off
// here and the face lines are explicit fmas, so every instance (fixed or
// runtime point count, single mesh or batch) gives the same bits.
__device__ __forceinline__ void llf_point(const MeshDev& m, double v, double sq, double w, double tLH,
                                          const double (&cLH)[3], double tRH, const double (&cRH)[3],
                                          double tLC, const double (&cLC)[3], double tRC,
                                          const double (&cRC)[3], double& fh, double& fc)
{
#pragma clang fp contract(off)
    auto line = [](const double (&c)[3], double x) { return fma(x, fma(x, c[2], c[1]), c[0]); };
    const double alpha = m.lf_mode ? m.global_speed : fabs(v);
    const double uLH = tLH + line(cLH, sq), uRH = tRH + line(cRH, sq);
    const double uLC = tLC + line(cLC, sq), uRC = tRC + line(cRC, sq);
    fh += w * (0.5 * (v * uLH + v * uRH - alpha * (uRH - uLH)));
    fc += w * (0.5 * (v * uLC + v * uRC - alpha * (uRC - uLC)));
}

// Both fields' LLF fluxes through one interior face, per Gauss point, in FP64
// (the plain drop-in form).  NQ > 0 fixes the point count at compile time.
template <int NQ>
__device__ __forceinline__ void llf_pair(const MeshDev& m, const double* vn, size_t vstride,
                                         const double* s, const double* W,
                                         double tLH, const double (&cLH)[3], double tRH,
                                         const double (&cRH)[3], double tLC,
                                         const double (&cLC)[3], double tRC,
                                         const double (&cRC)[3], double& FH, double& FC)
{
    double fh = 0.0, fc = 0.0;
    const int nq = NQ > 0 ? NQ : m.q;
#pragma unroll
    for (int q = 0; q < (NQ > 0 ? NQ : kMaxQ); ++q) {
        if (NQ == 0 && q >= nq) break;
        llf_point(m, vn[q * vstride], s[q], W[q], tLH, cLH, tRH, cRH, tLC, cLC, tRC, cRC, fh, fc);
    }
    FH = fh;
    FC = fc;
}

// llf_pair on the face's samples already in registers: the sweep issues a
// row's interior samples before its reconstruction, so their load latency is
// covered by the reconstruction instead of being waited for at the flux.
template <int NQ>
__device__ __forceinline__ void llf_pair_at(const MeshDev& m, const double (&vq)[NQ], const double* s,
                                            const double* W, double tLH, const double (&cLH)[3], double tRH,
                                            const double (&cRH)[3], double tLC, const double (&cLC)[3],
                                            double tRC, const double (&cRC)[3], double& FH, double& FC)
{
    double fh = 0.0, fc = 0.0;
#pragma unroll
    for (int q = 0; q < NQ; ++q)
        llf_point(m, vq[q], s[q], W[q], tLH, cLH, tRH, cRH, tLC, cLC, tRC, cRC, fh, fc);
    FH = fh;
    FC = fc;
}

// Both fields through one interior face from the precomputed flux moments:
// the first-order part t*V in FP64, the high-order correction in FP32.
__device__ __forceinline__ void moment_pair(double Vp, double Vm, float Mp1, float Mp2, float Mm1,
                                            float Mm2, double tLH, const float (&cLH)[3],
                                            double tRH, const float (&cRH)[3], double tLC,
                                            const float (&cLC)[3], double tRC,
                                            const float (&cRC)[3], double& FH, double& FC)
{
    const float vp = float(Vp), vm = float(Vm);
    const float ch = cLH[0] * vp + cLH[1] * Mp1 + cLH[2] * Mp2 + cRH[0] * vm + cRH[1] * Mm1 + cRH[2] * Mm2;
    const float cc = cLC[0] * vp + cLC[1] * Mp1 + cLC[2] * Mp2 + cRC[0] * vm + cRC[1] * Mm1 + cRC[2] * Mm2;
    FH = fma(tLH, Vp, tRH * Vm) + double(ch);
    FC = fma(tLC, Vp, tRC * Vm) + double(cc);
}

// Flux moments of every face (one thread per face), once per velocity refresh:
// the full record and its one-sided 16-byte form.
__global__ void face_prep_kernel(const MeshDev* __restrict__ mesh, FaceMom* fx, FaceMom* fy,
                                 FaceRec* rx, FaceRec* ry)
{
    const MeshDev& m = *mesh;
    const int64_t nxf = int64_t(m.ny) * (m.nx + 1), nyf = int64_t(m.ny + 1) * m.nx;
    int64_t f = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (f >= nxf + nyf) return;
    const bool vert = f < nxf;
    if (!vert) f -= nxf;
    const int64_t nf = vert ? nxf : nyf;
    const double* vn = (vert ? m.vnx : m.vny) + f;
    const double* s = vert ? m.sv : m.sh;
    const double* W = vert ? m.wv : m.wh;
    double vp = 0, vm = 0, p1 = 0, p2 = 0, m1 = 0, m2 = 0;
    for (int q = 0; q < m.q; ++q) {
        const double v = vn[q * nf];
        const double alpha = m.lf_mode ? m.global_speed : fabs(v);
        const double wl = W[q] * 0.5 * (v + alpha), wr = W[q] * 0.5 * (v - alpha);
        vp += wl;
        vm += wr;
        p1 += wl * s[q];
        p2 += wl * s[q] * s[q];
        m1 += wr * s[q];
        m2 += wr * s[q] * s[q];
    }
    FaceMom o;
    o.vp = vp;
    o.vm = vm;
    o.p1 = float(p1);
    o.p2 = float(p2);
    o.m1 = float(m1);
    o.m2 = float(m2);
    (vert ? fx : fy)[f] = o;
    // vp and the p moments are sums of +0 and positive terms, vm and the m
    // moments of +0 and negative ones, so "vm == 0" means the m side is +0
    // throughout (likewise vp); under the global LF speed both sides carry weight.
    FaceRec c;
    if (!m.lf_mode && vm == 0.0) {
        c.v = vp; c.a = o.p1; c.b = o.p2;
    } else if (!m.lf_mode && vp == 0.0) {
        c.v = vm; c.a = o.m1; c.b = o.m2;
    } else {
        c.v = 0.0; c.a = 0.0f; c.b = __builtin_nanf("");
    }
    (vert ? rx : ry)[f] = c;
}

// The FP64 sweep's one-sided face records (one thread per face), once per
// velocity refresh: the moments v, a, b of the face's samples, flagged with
// b = NaN where the points change sign or the LF speed is global.
__global__ void face_rec64_kernel(const MeshDev* __restrict__ mesh, FaceRec64* rx, FaceRec64* ry)
{
    const MeshDev& m = *mesh;
    const int64_t nxf = int64_t(m.ny) * (m.nx + 1), nyf = int64_t(m.ny + 1) * m.nx;
    int64_t f = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (f >= nxf + nyf) return;
    const bool vert = f < nxf;
    if (!vert) f -= nxf;
    const int64_t nf = vert ? nxf : nyf;
    const double* vn = (vert ? m.vnx : m.vny) + f;
    const double* s = vert ? m.sv : m.sh;
    const double* W = vert ? m.wv : m.wh;
    double v = 0, a = 0, b = 0;
    bool pos = false, neg = false;
    for (int q = 0; q < m.q; ++q) {
        const double x = vn[q * nf];
        pos = pos || x > 0.0;
        neg = neg || x < 0.0;
        const double w = W[q] * x;
        v += w;
        a += w * s[q];
        b += w * s[q] * s[q];
    }
    FaceRec64 r;
    r.v = v;
    r.a = a;
    r.b = (m.lf_mode || (pos && neg)) ? __builtin_nan("") : b;
    r.pad = 0.0;
    (vert ? rx : ry)[f] = r;
}

// Both fields through one interior face from its FP64 one-sided record: the
// upwind side's centre and face line (left/bottom where v >= 0) dotted with
// the record, as explicit fma chains so every kernel instance rounds alike.
__device__ __forceinline__ void rec64_pair(const FaceRec64& r, double tLH, const double (&cLH)[3], double tRH,
                                           const double (&cRH)[3], double tLC, const double (&cLC)[3],
                                           double tRC, const double (&cRC)[3], double& FH, double& FC)
{
#pragma clang fp contract(off)
    const bool up = r.v >= 0.0;
    const double uH = up ? tLH + cLH[0] : tRH + cRH[0];
    const double uC = up ? tLC + cLC[0] : tRC + cRC[0];
    FH = fma(up ? cLH[2] : cRH[2], r.b, fma(up ? cLH[1] : cRH[1], r.a, uH * r.v));
    FC = fma(up ? cLC[2] : cRC[2], r.b, fma(up ? cLC[1] : cRC[1], r.a, uC * r.v));
}

// Flux of both fields through one interior face; DIFF adds the thermal
// diffusion of H.
template <int MODE, int NQ, bool DIFF>
__device__ __forceinline__ void interior_pair(const MeshDev& m, bool vertical, size_t face,
                                              double tLH, const RealT<MODE> (&cLH)[3], double tRH,
                                              const RealT<MODE> (&cRH)[3], double tLC,
                                              const RealT<MODE> (&cLC)[3], double tRC,
                                              const RealT<MODE> (&cRC)[3], double& FH,
                                              double& FC)
{
    if constexpr (MODE == 0) {
        const double* vn = vertical ? m.vnx + face : m.vny + face;
        const size_t vs = vertical ? size_t(m.ny) * (m.nx + 1) : size_t(m.ny + 1) * m.nx;
        llf_pair<NQ>(m, vn, vs, vertical ? m.sv : m.sh, vertical ? m.wv : m.wh, tLH, cLH, tRH, cRH,
                     tLC, cLC, tRC, cRC, FH, FC);
    } else {
        const FaceMom f = (vertical ? m.fx : m.fy)[face];
        moment_pair(f.vp, f.vm, f.p1, f.p2, f.m1, f.m2, tLH, cLH, tRH, cRH, tLC, cLC, tRC, cRC,
                    FH, FC);
    }
    if constexpr (DIFF) FH = fma(vertical ? m.dif_x : m.dif_y, tLH - tRH, FH);
}

// The FP64 records of row r's bottom face y_r and right face x_{i+1}, from
// clamped in-bounds addresses (boundary faces and lanes without an interior
// face read them but do not use them); without records, b = NaN sends both
// faces point by point.
__device__ __forceinline__ void fetch_rec64(const MeshDev& m, int r, int ic, int i, FaceRec64& rb,
                                            FaceRec64& rr)
{
    if (m.ry64) {
        rb = gload(m.ry64 + size_t(clampi(r, 0, m.ny)) * m.nx + ic);
        rr = gload(m.rx64 + size_t(clampi(r, 0, m.ny - 1)) * (m.nx + 1) + clampi(i + 1, 0, m.nx));
    } else {
        rb.b = rr.b = __builtin_nan("");
    }
}

// interior_pair of the FP64 sweep point by point, its NQ samples issued
// together and waited for once.
template <int NQ, bool DIFF>
__device__ __forceinline__ void interior_pair_pts(const MeshDev& m, bool vertical, size_t face, double tLH,
                                                  const double (&cLH)[3], double tRH, const double (&cRH)[3],
                                                  double tLC, const double (&cLC)[3], double tRC,
                                                  const double (&cRC)[3], double& FH, double& FC)
{
    const size_t vs = vertical ? size_t(m.ny) * (m.nx + 1) : size_t(m.ny + 1) * m.nx;
    const auto* p = gmem(vertical ? m.vnx : m.vny) + face;
    double vq[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) vq[q] = p[q * vs];
    llf_pair_at<NQ>(m, vq, vertical ? m.sv : m.sh, vertical ? m.wv : m.wh, tLH, cLH, tRH, cRH, tLC, cLC,
                    tRC, cRC, FH, FC);
    if constexpr (DIFF) FH = fma(vertical ? m.dif_x : m.dif_y, tLH - tRH, FH);
}

// interior_pair of the FP64 sweep from the face's one-sided record, or point
// by point where the record is flagged.
template <int NQ, bool DIFF>
__device__ __forceinline__ void interior_pair_rec(const MeshDev& m, bool vertical, size_t face,
                                                  const FaceRec64& rec, double tLH, const double (&cLH)[3],
                                                  double tRH, const double (&cRH)[3], double tLC,
                                                  const double (&cLC)[3], double tRC, const double (&cRC)[3],
                                                  double& FH, double& FC)
{
    if (!isnan(rec.b)) {
        rec64_pair(rec, tLH, cLH, tRH, cRH, tLC, cLC, tRC, cRC, FH, FC);
        if constexpr (DIFF) FH = fma(vertical ? m.dif_x : m.dif_y, tLH - tRH, FH);
    } else if constexpr (NQ > 0) {
        interior_pair_pts<NQ, DIFF>(m, vertical, face, tLH, cLH, tRH, cRH, tLC, cLC, tRC, cRC, FH, FC);
    } else {
        interior_pair<0, 0, DIFF>(m, vertical, face, tLH, cLH, tRH, cRH, tLC, cLC, tRC, cRC, FH, FC);
    }
}

// LIM: apply the bound limiter to both fields' reconstructions (FP64 only).
// DIFF: the mesh has thermal diffusion or thermal boundary data; without it
// the diffusion terms are compiled out, which keeps the sweep within its
// register budget (with them it spills 64 B more per lane).
template <int MODE, int NQ, bool LIM = false, bool DIFF = true>
__device__ __forceinline__ void sweep_item(const MeshDev& m, const StagePtrs& sp, int item, int strip, int j0,
                           int j1)
{
    using R = RealT<MODE>;
    static_assert(!LIM || MODE == 0, "the limiter runs in the FP64 sweep");
    const int lane = threadIdx.x & (kWave - 1);
    const int nx = m.nx, ny = m.ny;
    const int i = strip * kStripCells - 1 + lane;
    const bool col_ok = i >= 0 && i < nx;
    const bool owned = col_ok && lane >= 1 && lane <= kStripCells;
    const int ic = clampi(i, 0, nx - 1), il = clampi(i - 1, 0, nx - 1), ir = clampi(i + 1, 0, nx - 1);
    const size_t vxs = size_t(ny) * (nx + 1);    // q stride of vnx
    const size_t vys = size_t(ny + 1) * nx;      // q stride of vny
    uint32_t err = 0;

    double wH[3][3], wC[3][3];
    auto load_row = [&](const double* base, int r, double (&row)[3]) {
        const double* p = base + size_t(clampi(r, 0, ny - 1)) * nx;
        row[0] = p[il];
        row[1] = p[ic];
        row[2] = p[ir];
    };
    int r = j0 > 0 ? j0 - 1 : 0;
    load_row(sp.Hk, r - 1, wH[0]); load_row(sp.Hk, r, wH[1]); load_row(sp.Hk, r + 1, wH[2]);
    load_row(sp.Ck, r - 1, wC[0]); load_row(sp.Ck, r, wC[1]); load_row(sp.Ck, r + 1, wC[2]);

    R topH[3] = {R(0), R(0), R(0)}, topC[3] = {R(0), R(0), R(0)};
    double tTopH = 0.0, tTopC = 0.0;
    double pFlH = 0, pFrH = 0, pFbH = 0, pFlC = 0, pFrC = 0, pFbC = 0;
    double bH = 0, bC = 0;

    for (; r <= j1; ++r) {
        const bool have = r < ny;
        const bool preroll = r == j0 - 1;
        double nH[3], nC[3];
        const bool shift = r + 1 <= j1;
        if (shift) {
            load_row(sp.Hk, r + 2, nH);
            load_row(sp.Ck, r + 2, nC);
        }
        // In the FP64 sweep, the one-sided records of this row's bottom face
        // y_r and right face x_{i+1}, issued before the reconstruction, which
        // covers their latency.
        constexpr bool RECS = MODE == 0;
        FaceRec64 rbot{}, rrgt{};
        if constexpr (RECS) fetch_rec64(m, r, ic, i, rbot, rrgt);
        FaceLines<R> LH, LC;
        if (have) {
            const unsigned mask = col_ok ? candidate_mask(i, r, nx, ny) : 2u;
            uint32_t e2 = 0;
            if constexpr (LIM) {
                const unsigned sides = (r == 0 ? 1u : 0u) | (i == nx - 1 ? 2u : 0u) |
                                       (r == ny - 1 ? 4u : 0u) | (i == 0 ? 8u : 0u);
                // one field at a time: its blend, limit factor and face lines
                // (both blends live at once spilled to scratch)
                auto limited = [&](const double (&w)[3][3], double cap, FaceLines<R>& L) {
                    double B[3][3];
                    blend<double>(m, w, mask, B, e2);
                    const double th = limit_factor<NQ>(m, w, B, sides, cap, e2);
#pragma unroll
                    for (int b = 0; b < 3; ++b)
#pragma unroll
                        for (int a = 0; a < 3; ++a) B[b][a] *= th;
                    face_lines<double>(m, B, L);
                };
                limited(wH, INFINITY, LH);
                limited(wC, m.lim_cupper, LC);
            } else {
                reconstruct<R>(m, wH, mask, LH, e2);
                // H's reconstruction is done before C's starts: interleaved,
                // the two spill past 256 registers with the diffusion terms
                __builtin_amdgcn_sched_barrier(0);
                reconstruct<R>(m, wC, mask, LC, e2);
            }
            if (col_ok) err |= e2;
        } else {
#pragma unroll
            for (int f = 0; f < 4; ++f)
#pragma unroll
                for (int k = 0; k < 3; ++k) { LH.c[f][k] = R(0); LC.c[f][k] = R(0); }
        }
        const double tH = wH[1][1], tC = wC[1][1];

        // Horizontal face y_r: top of row r-1, bottom of row r.
        double FvH = 0.0, FvC = 0.0;
        if (!preroll && col_ok) {
            const double* vn = m.vny + size_t(r) * nx + ic;
            if (r == 0) {
                const double* inH = m.inflowH[0] ? m.inflowH[0] + size_t(ic) * m.q : nullptr;
                const double* inC = m.inflowC[0] ? m.inflowC[0] + size_t(ic) * m.q : nullptr;
                const uint8_t pol = m.bc[0] ? m.bc[0][ic] : 0;
                const double oH = boundary_flux<R>(m, vn, vys, -1.0, tH, LH.c[0], m.sh, m.wh, pol, inH, err) +
                                  (DIFF ? thermal_flux(m, 0, ic, tH, m.wh, m.dif_by) : 0.0);
                const double oC = boundary_flux<R>(m, vn, vys, -1.0, tC, LC.c[0], m.sh, m.wh, pol, inC, err);
                FvH = -oH; FvC = -oC;
                if (owned) { bH += oH; bC += oC; }
            } else if (r == ny) {
                const double* inH = m.inflowH[2] ? m.inflowH[2] + size_t(ic) * m.q : nullptr;
                const double* inC = m.inflowC[2] ? m.inflowC[2] + size_t(ic) * m.q : nullptr;
                const uint8_t pol = m.bc[2] ? m.bc[2][ic] : 0;
                const double oH = boundary_flux<R>(m, vn, vys, 1.0, tTopH, topH, m.sh, m.wh, pol, inH, err) +
                                  (DIFF ? thermal_flux(m, 2, ic, tTopH, m.wh, m.dif_by) : 0.0);
                const double oC = boundary_flux<R>(m, vn, vys, 1.0, tTopC, topC, m.sh, m.wh, pol, inC, err);
                FvH = oH; FvC = oC;
                if (owned) { bH += oH; bC += oC; }
            } else if constexpr (RECS) {
                interior_pair_rec<NQ, DIFF>(m, false, size_t(r) * nx + ic, rbot, tTopH, topH, tH, LH.c[0],
                                            tTopC, topC, tC, LC.c[0], FvH, FvC);
            } else {
                interior_pair<MODE, NQ, DIFF>(m, false, size_t(r) * nx + ic, tTopH, topH, tH, LH.c[0],
                                        tTopC, topC, tC, LC.c[0], FvH, FvC);
            }
        }

        // Finish row r-1 now that its top face is known.
        if (!preroll && r > j0) {
            if (owned) {
                const size_t k = size_t(r - 1) * nx + i;
                const double rhsH = (pFlH - pFrH + pFbH - FvH) * m.inv_area;
                const double rhsC = (pFlC - pFrC + pFbC - FvC) * m.inv_area;
                const double wkH = wH[0][1] + sp.dt * rhsH;
                const double wkC = wC[0][1] + sp.dt * rhsC;
                double oH = sp.b * wkH, oC = sp.b * wkC;
                if (sp.a != 0.0) {
                    oH = sp.a * sp.Hn[k] + oH;
                    oC = sp.a * sp.Cn[k] + oC;
                }
                if (!isfinite(oH) || !isfinite(oC)) err |= kErrNonFinite;
                sp.Ho[k] = oH;
                sp.Co[k] = oC;
            }
        }

        // Vertical faces x_{i+1} of row r; neighbour lane i+1 supplies its left face.
        if (!preroll && r < j1) {
            const double nbtH = __shfl_down(tH, 1);
            const double nbtC = __shfl_down(tC, 1);
            R nbH[3], nbC[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                nbH[k] = __shfl_down(LH.c[3][k], 1);
                nbC[k] = __shfl_down(LC.c[3][k], 1);
            }
            double FrH = 0.0, FrC = 0.0;
            if (i >= -1 && i <= nx - 1 && have) {
                const size_t face = size_t(r) * (nx + 1) + (i + 1);
                const double* vn = m.vnx + face;
                if (i == -1) {
                    const double* inH = m.inflowH[3] ? m.inflowH[3] + size_t(r) * m.q : nullptr;
                    const double* inC = m.inflowC[3] ? m.inflowC[3] + size_t(r) * m.q : nullptr;
                    const uint8_t pol = m.bc[3] ? m.bc[3][r] : 0;
                    const double oH = boundary_flux<R>(m, vn, vxs, -1.0, nbtH, nbH, m.sv, m.wv, pol, inH, err) +
                                      (DIFF ? thermal_flux(m, 3, r, nbtH, m.wv, m.dif_bx) : 0.0);
                    const double oC = boundary_flux<R>(m, vn, vxs, -1.0, nbtC, nbC, m.sv, m.wv, pol, inC, err);
                    FrH = -oH; FrC = -oC;
                    bH += oH; bC += oC;
                } else if (i == nx - 1) {
                    const double* inH = m.inflowH[1] ? m.inflowH[1] + size_t(r) * m.q : nullptr;
                    const double* inC = m.inflowC[1] ? m.inflowC[1] + size_t(r) * m.q : nullptr;
                    const uint8_t pol = m.bc[1] ? m.bc[1][r] : 0;
                    const double oH = boundary_flux<R>(m, vn, vxs, 1.0, tH, LH.c[1], m.sv, m.wv, pol, inH, err) +
                                      (DIFF ? thermal_flux(m, 1, r, tH, m.wv, m.dif_bx) : 0.0);
                    const double oC = boundary_flux<R>(m, vn, vxs, 1.0, tC, LC.c[1], m.sv, m.wv, pol, inC, err);
                    FrH = oH; FrC = oC;
                    if (owned) { bH += oH; bC += oC; }
                } else if constexpr (RECS) {
                    interior_pair_rec<NQ, DIFF>(m, true, face, rrgt, tH, LH.c[1], nbtH, nbH, tC, LC.c[1],
                                                nbtC, nbC, FrH, FrC);
                } else {
                    interior_pair<MODE, NQ, DIFF>(m, true, face, tH, LH.c[1], nbtH, nbH, tC, LC.c[1],
                                            nbtC, nbC, FrH, FrC);
                }
            }
            pFlH = __shfl_up(FrH, 1);
            pFlC = __shfl_up(FrC, 1);
            pFrH = FrH; pFrC = FrC;
            pFbH = FvH; pFbC = FvC;
        }

#pragma unroll
        for (int k = 0; k < 3; ++k) { topH[k] = LH.c[2][k]; topC[k] = LC.c[2][k]; }
        tTopH = tH; tTopC = tC;
        if (shift) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                wH[0][k] = wH[1][k]; wH[1][k] = wH[2][k]; wH[2][k] = nH[k];
                wC[0][k] = wC[1][k]; wC[1][k] = wC[2][k]; wC[2][k] = nC[k];
            }
        }
    }

    item_partials(sp, item, 0.0, 0.0, 0.0, 0.0, bH, bC, err);
}

// Register budget: two waves per SIMD (256 registers), one wave per
// workgroup, so a finished item frees its slot at once.  The mixed path has
// its own sweep (mp_sweep32.hip).
template <int MODE, int NQ, bool LIM, bool DIFF>
__global__ __launch_bounds__(64, MODE == 0 ? 2 : 3) void stage_sweep_kernel(const MeshDev* __restrict__ mesh,
                                                          StagePtrs sp, int nitems)
{
    if (int(blockIdx.x) >= nitems) return;
    const int item = item_of_block(sp, int(blockIdx.x), nitems);
    const MeshDev& m = *mesh;
    int strip, j0, j1;
    item_region(sp, m.ny, item, strip, j0, j1);
    sweep_item<MODE, NQ, LIM, DIFF>(m, sp, item, strip, j0, j1);
}


// Step totals from the stage partials in a fixed order.  The item list is cut
// into contiguous slices (step_chunk); one wave reduces a slice (lanes take
// strided items, then the item_partials butterfly, compensated for the state
// sums), and a fixed LDS tree combines the slices.  The single-mesh form runs
// the slices as waves of one launch, so the partials (2.5 MB at 2048x4096)
// stream through many CUs; the batch form runs one mesh per block, its waves
// taking the slices in turn.  Both make the same operations in the same
// order, so a mesh's totals in a batch are bitwise its single-mesh totals.
// out[0..3] = sum H, sum C (times area), stage-weighted boundary H, C.

// Slices of an n-item list: about 64 items each (one per lane and stage), at
// least 4 and at most kStepSlices; slice c covers [c*chunk, (c+1)*chunk).
__device__ __forceinline__ int step_chunk(int n)
{
    const int slices = min(kStepSlices, max(4, (n + kWave - 1) / kWave));
    return max(1, (n + slices - 1) / slices);
}

// Slice [lo, hi) of the item list; every lane returns the slice's six sums.
__device__ __forceinline__ void step_slice(const double* __restrict__ partials, int nitems,
                                           int nstages, const double (&wst)[3], int lo, int hi,
                                           double (&v)[6])
{
    const int lane = threadIdx.x & (kWave - 1);
    double sH = 0, eH = 0, sC = 0, eC = 0, bH = 0, bC = 0;
    for (int st = 0; st < nstages; ++st) {
        const double* p = partials + size_t(st) * nitems * kPartials;
        const double w = wst[st];
        const bool last = st == nstages - 1;
        for (int it = lo + lane; it < hi; it += kWave) {
            const double* q = p + size_t(it) * kPartials;
            if (last) {
                two_sum_acc(sH, eH, q[0]);
                eH += q[1];
                two_sum_acc(sC, eC, q[2]);
                eC += q[3];
            }
            bH += w * q[4];
            bC += w * q[5];
        }
    }
    v[0] = sH; v[1] = eH; v[2] = sC; v[3] = eC; v[4] = bH; v[5] = bC;
    for (int off = 32; off >= 1; off >>= 1) {
        double o[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) o[k] = __shfl_xor(v[k], off);
        double s0 = v[0], e0 = v[1] + o[1];
        two_sum_acc(s0, e0, o[0]);
        double s1 = v[2], e1 = v[3] + o[3];
        two_sum_acc(s1, e1, o[2]);
        v[0] = s0; v[1] = e0; v[2] = s1; v[3] = e1;
        v[4] += o[4];
        v[5] += o[5];
    }
}

// Fixed-order tree over sh[6][kStepSlices] (all threads of the block call it):
// (0,1) and (2,3) compensated pairs, 4 and 5 plain; the result is in column 0.
__device__ __forceinline__ void step_tree(double (&sh)[6][kStepSlices], int tid)
{
    for (int half = kStepSlices / 2; half >= 1; half >>= 1) {
        if (tid < half) {
            double s0 = sh[0][tid], e0 = sh[1][tid] + sh[1][tid + half];
            two_sum_acc(s0, e0, sh[0][tid + half]);
            double s1 = sh[2][tid], e1 = sh[3][tid] + sh[3][tid + half];
            two_sum_acc(s1, e1, sh[2][tid + half]);
            sh[0][tid] = s0; sh[1][tid] = e0; sh[2][tid] = s1; sh[3][tid] = e1;
            sh[4][tid] += sh[4][tid + half];
            sh[5][tid] += sh[5][tid + half];
        }
        __syncthreads();
    }
}

__device__ __forceinline__ void step_out(const double (&sh)[6][kStepSlices], double area, double* out)
{
    out[0] = (sh[0][0] + sh[1][0]) * area;
    out[1] = (sh[2][0] + sh[3][0]) * area;
    out[2] = sh[4][0];
    out[3] = sh[5][0];
}

// One wave per slice and per workgroup (kStepSlices blocks of 64 threads): on
// the side stream this runs beside the next step's sweep, whose one-wave
// workgroups leave few CUs with room for a larger workgroup.
__global__ __launch_bounds__(64) void reduce_step_kernel(
    const double* __restrict__ partials, int nitems, int nstages, double w0, double w1,
    double w2, double* __restrict__ part)
{
    const int slice = int(blockIdx.x);
    const int chunk = step_chunk(nitems);
    const int lo = min(nitems, slice * chunk), hi = min(nitems, lo + chunk);
    const double wst[3] = {w0, w1, w2};
    double v[6] = {0, 0, 0, 0, 0, 0};
    if (lo < hi) step_slice(partials, nitems, nstages, wst, lo, hi, v);
    if ((threadIdx.x & (kWave - 1)) == 0)
#pragma unroll
        for (int k = 0; k < 6; ++k) part[size_t(slice) * 6 + k] = v[k];
}

// The preheat's balance check of each field (src/couple/preheat.cpp:268-273):
// |after - before + dt*outward flux| <= tol*(1 + |after| + |before| + |dt*flux|).
// out and prev: sum H, sum C (times area), outward boundary flux H, C.
__device__ __forceinline__ bool balance_holds(const double* out, const double* prev, double dt, double tol)
{
    bool ok = true;
    for (int f = 0; f < 2; ++f) {
        const double after = out[f], before = prev[f], flux = dt * out[2 + f];
        const double bal = after - before + flux;
        ok = ok && fabs(bal) <= tol * (1.0 + fabs(after) + fabs(before) + fabs(flux));
    }
    return ok;
}

// The step's totals; with prev (the totals before the step) also the
// balance check, kErrBalance when it fails.
__global__ __launch_bounds__(kStepSlices) void reduce_step_final_kernel(const double* __restrict__ part,
                                                                       double area, double* out,
                                                                       const double* prev, double dt,
                                                                       double tol, uint32_t* error)
{
    __shared__ double sh[6][kStepSlices];
    const int tid = threadIdx.x;
#pragma unroll
    for (int k = 0; k < 6; ++k) sh[k][tid] = part[size_t(tid) * 6 + k];
    __syncthreads();
    step_tree(sh, tid);
    if (tid == 0) {
        step_out(sh, area, out);
        if (prev && !balance_holds(out, prev, dt, tol)) atomicOr(error, kErrBalance);
    }
}

// Compensated sums of a state (H, C over n cells) as a step's slices:
// part[slice][6] with the boundary slots zero.
__global__ __launch_bounds__(256) void state_sums_kernel(const double* __restrict__ H,
                                                         const double* __restrict__ C, int64_t n,
                                                         double* __restrict__ part)
{
    // Four waves per slice, four independent compensated sums per lane so that
    // eight loads are in flight per lane; merged in a fixed order (the result
    // does not depend on timing).
    constexpr int kThreads = 256, kUnroll = 4;
    const int slice = int(blockIdx.x), t = int(threadIdx.x);
    const int64_t chunk = (n + kStepSlices - 1) / kStepSlices;
    const int64_t lo = min(n, int64_t(slice) * chunk), hi = min(n, lo + chunk);
    double v[kUnroll][4] = {};
    int64_t k = lo + t;
    for (; k + (kUnroll - 1) * kThreads < hi; k += kUnroll * kThreads) {
        double h[kUnroll], c[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            h[u] = H[k + u * kThreads];
            c[u] = C[k + u * kThreads];
        }
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            two_sum_acc(v[u][0], v[u][1], h[u]);
            two_sum_acc(v[u][2], v[u][3], c[u]);
        }
    }
    for (; k < hi; k += kThreads) {
        two_sum_acc(v[0][0], v[0][1], H[k]);
        two_sum_acc(v[0][2], v[0][3], C[k]);
    }
    auto merge = [](double* a, const double* b) {
        double s0 = a[0], e0 = a[1] + b[1];
        two_sum_acc(s0, e0, b[0]);
        double s1 = a[2], e1 = a[3] + b[3];
        two_sum_acc(s1, e1, b[2]);
        a[0] = s0; a[1] = e0; a[2] = s1; a[3] = e1;
    };
#pragma unroll
    for (int u = 1; u < kUnroll; ++u) merge(v[0], v[u]);
    for (int off = 32; off >= 1; off >>= 1) {
        double o[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) o[q] = __shfl_xor(v[0][q], off);
        merge(v[0], o);
    }
    __shared__ double wave_sums[kThreads / kWave][4];
    if ((t & (kWave - 1)) == 0)
#pragma unroll
        for (int q = 0; q < 4; ++q) wave_sums[t / kWave][q] = v[0][q];
    __syncthreads();
    if (t == 0) {
        for (int w = 1; w < kThreads / kWave; ++w) merge(v[0], wave_sums[w]);
#pragma unroll
        for (int q = 0; q < 4; ++q) part[size_t(slice) * 6 + q] = v[0][q];
        part[size_t(slice) * 6 + 4] = 0.0;
        part[size_t(slice) * 6 + 5] = 0.0;
    }
}

// ----------------------------------------------------------- batched sweep
//
// Many independent meshes in one persistent launch: each wave strides over a
// work list of (mesh, item) pairs, so meshes of any size share the chip and
// a small mesh never occupies a launch of its own.  Every mesh keeps its own
// MeshDev record, stage pointers, partials and error word; its results do not
// depend on which other meshes share the launch.
//
// Persistent form (queue != nullptr): every wave takes the next work item from
// a counter until the list is exhausted; otherwise one wave per item.
// LIM: meshes whose limiter is on (m.limit, wave-uniform) take the limited
// sweep, the others the plain one, in the same launch; DIFF likewise for
// meshes with diffusion or thermal boundary data (has_diffusion).
template <int MODE, int NQ, bool LIM = false, bool DIFF = false>
__global__ __launch_bounds__(256, MODE == 0 ? 2 : 3) void batch_sweep_kernel(
    const MeshDev* __restrict__ meshes, const StagePtrs* __restrict__ stages,
    const int2* __restrict__ items, int nitems, int* queue)
{
    // The work index is wave-uniform; saying so keeps every per-mesh table
    // load scalar (SGPRs) instead of replicating it across lanes.
    const int wave = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
    int it = blockIdx.x * int(blockDim.x >> 6) + wave;
    for (;;) {
        if (queue) {
            int t = 0;
            if ((threadIdx.x & (kWave - 1)) == 0) t = atomicAdd(queue, 1);
            it = __builtin_amdgcn_readfirstlane(t);
        }
        if (it >= nitems) break;
        const int2 w = items[it];
        const int slot = __builtin_amdgcn_readfirstlane(w.x);
        const int local = __builtin_amdgcn_readfirstlane(w.y);
        if (slot >= 0) {               // (-1, 0) pads an XCD slab
            const MeshDev& m = meshes[slot];
            const StagePtrs& sp = stages[slot];
            int strip, j0, j1;
            item_region(sp, m.ny, local, strip, j0, j1);
            if (j0 < m.ny) {           // always true with a current work list
                // the kernel's own flags as template arguments: a branch that
                // cannot run instantiates nothing new
                const bool lim = LIM && m.limit, dif = DIFF && has_diffusion(m);
                if (lim && dif) sweep_item<MODE, NQ, LIM, DIFF>(m, sp, local, strip, j0, j1);
                else if (lim) sweep_item<MODE, NQ, LIM, false>(m, sp, local, strip, j0, j1);
                else if (dif) sweep_item<MODE, NQ, false, DIFF>(m, sp, local, strip, j0, j1);
                else sweep_item<MODE, NQ, false, false>(m, sp, local, strip, j0, j1);
            }
        }
        if (!queue) break;
    }
}

// Per-mesh step totals: one block per mesh, its four waves taking the slices
// of the single-mesh reduction in turn, then the same tree.  Partials of mesh
// k start at stages[k].partials with nitems[k] items per stage; out[k*4 + 0..3]
// as in reduce_step_final_kernel.
// With prev (the totals before the step, prev[k*4 + 0..3]), each mesh's
// balance is checked at tolerance tol with its own dt and error word.
__global__ __launch_bounds__(256) void batch_reduce_kernel(
    const StagePtrs* __restrict__ stages, const int* __restrict__ nitems,
    const MeshDev* __restrict__ meshes, const int* __restrict__ slots, int nstages, double w0,
    double w1, double w2, double* out, const double* prev, double tol)
{
    __shared__ double sh[6][kStepSlices];
    const int k = slots[blockIdx.x];
    const int tid = threadIdx.x;
    const int n = nitems[k];
    const double* base = stages[k].partials;
    const double wst[3] = {w0, w1, w2};
    const int chunk = step_chunk(n);
    for (int slice = tid >> 6; slice < kStepSlices; slice += 4) {
        const int lo = min(n, slice * chunk), hi = min(n, lo + chunk);
        double v[6] = {0, 0, 0, 0, 0, 0};
        if (lo < hi) step_slice(base, n, nstages, wst, lo, hi, v);
        if ((tid & (kWave - 1)) == 0)
#pragma unroll
            for (int q = 0; q < 6; ++q) sh[q][slice] = v[q];
    }
    __syncthreads();
    step_tree(sh, tid);
    if (tid == 0) {
        double* o = out + size_t(k) * 4;
        step_out(sh, meshes[k].area, o);
        if (prev && !balance_holds(o, prev + size_t(k) * 4, stages[k].dt, tol))
            atomicOr(stages[k].error, kErrBalance);
    }
}

// Totals of each mesh's state before a call's first step (the state that the
// first step's stage-0 record names as u^n): compensated sums of H and C times
// the area in out[k*4 + 0..1], zero boundary slots.  Pass 1: kStateSlices
// blocks per mesh (grid x: the mesh, y: the slice), each summing a contiguous slice with four
// independent sums per lane, so that many loads are in flight; pass 2 merges
// a mesh's slices in a fixed order.
constexpr int kStateSlices = kStateSlicesMax;

__device__ __forceinline__ void merge_sums(double* a, const double* b)
{
    double s0 = a[0], e0 = a[1] + b[1];
    two_sum_acc(s0, e0, b[0]);
    double s1 = a[2], e1 = a[3] + b[3];
    two_sum_acc(s1, e1, b[2]);
    a[0] = s0; a[1] = e0; a[2] = s1; a[3] = e1;
}

__global__ __launch_bounds__(256) void batch_state_sums_kernel(const StagePtrs* __restrict__ stages,
                                                               const MeshDev* __restrict__ meshes,
                                                               const int* __restrict__ slots,
                                                               double* __restrict__ part)
{
    constexpr int kThreads = 256, kUnroll = 4;
    __shared__ double sh[kThreads / kWave][4];
    const int k = slots[blockIdx.x];
    const int tid = threadIdx.x;
    const MeshDev& m = meshes[k];
    const double* H = stages[k].Hn;
    const double* C = stages[k].Cn;
    const int64_t n = int64_t(m.nx) * m.ny;
    const int64_t chunk = (n + kStateSlices - 1) / kStateSlices;
    const int64_t lo = min(n, int64_t(blockIdx.y) * chunk), hi = min(n, lo + chunk);
    double v[kUnroll][4] = {};
    int64_t c = lo + tid;
    for (; c + (kUnroll - 1) * kThreads < hi; c += kUnroll * kThreads) {
        double h[kUnroll], q[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            h[u] = H[c + u * kThreads];
            q[u] = C[c + u * kThreads];
        }
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            two_sum_acc(v[u][0], v[u][1], h[u]);
            two_sum_acc(v[u][2], v[u][3], q[u]);
        }
    }
    for (; c < hi; c += kThreads) {
        two_sum_acc(v[0][0], v[0][1], H[c]);
        two_sum_acc(v[0][2], v[0][3], C[c]);
    }
#pragma unroll
    for (int u = 1; u < kUnroll; ++u) merge_sums(v[0], v[u]);
    for (int off = 32; off >= 1; off >>= 1) {
        double o[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) o[q] = __shfl_xor(v[0][q], off);
        merge_sums(v[0], o);
    }
    if ((tid & (kWave - 1)) == 0)
#pragma unroll
        for (int q = 0; q < 4; ++q) sh[tid / kWave][q] = v[0][q];
    __syncthreads();
    if (tid == 0) {
        for (int w = 1; w < kThreads / kWave; ++w) merge_sums(v[0], sh[w]);
        double* p = part + (size_t(blockIdx.x) * kStateSlices + blockIdx.y) * 4;
#pragma unroll
        for (int q = 0; q < 4; ++q) p[q] = v[0][q];
    }
}

__global__ void batch_state_sums_final_kernel(const MeshDev* __restrict__ meshes,
                                              const int* __restrict__ slots,
                                              const double* __restrict__ part, double* out)
{
    if (threadIdx.x != 0) return;
    const int k = slots[blockIdx.x];
    const double* p = part + size_t(blockIdx.x) * kStateSlices * 4;
    double v[4] = {p[0], p[1], p[2], p[3]};
    for (int s = 1; s < kStateSlices; ++s) merge_sums(v, p + size_t(s) * 4);
    double* o = out + size_t(k) * 4;
    o[0] = (v[0] + v[1]) * meshes[k].area;
    o[1] = (v[2] + v[3]) * meshes[k].area;
    o[2] = 0.0;
    o[3] = 0.0;
}

// Zero a small device array with a kernel (graph-safe: no memset node).
__global__ void clear_kernel(uint32_t* p, int n)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = 0u;
}

// ------------------------------------------------------------ diagnostics

template <int MODE>
__device__ __forceinline__ void load_window(const MeshDev& m, const double* u, int i, int j,
                                            double (&w)[3][3])
{
#pragma unroll
    for (int dj = -1; dj <= 1; ++dj)
#pragma unroll
        for (int di = -1; di <= 1; ++di)
            w[dj + 1][di + 1] = u[size_t(clampi(j + dj, 0, m.ny - 1)) * m.nx + clampi(i + di, 0, m.nx - 1)];
}

// Traces of one field (field 1 = C bounds the limiter by lim_cupper).
template <int MODE>
__global__ void recon_diag_kernel(const MeshDev* __restrict__ mesh, const double* __restrict__ u,
                                  int field, double* sigma, double* omega, double* traces,
                                  uint32_t* error)
{
    using R = RealT<MODE>;
    const MeshDev& m = *mesh;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.nx * m.ny) return;
    const int i = c % m.nx, j = c / m.nx;
    double w[3][3];
    load_window<MODE>(m, u, i, j, w);
    FaceLines<R> L;
    uint32_t err = 0;
    R B[3][3];
    blend<R>(m, w, candidate_mask(i, j, m.nx, m.ny), B, err, sigma + size_t(c) * 5,
             omega + size_t(c) * 5);
    if (m.limit) {
        const unsigned sides = (j == 0 ? 1u : 0u) | (i == m.nx - 1 ? 2u : 0u) |
                               (j == m.ny - 1 ? 4u : 0u) | (i == 0 ? 8u : 0u);
        double Bd[3][3];
#pragma unroll
        for (int b = 0; b < 3; ++b)
#pragma unroll
            for (int a = 0; a < 3; ++a) Bd[b][a] = double(B[b][a]);
        const double th = limit_factor<0>(m, w, Bd, sides, field ? m.lim_cupper : INFINITY, err);
#pragma unroll
        for (int b = 0; b < 3; ++b)
#pragma unroll
            for (int a = 0; a < 3; ++a) B[b][a] = R(double(B[b][a]) * th);
    }
    face_lines<R>(m, B, L);
    for (int f = 0; f < 4; ++f)
        for (int q = 0; q < m.q; ++q) {
            const double s = (f & 1) ? m.sv[q] : m.sh[q];
            traces[(size_t(c) * 4 + f) * m.q + q] = w[1][1] + double(line_at<R>(L.c[f], R(s)));
        }
    if (err) atomicOr(error, err);
}

// Limiter factors of both fields per cell (FP64 blend), for diagnostics.
__global__ __launch_bounds__(128) void limit_diag_kernel(const MeshDev* __restrict__ mesh, const double* __restrict__ H,
                                  const double* __restrict__ C, double* thetaH, double* thetaC,
                                  uint32_t* error)
{
    const MeshDev& m = *mesh;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.nx * m.ny) return;
    const int i = c % m.nx, j = c / m.nx;
    const unsigned mask = candidate_mask(i, j, m.nx, m.ny);
    const unsigned sides = (j == 0 ? 1u : 0u) | (i == m.nx - 1 ? 2u : 0u) |
                           (j == m.ny - 1 ? 4u : 0u) | (i == 0 ? 8u : 0u);
    uint32_t err = 0;
#pragma unroll
    for (int field = 0; field < 2; ++field) {
        double w[3][3], B[3][3];
        load_window<0>(m, field ? C : H, i, j, w);
        blend<double>(m, w, mask, B, err);
        (field ? thetaC : thetaH)[c] =
            limit_factor<0>(m, w, B, sides, field ? m.lim_cupper : INFINITY, err);
    }
    if (err) atomicOr(error, err);
}

// The reconstructions of H and C at points: query k is cell cell[k] and the
// point's offset from the cell centre in units of h = sqrt(area), (X[k], Y[k]).
// Value t + theta*P(X, Y), P the centred blend (Reconstruction::Evaluate,
// src/mlweno/reconstruction.cpp:474-502) and theta the limiter factor when the
// limiter is on (StateReconstruction::Evaluate, src/couple/state_reconstruction.cpp:121-134).
template <int MODE>
__global__ __launch_bounds__(256) void point_eval_kernel(const MeshDev* __restrict__ mesh, const double* __restrict__ H,
                                  const double* __restrict__ C, int64_t n,
                                  const int32_t* __restrict__ cell, const double* __restrict__ X,
                                  const double* __restrict__ Y, double* outH, double* outC,
                                  uint32_t* error)
{
    using R = RealT<MODE>;
    const MeshDev& m = *mesh;
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int c = cell[k];
    const int i = c % m.nx, j = c / m.nx;
    const unsigned mask = candidate_mask(i, j, m.nx, m.ny);
    const unsigned sides = (j == 0 ? 1u : 0u) | (i == m.nx - 1 ? 2u : 0u) |
                           (j == m.ny - 1 ? 4u : 0u) | (i == 0 ? 8u : 0u);
    const double x = X[k], y = Y[k];
    uint32_t err = 0;
#pragma unroll
    for (int field = 0; field < 2; ++field) {
        double w[3][3];
        load_window<MODE>(m, field ? C : H, i, j, w);
        R B[3][3];
        blend<R>(m, w, mask, B, err);
        double th = 1.0;
        if (m.limit) {
            double Bd[3][3];
#pragma unroll
            for (int b = 0; b < 3; ++b)
#pragma unroll
                for (int a = 0; a < 3; ++a) Bd[b][a] = double(B[b][a]);
            th = limit_factor<0>(m, w, Bd, sides, field ? m.lim_cupper : INFINITY, err);
        }
        double p = 0.0;
#pragma unroll
        for (int b = 2; b >= 0; --b)
            p = p * y + ((double(B[b][2]) * x + double(B[b][1])) * x + double(B[b][0]));
        (field ? outC : outH)[k] = w[1][1] + th * p;
    }
    if (err) atomicOr(error, err);
}

// L(u) per cell by reconstructing the cell and its four neighbours (5x the
// work of the sweep; diagnostics only).  bnd[c*2+f] = outward boundary flux.
template <int MODE>
__global__ __launch_bounds__(128) void rhs_diag_kernel(const MeshDev* __restrict__ mesh, const double* __restrict__ H,
                                const double* __restrict__ C, double* rhsH, double* rhsC,
                                double* bnd, uint32_t* error)
{
    using R = RealT<MODE>;
    const MeshDev& m = *mesh;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.nx * m.ny) return;
    const int nx = m.nx, ny = m.ny;
    const int i = c % nx, j = c / nx;
    const size_t vxs = size_t(ny) * (nx + 1), vys = size_t(ny + 1) * nx;
    uint32_t err = 0;
#pragma unroll
    for (int field = 0; field < 2; ++field) {
        const double* u = field ? C : H;
        double w[3][3];
        load_window<MODE>(m, u, i, j, w);
        FaceLines<R> L;
        reconstruct<R>(m, w, candidate_mask(i, j, nx, ny), L, err);
        const double t = w[1][1];
        double Fin = 0.0, Fout_b = 0.0;   // net inflow, outward boundary flux
        // neighbours: 0 bottom (i,j-1), 1 right (i+1,j), 2 top (i,j+1), 3 left (i-1,j)
        const int ni[4] = {i, i + 1, i, i - 1}, nj[4] = {j - 1, j, j + 1, j};
#pragma unroll
        for (int f = 0; f < 4; ++f) {
            const bool vert = f & 1;
            const double* vn = vert ? m.vnx + size_t(j) * (nx + 1) + (f == 1 ? i + 1 : i)
                                    : m.vny + size_t(f == 2 ? j + 1 : j) * nx + i;
            const size_t vs = vert ? vxs : vys;
            const double* s = vert ? m.sv : m.sh;
            const double* W = vert ? m.wv : m.wh;
            const bool outside = ni[f] < 0 || ni[f] >= nx || nj[f] < 0 || nj[f] >= ny;
            const double sign = (f == 1 || f == 2) ? 1.0 : -1.0;   // outward vs canonical
            if (outside) {
                const int side = f;   // bottom, right, top, left
                const int k = vert ? j : i;
                const double* const* inflow = field ? m.inflowC : m.inflowH;
                const double* in = inflow[side] ? inflow[side] + size_t(k) * m.q : nullptr;
                const uint8_t pol = m.bc[side] ? m.bc[side][k] : 0;
                double o = boundary_flux<R>(m, vn, vs, sign, t, L.c[f], s, W, pol, in, err);
                if (field == 0) o += thermal_flux(m, side, k, t, W, vert ? m.dif_bx : m.dif_by);
                Fin -= o;
                Fout_b += o;
            } else {
                double wn[3][3];
                load_window<MODE>(m, u, ni[f], nj[f], wn);
                FaceLines<R> N;
                reconstruct<R>(m, wn, candidate_mask(ni[f], nj[f], nx, ny), N, err);
                const int opp = (f + 2) & 3;
                double F;
                const double D = field == 0 ? (vert ? m.dif_x : m.dif_y) : 0.0;
                if (f == 1 || f == 2)
                    F = fma(D, t - wn[1][1], interior_flux<MODE>(m, vn, vs, t, L.c[f], wn[1][1], N.c[opp], s, W));
                else
                    F = fma(D, wn[1][1] - t, interior_flux<MODE>(m, vn, vs, wn[1][1], N.c[opp], t, L.c[f], s, W));
                Fin += sign > 0 ? -F : F;
            }
        }
        (field ? rhsC : rhsH)[c] = Fin * m.inv_area;
        bnd[size_t(c) * 2 + field] = Fout_b;
    }
    if (err) atomicOr(error, err);
}

// Per-cell advective rate (src/couple/preheat.cpp:125-152) of cell c.
__device__ __forceinline__ double cell_rate(const MeshDev& m, int c)
{
    const int nx = m.nx, ny = m.ny, i = c % nx, j = c / nx;
    const size_t vxs = size_t(ny) * (nx + 1), vys = size_t(ny + 1) * nx;
    double sx = 0.0, sy = 0.0;
    for (int q = 0; q < m.q; ++q) {
        const double a = m.wv[q] * fabs(m.vnx[q * vxs + size_t(j) * (nx + 1) + i]);
        const double b = m.wv[q] * fabs(m.vnx[q * vxs + size_t(j) * (nx + 1) + i + 1]);
        const double cc = m.wh[q] * fabs(m.vny[q * vys + size_t(j) * nx + i]);
        const double d = m.wh[q] * fabs(m.vny[q * vys + size_t(j + 1) * nx + i]);
        sx += a + b;
        sy += cc + d;
    }
    // thermal diffusion: |sum of sample partials| per face side
    // (src/couple/preheat.cpp:137-148)
    double dif = (i > 0 ? m.dif_x : 0.0) + (i < nx - 1 ? m.dif_x : 0.0) +
                 (j > 0 ? m.dif_y : 0.0) + (j < ny - 1 ? m.dif_y : 0.0);
    const int bside[4] = {j == 0 ? 0 : -1, i == nx - 1 ? 1 : -1, j == ny - 1 ? 2 : -1, i == 0 ? 3 : -1};
    for (int f = 0; f < 4; ++f) {
        if (bside[f] < 0 || !m.tbc[f]) continue;
        const int k = (f & 1) ? j : i;
        const double* W = (f & 1) ? m.wv : m.wh;
        const double coef = (f & 1) ? m.dif_bx : m.dif_by;
        for (int q = 0; q < m.q; ++q)
            if (m.tbc[f][size_t(k) * m.q + q] == 1) dif += W[q] * coef;
    }
    return (sx + sy + dif) * m.inv_area;
}

// Maximum cell rate: a grid-stride loop, then one atomic per block on the
// ordered bit pattern of the non-negative double (one atomic per wave on a
// single address serialised at 12 ms for 67M cells).
__global__ __launch_bounds__(256) void rate_kernel(const MeshDev* __restrict__ mesh,
                                                   unsigned long long* max_bits)
{
    __shared__ double wmax[4];
    const MeshDev& m = *mesh;
    const int n = m.nx * m.ny;
    double rate = 0.0;
    for (int c = int(blockIdx.x) * 256 + int(threadIdx.x); c < n; c += int(gridDim.x) * 256)
        rate = fmax(rate, cell_rate(m, c));
    for (int off = 32; off >= 1; off >>= 1) rate = fmax(rate, __shfl_xor(rate, off));
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = rate;
    __syncthreads();
    if (threadIdx.x == 0) {
        rate = fmax(fmax(wmax[0], wmax[1]), fmax(wmax[2], wmax[3]));
        atomicMax(max_bits, (unsigned long long)__double_as_longlong(rate));
    }
}

// EutecticModel::evaluate over arbitrary arrays.  P (optional) per element:
// Tep = TDe0 + P/nu/dT formed in double-double like the long double original.
__global__ void phase_kernel(PhaseDev ph, double TDe0, double nu, double dT, int64_t n,
                             const double* __restrict__ H, const double* __restrict__ C,
                             const double* __restrict__ P, double* T, double* phi, double* Cl,
                             double* Cs, uint8_t* region, double* phi1, double* phi2,
                             double* dTdC, double* dTdH, uint32_t* error)
{
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    uint32_t err = 0;
    PhaseDev p = ph;
    if (P) {
        const double Pk = P[k];
        if (!isfinite(Pk)) err |= kErrNonFinite;
        DD a = dd_div(Pk, nu);
        const double q1 = a.hi / dT;
        const double r1 = (fma(-q1, dT, a.hi) + a.lo) / dT;
        const DD shift = dd_two_sum(q1, r1);
        p.Tep = dd_add({TDe0, 0.0}, shift).hi;
        p.Tmp = p.Tep + 1.0;
    }
    const PhaseOut o = phase_eval(p, H[k], C[k], err);
    if (T) T[k] = o.T;
    if (phi) phi[k] = o.phil;
    if (Cl) Cl[k] = o.cl;
    if (Cs) Cs[k] = o.cs;
    if (region) region[k] = uint8_t(o.region);
    if (phi1) phi1[k] = o.phi1;
    if (phi2) phi2[k] = o.phi2;
    if (dTdC) dTdC[k] = o.dTdC;
    if (dTdH) dTdH[k] = o.dTdH;
    if (err) atomicOr(error, err);
}

// Phase of a state (used when no step has produced one yet).
__global__ void state_phase_kernel(const MeshDev* __restrict__ mesh, const double* H,
                                   const double* C, double* T, double* phi, double* Cl,
                                   double* Cs, uint8_t* region, uint32_t* error)
{
    const MeshDev& m = *mesh;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.nx * m.ny) return;
    uint32_t err = 0;
    const PhaseOut o = phase_eval(phase_at(m.ph, m.tep_row, c / m.nx), H[c], C[c], err);
    T[c] = o.T; phi[c] = o.phil; Cl[c] = o.cl; Cs[c] = o.cs; region[c] = uint8_t(o.region);
    if (err) atomicOr(error, err);
}

// One wave per workgroup and per item, in the sweep's dispatch order (10%
// faster than 4-wave workgroups at 2048x4096).
__global__ __launch_bounds__(64) void phase_items_kernel(const MeshDev* __restrict__ mesh,
                                                         StagePtrs sp, int nitems)
{
    if (int(blockIdx.x) >= nitems) return;
    const int item = item_of_block(sp, int(blockIdx.x), nitems);
    const MeshDev& m = *mesh;
    int strip, j0, j1;
    item_region(sp, m.ny, item, strip, j0, j1);
    phase_item(m, sp, item, strip, j0, j1);
}

// The same over a batch work list of (mesh slot, item) pairs.
__global__ __launch_bounds__(64) void batch_phase_items_kernel(const MeshDev* __restrict__ meshes,
                                                                const StagePtrs* __restrict__ stages,
                                                                const int2* __restrict__ items,
                                                                int nitems)
{
    const int it = int(blockIdx.x);
    if (it >= nitems) return;
    const int2 w = items[it];
    const int slot = __builtin_amdgcn_readfirstlane(w.x);
    const int local = __builtin_amdgcn_readfirstlane(w.y);
    if (slot < 0) return;              // pads an XCD slab
    const MeshDev& m = meshes[slot];
    const StagePtrs& sp = stages[slot];
    int strip, j0, j1;
    item_region(sp, m.ny, local, strip, j0, j1);
    if (j0 < m.ny) phase_item(m, sp, local, strip, j0, j1);
}

// ---------------------------------------------------------------- launchers

hipError_t launch_phase_items(const MeshDev* d_mesh, const StagePtrs& sp, int nitems,
                              hipStream_t stream)
{
    const dim3 grid(nitems), block(kWave);
    phase_items_kernel<<<grid, block, 0, stream>>>(d_mesh, sp, nitems);
    return hipGetLastError();
}

hipError_t launch_batch_phase_items(const MeshDev* meshes, const StagePtrs* stages,
                                    const int2* items, int nitems, hipStream_t stream)
{
    if (nitems <= 0) return hipSuccess;
    const dim3 grid(nitems), block(kWave);
    batch_phase_items_kernel<<<grid, block, 0, stream>>>(meshes, stages, items, nitems);
    return hipGetLastError();
}

// The FP64 sweep's instances: the limiter and the diffusion terms each
// compiled in only when the mesh uses them.
template <int NQ>
static void launch_fp64(const MeshDev* d_mesh, const StagePtrs& sp, int nitems, hipStream_t stream,
                        bool limit, bool diffusion)
{
    const dim3 grid(nitems), block(kWave);
    if (limit && diffusion) stage_sweep_kernel<0, NQ, true, true><<<grid, block, 0, stream>>>(d_mesh, sp, nitems);
    else if (limit) stage_sweep_kernel<0, NQ, true, false><<<grid, block, 0, stream>>>(d_mesh, sp, nitems);
    else if (diffusion) stage_sweep_kernel<0, NQ, false, true><<<grid, block, 0, stream>>>(d_mesh, sp, nitems);
    else stage_sweep_kernel<0, NQ, false, false><<<grid, block, 0, stream>>>(d_mesh, sp, nitems);
}

hipError_t launch_stage(int mode, int q, const MeshDev* d_mesh, const StagePtrs& sp,
                        int nitems, hipStream_t stream, int limit, bool diffusion, bool global_lf)
{
    if (mode != 0) return launch_stage32(d_mesh, sp, nitems, stream, diffusion, global_lf, limit);
    const bool lim = limit != 0;
    if (q == 5) launch_fp64<5>(d_mesh, sp, nitems, stream, lim, diffusion);
    else launch_fp64<0>(d_mesh, sp, nitems, stream, lim, diffusion);
    return hipGetLastError();
}



hipError_t launch_limit_diag(const MeshDev* d_mesh, int ncells, const double* H, const double* C,
                             double* thetaH, double* thetaC, uint32_t* error, hipStream_t stream)
{
    const dim3 grid((ncells + 127) / 128), block(128);
    limit_diag_kernel<<<grid, block, 0, stream>>>(d_mesh, H, C, thetaH, thetaC, error);
    return hipGetLastError();
}

hipError_t launch_face_prep(const MeshDev* d_mesh, int nx, int ny, FaceMom* fx, FaceMom* fy,
                            FaceRec* rx, FaceRec* ry, hipStream_t stream)
{
    const int64_t n = int64_t(ny) * (nx + 1) + int64_t(ny + 1) * nx;
    const dim3 grid(unsigned((n + 255) / 256)), block(256);
    face_prep_kernel<<<grid, block, 0, stream>>>(d_mesh, fx, fy, rx, ry);
    return hipGetLastError();
}

hipError_t launch_face_rec64(const MeshDev* d_mesh, int nx, int ny, FaceRec64* rx, FaceRec64* ry,
                             hipStream_t stream)
{
    const int64_t n = int64_t(ny) * (nx + 1) + int64_t(ny + 1) * nx;
    const dim3 grid(unsigned((n + 255) / 256)), block(256);
    face_rec64_kernel<<<grid, block, 0, stream>>>(d_mesh, rx, ry);
    return hipGetLastError();
}

hipError_t launch_recon_diag(int mode, const MeshDev* d_mesh, int ncells, const double* u,
                             int field, double* sigma, double* omega, double* traces,
                             uint32_t* error, hipStream_t stream)
{
    const dim3 grid((ncells + 127) / 128), block(128);
    if (mode == 0) recon_diag_kernel<0><<<grid, block, 0, stream>>>(d_mesh, u, field, sigma, omega, traces, error);
    else recon_diag_kernel<1><<<grid, block, 0, stream>>>(d_mesh, u, field, sigma, omega, traces, error);
    return hipGetLastError();
}

hipError_t launch_rhs_diag(int mode, const MeshDev* d_mesh, int ncells, const double* H,
                           const double* C, double* rhsH, double* rhsC, double* bnd,
                           uint32_t* error, hipStream_t stream)
{
    const dim3 grid((ncells + 127) / 128), block(128);
    if (mode == 0) rhs_diag_kernel<0><<<grid, block, 0, stream>>>(d_mesh, H, C, rhsH, rhsC, bnd, error);
    else rhs_diag_kernel<1><<<grid, block, 0, stream>>>(d_mesh, H, C, rhsH, rhsC, bnd, error);
    return hipGetLastError();
}

hipError_t launch_point_eval(int mode, const MeshDev* d_mesh, const double* H, const double* C,
                             int64_t n, const int32_t* cell, const double* X, const double* Y,
                             double* outH, double* outC, uint32_t* error, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    const dim3 grid(unsigned((n + 255) / 256)), block(256);
    if (mode == 0) point_eval_kernel<0><<<grid, block, 0, stream>>>(d_mesh, H, C, n, cell, X, Y, outH, outC, error);
    else point_eval_kernel<1><<<grid, block, 0, stream>>>(d_mesh, H, C, n, cell, X, Y, outH, outC, error);
    return hipGetLastError();
}

hipError_t launch_rate(const MeshDev* d_mesh, int ncells, unsigned long long* max_bits,
                       hipStream_t stream)
{
    const dim3 grid(std::min((ncells + 255) / 256, 2048)), block(256);
    rate_kernel<<<grid, block, 0, stream>>>(d_mesh, max_bits);
    return hipGetLastError();
}

hipError_t launch_phase(const PhaseDev& ph, double TDe0, double nu, double dT, int64_t n,
                        const double* H, const double* C, const double* P, double* T,
                        double* phi, double* Cl, double* Cs, uint8_t* region, double* phi1,
                        double* phi2, double* dTdC, double* dTdH, uint32_t* error,
                        hipStream_t stream)
{
    const dim3 grid(unsigned((n + 255) / 256)), block(256);
    phase_kernel<<<grid, block, 0, stream>>>(ph, TDe0, nu, dT, n, H, C, P, T, phi, Cl, Cs, region,
                                             phi1, phi2, dTdC, dTdH, error);
    return hipGetLastError();
}

hipError_t launch_state_phase(const MeshDev* d_mesh, int ncells, const double* H, const double* C,
                              double* T, double* phi, double* Cl, double* Cs, uint8_t* region,
                              uint32_t* error, hipStream_t stream)
{
    const dim3 grid((ncells + 255) / 256), block(256);
    state_phase_kernel<<<grid, block, 0, stream>>>(d_mesh, H, C, T, phi, Cl, Cs, region, error);
    return hipGetLastError();
}

hipError_t launch_reduce_step(const double* partials, int nitems, int nstages, const double* w,
                              double area, double* out, double* scratch, hipStream_t stream,
                              const double* prev, double dt, double tol, uint32_t* error)
{
    reduce_step_kernel<<<kStepSlices, kWave, 0, stream>>>(partials, nitems, nstages, w[0], w[1], w[2],
                                                           scratch);
    reduce_step_final_kernel<<<1, kStepSlices, 0, stream>>>(scratch, area, out, prev, dt, tol, error);
    return hipGetLastError();
}

hipError_t launch_state_sums(const double* H, const double* C, int64_t n, double area, double* out,
                             double* scratch, hipStream_t stream)
{
    state_sums_kernel<<<kStepSlices, 256, 0, stream>>>(H, C, n, scratch);
    reduce_step_final_kernel<<<1, kStepSlices, 0, stream>>>(scratch, area, out, nullptr, 0.0, 0.0, nullptr);
    return hipGetLastError();
}

hipError_t launch_batch_stage(int mode, int q, const MeshDev* meshes,
                              const StagePtrs* stages, const int2* items, int nitems, int blocks,
                              int* queue, hipStream_t stream, bool global_lf, int limit, bool diffusion)
{
    if (nitems <= 0) return hipSuccess;
    // A persistent grid (blocks > 0) runs 4-wave workgroups that take items from
    // the queue; otherwise one wave per workgroup and per item.
    int* qp = blocks > 0 ? queue : nullptr;
    const dim3 grid(blocks > 0 ? blocks : nitems), block(blocks > 0 ? 256 : kWave);
    if (mode != 0)
        return launch_batch_stage32(meshes, stages, items, nitems, blocks, queue, global_lf, stream, limit,
                                    diffusion);
#define MP_BATCH64(N, L, D) \
    batch_sweep_kernel<0, N, L, D><<<grid, block, 0, stream>>>(meshes, stages, items, nitems, qp)
#define MP_BATCH64_LD(N)                                                     \
    do {                                                                     \
        if (limit && diffusion) MP_BATCH64(N, true, true);                   \
        else if (limit) MP_BATCH64(N, true, false);                          \
        else if (diffusion) MP_BATCH64(N, false, true);                      \
        else MP_BATCH64(N, false, false);                                    \
    } while (0)
    if (q == 5) MP_BATCH64_LD(5);
    else MP_BATCH64_LD(0);
#undef MP_BATCH64_LD
#undef MP_BATCH64
    return hipGetLastError();
}

int batch_blocks_per_cu(int mode)
{
    if (mode != 0) return batch_blocks_per_cu32();
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, batch_sweep_kernel<0, 0>, 256, 0) != hipSuccess)
        return 0;
    return n;
}

hipError_t launch_batch_reduce(const StagePtrs* stages, const int* nitems, const MeshDev* meshes,
                               const int* slots, int nslots, int nstages, const double* w,
                               double* out, hipStream_t stream, const double* prev, double tol)
{
    if (nslots <= 0) return hipSuccess;
    batch_reduce_kernel<<<nslots, 256, 0, stream>>>(stages, nitems, meshes, slots, nstages, w[0], w[1],
                                                    w[2], out, prev, tol);
    return hipGetLastError();
}

hipError_t launch_batch_state_sums(const StagePtrs* stages, const MeshDev* meshes, const int* slots,
                                   int nslots, double* out, double* part, hipStream_t stream)
{
    if (nslots <= 0) return hipSuccess;
    batch_state_sums_kernel<<<dim3(unsigned(nslots), kStateSlices), 256, 0, stream>>>(stages, meshes, slots,
                                                                                      part);
    batch_state_sums_final_kernel<<<nslots, 64, 0, stream>>>(meshes, slots, part, out);
    return hipGetLastError();
}

hipError_t launch_clear(uint32_t* p, int n, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    clear_kernel<<<(n + 255) / 256, 256, 0, stream>>>(p, n);
    return hipGetLastError();
}

}  // namespace mp
