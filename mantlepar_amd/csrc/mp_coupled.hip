// Coupled transport law on the device: one RHS of MantlePar's phase-coupled
// evolution (src/couple/evolve.cpp:279-384) on a uniform rectangular mesh.
//
// Three launches per RHS, all FP64:
//   coupled_cell_kernel   one thread per cell: limited blends of H and C
//                         (the reconstruction + StateReconstruction::Limit),
//                         stored as 20 doubles per cell, and the cell
//                         quadrature (cooling source, cell-averaged T and phi)
//   coupled_face_kernel   one thread per face: phase at the face Gauss points
//                         on both sides, PhaseTransportVelocities, the T flux
//                         with its latent term and the C flux with the
//                         effective velocity (LLF), boundary policies, the
//                         CFL rate with the phase sensitivity, and the
//                         normal-line diffusion of T from phase evaluations at
//                         the in-cell samples
//   coupled_gather_kernel one thread per cell: the face fluxes into rhs_H,
//                         rhs_C, rate and mixture divergence (a gather, no
//                         atomics on state)
// followed by a fixed-order reduction of the totals.  Each face is evaluated
// once, so every phase evaluation of the reference (about 85 per cell) is
// made once.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "mp_common.h"
#include "mp_coupled.h"
#include "mp_device.h"
#include "mp_limit.h"
#include "mp_phase.h"
#include "mp_recon.h"

namespace mp {

namespace {

__device__ __forceinline__ double poly_at(const double* B, double x, double y)
{
    // B[b*3 + a]: coefficient of x^a y^b
    const double q0 = B[0] + x * (B[1] + x * B[2]);
    const double q1 = B[3] + x * (B[4] + x * B[5]);
    const double q2 = B[6] + x * (B[7] + x * B[8]);
    return q0 + y * (q1 + y * q2);
}

// Limited state of cell c at normalised offset (x, y): mean + theta*P.
struct CellPoly {
    double tH, tC, BH[9], BC[9];
};

__device__ __forceinline__ void load_cell(const CoupledDev& d, size_t c, size_t N, CellPoly& p)
{
    p.tH = d.blend[c];
    p.tC = d.blend[N + c];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        p.BH[k] = d.blend[(2 + k) * N + c];
        p.BC[k] = d.blend[(11 + k) * N + c];
    }
}

// Phase of the limited reconstruction at a point, with the coupled law's
// range check (src/couple/evolve.cpp:270-277); (tab, k): the point's
// lithostatic entry (tab null: the constant pressure of ph).  The override is
// formed here, in registers, not at the call site.  Inlined: as a call, every
// evaluation saved and restored the caller's registers through 512 B of
// scratch per lane, which bound the face kernel (VALU busy ~10%).
template <bool DERIV>
__device__ __forceinline__ PhaseOut point_phase(const PhaseDev& ph, const TepPair* tab, size_t k,
                                             const CellPoly& p, double x, double y, double& H,
                                             double& C, uint32_t& err)
{
    H = p.tH + poly_at(p.BH, x, y);
    C = p.tC + poly_at(p.BC, x, y);
    const PhaseOut o = phase_eval<DERIV>(phase_at(ph, tab, k), H, C, err);
    if (!(o.T >= 0.0 && o.phil < 1.0)) err |= kErrCoupledRange;
    return o;
}

// Lithostatic entry of normal-line sample k of face point fq: a vertical
// face's samples share the point's ordinate, a horizontal face's have their own.
__device__ __forceinline__ const TepPair* sample_table(const CoupledDev& d, bool vert)
{
    return vert ? d.tep_v : d.tep_hs;
}
__device__ __forceinline__ size_t sample_index(const CoupledDev& d, bool vert, size_t fq, int k)
{
    return vert ? fq : fq * d.n + k;
}

// The per-thread bodies of the three RHS kernels, the reductions and the stage
// update, shared by the single-mesh kernels and their batched forms, so a
// mesh gives the same bits in both.
__device__ __forceinline__ void cell_body(const CoupledDev& d, uint32_t* error, const double* __restrict__ H,
                                          const double* __restrict__ C, int c)
{
    const MeshDev& m = *d.m;
    const int nx = m.nx, ny = m.ny;
    if (c >= nx * ny) return;
    const size_t N = size_t(nx) * ny;
    const int i = c % nx, j = c / nx;
    uint32_t err = 0;
    const unsigned mask = candidate_mask(i, j, nx, ny);
    const unsigned sides = (j == 0 ? 1u : 0u) | (i == nx - 1 ? 2u : 0u) | (j == ny - 1 ? 4u : 0u) |
                           (i == 0 ? 8u : 0u);
    CellPoly p;
    // Each field writes its own members: a destination chosen at run time
    // would put p in scratch memory.
    auto limited = [&](const double* U, double cap, double& mean, double (&out)[9]) {
        double w[3][3], B[3][3];
#pragma unroll
        for (int dj = -1; dj <= 1; ++dj)
#pragma unroll
            for (int di = -1; di <= 1; ++di)
                w[dj + 1][di + 1] = U[size_t(clampi(j + dj, 0, ny - 1)) * nx + clampi(i + di, 0, nx - 1)];
        blend<double>(m, w, mask, B, err);
        const double th = limit_factor<0>(m, w, B, sides, cap, err);
#pragma unroll
        for (int b = 0; b < 3; ++b)
#pragma unroll
            for (int a = 0; a < 3; ++a) out[b * 3 + a] = th * B[b][a];
        mean = w[1][1];
    };
    limited(H, INFINITY, p.tH, p.BH);
    limited(C, m.lim_cupper, p.tC, p.BC);
    d.blend[c] = p.tH;
    d.blend[N + c] = p.tC;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        d.blend[(2 + k) * N + c] = p.BH[k];
        d.blend[(11 + k) * N + c] = p.BC[k];
    }
    // cell quadrature (src/couple/evolve.cpp:357-370): b along y outer, a along x
    double heat = 0, T = 0, phi = 0, srate = 0;
    const int n = m.lim_ncp;
#pragma unroll 1
    for (int b = 0; b < n; ++b)
#pragma unroll 1
        for (int a = 0; a < n; ++a) {
            double Hp, Cp;
            const PhaseOut o = point_phase<false>(m.ph, d.tep_cell, (size_t(j) * n + b) * n + a, p,
                                                  m.lim_cx[a], m.lim_cy[b], Hp, Cp, err);
            const double wab = m.lim_cw[a] * m.lim_cw[b] * 0.25;      // weight / area
            const double mixy = d.cmix[size_t(b * n + a) * N + c];
            const double source = -d.cooling * o.T * mixy;
            heat += wab * source;
            T += wab * o.T;
            phi += wab * o.phil;
            srate += wab * (d.cooling * fabs(mixy));
        }
    d.cell_out[c] = heat;
    d.cell_out[N + c] = T;
    d.cell_out[2 * N + c] = phi;
    d.cell_out[3 * N + c] = srate;
    if (err) atomicOr(error, err);
}

// One face.  Vertical faces f < ny*(nx+1) at x_i (i = f % (nx+1), j = f / (nx+1)),
// horizontal faces after them at y_j.  Outputs in the canonical direction
// (+x / +y): flux_H, flux_C, mixture flux, and the rate (same for both sides).
__device__ __forceinline__ void face_body(const CoupledDev& d, uint32_t* error, int64_t f)
{
    const MeshDev& m = *d.m;
    const int nx = m.nx, ny = m.ny, nq = m.q;
    const size_t N = size_t(nx) * ny;
    const int64_t nxf = int64_t(ny) * (nx + 1), nyf = int64_t(ny + 1) * nx;
    if (f >= nxf + nyf) return;
    const bool vert = f < nxf;
    const int64_t fl = vert ? f : f - nxf;
    int i, j;
    if (vert) { i = int(fl % (nx + 1)); j = int(fl / (nx + 1)); }
    else { i = int(fl % nx); j = int(fl / nx); }
    const int along = vert ? i : j, limit = vert ? nx : ny;
    const bool boundary = along == 0 || along == limit;
    // boundary side (0 bottom, 1 right, 2 top, 3 left), its face index, the
    // outward sign against the canonical direction and the interior cell
    const int side = vert ? (i == 0 ? 3 : 1) : (j == 0 ? 0 : 2);
    const int kface = vert ? j : i;
    const double sgn = (along == 0) ? -1.0 : 1.0;
    // cells: L behind the canonical normal, R ahead; on the boundary A is the
    // interior cell
    const int li = vert ? i - 1 : i, lj = vert ? j : j - 1;
    uint32_t err = 0;
    CellPoly P, Q;          // P: left (or the boundary's interior cell), Q: right
    const int ai = boundary ? (vert ? (i == 0 ? 0 : nx - 1) : i) : li;
    const int aj = boundary ? (vert ? j : (j == 0 ? 0 : ny - 1)) : lj;
    load_cell(d, size_t(aj) * nx + ai, N, P);
    if (!boundary) load_cell(d, size_t(j) * nx + i, N, Q);
    // trace offsets: the face sits at +h (interior left cell, or boundary with
    // outward +) / -h of the cell holding the point
    const double hf = vert ? m.hx : m.hy;
    const double pf = boundary ? sgn * hf : hf;     // in P's coordinates
    const double* s = vert ? m.sv : m.sh;
    const double* W = vert ? m.wv : m.wh;
    const int64_t fstride = vert ? nxf : nyf;
    const double* vsa = (vert ? d.xvs : d.yvs) + fl;
    const double* qda = (vert ? d.xqd : d.yqd) + fl;
    const double* pha = (vert ? d.xphi : d.yphi) + fl;
    // lithostatic tables of this face's points (both sides share a point)
    const TepPair* tpt = vert ? d.tep_v : d.tep_h;
    const size_t trow = size_t(j) * nq;
    double fluxH = 0, fluxC = 0, fluxMix = 0, rate = 0;
    uint8_t polH = 0, polC = 0;
    const double *inH = nullptr, *inC = nullptr;
    if (boundary) {
        polH = d.bcH[side] ? d.bcH[side][kface] : 0;
        polC = d.bcC[side] ? d.bcC[side][kface] : 0;
        inH = d.inH[side] ? d.inH[side] + size_t(kface) * nq : nullptr;
        inC = d.inC[side] ? d.inC[side] + size_t(kface) * nq : nullptr;
    }
#pragma unroll 1
    for (int q = 0; q < nq; ++q) {
        // outward (boundary) or canonical (interior) normal components
        const double vs = (boundary ? sgn : 1.0) * vsa[q * fstride];
        const double qd = (boundary ? sgn : 1.0) * qda[q * fstride];
        const double ephi = pha[q * fstride];
        const double xa = vert ? pf : s[q], ya = vert ? s[q] : pf;
        double hL, cL, hR, cR;
        const PhaseOut a = point_phase<true>(m.ph, tpt, trow + q, P, xa, ya, hL, cL, err);
        PhaseOut b = a;
        cR = cL;
        if (!boundary) {
            const double xb = vert ? -hf : s[q], yb = vert ? s[q] : -hf;
            b = point_phase<true>(m.ph, tpt, trow + q, Q, xb, yb, hR, cR, err);
        }
        (void)hR;
        // PhaseTransportVelocities (src/couple/evolve.cpp:36-52) along the normal
        const double cl = .5 * (a.cl + b.cl), cs = .5 * (a.cs + b.cs);
        const double bulk = ephi * cl + (1 - ephi) * cs;
        const double vn = vs + qd;
        const double ve = vs + (bulk > 0 ? cl / bulk * qd : 0);
        const double alphaH = d.lf_mode ? d.global_speed : fabs(vn);
        const double alphaC = d.lf_mode ? d.global_speed : fabs(ve);
        if (!(alphaH + 1e-14 >= fabs(vn) && alphaC + 1e-14 >= fabs(ve))) err |= kErrLFSpeed;
        const double LD = m.ph.LD;
        double fh = .5 * (vn * (a.T + b.T) - alphaH * (b.T - a.T)) - LD * (1 - ephi) * vs;
        double fc = .5 * (ve * (cL + cR) - alphaC * (cR - cL));
        if (boundary) {
            // EvaluateAdvectionBoundary (src/couple/initialization.cpp:469-489)
            // for H along vn and for C along ve; 0 prescribed, 1 extrapolated, 2 zero flux
            int tH = 2, tC = 2;
            double valH = 0, valC = 0;
            if (!(polH == 0 || vn == 0)) {
                if (vn < 0) {
                    if (polH != 2) err |= kErrBackflow;
                    else if (!inH) err |= kErrInflow;
                    else { tH = 0; valH = inH[q]; }
                } else tH = 1;
            }
            if (!(polC == 0 || ve == 0)) {
                if (ve < 0) {
                    if (polC != 2) err |= kErrBackflow;
                    else if (!inC) err |= kErrInflow;
                    else { tC = 0; valC = inC[q]; }
                } else tC = 1;
            }
            const double inflowC = tC == 0 ? valC : cL;
            double T = a.T;
            if (tH == 0) {
                uint32_t e2 = 0;
                T = phase_eval<false>(phase_at(m.ph, tpt, trow + q), valH, inflowC, e2).T;
                err |= e2;
            }
            fh = tH == 2 ? 0 : vn * T - LD * (1 - ephi) * vs;
            fc = tC == 2 ? 0 : ve * (tC == 0 ? valC : cL);
        }
        const double wq = W[q];
        fluxH += wq * fh;
        fluxC += wq * fc;
        fluxMix += wq * vn;
        const double sens = 1 + fmax(fabs(a.dTdC), fabs(b.dTdC));
        if (!isfinite(sens)) err |= kErrNonFinite;
        rate += wq * (alphaC + alphaH + 2 * fabs(vs)) * sens;
    }
    // normal-line diffusion of T (src/couple/evolve.cpp:334-350); the samples'
    // phases are evaluated (and range-checked) also when kappa is 0
    {
        const int n = d.n;
        const double ds = boundary ? (vert ? d.dsbx : d.dsby) : (vert ? d.dsx : d.dsy);
        const double* depth = boundary ? (vert ? m.lim_xb : m.lim_yb) : (vert ? m.lim_xi : m.lim_yi);
        double dflux = 0;
#pragma unroll 1
        for (int q = 0; q < nq; ++q) {
            const double wq = W[q];
            int pol = 1;
            double bval = 0;
            if (boundary) {
                pol = m.tbc[side] ? m.tbc[side][size_t(kface) * nq + q] : 0;
                bval = m.tval[side] ? m.tval[side][size_t(kface) * nq + q] : 0;
                if (pol == 0) continue;
                if (pol == 2) { dflux += wq * bval; continue; }
            }
            // samples k = 0..count-1 from the far side of the left (interior) cell
            double Tk[2 * kMaxLimS];
            const int nin = boundary ? n : n / 2;
#pragma unroll 1
            for (int k = 0; k < nin; ++k) {
                const double off = boundary ? sgn * depth[k] : depth[k];
                double h_, c_;
                Tk[k] = point_phase<false>(m.ph, sample_table(d, vert), sample_index(d, vert, trow + q, k),
                                           P, vert ? off : s[q], vert ? s[q] : off, h_, c_, err).T;
            }
            double last;
            if (boundary) {
                last = bval;
            } else {
#pragma unroll 1
                for (int k = n / 2; k < n; ++k) {
                    const double off = -depth[n - 1 - k];
                    double h_, c_;
                    Tk[k] = point_phase<false>(m.ph, sample_table(d, vert),
                                               sample_index(d, vert, trow + q, k), Q,
                                               vert ? off : s[q], vert ? s[q] : off, h_, c_, err).T;
                }
                last = Tk[n - 1];
            }
            const double* cw = boundary ? d.ce : d.cm;
            const int count = boundary ? n + 1 : n;
            double g = 0;
            for (int k = 0; k + 1 < count; ++k) g += cw[k] * (Tk[k] - last);
            g /= ds;
            dflux -= wq * d.kappa * g;
            rate += wq * d.kappa / ds * (boundary ? d.ce_abs : d.cm_abs);
        }
        fluxH += dflux;
    }
    // canonical orientation for storage
    const double o = boundary ? sgn : 1.0;
    d.face[f] = o * fluxH;
    d.face[(nxf + nyf) + f] = o * fluxC;
    d.face[2 * (nxf + nyf) + f] = o * fluxMix;
    d.face[3 * (nxf + nyf) + f] = rate;
    if (err) atomicOr(error, err);
}

__device__ __forceinline__ void gather_body(const CoupledDev& d, double* rhsH, double* rhsC, double* rate,
                                            double* div, double* bnd, int c)
{
    const MeshDev& m = *d.m;
    const int nx = m.nx, ny = m.ny;
    if (c >= nx * ny) return;
    const size_t N = size_t(nx) * ny;
    const int i = c % nx, j = c / nx;
    const int64_t nxf = int64_t(ny) * (nx + 1), nf = nxf + int64_t(ny + 1) * nx;
    const int64_t fl = int64_t(j) * (nx + 1) + i, fr = fl + 1;
    const int64_t fb = nxf + int64_t(j) * nx + i, ft = fb + nx;
    const double* F = d.face;
    const double ia = m.inv_area;
    rhsH[c] = (F[fl] - F[fr] + F[fb] - F[ft]) * ia + d.cell_out[c];
    rhsC[c] = (F[nf + fl] - F[nf + fr] + F[nf + fb] - F[nf + ft]) * ia;
    div[c] = (F[2 * nf + fr] - F[2 * nf + fl] + F[2 * nf + ft] - F[2 * nf + fb]) * ia;
    rate[c] = (F[3 * nf + fl] + F[3 * nf + fr] + F[3 * nf + fb] + F[3 * nf + ft]) * ia +
              d.cell_out[3 * N + c];
    // outward boundary fluxes of this cell (H, C) and its source total
    double bh = 0, bc = 0;
    if (i == 0) { bh -= F[fl]; bc -= F[nf + fl]; }
    if (i == nx - 1) { bh += F[fr]; bc += F[nf + fr]; }
    if (j == 0) { bh -= F[fb]; bc -= F[nf + fb]; }
    if (j == ny - 1) { bh += F[ft]; bc += F[nf + ft]; }
    bnd[c] = bh;
    bnd[N + c] = bc;
    bnd[2 * N + c] = d.cell_out[c] * m.area;
}

}  // namespace

// Launched with 128 threads per block; the bound lets the compiler use the
// registers that a 1024-thread default would deny (202 VGPRs in the face
// kernel, no scratch).
__global__ __launch_bounds__(128) void coupled_cell_kernel(CoupledDev d, const double* __restrict__ H,
                                    const double* __restrict__ C)
{
    cell_body(d, d.error, H, C, blockIdx.x * blockDim.x + threadIdx.x);
}

__global__ __launch_bounds__(128) void coupled_face_kernel(CoupledDev d)
{
    face_body(d, d.error, int64_t(blockIdx.x) * blockDim.x + threadIdx.x);
}

__global__ void coupled_gather_kernel(CoupledDev d, double* rhsH, double* rhsC, double* rate,
                                      double* div, double* bnd)
{
    gather_body(d, rhsH, rhsC, rate, div, bnd, blockIdx.x * blockDim.x + threadIdx.x);
}

// Fixed-order reductions of k arrays of n values: sum (Neumaier), or max |x|,
// min, max.  Two passes: block b reduces the contiguous chunk b of every
// array in a fixed tree, then one block combines the partials in order, so the
// result does not depend on scheduling.
__device__ __forceinline__ void red_combine(int mode, double& s, double& e, double o, double oe)
{
    if (mode == 0) {
        e += oe;
        two_sum_acc(s, e, o);
    } else if (mode == 1) s = fmax(s, fabs(o)) + 0.0 * o;   // a NaN propagates
    else if (mode == 2) s = (o < s || o != o) ? o : s;
    else s = (o > s || o != o) ? o : s;
}

__device__ __forceinline__ double red_init(int mode)
{
    return mode == 2 ? INFINITY : (mode == 3 ? -INFINITY : 0.0);
}

__device__ void block_reduce(int mode, double& s, double& e, double* ss, double* se)
{
    ss[threadIdx.x] = s;
    se[threadIdx.x] = e;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            double a = ss[threadIdx.x], ae = se[threadIdx.x];
            red_combine(mode, a, ae, ss[threadIdx.x + w], se[threadIdx.x + w]);
            ss[threadIdx.x] = a;
            se[threadIdx.x] = ae;
        }
        __syncthreads();
    }
    s = ss[0];
    e = se[0];
    __syncthreads();
}

// Every array of the chunk is read in the same loop, so a thread has loads
// of all k arrays in flight at once (k <= kReduceMax).
constexpr int kReduceMax = 6;

// Block b of nb reduces chunk b of every array into part.
__device__ __forceinline__ void reduce_partial_body(const double* __restrict__ x, int64_t n, int k,
                                                    const int* kind, const int* src, double* part,
                                                    int nb, int b, double* ss, double* se)
{
    const int64_t chunk = (n + nb - 1) / nb;
    const int64_t b0 = int64_t(b) * chunk, b1 = min(n, b0 + chunk);
    const double* v[kReduceMax];
    int mode[kReduceMax];
    double s[kReduceMax], e[kReduceMax];
#pragma unroll
    for (int a = 0; a < kReduceMax; ++a) {
        const int aa = a < k ? a : 0;
        v[a] = x + (src ? src[aa] : aa) * n;
        mode[a] = kind[aa];
        s[a] = red_init(mode[a]);
        e[a] = 0.0;
    }
    for (int64_t t = b0 + threadIdx.x; t < b1; t += blockDim.x) {
        double y[kReduceMax];
#pragma unroll
        for (int a = 0; a < kReduceMax; ++a) y[a] = a < k ? v[a][t] : 0.0;
#pragma unroll
        for (int a = 0; a < kReduceMax; ++a) {
            if (a >= k) break;
            if (mode[a] == 0) two_sum_acc(s[a], e[a], y[a]);
            else red_combine(mode[a], s[a], e[a], y[a], 0.0);
        }
    }
#pragma unroll
    for (int a = 0; a < kReduceMax; ++a) {
        if (a >= k) break;
        double sa = s[a], ea = e[a];
        block_reduce(mode[a], sa, ea, ss, se);
        if (threadIdx.x == 0) {
            part[(size_t(a) * nb + b) * 2] = sa;
            part[(size_t(a) * nb + b) * 2 + 1] = ea;
        }
    }
}

__global__ __launch_bounds__(256) void reduce_partial_kernel(const double* __restrict__ x, int64_t n,
                                                             int k, const int* kind, const int* src,
                                                             double* part)
{
    __shared__ double ss[256], se[256];
    reduce_partial_body(x, n, k, kind, src, part, int(gridDim.x), int(blockIdx.x), ss, se);
}

__device__ __forceinline__ void reduce_final_body(const double* __restrict__ part, int nb, int k,
                                                  const int* kind, double* out, double* ss, double* se)
{
    int mode[kReduceMax];
    double s[kReduceMax], e[kReduceMax];
#pragma unroll
    for (int a = 0; a < kReduceMax; ++a) {
        mode[a] = kind[a < k ? a : 0];
        s[a] = red_init(mode[a]);
        e[a] = 0.0;
    }
    for (int t = threadIdx.x; t < nb; t += blockDim.x) {
#pragma unroll
        for (int a = 0; a < kReduceMax; ++a) {
            if (a >= k) break;
            red_combine(mode[a], s[a], e[a], part[(size_t(a) * nb + t) * 2], part[(size_t(a) * nb + t) * 2 + 1]);
        }
    }
#pragma unroll
    for (int a = 0; a < kReduceMax; ++a) {
        if (a >= k) break;
        double sa = s[a], ea = e[a];
        block_reduce(mode[a], sa, ea, ss, se);
        if (threadIdx.x == 0) out[a] = mode[a] == 0 ? sa + ea : sa;
    }
}

__global__ __launch_bounds__(256) void reduce_final_kernel(const double* __restrict__ part, int nb,
                                                           int k, const int* kind, double* out)
{
    __shared__ double ss[256], se[256];
    reduce_final_body(part, nb, k, kind, out, ss, se);
}

// SSPRK2 stage combinations of src/couple/evolve.cpp:469-479:
// stage 1: u = u_old + dt*r1;  stage 2: u = 0.5*u_old + 0.5*(u + dt*r2).
__device__ __forceinline__ void stage_body(int64_t n, int stage, double dt, const double* __restrict__ oldH,
                                           const double* __restrict__ oldC, const double* __restrict__ rH,
                                           const double* __restrict__ rC, double* H, double* C, int64_t k)
{
    if (k >= n) return;
    if (stage == 1) {
        H[k] = dt * rH[k] + oldH[k];
        C[k] = dt * rC[k] + oldC[k];
    } else {
        const double h = H[k] + dt * rH[k], c = C[k] + dt * rC[k];
        H[k] = .5 * oldH[k] + .5 * h;
        C[k] = .5 * oldC[k] + .5 * c;
    }
}

__global__ void coupled_stage_kernel(int64_t n, int stage, double dt, const double* __restrict__ oldH,
                                     const double* __restrict__ oldC, const double* __restrict__ rH,
                                     const double* __restrict__ rC, double* H, double* C)
{
    stage_body(n, stage, dt, oldH, oldC, rH, rC, H, C, int64_t(blockIdx.x) * blockDim.x + threadIdx.x);
}

hipError_t launch_coupled_rhs(const CoupledDev& d, int nx, int ny, const double* H, const double* C,
                              double* rhsH, double* rhsC, double* rate, double* div, double* bnd,
                              hipStream_t stream)
{
    const int N = nx * ny;
    const int64_t nf = int64_t(ny) * (nx + 1) + int64_t(ny + 1) * nx;
    coupled_cell_kernel<<<(N + 127) / 128, 128, 0, stream>>>(d, H, C);
    coupled_face_kernel<<<unsigned((nf + 127) / 128), 128, 0, stream>>>(d);
    coupled_gather_kernel<<<(N + 255) / 256, 256, 0, stream>>>(d, rhsH, rhsC, rate, div, bnd);
    return hipGetLastError();
}

int coupled_reduce_blocks(int64_t n)
{
    return int(std::min<int64_t>(kReducePartials, std::max<int64_t>(1, (n + 2047) / 2048)));
}

hipError_t launch_reduce(const double* x, int64_t n, int k, const int* kind, const int* src,
                         double* out, double* part, hipStream_t stream)
{
    if (k < 1 || k > kReduceMax) return hipErrorInvalidValue;
    const int nb = coupled_reduce_blocks(n);
    reduce_partial_kernel<<<nb, 256, 0, stream>>>(x, n, k, kind, src, part);
    reduce_final_kernel<<<1, 256, 0, stream>>>(part, nb, k, kind, out);
    return hipGetLastError();
}

hipError_t launch_coupled_stage(int64_t n, int stage, double dt, const double* oldH,
                                const double* oldC, const double* rH, const double* rC, double* H,
                                double* C, hipStream_t stream)
{
    coupled_stage_kernel<<<unsigned((n + 255) / 256), 256, 0, stream>>>(n, stage, dt, oldH, oldC,
                                                                        rH, rC, H, C);
    return hipGetLastError();
}

// ============================================================ batched form
// Each kernel covers the workgroups of every handle of its span; a workgroup
// finds its handle by bisecting the span's base table (wave-uniform scalar
// loads) and runs the single-mesh body on its local index.

namespace {

// The k < n with base[k] <= b < base[k+1].
__device__ __forceinline__ int span_of(const int* __restrict__ base, int n, int b)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (base[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

}  // namespace

__global__ void coupled_batch_clear_kernel(const CoupledBatchMesh* __restrict__ M, const int* __restrict__ act,
                                           int nact, int first, int count)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nact * count) return;
    M[act[t / count]].d.error[first + t % count] = 0u;
}

__global__ void coupled_batch_copy_kernel(const CoupledBatchMesh* __restrict__ M, const int* __restrict__ seg,
                                          const int* __restrict__ base, int nseg)
{
    const int s = span_of(base, nseg, int(blockIdx.x));
    const CoupledBatchMesh& B = M[seg[2 * s]];
    const int kind = seg[2 * s + 1];
    const double* src = kind == 0 ? B.R : (kind == 1 ? B.S : B.O);
    double* dst = kind == 0 ? B.F : (kind == 1 ? B.O : B.S);
    const int64_t k = int64_t(int(blockIdx.x) - base[s]) * blockDim.x + threadIdx.x;
    if (k < 2 * B.N) dst[k] = src[k];
}

__global__ void coupled_batch_stage_kernel(const CoupledBatchMesh* __restrict__ M, CoupledBatchSpan sp, int stage,
                                           const double* __restrict__ dt)
{
    const int k = span_of(sp.base, sp.nact, int(blockIdx.x));
    const CoupledBatchMesh& B = M[sp.act[k]];
    const double* r = stage == 1 ? B.F : B.R;
    stage_body(B.N, stage, dt[k], B.O, B.O + B.N, r, r + B.N, B.S, B.S + B.N,
               int64_t(int(blockIdx.x) - sp.base[k]) * blockDim.x + threadIdx.x);
}

__global__ __launch_bounds__(128) void coupled_batch_cell_kernel(const CoupledBatchMesh* __restrict__ M,
                                                                 CoupledBatchSpan sp, int slot)
{
    const int k = span_of(sp.base, sp.nact, int(blockIdx.x));
    const CoupledBatchMesh& B = M[sp.act[k]];
    cell_body(B.d, B.d.error + slot, B.S, B.S + B.N, (int(blockIdx.x) - sp.base[k]) * blockDim.x + threadIdx.x);
}

__global__ __launch_bounds__(128) void coupled_batch_face_kernel(const CoupledBatchMesh* __restrict__ M,
                                                                 CoupledBatchSpan sp, int slot)
{
    const int k = span_of(sp.base, sp.nact, int(blockIdx.x));
    const CoupledBatchMesh& B = M[sp.act[k]];
    face_body(B.d, B.d.error + slot, int64_t(int(blockIdx.x) - sp.base[k]) * blockDim.x + threadIdx.x);
}

__global__ void coupled_batch_gather_kernel(const CoupledBatchMesh* __restrict__ M, CoupledBatchSpan sp)
{
    const int k = span_of(sp.base, sp.nact, int(blockIdx.x));
    const CoupledBatchMesh& B = M[sp.act[k]];
    double* b = B.bnd;
    const int64_t N = B.N;
    gather_body(B.d, B.R, B.R + N, b + 3 * N, b + 4 * N, b,
                (int(blockIdx.x) - sp.base[k]) * blockDim.x + threadIdx.x);
}

// which 0: the RHS totals of bnd (5 arrays), 1: the state checks of S (6)
__global__ __launch_bounds__(256) void coupled_batch_reduce_partial_kernel(const CoupledBatchMesh* __restrict__ M,
                                                                           CoupledBatchSpan sp, int which)
{
    __shared__ double ss[256], se[256];
    const int k = span_of(sp.base, sp.nact, int(blockIdx.x));
    const CoupledBatchMesh& B = M[sp.act[k]];
    if (which == 0)
        reduce_partial_body(B.bnd, B.N, 5, B.kinds, nullptr, B.part, B.nb, int(blockIdx.x) - sp.base[k], ss, se);
    else
        reduce_partial_body(B.S, B.N, 6, B.kinds + 5, B.kinds + 11, B.part, B.nb, int(blockIdx.x) - sp.base[k],
                            ss, se);
}

__global__ __launch_bounds__(256) void coupled_batch_reduce_final_kernel(const CoupledBatchMesh* __restrict__ M,
                                                                         const int* __restrict__ act, int which,
                                                                         int off)
{
    __shared__ double ss[256], se[256];
    const CoupledBatchMesh& B = M[act[blockIdx.x]];
    reduce_final_body(B.part, B.nb, which == 0 ? 5 : 6, which == 0 ? B.kinds : B.kinds + 5, B.red + off, ss, se);
}

__global__ void coupled_batch_readback_kernel(const CoupledBatchMesh* __restrict__ M, const int* __restrict__ act,
                                              int nact, double* __restrict__ out_red, uint32_t* __restrict__ out_err)
{
    const int k = blockIdx.x;
    if (k >= nact) return;
    const CoupledBatchMesh& B = M[act[k]];
    for (int t = threadIdx.x; t < kCoupledRedSize; t += blockDim.x) out_red[size_t(k) * kCoupledRedSize + t] = B.red[t];
    if (int(threadIdx.x) < kCoupledErrSlots) out_err[size_t(k) * kCoupledErrSlots + threadIdx.x] = B.d.error[threadIdx.x];
}

hipError_t launch_coupled_batch_clear(const CoupledBatchMesh* M, const int* act, int nact, int first, int count,
                                      hipStream_t stream)
{
    if (nact <= 0) return hipSuccess;
    coupled_batch_clear_kernel<<<(nact * count + 255) / 256, 256, 0, stream>>>(M, act, nact, first, count);
    return hipGetLastError();
}

hipError_t launch_coupled_batch_copy(const CoupledBatchMesh* M, const int* seg, const int* base, int nseg,
                                     int blocks, hipStream_t stream)
{
    if (nseg <= 0) return hipSuccess;
    coupled_batch_copy_kernel<<<blocks, 256, 0, stream>>>(M, seg, base, nseg);
    return hipGetLastError();
}

hipError_t launch_coupled_batch_stage(const CoupledBatchMesh* M, CoupledBatchSpan span, int blocks, int stage,
                                      const double* dt, hipStream_t stream)
{
    if (span.nact <= 0) return hipSuccess;
    coupled_batch_stage_kernel<<<blocks, 256, 0, stream>>>(M, span, stage, dt);
    return hipGetLastError();
}

hipError_t launch_coupled_batch_rhs(const CoupledBatchMesh* M, CoupledBatchSpan cell, int cell_blocks,
                                    CoupledBatchSpan face, int face_blocks, CoupledBatchSpan gather,
                                    int gather_blocks, CoupledBatchSpan red, int red_blocks, int slot,
                                    int off, hipStream_t stream)
{
    if (cell.nact <= 0) return hipSuccess;
    coupled_batch_cell_kernel<<<cell_blocks, 128, 0, stream>>>(M, cell, slot);
    coupled_batch_face_kernel<<<face_blocks, 128, 0, stream>>>(M, face, slot);
    coupled_batch_gather_kernel<<<gather_blocks, 256, 0, stream>>>(M, gather);
    coupled_batch_reduce_partial_kernel<<<red_blocks, 256, 0, stream>>>(M, red, 0);
    coupled_batch_reduce_final_kernel<<<red.nact, 256, 0, stream>>>(M, red.act, 0, off);
    return hipGetLastError();
}

hipError_t launch_coupled_batch_admissible(const CoupledBatchMesh* M, CoupledBatchSpan red, int red_blocks,
                                           int off, hipStream_t stream)
{
    if (red.nact <= 0) return hipSuccess;
    coupled_batch_reduce_partial_kernel<<<red_blocks, 256, 0, stream>>>(M, red, 1);
    coupled_batch_reduce_final_kernel<<<red.nact, 256, 0, stream>>>(M, red.act, 1, off);
    return hipGetLastError();
}

hipError_t launch_coupled_batch_readback(const CoupledBatchMesh* M, const int* act, int nact, double* out_red,
                                         uint32_t* out_err, hipStream_t stream)
{
    if (nact <= 0) return hipSuccess;
    coupled_batch_readback_kernel<<<nact, 64, 0, stream>>>(M, act, nact, out_red, out_err);
    return hipGetLastError();
}

}  // namespace mp
