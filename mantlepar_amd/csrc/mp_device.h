// Device helpers shared by the stage sweeps (FP64 and mixed) and the
// diagnostics kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "mp_common.h"
#include "mp_recon.h"

namespace mp {

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// The dispatcher places workgroup b on XCD b % 8.  Renumber the workgroups so
// each XCD takes one contiguous run of work items (a slab of row bands): the
// halo rows a band shares with the band above then sit in the same L2.  A
// bijection of [0, G) for any G.
__device__ __forceinline__ int xcd_slab_block(int b, int G)
{
    const int full = G / 8, rem = G % 8, x = b % 8, k = b / 8;
    return x * full + min(x, rem) + k;
}

// Strip and rows [j0, j1) of work item `item` of a sweep.  Items below row
// sp.split are sp.rows tall, the tail items above it sp.tail_rows, each band
// of items ordered by strip.
__device__ __forceinline__ void item_region(const StagePtrs& sp, int ny, int item, int& strip, int& j0,
                                            int& j1)
{
    const int S = sp.strips;
    if (item < sp.nlong) {
        strip = item % S;
        j0 = (item / S) * sp.rows;
        j1 = min(j0 + sp.rows, sp.split);
    } else {
        const int k = item - sp.nlong;
        strip = k % S;
        j0 = sp.split + (k / S) * sp.tail_rows;
        j1 = min(j0 + sp.tail_rows, ny);
    }
}

// Work item of workgroup b in a one-wave-per-workgroup launch over one mesh:
// each segment in XCD slabs, and the short tail items after the others, so
// they are the last to be dispatched and the launch drains on short items.
__device__ __forceinline__ int item_of_block(const StagePtrs& sp, int b, int nitems)
{
    return b < sp.nlong ? xcd_slab_block(b, sp.nlong)
                        : sp.nlong + xcd_slab_block(b - sp.nlong, nitems - sp.nlong);
}

// Compensated accumulation s + e += x (Neumaier); contraction would break it.
__device__ __forceinline__ void two_sum_acc(double& s, double& e, double x)
{
#pragma clang fp contract(off)
    const double y = s + x;
    e += fabs(s) >= fabs(x) ? (s - y) + x : (x - y) + s;
    s = y;
}

// Outward boundary flux of one face (src/mlweno/advectiveflux.cpp:229-268 with
// the couple policy src/couple/initialization.cpp:469-489).  sign turns the
// canonical (+x/+y) normal velocity into the outward one.
template <typename R>
__device__ __forceinline__ double boundary_flux(const MeshDev& m, const double* vn, size_t vstride,
                                                double sign, double t, const R (&c)[3],
                                                const double* s, const double* W, uint8_t policy,
                                                const double* inflow, uint32_t& err)
{
    double F = 0.0;
    for (int q = 0; q < m.q; ++q) {
        const double v = sign * vn[q * vstride];
        double f = 0.0;
        if (policy != 0 && v != 0.0) {
            if (v > 0.0) {
                f = v * (t + double(line_at<R>(c, R(s[q]))));
            } else if (policy == 2) {
                if (inflow) f = v * inflow[q];
                else err |= kErrInflow;
            } else {
                err |= kErrBackflow;
            }
        }
        F += W[q] * f;
    }
    return F;
}

// Outward thermal flux of boundary face k of `side` (legacy preheat
// diffusion, src/couple/preheat.cpp:176-190): t is the boundary cell's
// average of H, W the face's Gauss weights, coef dif_bx or dif_by.
__device__ __forceinline__ double thermal_flux(const MeshDev& m, int side, int k, double t,
                                               const double* W, double coef)
{
    const uint8_t* pol = m.tbc[side];
    if (!pol) return 0.0;
    pol += size_t(k) * m.q;
    const double* v = m.tval[side] + size_t(k) * m.q;
    double F = 0.0;
    for (int q = 0; q < m.q; ++q) {
        if (pol[q] == 1) F += W[q] * (coef * (t - v[q]));
        else if (pol[q] == 2) F += W[q] * v[q];
    }
    return F;
}

// Wave reduction of one work item's partial sums in a fixed butterfly order
// (compensated for the state sums), stored by lane 0; error bits are or-ed
// into the launch's error word.
// SUMS_ONLY writes the state sums (slots 0-3) and leaves the boundary fluxes
// (4-5) of the item's stage as they are.
template <bool SUMS_ONLY = false>
__device__ __forceinline__ void item_partials(const StagePtrs& sp, int item, double sH, double eH,
                                              double sC, double eC, double bH, double bC,
                                              uint32_t err)
{
    double v[6] = {sH, eH, sC, eC, bH, bC};
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
    if ((threadIdx.x & (kWave - 1)) == 0 && sp.partials) {
        double* p = sp.partials + size_t(item) * kPartials;
#pragma unroll
        for (int k = 0; k < (SUMS_ONLY ? 4 : 6); ++k) p[k] = v[k];
    }
    if (err) atomicOr(sp.error, err);
}

}  // namespace mp
