// Device helpers shared by the stage sweeps (FP64 and mixed) and the
// diagnostics kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "mp_common.h"
#include "mp_recon.h"

namespace mp {

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// The global-memory view of a pointer into device memory.  Pointers read from
// device-memory tables (MeshDev, StagePtrs) are generic to the compiler, so
// accesses through them become flat_* instructions, which also count against
// the scalar loads' wait counter; through this view they are global_*.
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T* gmem(T* p)
{
    return (__attribute__((address_space(1))) T*)p;
}

// A record of 16-byte multiples loaded through the global view.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
template <typename T>
__device__ __forceinline__ T gload(const T* p)
{
    static_assert(sizeof(T) % 16 == 0, "16-byte records only");
    u32x4 w[sizeof(T) / 16];
    const auto* g = gmem(reinterpret_cast<const u32x4*>(p));
#pragma unroll
    for (int k = 0; k < int(sizeof(T) / 16); ++k) w[k] = g[k];
    return __builtin_bit_cast(T, w);
}

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

// Row band of row r in a sweep's layout: bands of sp.rows rows below sp.split
// (the last may be shorter), of sp.tail_rows above.  Item (strip, band) is
// band * strips + strip in both segments.
__device__ __forceinline__ int band_of_row(const StagePtrs& sp, int r)
{
    if (r < sp.split) return r / sp.rows;
    return (sp.split + sp.rows - 1) / sp.rows + (r - sp.split) / sp.tail_rows;
}

// One launch per step runs the stages of a step as consecutive segments of
// its workgroups.  An item of stage k reads, of stage k-1's output, rows
// j0-2 .. j1+1 (clamped to the mesh) and columns strip*62-2 .. strip*62+63:
// the outputs of the items of strips strip-1 .. strip+1 and of the bands that
// hold those rows (at most four: only the last band below the tail items can
// be shorter than two rows).  Each item's completion is a flag set to the
// launch's epoch with release semantics at agent scope; a waiting item polls
// the flags it needs (one lane each) and then acquires, so the stores of the
// items it depends on are visible on its XCD.
//
// Forward progress: an item only waits on items of the previous stage, whose
// workgroups have lower ids.  If each hardware dispatch queue issues its
// workgroups in id order, the lowest stage with unfinished items always has a
// resident item whose inputs are done, so the launch drains.  The wait is
// still bounded: after spin_limit polls (or once another item has given up)
// the item sets kErrStall and returns without writing, the step fails, and
// the caller's state is left as it was.
__device__ __forceinline__ bool wait_inputs(const StagePtrs& sp, int ny, int strip, int j0, int j1,
                                            const uint32_t* prev, uint32_t epoch, int spin_limit)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int blo = band_of_row(sp, j0 >= 2 ? j0 - 2 : 0);
    const int bhi = band_of_row(sp, j1 + 1 <= ny - 1 ? j1 + 1 : ny - 1);
    int idx = -1;
    if (lane < 3 * (bhi - blo + 1)) {
        const int s = strip - 1 + lane % 3;
        if (s >= 0 && s < sp.strips) idx = (blo + lane / 3) * sp.strips + s;
    }
    bool ok = idx < 0;
    for (int spin = 0;; ++spin) {
        if (!ok) ok = __hip_atomic_load(prev + idx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == epoch;
        if (__all(ok)) break;
        if ((spin & 63) == 63) {
            const uint32_t e = __hip_atomic_load(sp.error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if ((e & kErrStall) || spin >= spin_limit) {
                if (lane == 0) atomicOr(sp.error, kErrStall);
                return false;
            }
        }
        __builtin_amdgcn_s_sleep(2);
    }
    // the shared rows are read with device-coherent loads issued after the
    // flags were seen; keep the compiler from moving them above this point
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    return true;
}

// Mark an item done for the items of the next stage: its shared rows were
// written with device-coherent stores; wait until every memory operation of
// the wave is acknowledged, then set the flag.
__device__ __forceinline__ void publish_item(uint32_t* flag, uint32_t epoch)
{
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    if ((threadIdx.x & (kWave - 1)) == 0)
        __hip_atomic_store(flag, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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
    // a zero-flux face carries +0 whatever its velocity: skip the loads
    if (policy == 0) return F;
    for (int q = 0; q < m.q; ++q) {
        const double v = sign * gmem(vn)[q * vstride];
        double f = 0.0;
        if (policy != 0 && v != 0.0) {
            if (v > 0.0) {
                f = v * (t + double(line_at<R>(c, R(s[q]))));
            } else if (policy == 2) {
                if (inflow) f = v * gmem(inflow)[q];
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
        auto* p = gmem(sp.partials) + size_t(item) * kPartials;
#pragma unroll
        for (int k = 0; k < (SUMS_ONLY ? 4 : 6); ++k) p[k] = v[k];
    }
    if (err) atomicOr(sp.error, err);
}

}  // namespace mp
