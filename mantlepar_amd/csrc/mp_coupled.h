// Device records and launchers of the coupled transport law (mp_coupled.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "mp_common.h"

namespace mp {

// Everything one coupled RHS reads, by value as a kernel argument.
struct CoupledDev {
    const MeshDev* m;            // mesh, limiter tables (lim_*), thermal data (tbc/tval), phase
    double* blend;               // [20][N]: mean H, mean C, limited blends BH[9], BC[9]
    double* cell_out;            // [4][N]: heat source, <T>, <phi>, source rate
    double* face;                // [4][faces]: flux H, flux C, mixture flux (canonical), rate
    uint32_t* error;
    // flow along the canonical normal: Stokes v.n, Darcy flux phi^(1+theta) u_d.n, flow
    // porosity; vertical faces [q][ny][nx+1], horizontal faces [q][ny+1][nx]
    const double *xvs, *xqd, *xphi, *yvs, *yqd, *yphi;
    const double* cmix;          // [cell point b*n+a][N]: mixture velocity, y component
    double cooling;              // alpha0*l0*g/cp
    double kappa;                // thermal diffusivity
    double global_speed;
    int32_t lf_mode;
    int32_t n;                   // diffusion samples per interior edge point
    double dsx, dsy, dsbx, dsby; // sample spacings: interior / boundary, vertical / horizontal faces
    double cm[kMaxLimS + 1];     // midpoint derivative weights (n nodes)
    double ce[kMaxLimS + 1];     // end derivative weights (n + 1 nodes)
    double cm_abs, ce_abs;       // sums of |sample partial coefficients|
    const uint8_t* bcH[4];       // advection policies per boundary face, H and C
    const uint8_t* bcC[4];
    const double* inH[4];        // inflow states [face][q]
    const double* inC[4];
    // Lithostatic pressure (null: the constant pressure of m->ph): {Tep, Tmp}
    // at each point's ordinate -- cell points [(j*n + b)*n + a], vertical face
    // points (and their normal-line samples) [j*q + qi], horizontal face points
    // [j*q + qi] (j = 0..ny) and their samples [(j*q + qi)*n + k]
    const TepPair* tep_cell;
    const TepPair* tep_v;
    const TepPair* tep_h;
    const TepPair* tep_hs;
};

hipError_t launch_coupled_rhs(const CoupledDev& d, int nx, int ny, const double* H, const double* C,
                              double* rhsH, double* rhsC, double* rate, double* div, double* bnd,
                              hipStream_t stream);
// k reductions (1 <= k <= 6) of length-n arrays: array a is x + src[a]*n (src null: x + a*n),
// reduced by kind[a]: 0 compensated sum, 1 max |x|, 2 min, 3 max (NaN
// propagates); kind and src in device memory; part: scratch of
// 2 * k * kReducePartials doubles.
constexpr int kReducePartials = 1024;
hipError_t launch_reduce(const double* x, int64_t n, int k, const int* kind, const int* src,
                         double* out, double* part, hipStream_t stream);
hipError_t launch_coupled_stage(int64_t n, int stage, double dt, const double* oldH,
                                const double* oldC, const double* rH, const double* rC, double* H,
                                double* C, hipStream_t stream);

// ------------------------------------------------------------ batched form
// Totals and error words of one handle's step attempt: offsets of the RHS
// totals (5 doubles) and state checks (6 doubles) in its totals buffer.
constexpr int kCoupledRedRhs0 = 0, kCoupledRedAdm0 = 8, kCoupledRedAdm1 = 16, kCoupledRedRhs1 = 24,
              kCoupledRedAdm2 = 32, kCoupledRedRhs2 = 40, kCoupledRedSize = 48;
constexpr int kCoupledErrSlots = 3;

// One handle of a batched coupled run, in device memory: its RHS record and
// the buffers of its step loop.
struct CoupledBatchMesh {
    CoupledDev d;                // d.error: the handle's kCoupledErrSlots error words
    double *S, *O, *R, *F;       // state, last accepted state, RHS, first RHS of the step: [2][N]
    double* bnd;                 // [5][N] boundary fluxes, source, rate, divergence
    double* red;                 // [kCoupledRedSize] totals
    double* part;                // reduction partials
    const int* kinds;            // reduction kinds and array maps (RHS 0..4, state 5..10, 11..16)
    int64_t N, nf;               // cells and faces
    int nb;                      // reduction blocks of a length-N array
    int pad;
};

// Work of one launch over several handles: handle act[k] (k < nact) owns the
// workgroups [base[k], base[k+1]).
struct CoupledBatchSpan {
    const int* act;
    const int* base;
    int nact;
};

// Blocks of each kernel for one handle.
inline int coupled_cell_blocks(int64_t N) { return int((N + 127) / 128); }
inline int coupled_face_blocks(int64_t nf) { return int((nf + 127) / 128); }
inline int coupled_gather_blocks(int64_t N) { return int((N + 255) / 256); }
inline int coupled_stage_blocks(int64_t N) { return int((N + 255) / 256); }
inline int coupled_copy_blocks(int64_t N) { return int((2 * N + 255) / 256); }
int coupled_reduce_blocks(int64_t n);

// Error words first..first+count-1 of every handle of the span set to zero
// (a fill kernel, not a memset).
hipError_t launch_coupled_batch_clear(const CoupledBatchMesh* M, const int* act, int nact, int first, int count,
                                      hipStream_t stream);
// Copies of [2][N] arrays: segment s copies, for handle seg[2s], kind seg[2s+1]
// (0: F <- R, 1: O <- S, 2: S <- O); span.base over segments in copy blocks.
hipError_t launch_coupled_batch_copy(const CoupledBatchMesh* M, const int* seg, const int* base, int nseg,
                                     int blocks, hipStream_t stream);
// SSPRK2 stage `stage` of every handle of the span with its own dt[k]
// (stage 1 from F, stage 2 from R); span.base in stage blocks.
hipError_t launch_coupled_batch_stage(const CoupledBatchMesh* M, CoupledBatchSpan span, int blocks, int stage,
                                      const double* dt, hipStream_t stream);
// One RHS of every handle of the span from S into R and bnd, error bits into
// word `slot`, totals into red + off; spans in cell, face, gather and
// reduction blocks.
hipError_t launch_coupled_batch_rhs(const CoupledBatchMesh* M, CoupledBatchSpan cell, int cell_blocks,
                                    CoupledBatchSpan face, int face_blocks, CoupledBatchSpan gather,
                                    int gather_blocks, CoupledBatchSpan red, int red_blocks, int slot,
                                    int off, hipStream_t stream);
// The state checks of every handle of the span into red + off.
hipError_t launch_coupled_batch_admissible(const CoupledBatchMesh* M, CoupledBatchSpan red, int red_blocks,
                                           int off, hipStream_t stream);
// Every handle's totals and error words into out_red[k][kCoupledRedSize] and
// out_err[k][kCoupledErrSlots].
hipError_t launch_coupled_batch_readback(const CoupledBatchMesh* M, const int* act, int nact, double* out_red,
                                         uint32_t* out_err, hipStream_t stream);

}  // namespace mp
