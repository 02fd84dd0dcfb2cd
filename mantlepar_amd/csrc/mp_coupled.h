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

}  // namespace mp
