// Shared host/device layouts of the MI355X transport library.
//
// Everything a stage kernel reads about one mesh lives in one MeshDev record
// in device memory; kernels address it through a wave-uniform pointer so the
// compiler keeps the tables in scalar registers (s_load), not VGPRs.
#pragma once

#include <cstdint>

namespace mp {

constexpr int kMaxQ = 8;          // edge Gauss points supported on device
constexpr int kWave = 64;         // CDNA wavefront
constexpr int kStripCells = 62;   // owned columns per wave (2 halo lanes)
constexpr int kPartials = 8;      // doubles per work item in the partials buffer
constexpr int kMaxLimS = 8;       // limiter: normal-line samples per edge point

// Error word bits (device side); the host maps them to mp_status codes.
enum : uint32_t {
    kErrNonFinite = 1u << 0,      // state, trace or flux not finite
    kErrComposition = 1u << 1,    // C outside [0, Xe] (or [0,1] with the mirror)
    kErrFraction = 1u << 2,       // unitFraction outside [0,1] beyond 64 eps
    kErrBackflow = 1u << 3,       // inflow on an outflow-only boundary face
    kErrInflow = 1u << 4,         // inflow without prescribed boundary data
    kErrWeights = 1u << 5,        // no finite active log weight
    kErrPhaseRange = 1u << 6,     // phase boundaries not resolvable
    kErrInadmissible = 1u << 7,   // limiter: average outside H >= 0, 0 <= C <= Xe
    kErrCoupledRange = 1u << 8,   // coupled law: T < 0 or phi >= 1 at a point
    kErrLFSpeed = 1u << 9,        // coupled law: global LF speed below a flow speed
    kErrStall = 1u << 10,         // one-launch step: an item's inputs were not done in time
    kErrBalance = 1u << 11,       // a step's conservation balance failed (PETSC_ERR_PLIB)
};

// Binary eutectic parameters as the device uses them.  Tep/Tmp are derived
// on the host in long double exactly as EutecticModel::boundaries does
// (src/phase/eutectic_rescaled.cpp:163-170): P is uniform per mesh, or, for a
// lithostatic pressure, uniform per table entry (TepPair).  Reciprocals are
// double-doubles (hi + lo) formed on the host, so the device multiplies where
// the reference divides, at double-double accuracy.
struct PhaseDev {
    double LD;        // L/h0
    double Xe;        // eutectic composition
    double Tep;       // TDe0 + P/nu/dT, rounded once
    double Tmp;       // Tep + 1, rounded once
    int32_t hyper;    // evaluate C > Xe on the mirrored diagram
    int32_t pad;
    double ixh, ixl;  // 1/Xe
    double imh, iml;  // 1/(1 - Xe), of the double 1 - Xe the mirror uses
    double ilh, ill;  // 1/LD
};

// Pressure-corrected eutectic and melting temperatures of one lithostatic
// pressure P = staticPressure(surface_y - y) (src/phase/eutectic_rescaled.cpp:
// 143-161, src/couple/initialization.cpp:155-162), rounded on the host as the
// reference rounds them; they replace PhaseDev's Tep and Tmp where the phase
// is evaluated at that ordinate.
struct __attribute__((aligned(16))) TepPair {
    double tep, tmp;
};

// Per-face flux moments of the mixed path, built once per velocity refresh.
// With W_q the Gauss weights times half the face length, alpha_q the LF speed
// and s_q the tangential point offsets:
//   vp = sum_q W_q (v_q + alpha_q)/2,  vm = sum_q W_q (v_q - alpha_q)/2      (FP64)
//   p_m = sum_q W_q (v_q + alpha_q)/2 s_q^m,  m_m likewise with -alpha  (FP32)
// so an interior face flux is tL*vp + tR*vm + sum_m (cL_m p_m + cR_m m_m),
// c being the face-line coefficients of the centred reconstructions.  One
// 32-byte record per face: two 16-byte loads per face in the sweep.
struct __attribute__((aligned(16))) FaceMom {
    double vp, vm;
    float p1, p2, m1, m2;
};

// The same moments of a face whose LF weights are one-sided, in 16 bytes.
// With the local LF speed alpha_q = |v_q|, (v_q + alpha_q)/2 = max(v_q, 0) and
// (v_q - alpha_q)/2 = min(v_q, 0), so where no point of the face changes sign
// one side's moments are exactly +0: v >= 0 holds (vp, p1, p2) = (v, a, b),
// v < 0 holds (vm, m1, m2) = (v, a, b), and the other three are +0.  A face
// whose points change sign, and every face under the global LF speed, has
// b = NaN: its moments are the FaceMom record at the same face index.
struct __attribute__((aligned(16))) FaceRec {
    double v;
    float a, b;
};

// The FP64 sweep's one-sided record of a face: under the local LF speed a
// face whose points all have v >= 0 (or all v < 0) has flux
// (t + c0)*v + c1*a + c2*b from its upwind side's centre t and face line c,
// with v = sum_q W_q v_q, a = sum_q W_q v_q s_q, b = sum_q W_q v_q s_q^2 in
// FP64.  A face whose points change sign, and every face under the global LF
// speed, has b = NaN and is integrated point by point from its samples.
struct __attribute__((aligned(16))) FaceRec64 {
    double v, a, b, pad;
};

// One uniform rectangular mesh.  All reconstruction tables are the closed
// forms of the reference's moment / smoothness construction on rectangles,
// in coordinates normalised by h = sqrt(dx*dy) about the target centre.
struct MeshDev {
    int32_t nx, ny, q, lf_mode;
    double dx, dy, area, inv_area;
    double aw, bw;              // dx/h, dy/h
    double hx, hy;              // aw/2, bw/2: face offsets in normalised units
    // 1-D quadratic through three cell averages at offsets -1,0,1:
    //   A2 = (r + l - 2m)*k2,  A1 = (r - l)*k1,  A0 = m - A2*k0
    double kx0, kx1, kx2, ky0, ky1, ky2;
    // 2x2 bilinear candidates: slopes 1/aw, 1/bw, 1/(aw*bw) and the
    // target-cell mean of the xy-term (aw^2+bw^2)/12.
    double inv_aw, inv_bw, inv_awbw, q11w;
    // the mixed sweep's indicator works on q11 * sqrt(q11w): its scale
    // inv_awbw * sqrt(q11w) and the inverse 1 / sqrt(q11w)
    double q11s, q11u;
    // ML-WENO weights (src/mlweno/reconstruction.cpp:212-265)
    double epsA;                // epsilon * area
    double pL, pS, pC;          // exponents s*r + eta(r): large, small, constant
    double ldL, ldS[4], ldC;    // log of linear weights (-inf when zero)
    int32_t use_constant, pad0;
    // The same weights in direct form for the FP64 reconstruction: the
    // exponents are integers, so alpha_k = d_k * (1/(sigma_k + epsA))^p_k by a
    // reciprocal and multiplications.  A target takes it when every active
    // denominator lies in [wlo, whi], which keeps each alpha and their sum
    // normal and finite; wdirect = 0 when the exponents or linear weights are
    // out of that range for every target.
    double dL, dS[4];           // linear weights (0 when off)
    double aC;                  // alpha of the constant candidate (0 when off)
    double wlo, whi;
    int32_t ipL, ipS, wdirect, pad1;
    double global_speed;
    // edge quadrature
    double sv[kMaxQ];           // vertical faces: point offsets hy*xi_q
    double sh[kMaxQ];           // horizontal faces: point offsets hx*xi_q
    double wv[kMaxQ];           // w_q * dy/2
    double wh[kMaxQ];           // w_q * dx/2
    PhaseDev ph;
    // Lithostatic pressure: {Tep, Tmp} of each cell row at the cell centre
    // (null: the constant pressure of ph).
    const TepPair* tep_row;
    // Bound limiter of the reconstruction (limit = 0: off).  Point offsets in
    // the same normalised units as hx, hy: cell Gauss points lim_cx x lim_cy;
    // in-cell normal-line samples of the right face at x = lim_x*[k] and of the
    // top face at y = lim_y*[k] (left and bottom faces mirror them), for an
    // interior edge (lim_ni = n/2 samples) and a boundary edge (lim_nb = n).
    int32_t limit, lim_ncp, lim_ni, lim_nb;
    double lim_cupper;          // upper bound of C: Xe (1 with the mirror)
    double lim_cx[kMaxQ], lim_cy[kMaxQ];
    double lim_cw[kMaxQ];       // cell Gauss weights on [-1, 1]
    double lim_xi[kMaxLimS], lim_xb[kMaxLimS];
    double lim_yi[kMaxLimS], lim_yb[kMaxLimS];
    // The same point offsets rounded to FP32 for the mixed sweep's limiter,
    // with the edge point offsets sv, sh.
    float lim32_cx[kMaxQ], lim32_cy[kMaxQ], lim32_sv[kMaxQ], lim32_sh[kMaxQ];
    float lim32_xi[kMaxLimS], lim32_xb[kMaxLimS], lim32_yi[kMaxLimS], lim32_yb[kMaxLimS];
    // Thermal diffusion of H in the legacy preheat form (samples are the cell
    // averages of each side): interior face flux dif_x*(H_L - H_R) along +x on
    // vertical faces, dif_y likewise on horizontal ones (0: off); a boundary
    // point with prescribed temperature b adds W_q*dif_b*(H_in - b) outward,
    // one with prescribed outward flux g adds W_q*g.
    double dif_x, dif_y, dif_bx, dif_by;
    const uint8_t* tbc[4];      // thermal policy [face][q]: 0 zero flux, 1 temperature, 2 flux
    const double* tval[4];      // thermal data [face][q]
    // boundary data (per face policy and optional inflow [face][q])
    const uint8_t* bc[4];       // bottom(nx) right(ny) top(nx) left(ny)
    const double* inflowH[4];
    const double* inflowC[4];
    // face velocities [q][ny][nx+1] and [q][ny+1][nx]
    const double* vnx;
    const double* vny;
    // Per-face flux moments (mixed path): [ny][nx+1] and [ny+1][nx], full
    // records and their 16-byte one-sided form.
    const FaceMom* fx;
    const FaceMom* fy;
    const FaceRec* rx;
    const FaceRec* ry;
    // FP64 one-sided records ([ny][nx+1] and [ny+1][nx]); null: every face of
    // the FP64 sweep is integrated point by point.
    const FaceRec64* rx64;
    const FaceRec64* ry64;
};

// Per launch, per mesh: the stage u_out = a*u_n + b*(u_k + dt*L(u_k)).
struct StagePtrs {
    const double* Hn;
    const double* Cn;
    const double* Hk;
    const double* Ck;
    double* Ho;
    double* Co;
    double a, b, dt;
    double* T;          // phase outputs of a step's last stage (written by the phase kernel)
    double* phi;
    double* Cl;
    double* Cs;
    uint8_t* region;
    double* partials;   // [item][kPartials]: sumH, errH, sumC, errC, boundaryH, boundaryC
    uint32_t* error;
    int32_t rows;       // rows per work item below row `split`
    int32_t strips;     // work items per row band
    int32_t split;      // first row of the tail bands (ny: none)
    int32_t tail_rows;  // rows per work item from `split` up
    int32_t nlong;      // work items below `split`
    int32_t pad;
};

}  // namespace mp
