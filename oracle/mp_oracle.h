/*
 * mp_oracle.h -- CPU restatement of MantlePar's H/C transport step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the correctness checker and the
 * timed CPU baseline for the HIP transport library.  Only tests/, the smoke
 * check in __graft_entry__.py and the cpu_baseline leg of bench.py may load
 * it.  The product path (mantlepar_amd) never links or calls it.
 *
 * Every function restates the reference algorithm from its behaviour (C11,
 * long double where the reference accumulates in long double).  Citations are
 * path:line in aband/MantlePar.
 *
 *   geometry      src/core/integral.cpp:7-121, src/core/quad_validation.h:15-46
 *   stencil       src/mlweno/tensorstencilpoly.cpp:61-459
 *   weno          src/mlweno/reconstruction.cpp:43-355, 474-502
 *   face flux     src/mlweno/advectiveflux.cpp:20-164, 192-268, 394-437
 *   operator      src/mlweno/test/transport_test_support.h:253-345 (face loop),
 *                 src/couple/preheat.cpp:125-203 (rates, RHS scatter),
 *                 src/couple/state_reconstruction.cpp:64-75 (production options)
 *   phase         src/phase/eutectic_rescaled.cpp:42-282
 *
 * Status codes: 0 = success; nonzero = the error category below.  Outputs are
 * written only on success, as in the reference ("checked output arguments are
 * assigned only on success", src/mlweno/reconstruction.h:61-62).
 */
#ifndef MP_ORACLE_H
#define MP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    MPO_OK = 0,
    MPO_ERR_ARG = 1,        /* invalid argument (PETSC_ERR_ARG_*, std::invalid_argument) */
    MPO_ERR_FP = 2,         /* non-finite / unrepresentable (PETSC_ERR_FP, overflow_error) */
    MPO_ERR_STATE = 3,      /* wrong object state */
    MPO_ERR_RUNTIME = 4,    /* runtime_error (phase fraction outside [0,1]) */
    MPO_ERR_LIB = 5,        /* singular moment solve (PETSC_ERR_LIB) */
    MPO_ERR_MEM = 6,
    MPO_ERR_BACKFLOW = 7    /* backflow on an outflow-only boundary */
};

const char* mpo_last_error(void);

/* ---------------------------------------------------------------- geometry */
typedef struct { double p[2]; } mpo_point;

int mpo_gauss_rule(int n, double* points, double* weights);
mpo_point mpo_map_edge_point(double ref, const mpo_point edge[2]);
mpo_point mpo_map_cell_point(mpo_point ref, const mpo_point corners[4]);
double mpo_cell_jacobian(mpo_point ref, const mpo_point corners[4]);
int mpo_edge_geometry(const mpo_point edge[2], double* length, mpo_point* normal);
int mpo_validate_quad(const mpo_point corners[4]);

/* ------------------------------------------------------- TensorStencilPoly */
#define MPO_MAX_STENCIL 12

typedef struct {
    int start_i, start_j, size_i, size_j, count, order;
    mpo_point center;
    double h;
    double coef[MPO_MAX_STENCIL * MPO_MAX_STENCIL];   /* basis-major: coef[cell*n+k] */
    double sigma[MPO_MAX_STENCIL * MPO_MAX_STENCIL];  /* reference-square indicator */
    int target_index;                                 /* -1: no TargetCell matrix */
    double target_sigma[MPO_MAX_STENCIL * MPO_MAX_STENCIL];
    mpo_point corners[MPO_MAX_STENCIL][4];
} mpo_stencil;

int mpo_stencil_init(mpo_stencil* s, const mpo_point (*corners)[4], int start_i, int start_j,
                     int size_i, int size_j, int order, double h);
int mpo_stencil_set_target(mpo_stencil* s, int local_i, int local_j);
int mpo_stencil_smoothness(const mpo_stencil* s, const double* averages, int use_target,
                           double* value);
int mpo_stencil_eval(const mpo_stencil* s, const double* averages, mpo_point x, int dx,
                     int dy, double* value);

/* ----------------------------------------------------------- Reconstruction */
#define MPO_MAX_CAND 8

typedef struct {
    int size_i, size_j, order;
    int noffsets;
    int off_i[MPO_MAX_CAND], off_j[MPO_MAX_CAND];
    double linear_weight[MPO_MAX_CAND];   /* used when has_weights */
    int has_weights;
    int target_smoothness;                /* 0 ReferenceSquare, 1 TargetCell */
} mpo_family;

typedef struct {
    mpo_family large, small;
    int use_constant;
    double constant_linear_weight;
    double epsilon;
    int s;
} mpo_weno_options;

/* Production options: src/couple/state_reconstruction.cpp:64-68. */
void mpo_production_options(mpo_weno_options* o);

typedef struct {
    int family;             /* 0 large, 1 small */
    int configured_index;
    int linear_index;       /* index into the family's offsets */
    double linear_weight;
    const mpo_stencil* stencil;   /* geometry tables (owned elsewhere) */
    int start_i, start_j;         /* candidate stencil origin (ReconstructionCandidateInfo) */
    mpo_point center;             /* centre of this candidate's stencil */
    double h;
    double centered[MPO_MAX_STENCIL];
    double smoothness, weight;
} mpo_candidate;

typedef struct {
    int nx, ny;             /* mesh cell dimensions */
    int ti, tj;             /* target */
    int req_bi, req_bj, req_ei, req_ej;   /* RequiredCells (half open) */
    mpo_weno_options opt;
    int ncand;
    mpo_candidate cand[MPO_MAX_CAND];
    int has_weights;
    double target_average, constant_weight, area;
} mpo_recon;

/* Mesh: (ny+1)*(nx+1) vertices, coords[(j*(nx+1)+i)*2 + d]. */
typedef struct {
    int nx, ny;
    const double* coords;
} mpo_mesh;

void mpo_cell_corners(const mpo_mesh* m, int i, int j, mpo_point c[4]);
double mpo_cell_area(const mpo_mesh* m, int i, int j);

/* Initialize: stencil tables are built into `storage` (caller-provided, at
 * least mpo_recon_storage_count(opt) entries) unless `shared` is non-null, in
 * which case it uses shared[slot].  slot = k for large offset k and
 * large.noffsets + k for small offset k. */
int mpo_recon_storage_count(const mpo_weno_options* opt);
int mpo_recon_init(mpo_recon* r, const mpo_mesh* mesh, int ti, int tj,
                   const mpo_weno_options* opt, double h, mpo_stencil* storage,
                   const mpo_stencil* const* shared);
int mpo_recon_update(mpo_recon* r, const double* values, int bi, int bj, int ei, int ej,
                     double area);
int mpo_recon_eval(const mpo_recon* r, mpo_point x, int dx, int dy, double* value);

/* ------------------------------------------------------------ advective flux */
/* Local LF (mode 0) or global LF (mode 1) for the linear law F_n(u)=(v.n)u. */
int mpo_llf_point(double left, double right, mpo_point velocity, mpo_point normal, int mode,
                  double global_speed, double* flux, double* local_speed);
int mpo_integrate_edge_flux(const mpo_point edge[2], int nq, const double* qp,
                            const double* qw, const mpo_point* velocity, const double* left,
                            const double* right, int mode, double global_speed,
                            double* flux);
/* Boundary types: 0 prescribed, 1 extrapolated, 2 zero flux. */
int mpo_boundary_data(const mpo_point edge[2], int nq, const double* qp, const double* qw,
                      const mpo_point* velocity, const double* inflow_values, int* types,
                      double* values);
int mpo_integrate_boundary_flux(const mpo_point edge[2], int nq, const double* qp,
                                const double* qw, const mpo_point* velocity,
                                const double* interior, const int* types,
                                const double* values, double* flux);

/* -------------------------------------------------------------------- phase */
typedef struct {
    double Tm0, Te0, nu, L, cp, Xe, rho, rhor, mus, mul, k0, g, alpha0;
} mpo_material;

typedef struct {
    double dT, h0, gamma, LD, TDm0, TDe0, invk0, l0, u0, p0, t0;
} mpo_derived;

typedef struct {
    mpo_material p;
    mpo_derived d;
} mpo_phase_model;

typedef struct { double Tep, Tmp, H_eutectic_end, H_liquidus; } mpo_phase_bounds;

typedef struct {
    double phi1, phi2, phil, cl, cs;
    int region;           /* 1 PureSolid .. 6 PureMelting (PhaseRegion values) */
    double TDp, Tep, dTD_dCD, dTD_dHD;
    int has_solid, has_liquid, finite_temperature_derivatives;
} mpo_phase_state;

void mpo_material_defaults(mpo_material* p);
int mpo_phase_set_parameters(mpo_phase_model* m, const mpo_material* p);
int mpo_phase_boundaries(const mpo_phase_model* m, double C, double P, mpo_phase_bounds* b);
int mpo_phase_evaluate(const mpo_phase_model* m, double H, double C, double P,
                       mpo_phase_state* s);
int mpo_phase_classify(const mpo_phase_model* m, double H, double C, double P, int* region);
/* staticPressure, src/phase/eutectic_rescaled.cpp:151-161: rho*g*zD*length
 * in long double, rounded once; the unscaled form uses the model's l0. */
int mpo_static_pressure(const mpo_phase_model* m, double zD, double* P);
int mpo_static_pressure_scaled(const mpo_phase_model* m, double zD, double length, double* P);
/* Hypereutectic extension (no reference behaviour): for C > Xe evaluates the
 * mirrored diagram C -> 1-C, Xe -> 1-Xe and maps phases back. */
int mpo_phase_evaluate_ext(const mpo_phase_model* m, double H, double C, double P,
                           mpo_phase_state* s);
/* Batched form used by the transport operator and by tests. */
int mpo_phase_batch(const mpo_phase_model* m, int64_t n, const double* H, const double* C,
                    double P, int hyper, double* T, double* phi, double* Cl, double* Cs,
                    uint8_t* region);

/* ---------------------------------------------------------- diffusive flux */
#define MPO_MAX_SAMPLES 8
enum { MPO_LAGRANGE_MIDDLE = 0, MPO_LAGRANGE_END_LO = 1, MPO_LAGRANGE_END_HI = 2 };
enum { MPO_SIDE_LEFT = 0, MPO_SIDE_RIGHT = 1, MPO_SIDE_BOUNDARY = 2 };
/* DiffusiveBoundaryType order of src/mlweno/diffusiveflux.h */
enum { MPO_DIFF_STATE = 0, MPO_DIFF_QUANTITY = 1, MPO_DIFF_FLUX = 2, MPO_DIFF_ZERO_FLUX = 3 };

typedef struct {
    int nq, count, boundary;
    double length;
    mpo_point normal;
    mpo_point points[16];                       /* edge quadrature points */
    double spacing[16], weights[16];            /* physical ds, W_q = w_q*length/2 */
    double coefficients[MPO_MAX_SAMPLES + 1];   /* dimensionless derivative weights */
    mpo_point samples[16 * (MPO_MAX_SAMPLES + 1)];   /* [q][k] */
    int sides[16 * (MPO_MAX_SAMPLES + 1)];
} mpo_sampling;

/* src/mlweno/diffusiveflux.cpp:203-254 */
int mpo_lagrange_weights(int npts, int location, double* w);
/* src/mlweno/diffusiveflux.cpp:256-358 (right == NULL: boundary sampling) */
int mpo_diffusive_sampling(const mpo_point edge[2], const mpo_point left[4], const mpo_point* right,
                           int nq, const double* qp, const double* qw, int n, double extent,
                           mpo_sampling* out);
/* src/mlweno/diffusiveflux.cpp:102-200, identity diffused quantity */
int mpo_diffusive_flux(const mpo_sampling* s, const double* samples, const double* kappa,
                       const int* bc_type, const double* bc_value, double* flux);

/* -------------------------------------------------------- transport operator */
enum { MPO_BC_ZERO = 0, MPO_BC_OUTFLOW = 1, MPO_BC_INFLOW_OUTFLOW = 2 };

typedef struct mpo_transport mpo_transport;

/* vnx: [q][ny][nx+1] normal velocity on vertical faces along +x;
 * vny: [q][ny+1][nx] normal velocity on horizontal faces along +y.
 * Samples follow the ascending Gauss points along +y resp. +x.
 * bc_policy[side*n + k]: per boundary face (side 0 bottom, 1 right, 2 top,
 * 3 left; k runs along +x or +y).  inflow_{H,C}: [side][k][q] or NULL.
 * table_mode 0: per-cell stencil tables; 1: one table per candidate type
 * (valid for uniform rectangular meshes). */
int mpo_transport_create(mpo_transport** out, const mpo_mesh* mesh, int q,
                         const mpo_weno_options* opt, const double* vnx, const double* vny,
                         const uint8_t* bc_bottom, const uint8_t* bc_right,
                         const uint8_t* bc_top, const uint8_t* bc_left,
                         const double* inflow_H, const double* inflow_C, int table_mode,
                         int threads);
void mpo_transport_destroy(mpo_transport* t);
/* Per-cell CFL rate (src/couple/preheat.cpp:125-152) and dt = cfl/max rate. */
int mpo_transport_rates(mpo_transport* t, double* rate, double* max_rate);
/* One spatial operator evaluation L(u) for both fields, and boundary flux totals. */
int mpo_transport_rhs(mpo_transport* t, const double* H, const double* C, double* rhsH,
                      double* rhsC, double* bfluxH, double* bfluxC);
/* Reconstruction diagnostics per cell: sigma[5], omega[5] (candidate slots:
 * large, small offsets 0..3; 0 for discarded) and traces at the 4 faces' Gauss
 * points [N][4][q] in face order bottom,right,top,left, Gauss order along +x/+y. */
int mpo_transport_reconstruct(mpo_transport* t, const double* u, double* sigma,
                              double* omega, double* traces);
/* Lax-Friedrichs speed of the interior faces (src/mlweno/advectiveflux.cpp:129-164):
 * mode 0 local |v.n| per point (the default), 1 global_speed at every point. */
int mpo_transport_set_lf(mpo_transport* t, int mode, double global_speed);
/* Bound limiter of the coupled reconstruction (src/couple/state_reconstruction.cpp:136-192):
 * cell_points Gauss points per axis, `samples` normal-line diffusion samples
 * per edge point with extent fraction `extent`, composition bounded by c_upper
 * (Xe).  enable = 0 switches it off (the default). */
int mpo_transport_set_limiter(mpo_transport* t, int enable, int cell_points, int samples,
                              double extent, double c_upper);
/* Thermal diffusion of H as the legacy preheat applies it (src/couple/preheat.cpp:176-190):
 * normal-line samples of the cell averages, diffusivity kappa, `samples` per
 * interior edge point with extent fraction `extent`.  types[side][face][q]
 * (NULL: zero flux everywhere) 0 zero flux, 1 prescribed temperature, 2
 * prescribed outward flux; values likewise, canonical point order.  The rates
 * gain the diffusion terms of src/couple/preheat.cpp:137-148. */
int mpo_transport_set_diffusion(mpo_transport* t, double kappa, int samples, double extent,
                                const uint8_t* types, const double* values);
/* Traces of one field (0 H, 1 C: the limiter bounds C by c_upper) [N][4][q]. */
int mpo_transport_traces(mpo_transport* t, int field, const double* u, double* traces);
/* Limiter factors theta of both fields for one state [N] (1 with the limiter off). */
int mpo_transport_limits(mpo_transport* t, const double* H, const double* C, double* thetaH,
                         double* thetaC);
/* Advance nsteps with the given scheme (1 forward Euler, 2 SSPRK2, 3 SSPRK3)
 * at fixed dt.  totals[4*s+{0,1,2,3}] = sum A*H, sum A*C, step boundary flux H,
 * C (stage-weighted), for each step s (may be NULL). */
int mpo_transport_advance(mpo_transport* t, double* H, double* C, int nsteps, double dt,
                          int scheme, double* totals);
/* Compensated (Neumaier) sum of area*u in fixed order. */
double mpo_transport_integral(const mpo_transport* t, const double* u);

/* --------------------------------------------------- coupled transport law */
/* Parameters of the coupled RHS (src/couple/evolve.cpp:279-384). */
typedef struct {
    double LD;               /* c.phase.derived().LD */
    double theta;            /* c.flow.material.theta (Darcy flux phi^(1+theta)) */
    double cooling;          /* alpha0*l0*g/cp */
    double kappa;            /* c.thermalDiffusivity */
    int lf_mode;             /* 0 local, 1 global Lax-Friedrichs */
    double global_speed;
    int cell_points;         /* c.input.quadrature.cellPointsPerAxis */
    int samples;             /* c.sampling.numberOfSamples */
    double extent;           /* c.sampling.extentFraction */
    double pressure;         /* constant phase pressure [Pa] */
    int lithostatic;         /* 1: P = staticPressure(surface_y - y) at each point */
    double surface_y;        /* nondimensional surface ordinate (>= mesh top) */
    int hypereutectic;       /* evaluate C > Xe on the mirrored diagram */
    double admissible_c_max; /* Admissible: C <= this (Xe) */
} mpo_coupled_params;

/* Flow samples (FlowPointSample, src/couple/visualization.h:8-12), fixed over
 * the call.  Edge arrays by reference edge id (horizontal first, j*nx+i; then
 * vertical nx*(ny+1) + j*(nx+1) + i) and canonical q: stokes/darcy [e][q][2],
 * phi [e][q].  Cell arrays [cell][b*n+a][2] / [cell][b*n+a] for the cell
 * Gauss points (a along x, b along y, n = cell_points). */
typedef struct {
    const double *edge_stokes, *edge_darcy, *edge_phi;
    const double *cell_stokes, *cell_darcy, *cell_phi;
} mpo_flow;

typedef struct { double boundaryH, boundaryC, sourceH, maximumRate, divergence; } mpo_coupled_rates;

/* time.* controls of the coupled loop (src/input/input.h) */
typedef struct { double cfl, initial, maximum, minimum, end; int fixed; } mpo_coupled_time;

typedef struct {
    double dt, courant, balanceH, balanceC, boundaryH, boundaryC, sourceH, maximumRate;
    int retries;
} mpo_coupled_step;

/* PhaseTransportVelocities (src/couple/evolve.cpp:36-52) on two phase states
 * given by their cl, cs. */
void mpo_phase_transport_velocities(const double solid[2], const double relative[2], double phi,
                                    double theta, double cl_a, double cs_a, double cl_b,
                                    double cs_b, double mixture[2], double effective[2]);
/* Advection policies of C when they differ from those of H. */
int mpo_transport_set_composition_boundary(mpo_transport* t, const uint8_t* bottom,
                                           const uint8_t* right, const uint8_t* top,
                                           const uint8_t* left);
/* One coupled RHS (src/couple/evolve.cpp:279-384) on the limited
 * reconstruction (the transport's limiter must be on; C bound = c_upper):
 * rhsH, rhsC, per-cell rate and mixture divergence, cell-averaged T and phi
 * (any may be NULL) and the totals.  Thermal boundary data: set_diffusion's
 * types/values; inflow: the transport's inflow arrays. */
int mpo_coupled_rhs(mpo_transport* t, const mpo_phase_model* pm, const mpo_coupled_params* cp,
                    const mpo_flow* flow, const double* H, const double* C, double* rhsH,
                    double* rhsC, double* rate, double* div, double* T, double* phi,
                    mpo_coupled_rates* totals);
/* SSPRK2 steps of src/couple/evolve.cpp:428-511 with the flow held fixed. */
int mpo_coupled_run(mpo_transport* t, const mpo_phase_model* pm, const mpo_coupled_params* cp,
                    const mpo_flow* flow, const mpo_coupled_time* tc, double* H, double* C,
                    double* time, int nsteps, mpo_coupled_step* steps);

/* ------------------------------------------------------------ workloads */
/* SplitMix64 finaliser of src/core/mesh.cpp:12-18. */
uint64_t mpo_mix64(uint64_t v);

#ifdef __cplusplus
}
#endif
#endif
