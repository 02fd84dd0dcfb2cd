/*
 * mantlepar_amd.h -- C ABI of the MI355X (gfx950) H/C transport library.
 *
 * One handle advances one rank's share of a structured mesh: ML-WENO
 * reconstruction, smoothness-indicator weights, Lax-Friedrichs face fluxes and
 * the SSP Runge-Kutta stage update for enthalpy H and composition C, fused with
 * the per-cell binary-eutectic equilibrium (H,C) -> (T, phi, Cl, Cs, region).
 *
 * The entry points replace, for all owned cells at once, the reference's
 * per-object calls (citations are path:line in aband/MantlePar):
 *   StateReconstruction::Update + Enthalpy/Evaluate   src/couple/state_reconstruction.cpp:101-134
 *   Reconstruction::Initialize/Update/Evaluate        src/mlweno/reconstruction.h:78-137
 *   IntegrateAdvectiveFlux / ...BoundaryFlux          src/mlweno/advectiveflux.h:126-186
 *   the preheat RHS face loop and FE/SSPRK stages     src/couple/preheat.cpp:160-203, 253-284
 *   the SSPRK2 stage combination                      src/couple/evolve.cpp:465-492
 *   EutecticModel::evaluate/classify/boundaries       src/phase/eutectic_rescaled.h:103-136
 *
 * Conventions (mirroring the reference):
 *   - Every call returns mp_status: 0 on success, a category code otherwise,
 *     like PetscErrorCode.  mp_last_error() returns the thread's last message.
 *   - On error the handle's state and previously returned outputs are left
 *     unchanged ("failed calls preserve the previous valid state",
 *     src/mlweno/reconstruction.h:61-62).  State is double-buffered on device.
 *   - Host arrays are owned by the caller and copied; device buffers are owned
 *     by the handle.  A handle is single-threaded (like EutecticModel's
 *     setParameters contract, src/phase/eutectic_rescaled.h:100-102); use one
 *     handle per thread or external locking.
 *   - Cell arrays are row-major H[j*nx + i] (the reference's ordering,
 *     src/couple/initialization.cpp:786).
 *   - Face velocities are the normal velocity at the q edge Gauss points, in
 *     MapEdgePoint order along the canonical direction (src/core/integral.cpp:91-97,
 *     src/couple/flow_sampling.cpp:101-115):
 *       vnx[q][ny][nx+1]: vertical faces x_i, along +x, points ascending in y;
 *       vny[q][ny+1][nx]: horizontal faces y_j, along +y, points ascending in x.
 */
#ifndef MANTLEPAR_AMD_H
#define MANTLEPAR_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t mp_status;

enum {
    MP_OK = 0,
    MP_ERR_ARG = 1,        /* invalid argument (PETSC_ERR_ARG_*, std::invalid_argument) */
    MP_ERR_FP = 2,         /* non-finite or unrepresentable value (PETSC_ERR_FP) */
    MP_ERR_STATE = 3,      /* call out of order (PETSC_ERR_ARG_WRONGSTATE) */
    MP_ERR_PHASE = 4,      /* phase fraction outside [0,1] (std::runtime_error) */
    MP_ERR_DEVICE = 5,     /* HIP runtime failure */
    MP_ERR_MEMORY = 6,     /* allocation or HBM budget exhausted */
    MP_ERR_BACKFLOW = 7,   /* backflow on an outflow-only boundary (initialization.cpp:478-481) */
    MP_ERR_COMPOSITION = 8, /* C outside the admissible interval (eutectic_rescaled.cpp:88-93) */
    MP_ERR_PENDING = 9,     /* request not finished within the wait timeout */
    MP_ERR_BALANCE = 10     /* a step's conservation balance failed (PETSC_ERR_PLIB) */
};

enum { MP_PREC_FP64 = 0, MP_PREC_MIXED = 1 };

/* Per boundary face policy (src/couple/initialization.cpp:469-489). */
enum { MP_BC_ZERO_FLUX = 0, MP_BC_OUTFLOW = 1, MP_BC_INFLOW_OUTFLOW = 2 };

/* Time integrators: forward Euler (preheat.cpp:267), SSPRK2 (evolve.cpp:469-479),
 * SSPRK3 (Shu-Osher; not in the reference). */
enum { MP_SCHEME_FE = 1, MP_SCHEME_SSPRK2 = 2, MP_SCHEME_SSPRK3 = 3 };

/* Uniform rectangular mesh: nx*ny cells on [x0,x1]x[y0,y1], q edge Gauss points. */
typedef struct {
    int32_t nx, ny;
    double x0, x1, y0, y1;
    int32_t q;             /* 1..8; the reference default is 5 (src/input/input.h:64) */
} mp_mesh_desc;

/* ReconstructionOptions for the production families (3x3 order-2 large
 * candidate, four 2x2 order-1 TargetCell candidates;
 * src/couple/state_reconstruction.cpp:64-68, src/mlweno/reconstruction.h:28-34). */
typedef struct {
    double epsilon;                 /* 1e-2 */
    int32_t s;                      /* 1 */
    double large_weight;            /* linear weight d of the large candidate, 1 */
    double small_weight[4];         /* offsets (0,-1),(0,0),(-1,-1),(-1,0), 1 each */
    int32_t use_constant;           /* 0 */
    double constant_linear_weight;  /* 1e-5 */
    int32_t lf_mode;                /* 0 local LF, 1 global LF (advectiveflux.h:18-25) */
    double global_speed;            /* used when lf_mode == 1 */
} mp_weno_options;

/* MaterialParameters (src/phase/eutectic_rescaled.h:10-26) plus a uniform
 * pressure P [Pa] and the hypereutectic switch: 0 rejects C > Xe like the
 * reference, 1 evaluates the mirrored diagram for Xe < C <= 1. */
typedef struct {
    double Tm0, Te0, nu, L, cp, Xe, rho, rhor, mus, mul, k0, g, alpha0;
    double pressure;
    int32_t hypereutectic;
} mp_phase_params;

typedef struct {
    double dt;              /* step used */
    double max_rate;        /* max_K rate_K (src/couple/preheat.cpp:125-152) */
    double courant;         /* dt * max_rate */
    double total_H, total_C;          /* compensated sum |K|*u after the call */
    double boundary_H, boundary_C;    /* sum over steps of dt * stage-weighted outward boundary flux */
    uint32_t error_bits;
    int32_t steps_done;
} mp_step_report;

typedef struct mp_transport mp_transport;
typedef struct mp_batch mp_batch;
typedef struct mp_server mp_server;
typedef struct mp_coupled mp_coupled;

/* Defaults: production options and the reference's default material. */
void mp_default_weno_options(mp_weno_options* o);
void mp_default_phase_params(mp_phase_params* p);
const char* mp_last_error(void);
const char* mp_version(void);
/* Sets the thread's last error (as mp_last_error returns it) and returns
 * code: for wrappers layered on this ABI that reject a call themselves. */
mp_status mp_report_error(mp_status code, const char* message);

/* --------------------------------------------------------- single mesh --- */
/* stream: a hipStream_t, or NULL for a stream owned by the handle. */
mp_status mp_transport_create(const mp_mesh_desc* mesh, const mp_weno_options* weno,
                              const mp_phase_params* phase, int32_t precision, void* stream,
                              mp_transport** out);
mp_status mp_transport_destroy(mp_transport* t);
mp_status mp_transport_upload_state(mp_transport* t, const double* H, const double* C);
mp_status mp_transport_download_state(mp_transport* t, double* H, double* C);
mp_status mp_transport_set_face_velocity(mp_transport* t, const double* vnx, const double* vny);
/* Per boundary face policies; each array runs along +x (bottom, top) or +y
 * (left, right).  NULL means zero flux on that side. */
mp_status mp_transport_set_boundary(mp_transport* t, const uint8_t* bottom, const uint8_t* right,
                                    const uint8_t* top, const uint8_t* left);
/* Inflow states [side 0..3 = bottom,right,top,left][face][q], canonical point order. */
mp_status mp_transport_set_inflow(mp_transport* t, const double* inflow_H, const double* inflow_C);
/* dt = cfl / max_K rate_K for the current face velocities. */
mp_status mp_transport_stable_dt(mp_transport* t, double cfl, double* dt);
/* Advance nsteps at fixed dt; each step ends with the per-cell phase.  The
 * call enqueues every step, then waits once for the device (it reads the
 * error word to decide whether the new state is committed); report may be
 * NULL. */
mp_status mp_transport_step(mp_transport* t, int32_t nsteps, int32_t scheme, double dt,
                            mp_step_report* report);
/* mp_transport_step without device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The step's work memory (stage partials, the step reduction's block results,
 * per-step totals and the error word) lives in a static 32 MiB device array of
 * the library instead of buffers the handle allocates; the states, face data
 * and phase outputs are the handle's, allocated at create time.  Results are
 * bitwise those of mp_transport_step.  Beside a fixed 3 MiB of completion
 * flags, a call may use 29 MiB: 384 bytes per work item plus 32 per step (an
 * 8192 x 8192 mesh fits); a larger one fails with MP_ERR_MEMORY.  The
 * handle's device must be current.  Each GPU holds one such array, shared by
 * every call on it: calls that overlap on one GPU (several handles on
 * different threads or streams) race; serialise them. */
mp_status mp_transport_step_noalloc(mp_transport* t, int32_t nsteps, int32_t scheme, double dt,
                                    mp_step_report* report);
/* Phase of the latest step (or of the uploaded state, computed on demand). */
mp_status mp_transport_phase(mp_transport* t, double* T, double* phi, double* Cl, double* Cs,
                             uint8_t* region);
/* Diagnostics used by parity tests: L(u) of the current state, and per-cell
 * reconstruction data for field 0 (H) or 1 (C): sigma[N][5], omega[N][5]
 * (slots: large, small offsets 0..3; 0 for discarded candidates) and traces
 * [N][4][q] at the faces bottom,right,top,left in canonical point order. */
mp_status mp_transport_rhs(mp_transport* t, double* rhsH, double* rhsC, double* boundary);
mp_status mp_transport_reconstruct(mp_transport* t, int32_t field, double* sigma,
                                   double* omega, double* traces);
/* The reconstructions of the current state at n points, replacing
 * StateReconstruction::Evaluate (src/couple/state_reconstruction.cpp:121-134)
 * over Reconstruction::Evaluate (src/mlweno/reconstruction.cpp:474-502):
 * point k = (x[k], y[k]) in physical coordinates is evaluated on the
 * reconstruction of cell cells[k] (j*nx + i), limited by the cell's factor
 * theta when the bound limiter is on: H[k] = mean + theta*(R(p) - mean), C[k]
 * likewise.  Cells outside the mesh and non-finite points fail with
 * MP_ERR_ARG.  H or C may be NULL. */
mp_status mp_transport_evaluate(mp_transport* t, int64_t n, const int32_t* cells, const double* x,
                                const double* y, double* H, double* C);
/* Device pointers of the current state (valid until the next step/upload). */
mp_status mp_transport_device_state(mp_transport* t, double** H, double** C);
mp_status mp_transport_synchronize(mp_transport* t);
/* Time nsteps of the stage kernel with HIP events on the handle's stream:
 * ms_total = wall time of the steps; ms_stage = mean duration of one stage
 * launch; ms_phase_stage = mean duration of the final stage together with the
 * phase kernel that follows it (each launch timed alone, on one stream).  With
 * one launch per step (MP_LAUNCH_PER_STEP on a mixed handle), ms_stage is the
 * mean duration of a step's launch (all its stages) and ms_phase_stage that of
 * its phase kernel. */
mp_status mp_transport_profile(mp_transport* t, int32_t nsteps, int32_t scheme, double dt,
                               double* ms_total, double* ms_stage, double* ms_phase_stage);
/* Every step of mp_transport_step is checked against the preheat's
 * conservation balance (src/couple/preheat.cpp:268-273), for H and for C:
 * |total(after) - total(before) + dt*outward boundary flux| must not exceed
 * tol*(1 + |after| + |before| + |dt*flux|), the totals being the compensated
 * sums of |K|*u.  A step that fails is rejected with MP_ERR_BALANCE
 * ("Preheat heat balance failed; candidate rejected", PETSC_ERR_PLIB there)
 * and the state is left unchanged.  tol is 1e-10, the reference's value,
 * unless set here (finite, >= 0). */
mp_status mp_transport_set_balance_tolerance(mp_transport* t, double tol);
/* Sweep tile shape: rows per work item (>= 4). */
mp_status mp_transport_set_tile(mp_transport* t, int32_t rows_per_item);
/* How a mixed-precision handle launches its SSPRK2/SSPRK3 steps:
 * MP_LAUNCH_PER_STEP runs all stages of a step in one launch, each work item
 * of a stage starting when the items whose output it reads are done;
 * MP_LAUNCH_PER_STAGE runs one launch per stage; MP_LAUNCH_AUTO (the default)
 * takes the per-step launch, the faster one at every mesh size measured
 * (2048x4096 to 8192^2).  All give bitwise the same results.  FP64 handles, forward Euler and the
 * limiter always launch per stage.  MP_LAUNCH_STALL_CHECK is a diagnostic of
 * the per-step launch's bounded waits: items wait for a completion that is
 * never published and give up after 64 polls, so a step fails with
 * MP_ERR_DEVICE and leaves the state as it was. */
enum { MP_LAUNCH_PER_STAGE = 0, MP_LAUNCH_PER_STEP = 1, MP_LAUNCH_STALL_CHECK = 2, MP_LAUNCH_AUTO = 3 };
mp_status mp_transport_set_launch(mp_transport* t, int32_t mode);
/* The sweep's work layout (any output may be NULL): rows per work item, work
 * items per stage, the first row of the 4-row tail items (ny: none; it
 * depends on the GPU's CU count), and whether SSPRK3 steps run as one launch
 * each (1) or one launch per stage (0). */
mp_status mp_transport_layout(mp_transport* t, int32_t* rows_per_item, int32_t* items,
                              int32_t* tail_split, int32_t* one_launch_steps);
/* Bound limiter of every reconstruction the stage kernels build: replaces
 * StateReconstruction::Limit (src/couple/state_reconstruction.cpp:136-192;
 * the reference calls it per coupled RHS, src/couple/evolve.cpp:281), with
 * config.input.quadrature.cellPointsPerAxis, config.sampling.numberOfSamples
 * and extentFraction (src/mlweno/diffusiveflux.h:30-38; defaults 5, 4, 0.9)
 * as arguments.  Each cell's factor theta makes mean + theta*(R(p) - mean)
 * stay within its 3x3 neighbourhood's bounds, clamped to H >= 0 and
 * 0 <= C <= Xe (C <= 1 with the hypereutectic mirror).  Both precision
 * modes: the mixed sweep evaluates the blend's extremes at the same points in
 * FP32 and takes the bounds relative to the mean from its FP64 differences.
 * An average outside those ranges fails the step with MP_ERR_ARG
 * ("Inadmissible cell average in coupled reconstruction").  enable = 0
 * switches it off (the default). */
mp_status mp_transport_set_limiter(mp_transport* t, int32_t enable, int32_t cell_points_per_axis,
                                   int32_t diffusion_samples, double extent_fraction);
/* Thermal diffusion of H in the legacy preheat form, added to every stage's
 * face fluxes: replaces the diffusion half of the preheat RHS
 * (src/couple/preheat.cpp:176-190: IntegrateDiffusiveFlux /
 * IntegrateDiffusiveBoundaryFlux of src/mlweno/diffusiveflux.cpp:102-200 on
 * normal-line samples that take each side's cell average of H) with
 * kappa = config thermalDiffusivity (src/couple/initialization.cpp:284) and
 * config.sampling (numberOfSamples, extentFraction; defaults 4, 0.9).  The
 * stable-step rate gains the diffusion terms of src/couple/preheat.cpp:137-148.
 * C is not diffused.  kappa = 0 switches it off (the default). */
mp_status mp_transport_set_diffusion(mp_transport* t, double kappa, int32_t diffusion_samples,
                                     double extent_fraction);
/* Thermal boundary data per boundary Gauss point, replacing ThermalData /
 * EvaluateThermalBoundary (src/couple/preheat.cpp:52-62,
 * src/couple/initialization.cpp:490-506): types[side][face][q] 0 zero flux,
 * 1 prescribed temperature (the boundary node's value), 2 prescribed outward
 * flux per length; values likewise; sides bottom, right, top, left, faces
 * along +x/+y, points in canonical order.  types = NULL: zero flux everywhere
 * (the default). */
mp_status mp_transport_set_thermal_boundary(mp_transport* t, const uint8_t* types,
                                            const double* values);
/* Lithostatic phase pressure, replacing the pressure model "lithostatic"
 * (src/couple/initialization.cpp:155-162 with EutecticModel::staticPressure,
 * src/phase/eutectic_rescaled.cpp:151-161): the phase of each cell is
 * evaluated at P = rho*g*(surface_y - y_c)*l0, y_c the cell centre
 * (nondimensional), instead of the constant phase.pressure.  surface_y must
 * be at or above the mesh top y1 (MP_ERR_ARG otherwise).  enable = 0 returns
 * to the constant pressure. */
mp_status mp_transport_set_lithostatic(mp_transport* t, int32_t enable, double surface_y);
/* Limiter factors theta[N] of H and C for the current state (diagnostics). */
mp_status mp_transport_limits(mp_transport* t, double* thetaH, double* thetaC);

/* --------------------------------------------------------------- batch --- */
/* Many independent meshes advanced together.  The reference runs one mesh per
 * rank, each with its own StateReconstruction and stepping loop
 * (src/couple/state_reconstruction.cpp:35-120, src/couple/preheat.cpp:234-287);
 * a batch holds many such shares on one device and runs every stage as ONE
 * launch over all of them.  Each mesh keeps its own dt, boundary
 * data, error word and report, and its results are bitwise those of an
 * mp_transport handle on the same inputs.  No reference counterpart for the
 * batching itself.  Mesh ids are slot numbers in [0, capacity). */
enum { MP_BATCH_EAGER = 0, MP_BATCH_GRAPH = 1 };

/* hbm_budget_bytes bounds the device memory of all resident meshes (0: no bound). */
mp_status mp_batch_create(int32_t capacity, const mp_weno_options* weno,
                          const mp_phase_params* phase, int32_t precision,
                          uint64_t hbm_budget_bytes, void* stream, mp_batch** out);
mp_status mp_batch_destroy(mp_batch* b);
mp_status mp_batch_add_mesh(mp_batch* b, const mp_mesh_desc* mesh, int32_t* id);
mp_status mp_batch_remove_mesh(mp_batch* b, int32_t id);
mp_status mp_batch_upload_state(mp_batch* b, int32_t id, const double* H, const double* C);
mp_status mp_batch_download_state(mp_batch* b, int32_t id, double* H, double* C);
mp_status mp_batch_set_face_velocity(mp_batch* b, int32_t id, const double* vnx,
                                     const double* vny);
mp_status mp_batch_set_boundary(mp_batch* b, int32_t id, const uint8_t* bottom,
                                const uint8_t* right, const uint8_t* top, const uint8_t* left);
mp_status mp_batch_set_inflow(mp_batch* b, int32_t id, const double* inflow_H,
                              const double* inflow_C);
/* dt = cfl / max rate of mesh id (returned, not applied). */
mp_status mp_batch_stable_dt(mp_batch* b, int32_t id, double cfl, double* dt);
/* Fixed step of mesh id for the following mp_batch_step calls. */
mp_status mp_batch_set_dt(mp_batch* b, int32_t id, double dt);
/* Advance every resident mesh nsteps with its own dt.  flags: MP_BATCH_GRAPH
 * replays a hipGraph of the whole call (captured on first use, re-captured
 * only when membership, schedule or nsteps change).  reports[id] for each
 * resident id (array of capacity entries, may be NULL).  A mesh whose step
 * fails keeps its state and reports its error bits; the others advance.  The
 * status is that of the first failed mesh. */
mp_status mp_batch_step(mp_batch* b, int32_t nsteps, int32_t scheme, int32_t flags,
                        mp_step_report* reports);
/* Lithostatic phase pressure of mesh id, as mp_transport_set_lithostatic
 * (src/couple/initialization.cpp:155-162). */
mp_status mp_batch_set_lithostatic(mp_batch* b, int32_t id, int32_t enable, double surface_y);
/* Bound limiter of mesh id, as mp_transport_set_limiter
 * (src/couple/state_reconstruction.cpp:136-192), in both precisions.  Meshes
 * with and without it share the launches; while any mesh of the batch has it
 * on, mixed steps launch per stage. */
mp_status mp_batch_set_limiter(mp_batch* b, int32_t id, int32_t enable, int32_t cell_points_per_axis,
                               int32_t diffusion_samples, double extent_fraction);
/* Thermal diffusion of H and thermal boundary data of mesh id, as
 * mp_transport_set_diffusion / mp_transport_set_thermal_boundary (same array
 * layouts, checks and messages); both refresh the mesh's stable step.  Meshes
 * with and without them share the launches; while any mesh has them, mixed
 * steps launch per stage. */
mp_status mp_batch_set_diffusion(mp_batch* b, int32_t id, double kappa, int32_t diffusion_samples,
                                 double extent_fraction);
mp_status mp_batch_set_thermal_boundary(mp_batch* b, int32_t id, const uint8_t* types,
                                        const double* values);
mp_status mp_batch_phase(mp_batch* b, int32_t id, double* T, double* phi, double* Cl,
                         double* Cs, uint8_t* region);
/* Schedule: rows per work item (>= 8) and persistent grid size in blocks of
 * four waves (0: one wave per work item, the default with 12 rows). */
mp_status mp_batch_set_schedule(mp_batch* b, int32_t rows_per_item, int32_t blocks);
mp_status mp_batch_get_schedule(mp_batch* b, int32_t* rows_per_item, int32_t* blocks);
/* Mixed precision with one wave per work item: MP_LAUNCH_PER_STEP runs each
 * step's stages in one launch over the work list, each item waiting on the
 * items of the previous stage whose rows it reads (bounded waits; a stall
 * fails the step with MP_ERR_DEVICE); MP_LAUNCH_PER_STAGE launches per stage;
 * MP_LAUNCH_AUTO (the default) takes the autotuner's measured choice and,
 * before it has run, one launch per step when the work list gives each wave
 * slot of the chip fewer than eight items (the single-mesh rule; the server
 * applies it to every epoch step).  Results are bitwise the same in every
 * mode.
 * mp_batch_get_launch reports whether steps run in one launch now.  No
 * reference counterpart (launch scheduling). */
mp_status mp_batch_set_launch(mp_batch* b, int32_t mode);
/* Every step of every mesh is checked against the preheat's heat balance
 * (src/couple/preheat.cpp:268-273), as mp_transport_step does: a mesh whose
 * balance fails keeps its state and reports MP_ERR_BALANCE; the others
 * commit.  Relative tolerance, default 1e-10 (server requests use the
 * default). */
mp_status mp_batch_set_balance_tolerance(mp_batch* b, double tol);
mp_status mp_batch_get_launch(mp_batch* b, int32_t* per_step);
/* Time candidate schedules on the resident meshes (state is left unchanged),
 * apply the fastest and return it.  cache_path (may be NULL) is a JSON file
 * keyed by device, precision, scheme and batch shape: a hit skips the search,
 * a miss appends the result.  A mixed batch also times per-stage against
 * per-step launches (mp_batch_set_launch). */
mp_status mp_batch_autotune(mp_batch* b, int32_t nsteps, int32_t scheme, const char* cache_path,
                            int32_t* rows_per_item, int32_t* blocks, double* ms_per_step);
/* Time nsteps (state unchanged): ms_total, mean stage launch, mean
 * final stage launch with its phase kernel (HIP events; eager launches). */
mp_status mp_batch_profile(mp_batch* b, int32_t nsteps, int32_t scheme, double* ms_total,
                           double* ms_stage, double* ms_phase_stage);
/* Resident meshes, their cells, work items and arena bytes in use. */
mp_status mp_batch_info(mp_batch* b, int32_t* meshes, int64_t* cells, int64_t* items,
                        uint64_t* hbm_in_use);

/* ------------------------------------------------------------- serving --- */
/* Continuous batching of transport runs.  Requests arrive at any time from any
 * thread; a server thread admits them into slots of one device engine while
 * slots and the HBM budget allow, advances all resident runs together in
 * epochs of at most max_epoch_steps steps (one launch per stage),
 * and retires each run after its last step: outputs are written to the
 * caller's arrays and the slot and memory are released.  A run's results do
 * not depend on which other runs shared its launches.  No reference
 * counterpart (the reference is a batch MPI program). */
typedef struct {
    mp_mesh_desc mesh;
    const double* H;          /* initial cell averages, j*nx+i */
    const double* C;
    const double* vnx;        /* face velocities as mp_transport_set_face_velocity */
    const double* vny;
    const uint8_t* bc[4];     /* bottom, right, top, left policies (NULL: zero flux) */
    double cfl;               /* dt = cfl / max rate */
    int32_t nsteps;
    /* outputs (any may be NULL); written before the request completes */
    double* H_out;
    double* C_out;
    double* T;
    double* phi;
    double* Cl;
    double* Cs;
    uint8_t* region;
    /* phase pressure model: 0 constant (mp_phase_params.pressure), 1 lithostatic below
     * surface_y, as mp_transport_set_lithostatic */
    int32_t lithostatic;
    double surface_y;
    /* bound limiter as mp_transport_set_limiter: limiter = 0 off (zero-filled
     * requests), else with the given tables (defaults 5, 4, 0.9) */
    int32_t limiter;
    int32_t cell_points_per_axis;
    int32_t diffusion_samples;
    double extent_fraction;
    /* thermal diffusion of H as mp_transport_set_diffusion (0: off), with the
     * sampling above, and thermal boundary data as
     * mp_transport_set_thermal_boundary (NULL: zero flux) */
    double thermal_diffusivity;
    const uint8_t* thermal_types;
    const double* thermal_values;
} mp_request;

typedef struct {
    int64_t id;
    int32_t status;               /* mp_status of the run */
    uint32_t error_bits;
    int32_t steps_done;
    double dt;
    double total_H, total_C, boundary_H, boundary_C;
    double t_submit, t_admit, t_done;   /* seconds on the server clock */
} mp_request_result;

typedef struct {
    int64_t submitted, completed, failed, queued, resident;
    double cell_steps;            /* sum over completed runs of cells * steps */
    double elapsed_s;             /* first submit to last completion */
    double p50_ms, p99_ms, mean_ms, max_ms;   /* submit -> done latency, last 65536 runs */
    uint64_t hbm_in_use, hbm_peak, hbm_budget;
    int64_t epochs, max_resident;
} mp_server_stats;

/* Input and output arrays of a request must stay valid until it completes. */
mp_status mp_server_create(int32_t max_slots, const mp_weno_options* weno,
                           const mp_phase_params* phase, int32_t precision, int32_t scheme,
                           uint64_t hbm_budget_bytes, int32_t max_epoch_steps, mp_server** out);
mp_status mp_server_submit(mp_server* s, const mp_request* req, int64_t* id);
/* Wait up to timeout_s (< 0: forever) for request id; MP_ERR_PENDING on
 * timeout.  A finished request's result is handed back once: the server then
 * releases it, and a later wait on the same id fails with MP_ERR_ARG. */
mp_status mp_server_wait(mp_server* s, int64_t id, double timeout_s, mp_request_result* out);
/* Wait until every submitted request has completed. */
mp_status mp_server_drain(mp_server* s);
mp_status mp_server_stats_get(mp_server* s, mp_server_stats* out);
/* Rows per work item for runs admitted afterwards (>= 8; default 12).  A run's
 * reported boundary sums depend on it at rounding level; states and phase
 * outputs do not. */
mp_status mp_server_set_tile(mp_server* s, int32_t rows_per_item);
mp_status mp_server_destroy(mp_server* s);
/* Seconds on the server clock (the clock of t_submit / t_done). */
double mp_server_now(mp_server* s);
/* Progress of the device work, one sample per completed epoch (the most
 * recent 65536): t[i] on the server clock, and the cell-steps that all epochs
 * completed by then advanced (cumulative, successful runs).  Up to cap
 * samples are written, oldest first; *n receives the count written.  Steady
 * throughput is the slope of this curve away from the ramp and the drain. */
mp_status mp_server_progress(mp_server* s, double* t, double* cell_steps, int32_t cap, int32_t* n);

/* ------------------------------------------------------------- formats --- */
/* The H/C handoff between a preheat and an evolution: cellH1.dat and
 * cellC1.dat (text matrices, one row per j, 17 significant digits) and the
 * schema-1 transport_state.json manifest, as written by
 * src/couple/initialization.cpp:642-686 (matrices), :823-841 (manifest) and
 * checked by LoadPreheatState, src/couple/evolve.cpp:54-125.  Host only. */
typedef struct {
    char family[64];          /* "rectangular", "perturbed_quadrilateral", ... */
    int32_t nx, ny;
    double domain[4];         /* x0, x1, y0, y1 */
    double perturbation;
    uint64_t seed;
    double enthalpy_Jkg, length_m, time_scale_s;   /* scales h0, l0, t0 */
    double time;              /* nondimensional */
    int64_t accepted_steps;   /* -1 when the manifest has none */
} mp_handoff_meta;

/* Writes the three files into an existing directory, each through a temporary
 * file and a rename. */
mp_status mp_handoff_write(const char* directory, const mp_handoff_meta* meta, const double* H,
                           const double* C);
/* Reads and validates against expect (family, nx, ny, domain and scales to
 * 1e-12, perturbation and seed unless rectangular); cell averages must be
 * finite and >= 0, C <= c_max.  found (may be NULL) receives the manifest's
 * time, accepted steps, perturbation and seed.  H, C are assigned on success. */
mp_status mp_handoff_read(const char* directory, const mp_handoff_meta* expect, double c_max,
                          double* H, double* C, mp_handoff_meta* found);

/* One edge Gauss-point sample of the flow solve (FlowPointSample,
 * src/couple/visualization.h:8-12): edge id (horizontal edges first,
 * j*nx+i; then vertical, nx*(ny+1) + j*(nx+1) + i; src/core/mesh_info.h:79),
 * point index q in the edge's canonical (increasing i or j) order, and the
 * Stokes velocity. */
typedef struct {
    int64_t edge_id;
    int32_t q;
    double vx, vy;
} mp_edge_sample;

/* Face velocities for mp_transport_set_face_velocity from edge samples, as
 * the preheat builds its faces (src/couple/preheat.cpp:83-90): the last
 * sample of an (edge, q) wins.  Every face point must be sampled. */
mp_status mp_face_velocity_from_samples(int32_t nx, int32_t ny, int32_t q, int64_t nsamples,
                                        const mp_edge_sample* samples, double* vnx, double* vny);

/* ---------------------------------------------------- coupled transport --- */
/* The phase-coupled transport law of MantlePar's evolution (one rank's share
 * of a uniform mesh, FP64): the RHS of src/couple/evolve.cpp:279-384 and the
 * SSPRK2 loop with admissibility/CFL retries of :428-511.  Traces come from
 * the limited reconstruction (StateReconstruction::Update + Limit); H carries
 * the temperature flux with the latent term -LD*(1-phi)*v_s.n, C the
 * effective velocity of PhaseTransportVelocities (:36-52); T diffuses by
 * normal-line samples of the reconstruction; the cell Gauss points carry the
 * cooling source.  The Stokes-Darcy solve stays with the caller: the flow
 * samples are inputs, held fixed over a call (the reference re-solves the
 * flow inside every RHS, SolveReconstructedFlow at :282). */
typedef struct {
    double theta;                 /* c.flow.material.theta: Darcy flux phi^(1+theta) u_d */
    double cooling;               /* alpha0*l0*g/cp (evolve.cpp:285) */
    double thermal_diffusivity;   /* c.thermalDiffusivity */
    int32_t cell_points_per_axis; /* c.input.quadrature.cellPointsPerAxis (5) */
    int32_t diffusion_samples;    /* c.sampling.numberOfSamples (4) */
    double extent_fraction;       /* c.sampling.extentFraction (0.9) */
    int32_t lf_mode;              /* c.stabilizer.mode: 0 local, 1 global */
    double global_speed;          /* c.stabilizer.globalSpeed */
} mp_coupled_params;

typedef struct {
    double boundary_H, boundary_C;   /* outward boundary flux totals */
    double source_H;                 /* cooling source total */
    double maximum_rate;             /* max per-cell CFL rate */
    double divergence;               /* max |mixture divergence| */
} mp_coupled_rates;

/* input time controls (src/input/input.h: time.step.*) */
typedef struct {
    double cfl, initial_step, maximum_step, minimum_step, end_time;
    int32_t fixed;                /* control "fixed" (1) or "cfl" (0) */
} mp_coupled_time;

/* One accepted step, as the history row of src/couple/evolve.cpp:394-409. */
typedef struct {
    double dt, courant, balance_H, balance_C, boundary_H, boundary_C, source_H, maximum_rate;
    int32_t retries, pad;
} mp_coupled_step;

void mp_default_coupled_params(mp_coupled_params* p);
mp_status mp_coupled_create(const mp_mesh_desc* mesh, const mp_phase_params* phase,
                            const mp_coupled_params* params, void* stream, mp_coupled** out);
mp_status mp_coupled_destroy(mp_coupled* c);
/* FlowPointSample arrays (src/couple/visualization.h:8-12): edges by
 * reference edge id (horizontal first, j*nx+i; then vertical) and canonical q,
 * stokes/darcy [e][q][2], phi [e][q]; cell points [cell][b*n+a][2] / [..]. */
mp_status mp_coupled_set_flow(mp_coupled* c, const double* edge_stokes, const double* edge_darcy,
                              const double* edge_phi, const double* cell_stokes,
                              const double* cell_darcy, const double* cell_phi);
/* Advection policies (0 zero flux, 1 outflow only, 2 inflow/outflow) of H
 * (field 0, c.boundary.enthalpy) or C (field 1, c.boundary.composition). */
mp_status mp_coupled_set_boundary(mp_coupled* c, int32_t field, const uint8_t* bottom,
                                  const uint8_t* right, const uint8_t* top, const uint8_t* left);
mp_status mp_coupled_set_inflow(mp_coupled* c, const double* inflow_H, const double* inflow_C);
/* EvaluateThermalBoundary data per boundary point, as mp_transport_set_thermal_boundary. */
mp_status mp_coupled_set_thermal_boundary(mp_coupled* c, const uint8_t* types, const double* values);
/* Lithostatic phase pressure (src/couple/initialization.cpp:155-162): every
 * phase evaluation of the RHS -- cell Gauss points, face Gauss points (both
 * sides and the inflow temperature) and normal-line diffusion samples -- uses
 * P = staticPressure(surface_y - y) at its own point, as c.pressure(p, t)
 * does in src/couple/evolve.cpp:189, 273, 318.  surface_y >= mesh top.
 * enable = 0 returns to the constant phase.pressure. */
mp_status mp_coupled_set_lithostatic(mp_coupled* c, int32_t enable, double surface_y);
mp_status mp_coupled_upload_state(mp_coupled* c, const double* H, const double* C);
mp_status mp_coupled_download_state(mp_coupled* c, double* H, double* C);
/* RHS of the current state: rhs_H, rhs_C, rate, divergence, cell-averaged T
 * and phi [N] (any may be NULL) and the totals (RHS, evolve.cpp:279-384). */
mp_status mp_coupled_rhs(mp_coupled* c, double* rhsH, double* rhsC, double* rate, double* div,
                         double* T, double* phi, mp_coupled_rates* totals);
/* Up to nsteps accepted SSPRK2 steps from *time (Run, evolve.cpp:428-511):
 * on an error the state and *time are those of the last accepted step. */
mp_status mp_coupled_run(mp_coupled* c, const mp_coupled_time* time_controls, double* time,
                         int32_t nsteps, mp_coupled_step* steps, int32_t* accepted);
/* Batched form of mp_coupled_run (no counterpart in the reference, which
 * advances one mesh per rank): n handles advance together, each with its own
 * time controls time_controls[k] and time times[k].  Every round enqueues one
 * step attempt of each handle still running as one launch per kernel over all
 * of them and reads their totals back with one wait; each handle takes
 * mp_coupled_run's decisions, so it ends as its own mp_coupled_run call would
 * (state, time, step records).  steps: n*nsteps records, handle k's from
 * steps + k*nsteps (may be NULL); accepted[k], status[k] per handle (may be
 * NULL).  A handle that fails rolls back alone; the others go on.  Returns
 * the status of the first failed handle.  The handles must share a device
 * and not be used elsewhere during the call; it runs on handles[0]'s stream. */
mp_status mp_coupled_run_batch(mp_coupled* const* handles, int32_t n, const mp_coupled_time* time_controls,
                               double* times, int32_t nsteps, mp_coupled_step* steps, int32_t* accepted,
                               mp_status* status);

/* ------------------------------------------------------------- phase only --- */
/* EutecticModel::evaluate over n cells on the device (host arrays in/out).
 * P may be NULL (uses params->pressure).  Any output may be NULL. */
mp_status mp_phase_evaluate(const mp_phase_params* params, int64_t n, const double* H,
                            const double* C, const double* P, double* T, double* phi,
                            double* Cl, double* Cs, uint8_t* region, double* phi1,
                            double* phi2, double* dTdC, double* dTdH);
/* DerivedParameters of src/phase/eutectic_rescaled.h:29-41, in its order. */
typedef struct {
    double dT, h0, gamma, LD, TDm0, TDe0, invk0, l0, u0, p0, t0;
} mp_phase_derived;
/* EutecticModel's derived scales (derive(), src/phase/eutectic_rescaled.cpp:
 * 42-86), with its checks in its order: invalid parameters fail with
 * MP_ERR_ARG (std::invalid_argument there), unrepresentable scales with
 * MP_ERR_FP (std::overflow_error).  Host only. */
mp_status mp_phase_derive(const mp_phase_params* params, mp_phase_derived* out);
/* EutecticModel::staticPressure(zD) (use_length = 0: the length scale l0) and
 * staticPressure(zD, length_scale) (src/phase/eutectic_rescaled.cpp:151-161).
 * Host only. */
mp_status mp_phase_static_pressure(const mp_phase_params* params, double zD, double length_scale,
                                   int32_t use_length, double* P);
/* EutecticModel::pressureCorrectedTemperature (src/phase/eutectic_rescaled.cpp:
 * 143-149).  Host only. */
mp_status mp_phase_pressure_corrected_temperature(const mp_phase_params* params, double TD, double P,
                                                  double* out);
mp_status mp_phase_boundaries(const mp_phase_params* params, double C, double P,
                              double* Tep, double* Tmp, double* H_eutectic_end,
                              double* H_liquidus);

#ifdef __cplusplus
}
#endif
#endif
