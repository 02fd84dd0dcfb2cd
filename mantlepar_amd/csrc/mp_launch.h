// Host-side launchers of the device kernels (mp_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "mp_common.h"

namespace mp {

// One stage of either sweep; the phase of a step's last stage runs in its own
// kernel after it (launch_phase_items).
// limit: apply the bound limiter (FP64 sweep only).
// diffusion: the mesh has thermal diffusion or thermal boundary data (the
// mixed sweep compiles the terms out otherwise).
// global_lf: the mesh uses the global LF speed (mixed sweep's record form).
// limit: 0 off, else the bound limiter (mixed: the mode of limit_mode32).
hipError_t launch_stage(int mode, int q, const MeshDev* d_mesh, const StagePtrs& sp,
                        int nitems, hipStream_t stream, int limit = 0, bool diffusion = true,
                        bool global_lf = false);
hipError_t launch_limit_diag(const MeshDev* d_mesh, int ncells, const double* H, const double* C,
                             double* thetaH, double* thetaC, uint32_t* error, hipStream_t stream);
hipError_t launch_face_prep(const MeshDev* d_mesh, int nx, int ny, FaceMom* fx, FaceMom* fy,
                            FaceRec* rx, FaceRec* ry, hipStream_t stream);
hipError_t launch_face_rec64(const MeshDev* d_mesh, int nx, int ny, FaceRec64* rx, FaceRec64* ry,
                             hipStream_t stream);
hipError_t launch_recon_diag(int mode, const MeshDev* d_mesh, int ncells, const double* u,
                             int field, double* sigma, double* omega, double* traces,
                             uint32_t* error, hipStream_t stream);
hipError_t launch_rhs_diag(int mode, const MeshDev* d_mesh, int ncells, const double* H,
                           const double* C, double* rhsH, double* rhsC, double* bnd,
                           uint32_t* error, hipStream_t stream);
// Reconstructions of H and C at n points (cell index, offset from the cell
// centre in units of sqrt(area)).
hipError_t launch_point_eval(int mode, const MeshDev* d_mesh, const double* H, const double* C,
                             int64_t n, const int32_t* cell, const double* X, const double* Y,
                             double* outH, double* outC, uint32_t* error, hipStream_t stream);
hipError_t launch_rate(const MeshDev* d_mesh, int ncells, unsigned long long* max_bits,
                       hipStream_t stream);
hipError_t launch_phase(const PhaseDev& ph, double TDe0, double nu, double dT, int64_t n,
                        const double* H, const double* C, const double* P, double* T,
                        double* phi, double* Cl, double* Cs, uint8_t* region, double* phi1,
                        double* phi2, double* dTdC, double* dTdH, uint32_t* error,
                        hipStream_t stream);
hipError_t launch_state_phase(const MeshDev* d_mesh, int ncells, const double* H, const double* C,
                              double* T, double* phi, double* Cl, double* Cs, uint8_t* region,
                              uint32_t* error, hipStream_t stream);
// blocks > 0: persistent grid of that many blocks taking items from *queue
// (an int that must be zero at launch); blocks == 0: one wave per item.
// limit: 0 when no mesh of the list has its limiter on; else the limiter
// instance, in which each mesh takes the limited sweep only if its limiter is
// on (mixed: with the table instance of its single-mesh launch).
// diffusion: some mesh of the list has diffusion or thermal boundary data;
// each mesh then takes the diffusion terms only if it has them.
hipError_t launch_batch_stage(int mode, int q, const MeshDev* meshes,
                              const StagePtrs* stages, const int2* items, int nitems, int blocks,
                              int* queue, hipStream_t stream, bool global_lf = false, int limit = 0,
                              bool diffusion = false);
// prev (may be null): per-mesh totals before the step; with it each mesh's
// heat balance is checked at tolerance tol (kErrBalance in its error word).
hipError_t launch_batch_reduce(const StagePtrs* stages, const int* nitems, const MeshDev* meshes,
                               const int* slots, int nslots, int nstages, const double* w,
                               double* out, hipStream_t stream, const double* prev = nullptr,
                               double tol = 0.0);
// Per-mesh totals of the state that stages[k] names as u^n (out[k*4 + 0..3]);
// part: scratch of nslots * kStateSlicesMax * 4 doubles.
constexpr int kStateSlicesMax = 32;
hipError_t launch_batch_state_sums(const StagePtrs* stages, const MeshDev* meshes, const int* slots,
                                   int nslots, double* out, double* part, hipStream_t stream);
hipError_t launch_clear(uint32_t* p, int n, hipStream_t stream);
// Resident blocks per CU of the batched stage kernel (0 if the query fails).
int batch_blocks_per_cu(int mode);
// Mixed-precision sweeps (mp_sweep32.hip).
int batch_blocks_per_cu32();
// Neither sweep evaluates the phase: phase_items_kernel follows a step's last stage.
// diffusion: the mesh has thermal diffusion or thermal boundary data.
// global_lf: the mesh uses the global LF speed, so the sweep fetches full face
// records instead of their one-sided form.
// limit: the bound limiter (the mesh's limiter tables set): 0 off, 1 any
// tables, 2 the default tables (q = 5, 5 cell points per axis, 4 samples),
// whose point loops are unrolled at compile time.
hipError_t launch_stage32(const MeshDev* d_mesh, const StagePtrs& sp, int nitems, hipStream_t stream,
                          bool diffusion, bool global_lf, int limit = 0);
// Whether a mesh record carries thermal diffusion or thermal boundary data
// (the sweeps then compile in the diffusion terms).
__host__ __device__ inline bool has_diffusion(const MeshDev& m)
{
    bool on = m.dif_x != 0.0 || m.dif_y != 0.0 || m.dif_bx != 0.0 || m.dif_by != 0.0;
    for (int k = 0; k < 4; ++k) on = on || m.tbc[k] != nullptr;
    return on;
}

// The limit argument of launch_stage32 for a mesh record.
__host__ __device__ inline int limit_mode32(const MeshDev& m)
{
    if (!m.limit) return 0;
    return (m.q == 5 && m.lim_ncp == 5 && m.lim_ni == 2 && m.lim_nb == 4) ? 2 : 1;
}
// All stages of one mixed-precision step in one launch: tab[0..nstages) are
// the stage records, flags [nstages][nitems] completion flags that hold
// earlier epochs only (never `epoch`) when the launch starts.  Items publish
// `epoch` and wait for `wait_epoch` (equal to it except in the stall check),
// each wait giving up after spin_limit polls.
constexpr int kSpinLimit = 1 << 20;
constexpr uint32_t kEpochNever = 0xffffffffu;   // no flag ever holds it
// ptab (device memory, may be null) and nphase > 0: the launch also runs the
// nphase phase items of the stage record *ptab (the previous step's last
// stage) after its sweep items.
hipError_t launch_step32(const MeshDev* d_mesh, const StagePtrs* tab, int nitems, int nstages,
                         uint32_t* flags, uint32_t epoch, uint32_t wait_epoch, int spin_limit,
                         hipStream_t stream, bool diffusion, bool global_lf,
                         const StagePtrs* ptab = nullptr, int nphase = 0);
hipError_t launch_batch_stage32(const MeshDev* meshes, const StagePtrs* stages, const int2* items,
                                int nitems, int blocks, int* queue, bool global_lf, hipStream_t stream,
                                int limit = 0, bool diffusion = false);
// All stages of one mixed-precision step over a batch work list (one wave per
// entry, (-1, 0) pads skipped) in one launch.  Stage st of slot k reads
// tab[st * capacity + k]; completion flags are flags[st * nflags +
// flag_off[k] + item].  The launch first advances *epoch_ctr (a one-thread
// kernel) and its items publish and wait for the new value, so a captured
// graph uses a new epoch on every replay; flags must hold earlier values.
hipError_t launch_batch_step32(const MeshDev* meshes, const StagePtrs* tab, int capacity,
                               const int2* items, int nitems, int nstages, uint32_t* flags,
                               const int* flag_off, int nflags, uint32_t* epoch_ctr, int spin_limit,
                               bool global_lf, hipStream_t stream);
// Phase and state sums of a stage's output over the stage's work items
// (sp.Ho/sp.Co in, sp.T..region and partial slots 0-3 out).
hipError_t launch_phase_items(const MeshDev* d_mesh, const StagePtrs& sp, int nitems,
                              hipStream_t stream);
hipError_t launch_batch_phase_items(const MeshDev* meshes, const StagePtrs* stages,
                                    const int2* items, int nitems, hipStream_t stream);
// Step totals (two passes over kStepSlices wave-sized slices of the item
// list); scratch: kStepSlices * 6 doubles.
constexpr int kStepSlices = 128;
// prev (may be null): the totals before the step; the step is then checked
// against the conservation balance with relative tolerance tol (kErrBalance).
hipError_t launch_reduce_step(const double* partials, int nitems, int nstages, const double* w,
                              double area, double* out, double* scratch, hipStream_t stream,
                              const double* prev = nullptr, double dt = 0.0, double tol = 0.0,
                              uint32_t* error = nullptr);
// out[0..1]: compensated |K|-weighted sums of H and C (out[2..3] zero).
hipError_t launch_state_sums(const double* H, const double* C, int64_t n, double area, double* out,
                             double* scratch, hipStream_t stream);

}  // namespace mp
