// Mixed-precision steps in one launch each: one mesh, and a batch work list
// (the device code is in mp_sweep32.h).
#include "mp_sweep32.h"

namespace mp {
namespace {

// One launch per step: nstages * nitems one-wave workgroups.  Workgroup b runs
// stage b / nitems on the item that the single-stage launch gives its
// workgroup b % nitems (XCD slabs, tail items last), so each stage's segment
// is dispatched as a single-stage launch would be, and a stage's first items
// (the bottom bands) wait on the previous stage's first items, long done.  An
// item of stage k > 0 starts once the items whose output it reads are done
// (wait_inputs) and publishes its own completion; tab[k] is stage k's record,
// flags[k][item] its completion flags.  Items publish `epoch` and wait for
// `wait_epoch`, the same value except in the stall check, where the host
// passes a value no flag ever holds.  The ramp and drain of a launch are
// paid once per step instead of once per stage.
// nphase > 0: nphase more workgroups after the sweep items run the phase items
// of the previous step's output (*ptab, its last stage record), which this
// step only reads.  They wait on nothing and are dispatched last, so they run
// in the slots that the launch's drain leaves idle.
template <bool DIFF, bool FULL>
__global__ __launch_bounds__(64, 3) void step32_kernel(const MeshDev* __restrict__ mesh,
                                                       const StagePtrs* __restrict__ tab, int nitems,
                                                       int nstages, uint32_t* __restrict__ flags,
                                                       uint32_t epoch, uint32_t wait_epoch, int spin_limit,
                                                       const StagePtrs* __restrict__ ptab, int nphase)
{
    const int b = int(blockIdx.x);
    const int stage = b / nitems;
    if (stage >= nstages) {
        const int pb = b - nstages * nitems;
        if (pb >= nphase) return;
        const StagePtrs& sp = *ptab;
        const MeshDev& m = *mesh;
        const int item = item_of_block(sp, pb, nphase);
        int strip, j0, j1;
        item_region(sp, m.ny, item, strip, j0, j1);
        phase_item(m, sp, item, strip, j0, j1);
        return;
    }
    const StagePtrs& sp = tab[stage];
    const int item = item_of_block(sp, b - stage * nitems, nitems);
    const MeshDev& m = *mesh;
    int strip, j0, j1;
    item_region(sp, m.ny, item, strip, j0, j1);
    if (stage > 0 &&
        !wait_inputs(sp, m.ny, strip, j0, j1, flags + size_t(stage - 1) * nitems, wait_epoch, spin_limit))
        return;
    sweep32_item<DIFF, FULL, true>(m, sp, item, strip, j0, j1);
    if (stage + 1 < nstages) publish_item(flags + size_t(stage) * nitems + item, epoch);
}


// One launch per step over a batch work list: nstages x nitems one-wave
// workgroups, stage-major; workgroup b runs stage b / nitems on list entry
// b % nitems, so each stage's segment keeps the list's XCD slabs.  The
// dependencies are those of step32_kernel within each mesh.  (The batch keeps
// each step's phase items in their own launch on the side stream: carried
// in the next step's launch, as the single mesh does, they measured 1%
// slower on the 256-mesh batch.)
template <bool FULL>
__global__ __launch_bounds__(64, 3) void batch_step32_kernel(
    const MeshDev* __restrict__ meshes, const StagePtrs* __restrict__ tab, int capacity,
    const int2* __restrict__ items, int nitems, int nstages, uint32_t* __restrict__ flags,
    const int* __restrict__ flag_off, int nflags, const uint32_t* __restrict__ epoch_ctr, int spin_limit)
{
    const int b = int(blockIdx.x);
    const int stage = b / nitems;
    if (stage >= nstages) return;
    const int2 w = items[b - stage * nitems];
    const int slot = __builtin_amdgcn_readfirstlane(w.x);
    const int local = __builtin_amdgcn_readfirstlane(w.y);
    if (slot < 0) return;              // (-1, 0) pads an XCD slab
    // written by the preceding one-thread kernel, so a plain load sees it
    const uint32_t epoch = __builtin_amdgcn_readfirstlane(*epoch_ctr);
    const MeshDev& m = meshes[slot];
    const StagePtrs& sp = tab[size_t(stage) * capacity + slot];
    int strip, j0, j1;
    item_region(sp, m.ny, local, strip, j0, j1);
    const int off = __builtin_amdgcn_readfirstlane(flag_off[slot]);
    if (stage > 0 &&
        !wait_inputs(sp, m.ny, strip, j0, j1, flags + size_t(stage - 1) * nflags + off, epoch, spin_limit))
        return;
    sweep32_item<false, FULL, true>(m, sp, local, strip, j0, j1);
    if (stage + 1 < nstages) publish_item(flags + size_t(stage) * nflags + off + local, epoch);
}


__global__ void epoch_bump_kernel(uint32_t* ctr)
{
    if (threadIdx.x == 0) *ctr += 1u;
}

}  // namespace

hipError_t launch_step32(const MeshDev* d_mesh, const StagePtrs* tab, int nitems, int nstages,
                         uint32_t* flags, uint32_t epoch, uint32_t wait_epoch, int spin_limit,
                         hipStream_t stream, bool diffusion, bool global_lf, const StagePtrs* ptab,
                         int nphase)
{
    if (!ptab) nphase = 0;
    const dim3 grid(unsigned(nitems) * unsigned(nstages) + unsigned(nphase)), block(kWave);
#define MP_STEP32(D, F)                                                                                   \
    step32_kernel<D, F><<<grid, block, 0, stream>>>(d_mesh, tab, nitems, nstages, flags, epoch, wait_epoch, \
                                                    spin_limit, ptab, nphase)
    if (global_lf) {
        if (diffusion) MP_STEP32(true, true);
        else MP_STEP32(false, true);
    } else {
        if (diffusion) MP_STEP32(true, false);
        else MP_STEP32(false, false);
    }
#undef MP_STEP32
    return hipGetLastError();
}

hipError_t launch_batch_step32(const MeshDev* meshes, const StagePtrs* tab, int capacity,
                               const int2* items, int nitems, int nstages, uint32_t* flags,
                               const int* flag_off, int nflags, uint32_t* epoch_ctr, int spin_limit,
                               bool global_lf, hipStream_t stream)
{
    if (nitems <= 0) return hipSuccess;
    epoch_bump_kernel<<<1, 64, 0, stream>>>(epoch_ctr);
    const dim3 grid(unsigned(nitems) * unsigned(nstages)), block(kWave);
    if (global_lf)
        batch_step32_kernel<true><<<grid, block, 0, stream>>>(meshes, tab, capacity, items, nitems, nstages, flags,
                                                              flag_off, nflags, epoch_ctr, spin_limit);
    else
        batch_step32_kernel<false><<<grid, block, 0, stream>>>(meshes, tab, capacity, items, nitems, nstages, flags,
                                                               flag_off, nflags, epoch_ctr, spin_limit);
    return hipGetLastError();
}

}  // namespace mp
