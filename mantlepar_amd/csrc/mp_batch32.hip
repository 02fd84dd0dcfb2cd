// Mixed-precision stage sweep over a batch work list, one launch per stage
// (the device code is in mp_sweep32.h).
#include "mp_sweep32.h"

namespace mp {
namespace {

// Persistent form (queue != nullptr): every wave takes the next work item
// from a counter until the list is exhausted; otherwise one wave per item
// (and per workgroup).
// Batch meshes carry no thermal diffusion and share one LF mode (FULL: global).
// LIM != 0: meshes whose limiter is on take the limited sweep (with the
// table instance that their single-mesh launch takes), the others the plain
// one, at the limited instance's occupancy; DIFF likewise for meshes with
// diffusion or thermal boundary data.
template <bool FULL, int LIM = 0, bool DIFF = false>
__global__ __launch_bounds__(256, LIM != 0 || DIFF ? 2 : 3) void batch32_kernel(const MeshDev* __restrict__ meshes,
                                                         const StagePtrs* __restrict__ stages,
                                                         const int2* __restrict__ items, int nitems,
                                                         int* queue)
{
    const int wave = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
    int it = blockIdx.x * int(blockDim.x >> 6) + wave;
    for (;;) {
        if (queue) {
            int t = 0;
            if ((threadIdx.x & (kWave - 1)) == 0) t = atomicAdd(queue, 1);
            it = __builtin_amdgcn_readfirstlane(t);
        }
        if (it >= nitems) break;
        const int2 w = items[it];
        const int slot = __builtin_amdgcn_readfirstlane(w.x);
        const int local = __builtin_amdgcn_readfirstlane(w.y);
        if (slot >= 0) {               // (-1, 0) pads an XCD slab
            const MeshDev& m = meshes[slot];
            const StagePtrs& sp = stages[slot];
            int strip, j0, j1;
            item_region(sp, m.ny, local, strip, j0, j1);
            if (j0 < m.ny) {           // always true with a current work list
                // the kernel's own flags as template arguments: a branch that
                // cannot run instantiates nothing new
                constexpr int L = LIM != 0 ? 3 : 0;
                const bool lim = LIM != 0 && m.limit, dif = DIFF && has_diffusion(m);
                if (lim && dif) sweep32_item<DIFF, FULL, false, L>(m, sp, local, strip, j0, j1);
                else if (lim) sweep32_item<false, FULL, false, L>(m, sp, local, strip, j0, j1);
                else if (dif) sweep32_item<DIFF, FULL>(m, sp, local, strip, j0, j1);
                else sweep32_item<false, FULL>(m, sp, local, strip, j0, j1);
            }
        }
        if (!queue) break;
    }
}

}  // namespace

int batch_blocks_per_cu32()
{
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, batch32_kernel<false>, 256, 0) != hipSuccess) return 0;
    return n;
}

hipError_t launch_batch_stage32(const MeshDev* meshes, const StagePtrs* stages, const int2* items,
                                int nitems, int blocks, int* queue, bool global_lf, hipStream_t stream,
                                int limit, bool diffusion)
{
    if (nitems <= 0) return hipSuccess;
    int* q = blocks > 0 ? queue : nullptr;
    const dim3 grid(blocks > 0 ? blocks : nitems), block(blocks > 0 ? 256 : kWave);
#define MP_BATCH32(F, L, D) batch32_kernel<F, L, D><<<grid, block, 0, stream>>>(meshes, stages, items, nitems, q)
#define MP_BATCH32_LD(F)                                   \
    do {                                                   \
        if (limit && diffusion) MP_BATCH32(F, 1, true);    \
        else if (limit) MP_BATCH32(F, 1, false);           \
        else if (diffusion) MP_BATCH32(F, 0, true);        \
        else MP_BATCH32(F, 0, false);                      \
    } while (0)
    // a limited mesh takes the table instance its single-mesh launch takes
    if (global_lf) MP_BATCH32_LD(true);
    else MP_BATCH32_LD(false);
#undef MP_BATCH32_LD
#undef MP_BATCH32
    return hipGetLastError();
}

}  // namespace mp
