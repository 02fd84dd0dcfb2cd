// Mixed-precision stage sweep, one launch per stage over one mesh (the
// device code is in mp_sweep32.h).
#include "mp_sweep32.h"

namespace mp {
namespace {

// One wave per workgroup: a finished item frees its slot at once instead of
// waiting for the other waves of a larger workgroup (7% faster at 2048x4096).
// Register budget: three waves per SIMD, twelve workgroups per CU.
// LIM (the bound limiter): three waves per SIMD hold the limiter's point
// loops without scratch on their own; the unrolled default tables, the
// diffusion terms or the full face records would spill 12-20 B per lane
// there, so those instances run two waves.
template <bool DIFF, bool FULL, int LIM = 0>
__global__ __launch_bounds__(64, LIM == 2 || (LIM != 0 && (DIFF || FULL)) ? 2 : 3) void stage32_kernel(const MeshDev* __restrict__ mesh, StagePtrs sp,
                                                                  int nitems)
{
    if (int(blockIdx.x) >= nitems) return;
    const int item = item_of_block(sp, int(blockIdx.x), nitems);
    const MeshDev& m = *mesh;
    int strip, j0, j1;
    item_region(sp, m.ny, item, strip, j0, j1);
    sweep32_item<DIFF, FULL, false, LIM>(m, sp, item, strip, j0, j1);
}

}  // namespace

hipError_t launch_stage32(const MeshDev* d_mesh, const StagePtrs& sp, int nitems, hipStream_t stream,
                          bool diffusion, bool global_lf, int limit)
{
    const dim3 grid(nitems), block(kWave);
    if (limit == 2) {
        if (global_lf) {
            if (diffusion) stage32_kernel<true, true, 2><<<grid, block, 0, stream>>>(d_mesh, sp, nitems);
            else stage32_kernel<false, true, 2><<<grid, block, 0, stream>>>(d_mesh, sp, nitems);
        } else {
            if (diffusion) stage32_kernel<true, false, 2><<<grid, block, 0, stream>>>(d_mesh, sp, nitems);
            else stage32_kernel<false, false, 2><<<grid, block, 0, stream>>>(d_mesh, sp, nitems);
        }
    } else if (limit) {
        if (global_lf) {
            if (diffusion) stage32_kernel<true, true, 1><<<grid, block, 0, stream>>>(d_mesh, sp, nitems);
            else stage32_kernel<false, true, 1><<<grid, block, 0, stream>>>(d_mesh, sp, nitems);
        } else {
            if (diffusion) stage32_kernel<true, false, 1><<<grid, block, 0, stream>>>(d_mesh, sp, nitems);
            else stage32_kernel<false, false, 1><<<grid, block, 0, stream>>>(d_mesh, sp, nitems);
        }
    } else if (global_lf) {
        if (diffusion) stage32_kernel<true, true><<<grid, block, 0, stream>>>(d_mesh, sp, nitems);
        else stage32_kernel<false, true><<<grid, block, 0, stream>>>(d_mesh, sp, nitems);
    } else {
        if (diffusion) stage32_kernel<true, false><<<grid, block, 0, stream>>>(d_mesh, sp, nitems);
        else stage32_kernel<false, false><<<grid, block, 0, stream>>>(d_mesh, sp, nitems);
    }
    return hipGetLastError();
}

}  // namespace mp
