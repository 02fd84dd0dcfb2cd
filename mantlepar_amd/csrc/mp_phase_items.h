// The phase of a stage's output and its state sums, per work item of the
// stage sweep's layout; shared by the phase kernels and the one-launch mixed
// step, which runs the previous step's phase items after its own stages.
#pragma once

#include <hip/hip_runtime.h>

#include "mp_common.h"
#include "mp_device.h"
#include "mp_phase.h"

namespace mp {

// Phase of a stage's output and its state sums, per work item of the stage
// sweep: lane = column, rows in order, the same butterfly as the sweep's own
// partials.  Both precisions run their last stage without a phase epilogue
// and this kernel after it (beside the next step when stepping): the sweep
// keeps its register budget, and the phase work streams at full occupancy.
__device__ __forceinline__ void phase_item(const MeshDev& m, const StagePtrs& sp, int item,
                                           int strip, int j0, int j1)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int i = strip * kStripCells - 1 + lane;
    const bool owned = i >= 0 && i < m.nx && lane >= 1 && lane <= kStripCells;
    double sH = 0, eH = 0, sC = 0, eC = 0;
    uint32_t err = 0;
    if (owned) {
        for (int r = j0; r < j1; ++r) {
            const size_t k = size_t(r) * m.nx + i;
            const double h = sp.Ho[k], c = sp.Co[k];
            const PhaseOut po = phase_eval<false>(phase_at(m.ph, m.tep_row, r), h, c, err);
            sp.T[k] = po.T;
            sp.phi[k] = po.phil;
            sp.Cl[k] = po.cl;
            sp.Cs[k] = po.cs;
            sp.region[k] = uint8_t(po.region);
            two_sum_acc(sH, eH, h);
            two_sum_acc(sC, eC, c);
        }
    }
    item_partials<true>(sp, item, sH, eH, sC, eC, 0.0, 0.0, err);
}

}  // namespace mp
