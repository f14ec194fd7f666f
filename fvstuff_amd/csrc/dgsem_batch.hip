// Batched (multi-instance) stage launches: one workgroup per tile of the
// tile list, every instance and degree in the same launch.
// Compiled once per compute precision (-DDG_REAL=float | double).
#include "dgsem_launch.hpp"

#ifndef DG_REAL
#error "compile with -DDG_REAL=float or double"
#endif

namespace dg {


// One launch of the batched stage: one workgroup per tile (grid = ntiles),
// or the persistent form on `grid` workgroups (grid < ntiles).
#define DG_BATCH_LAUNCH(VOLV)                                                                              \
    {                                                                                                    \
        auto k = grid > 0 ? batch_stage_persistent<DG_REAL, false, VOLV> : batch_stage_kernel<DG_REAL, false, VOLV>; \
        static size_t done[2] = {0, 0};   /* raised once per kernel, never under stream capture */      \
        size_t& d = done[grid > 0];                                                                      \
        if (lds > 65536 && lds > d) {                                                                    \
            hipError_t e = hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds); \
            if (e != hipSuccess) return e;                                                               \
            d = lds;                                                                                     \
        }                                                                                                \
        hipLaunchKernelGGL(k, dim3(grid > 0 ? grid : ntiles), dim3(lanes), lds, s, args, tiles, ntiles, stage, sc); \
        return hipGetLastError();                                                                        \
    }

#define DG_CAT2(a, b) a##b
#define DG_CAT(a, b) DG_CAT2(a, b)

hipError_t DG_CAT(launch_batch_stage_, DG_REAL)(const KArgs* args, const int2* tiles, int ntiles, int stage,
                                                StepCtl sc, int vol, int lanes, size_t lds, int grid,
                                                hipStream_t s) {
    switch (vol) {
        case VOL_RANOCHA: DG_BATCH_LAUNCH(VOL_RANOCHA)
        case VOL_CHANDRASHEKAR: DG_BATCH_LAUNCH(VOL_CHANDRASHEKAR)
        default: DG_BATCH_LAUNCH(VOL_COLLOCATION)
    }
}

// Workgroups of the persistent form resident per CU (occupancy of the
// kernel at this size and LDS).
hipError_t DG_CAT(batch_persistent_per_cu_, DG_REAL)(int vol, int lanes, size_t lds, int* per_cu) {
    const void* k = vol == VOL_RANOCHA ? (const void*)batch_stage_persistent<DG_REAL, false, VOL_RANOCHA>
                  : vol == VOL_CHANDRASHEKAR ? (const void*)batch_stage_persistent<DG_REAL, false, VOL_CHANDRASHEKAR>
                  : (const void*)batch_stage_persistent<DG_REAL, false, VOL_COLLOCATION>;
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, k, lanes, lds);
}

}  // namespace dg
