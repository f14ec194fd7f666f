// Batched (multi-instance) kernels: the persistent stage kernel over a tile
// list, plus the per-tile wave-speed kernel that seeds each instance's dt.
// Compiled once per compute precision (-DDG_REAL=float | double).
#include "dgsem_launch.hpp"

#ifndef DG_REAL
#error "compile with -DDG_REAL=float or double"
#endif

namespace dg {


#define DG_BATCH_LAUNCH(VOLV)                                                                              \
    {                                                                                                    \
        auto k = batch_stage_kernel<DG_REAL, false, VOLV>;                                               \
        static size_t done = 0;   /* raised once, never under stream capture */                         \
        if (lds > 65536 && lds > done) {                                                                 \
            hipError_t e = hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds); \
            if (e != hipSuccess) return e;                                                               \
            done = lds;                                                                                  \
        }                                                                                                \
        hipLaunchKernelGGL(k, dim3(grid), dim3(BLOCK), lds, s, args, tiles, ntiles, stage, sc);          \
        return hipGetLastError();                                                                        \
    }

#define DG_CAT2(a, b) a##b
#define DG_CAT(a, b) DG_CAT2(a, b)

hipError_t DG_CAT(launch_batch_stage_, DG_REAL)(const KArgs* args, const int2* tiles, int ntiles, int stage,
                                                StepCtl sc, int vol, int grid, size_t lds, hipStream_t s) {
    switch (vol) {
        case VOL_RANOCHA: DG_BATCH_LAUNCH(VOL_RANOCHA)
        case VOL_CHANDRASHEKAR: DG_BATCH_LAUNCH(VOL_CHANDRASHEKAR)
        default: DG_BATCH_LAUNCH(VOL_COLLOCATION)
    }
}

hipError_t DG_CAT(batch_stage_occupancy_, DG_REAL)(int vol, size_t lds, int* blocks_per_cu) {
    const void* k = vol == VOL_RANOCHA ? (const void*)batch_stage_kernel<DG_REAL, false, VOL_RANOCHA>
                  : vol == VOL_CHANDRASHEKAR ? (const void*)batch_stage_kernel<DG_REAL, false, VOL_CHANDRASHEKAR>
                                             : (const void*)batch_stage_kernel<DG_REAL, false, VOL_COLLOCATION>;
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k, BLOCK, lds);
}

}  // namespace dg
