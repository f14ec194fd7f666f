// Batched (multi-instance) stage launches: one workgroup per tile of the
// tile list, every instance and degree in the same launch.
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
        hipLaunchKernelGGL(k, dim3(ntiles), dim3(lanes), lds, s, args, tiles, ntiles, stage, sc);          \
        return hipGetLastError();                                                                        \
    }

#define DG_CAT2(a, b) a##b
#define DG_CAT(a, b) DG_CAT2(a, b)

hipError_t DG_CAT(launch_batch_stage_, DG_REAL)(const KArgs* args, const int2* tiles, int ntiles, int stage,
                                                StepCtl sc, int vol, int lanes, size_t lds, hipStream_t s) {
    switch (vol) {
        case VOL_RANOCHA: DG_BATCH_LAUNCH(VOL_RANOCHA)
        case VOL_CHANDRASHEKAR: DG_BATCH_LAUNCH(VOL_CHANDRASHEKAR)
        default: DG_BATCH_LAUNCH(VOL_COLLOCATION)
    }
}

}  // namespace dg
