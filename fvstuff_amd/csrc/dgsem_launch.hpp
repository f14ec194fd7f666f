// Host launchers for the per-degree kernel instantiations (one object file per p).
#pragma once
#include <hip/hip_runtime.h>
#include "dgsem_kernels.hpp"

namespace dg {

// Launches stage_kernel<P, Real, FULL> (mfma: its p = 3 FP32 form with the AV
// of the shock-capturing kernels, or the lean collocation kernel's volume
// term, on the matrix cores).  Returns the launch's hipError_t.
typedef hipError_t (*StageLauncher)(const KArgs& a, bool fp32, bool full, bool mfma, int grid, int block,
                                    size_t lds, hipStream_t s);
// Limiter on E-element tiles in place.
typedef hipError_t (*LimitLauncher)(double* Q, const double* w64, int NEL, int E, double g, double eps_rho,
                                    double eps_p, hipStream_t s);
// Per-element shock sensor.
typedef hipError_t (*SensorLauncher)(const double* Q, double* sensor, const double* D64, int K, double h,
                                     double g, double ksensor, hipStream_t s);

// Batched stage of one degree: the tiles of every instance of that degree in
// one launch of `ntiles` workgroups of `block` threads, with the single-solve
// kernel's body and schedule (full: the shock-capturing body); xcd:
// XCD-aware block-to-tile order.
typedef hipError_t (*BatchDegreeLauncher)(const KArgs* args, const int2* tiles, int ntiles, int stage,
                                          StepCtl sc, bool fp32, bool full, int vol, int block, size_t lds,
                                          int xcd, hipStream_t s);

#define DG_DECLARE_P(P)                                                                              \
    hipError_t launch_stage_p##P(const KArgs&, bool, bool, bool, int, int, size_t, hipStream_t);    \
    hipError_t launch_limit_p##P(double*, const double*, int, int, double, double, double, hipStream_t); \
    hipError_t launch_sensor_p##P(const double*, double*, const double*, int, double, double, double, hipStream_t); \
    hipError_t launch_batch_degree_p##P(const KArgs*, const int2*, int, int, StepCtl, bool, bool, int, int, size_t, \
                                        int, hipStream_t);
DG_DECLARE_P(1)
DG_DECLARE_P(2)
DG_DECLARE_P(3)
DG_DECLARE_P(4)
DG_DECLARE_P(5)
DG_DECLARE_P(6)
DG_DECLARE_P(7)
#undef DG_DECLARE_P

// Batched stage launchers (dgsem_batch.hip, one object per precision): one
// workgroup of `lanes` threads per tile (grid = 0), or the persistent form on
// `grid` workgroups.
hipError_t launch_batch_stage_float(const KArgs* args, const int2* tiles, int ntiles, int stage, StepCtl sc,
                                    int vol, int lanes, size_t lds, int grid, hipStream_t s);
hipError_t launch_batch_stage_double(const KArgs* args, const int2* tiles, int ntiles, int stage, StepCtl sc,
                                     int vol, int lanes, size_t lds, int grid, hipStream_t s);
// Resident workgroups per CU of the persistent batched kernel.
hipError_t batch_persistent_per_cu_float(int vol, int lanes, size_t lds, int* per_cu);
hipError_t batch_persistent_per_cu_double(int vol, int lanes, size_t lds, int* per_cu);

constexpr int PMAX_GPU = 7;
constexpr int BATCH_BLOCK = 64;   // default lanes per workgroup of the batched stage kernel

}  // namespace dg
