// Per-degree instantiation unit: compiled once per p with -DDG_P=<p>.
#include "dgsem_launch.hpp"

#ifndef DG_P
#error "compile with -DDG_P=<degree>"
#endif

namespace dg {

// Limiter-only kernel: E elements per workgroup, state staged in LDS.
template <int P>
__global__ __launch_bounds__(BLOCK) void limit_kernel(double* Q, const double* w64, int NEL, int E, double g,
                                                      double eps_rho, double eps_p) {
    constexpr int N = P + 1, N2 = N * N;
    extern __shared__ __align__(16) unsigned char smem[];
    double* sU = (double*)smem;
    double* sEl = sU + (size_t)E * 4 * N2;
    const int el0 = blockIdx.x * E, nE = min(E, NEL - el0);
    const size_t S = (size_t)NEL * N2;
    for (int s = threadIdx.x; s < nE * N2; s += blockDim.x) {
        const int e = s / N2, nd = s - e * N2;
        const size_t o = (size_t)(el0 + e) * N2 + nd;
        for (int q = 0; q < 4; q++) sU[(size_t)e * 4 * N2 + q * N2 + nd] = Q[q * S + o];
    }
    limit_block<N>(sU, w64, nE, g, eps_rho, eps_p, sEl, sEl + 4 * E, (unsigned long long*)(sEl + 5 * E));
    for (int s = threadIdx.x; s < nE * N2; s += blockDim.x) {
        const int e = s / N2, nd = s - e * N2;
        const size_t o = (size_t)(el0 + e) * N2 + nd;
        for (int q = 0; q < 4; q++) Q[q * S + o] = sU[(size_t)e * 4 * N2 + q * N2 + nd];
    }
}

// Diagnostic per-element shock sensor, compute_sensor (:662-689): the
// divergence is not clipped to compressions here, unlike inside the RHS.
template <int P>
__global__ __launch_bounds__(BLOCK) void sensor_kernel(const double* Q, double* sensor, const double* D, int K,
                                                       double h, double g, double ksensor) {
    constexpr int N = P + 1, N2 = N * N;
    const int NEL = K * K;
    const size_t S = (size_t)NEL * N2;
    const double jac = 2.0 / h, len_p = h / (2 * P + 1), dx2inv = 1.0 / (len_p * len_p);
    for (int el = blockIdx.x * blockDim.x + threadIdx.x; el < NEL; el += blockDim.x * gridDim.x) {
        double u[N2], v[N2], c2[N2];
        for (int n = 0; n < N2; n++) {
            double W[4];
            load_prim64(Q, S, (size_t)el * N2 + n, g, W);
            u[n] = W[1]; v[n] = W[2]; c2[n] = g * W[3] / fmax(W[0], 1e-12);
        }
        double th_e = 0.0;
        for (int jn = 0; jn < N; jn++)
            for (int in = 0; in < N; in++) {
                double du = 0.0, dv = 0.0;
                for (int k = 0; k < N; k++) { du += D[in * N + k] * u[jn * N + k]; dv += D[jn * N + k] * v[k * N + in]; }
                double divu = jac * (du + dv), du2 = divu * divu;
                th_e = fmax(th_e, du2 / (du2 + ksensor * c2[jn * N + in] * dx2inv + 1e-30));
            }
        sensor[el] = th_e;
    }
}

// Raise a kernel's dynamic-LDS limit once (not per launch: the launches of
// the fused loop may be under stream capture).  `done` is per instantiation.
template <typename K>
static hipError_t set_lds(K kern, size_t lds, size_t& done) {
    if (lds <= 65536 || lds <= done) return hipSuccess;
    hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e == hipSuccess) done = lds;
    return e;
}

#define DG_CAT2(a, b) a##b
#define DG_CAT(a, b) DG_CAT2(a, b)

template <typename Real, bool FULL, int VOL, bool MFMA = false>
static hipError_t launch_one(const KArgs& a, int grid, int block, size_t lds, hipStream_t s) {
    static size_t done = 0;
    hipError_t e = set_lds(stage_kernel<DG_P, Real, FULL, VOL, MFMA>, lds, done);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((stage_kernel<DG_P, Real, FULL, VOL, MFMA>), dim3(grid), dim3(block), lds, s, a);
    return hipGetLastError();
}

template <typename Real, bool FULL>
static hipError_t launch_vol(const KArgs& a, int grid, int block, size_t lds, hipStream_t s) {
    switch (a.vol) {
        case VOL_RANOCHA: return launch_one<Real, FULL, VOL_RANOCHA>(a, grid, block, lds, s);
        case VOL_CHANDRASHEKAR: return launch_one<Real, FULL, VOL_CHANDRASHEKAR>(a, grid, block, lds, s);
        default: return launch_one<Real, FULL, VOL_COLLOCATION>(a, grid, block, lds, s);
    }
}

hipError_t DG_CAT(launch_stage_p, DG_P)(const KArgs& a, bool fp32, bool full, bool mfma, int grid, int block,
                                        size_t lds, hipStream_t s) {
#if DG_P == 3
    // the FP32 lean collocation kernel with its volume term on the matrix cores
    if (mfma && fp32 && !full && a.vol == VOL_COLLOCATION)
        return launch_one<float, false, VOL_COLLOCATION, true>(a, grid, block, lds, s);
    // the FP32 shock-capturing kernels with the AV on the matrix cores
    if (mfma && fp32 && full && a.vol != VOL_SD) {
        switch (a.vol) {
            case VOL_RANOCHA: return launch_one<float, true, VOL_RANOCHA, true>(a, grid, block, lds, s);
            case VOL_CHANDRASHEKAR: return launch_one<float, true, VOL_CHANDRASHEKAR, true>(a, grid, block, lds, s);
            default: return launch_one<float, true, VOL_COLLOCATION, true>(a, grid, block, lds, s);
        }
    }
#else
    if (mfma) return hipErrorInvalidValue;
#endif
    // the spectral-difference sibling has one (lean) kernel per precision
    if (a.vol == VOL_SD)
        return fp32 ? launch_one<float, false, VOL_SD>(a, grid, block, lds, s)
                    : launch_one<double, false, VOL_SD>(a, grid, block, lds, s);
    if (fp32) return full ? launch_vol<float, true>(a, grid, block, lds, s) : launch_vol<float, false>(a, grid, block, lds, s);
    return full ? launch_vol<double, true>(a, grid, block, lds, s) : launch_vol<double, false>(a, grid, block, lds, s);
}

// Batched stage of one degree: tiles[] lists (instance, tile) pairs of the
// instances of degree DG_P, an instance's tiles contiguous; one workgroup per
// entry, in the XCD-aware order when xcd is set (each XCD takes a contiguous
// range of the list, so an instance's tiles and the face nodes they read of
// each other share an L2).  The body is the single-solve lean kernel's, with
// its schedule and register budget for this degree alone (the all-degree
// batched kernel carries the largest degree's), so every instance gets the
// bits its own dg_advance would give.  The list and the argument blocks are
// read through the constant address space: the kernel stores to global
// memory, so the compiler cannot prove that plain global loads of them see
// no store of the launch, and it loads the fields with vector loads into
// VGPRs (p = 7 FP32: 168 VGPRs and 132 bytes of spills per lane); constant
// loads are scalar and the body gets the single-solve kernel's registers
// (164 VGPRs, no spills).  Both arrays are written only by the host.
// (The host pass compiles the kernel as a stub and has no address spaces.)
#ifdef __HIP_DEVICE_COMPILE__
typedef __attribute__((address_space(4))) const KArgs ConstKArgs;
typedef __attribute__((address_space(4))) const int2 ConstInt2;
#else
typedef const KArgs ConstKArgs;
typedef const int2 ConstInt2;
#endif
// (The degree is a template argument: every object file instantiates this
// kernel, and instantiations that differ only in DG_P would share one
// symbol, of which the link keeps one.)
// FULL: the shock-capturing body (AV and the Zhang-Shu limiter), as the
// single-solve stage_kernel<P, Real, true, VOL> runs it.
template <int P, typename Real, bool FULL, int VOL>
__global__ __launch_bounds__(BLOCK, (Occ<P, Real, VOL, FULL>::waves)) void batch_degree_kernel(
    const KArgs* __restrict__ args, const int2* __restrict__ tiles, int ntiles, int stage, StepCtl sc, int xcd) {
    const int t = tile_of_block(blockIdx.x, gridDim.x, xcd);
    if (t >= ntiles) return;
    const int2 tl = ((ConstInt2*)tiles)[t];
    const KArgs a = ((ConstKArgs*)args)[(size_t)tl.x * 3 + (stage - 1)];
    stage_tile<P, Real, FULL, VOL, (sizeof(Real) == 8), true, !FULL, true, false>(a, sc, tl.y, tl.y == 0,
                                                                                    (int)threadIdx.x);
}

template <typename Real, bool FULL, int VOL>
static hipError_t launch_degree_one(const KArgs* args, const int2* tiles, int ntiles, int stage, StepCtl sc,
                                    int block, size_t lds, int xcd, hipStream_t s) {
    static size_t done = 0;
    hipError_t e = set_lds(batch_degree_kernel<DG_P, Real, FULL, VOL>, lds, done);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((batch_degree_kernel<DG_P, Real, FULL, VOL>), dim3(ntiles), dim3(block), lds, s, args,
                       tiles, ntiles, stage, sc, xcd);
    return hipGetLastError();
}

template <typename Real, bool FULL>
static hipError_t launch_degree_vol(const KArgs* args, const int2* tiles, int ntiles, int stage, StepCtl sc,
                                    int vol, int block, size_t lds, int xcd, hipStream_t s) {
    switch (vol) {
        case VOL_RANOCHA:
            return launch_degree_one<Real, FULL, VOL_RANOCHA>(args, tiles, ntiles, stage, sc, block, lds, xcd, s);
        case VOL_CHANDRASHEKAR:
            return launch_degree_one<Real, FULL, VOL_CHANDRASHEKAR>(args, tiles, ntiles, stage, sc, block, lds, xcd, s);
        case VOL_SD:
            // the spectral-difference sibling has a lean kernel only
            if constexpr (!FULL)
                return launch_degree_one<Real, false, VOL_SD>(args, tiles, ntiles, stage, sc, block, lds, xcd, s);
            else
                return hipErrorInvalidValue;
        default:
            return launch_degree_one<Real, FULL, VOL_COLLOCATION>(args, tiles, ntiles, stage, sc, block, lds, xcd, s);
    }
}

hipError_t DG_CAT(launch_batch_degree_p, DG_P)(const KArgs* args, const int2* tiles, int ntiles, int stage,
                                               StepCtl sc, bool fp32, bool full, int vol, int block, size_t lds,
                                               int xcd, hipStream_t s) {
    if (ntiles < 1) return hipSuccess;
    if (fp32)
        return full ? launch_degree_vol<float, true>(args, tiles, ntiles, stage, sc, vol, block, lds, xcd, s)
                    : launch_degree_vol<float, false>(args, tiles, ntiles, stage, sc, vol, block, lds, xcd, s);
    return full ? launch_degree_vol<double, true>(args, tiles, ntiles, stage, sc, vol, block, lds, xcd, s)
                : launch_degree_vol<double, false>(args, tiles, ntiles, stage, sc, vol, block, lds, xcd, s);
}

hipError_t DG_CAT(launch_limit_p, DG_P)(double* Q, const double* w64, int NEL, int E, double g, double eps_rho,
                                        double eps_p, hipStream_t s) {
    constexpr int N2 = (DG_P + 1) * (DG_P + 1);
    size_t lds = ((size_t)E * 4 * N2 + 6 * E + 4) * sizeof(double);
    static size_t done = 0;
    hipError_t e = set_lds(limit_kernel<DG_P>, lds, done);
    if (e != hipSuccess) return e;
    int grid = (NEL + E - 1) / E;
    hipLaunchKernelGGL((limit_kernel<DG_P>), dim3(grid), dim3(BLOCK), lds, s, Q, w64, NEL, E, g, eps_rho, eps_p);
    return hipGetLastError();
}

hipError_t DG_CAT(launch_sensor_p, DG_P)(const double* Q, double* sensor, const double* D64, int K, double h,
                                         double g, double ksensor, hipStream_t s) {
    int NEL = K * K;
    int grid = (NEL + 127) / 128;
    if (grid > 65536) grid = 65536;
    hipLaunchKernelGGL((sensor_kernel<DG_P>), dim3(grid), dim3(128), 0, s, Q, sensor, D64, K, h, g, ksensor);
    return hipGetLastError();
}

}  // namespace dg
