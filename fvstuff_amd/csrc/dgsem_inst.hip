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
