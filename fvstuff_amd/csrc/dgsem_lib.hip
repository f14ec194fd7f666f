// C ABI of the MI355X DGSEM step (include/dgsem.h) and the small kernels that
// are not per-degree: IC, unfused RK stage, wave-speed reductions, totals.
#include <hip/hip_runtime.h>
#include <math.h>
#include <cmath>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
#include <algorithm>
#include <mutex>
#include <vector>

#include "../../include/dgsem.h"
#include "dgsem_launch.hpp"

using namespace dg;

struct dg_ctx {
    dg_config cfg;
    hipStream_t stream;
    int N, N2, NEL;
    double h;
    double x[16], w[16], D[256];        // host FP64 operator tables
    double xf[17], Isf[17 * 16], Dfs[16 * 17];   // spectral-difference tables (NF = N + 1)
    void* d_tab64 = nullptr;             // D, w, winv in FP64
    void* d_tab32 = nullptr;             // D, w, winv in FP32
    double* d_U1 = nullptr;              // stage buffers of the fused stepping
    double* d_U2 = nullptr;
    DevState* d_st = nullptr;
    unsigned long long* d_red = nullptr; // reduction slot for dg_cfl_dt
    double* d_part = nullptr;            // per-block partials of dg_totals / dg_error_norms
    double* d_xi = nullptr;              // LGL node coordinates (IC kernel)
    int part_blocks = 0;
    int E = 0, xcd = 0;
    bool mfma = false;                   // contraction on the matrix cores (p = 3, FP32; AV or collocation)
    int last_steps = 0;                  // steps taken by the last dg_advance
    // graph replay of the fused loop: two steps per replay, captured on `cap`
    hipStream_t cap = nullptr;
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    hipGraphExec_t gexec = nullptr;
    struct { const double* Q; double cfl, t_end; int E, xcd; bool mfma; } gkey = {nullptr, 0, 0, 0, 0, false};
    bool use_graph = true;
    char err[512] = {0};
};

// ── error plumbing ─────────────────────────────────────────────────────────
static int fail(dg_ctx* c, int code, const char* fmt, ...) {
    if (c) {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(c->err, sizeof(c->err), fmt, ap);
        va_end(ap);
    }
    return code;
}
// The stage kernels move pairs of adjacent nodes with 16-byte accesses: the
// state they read must be 16-byte aligned (hipMalloc'd memory is).
static bool aligned16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

#define HIPCK(c, x)                                                                         \
    do {                                                                                    \
        hipError_t _e = (x);                                                                \
        if (_e != hipSuccess)                                                               \
            return fail((c), DG_ERR_HIP, "%s:%d: %s", __FILE__, __LINE__, hipGetErrorString(_e)); \
    } while (0)

// ── operators: LGL nodes / GLL weights and D (replaces :145-179, :924-940) ──
// p <= 3 use the reference's literal tables (dgsem_lobatto_2d.cu:145-157);
// higher p: x_j are +-1 and the zeros of L_{p+1} - L_{p-1} (Newton in long
// double from Chebyshev-Lobatto guesses), w_j = 2/(p(p+1) L_p(x_j)^2).
static void legendre_pair(int n, long double x, long double* L, long double* dL) {
    long double a = 1, b = x, da = 0, db = 1;
    if (n == 0) { *L = 1; *dL = 0; return; }
    for (int k = 2; k <= n; k++) {
        long double c = ((2 * k - 1) * x * b - (k - 1) * a) / k;
        long double dc = da + (2 * k - 1) * b;
        a = b; b = c; da = db; db = dc;
    }
    *L = b; *dL = db;
}

static void lgl_nodes(int p, double* x, double* w) {
    static const double tx[3][4] = {{-1.0, 1.0}, {-1.0, 0.0, 1.0},
                                    {-1.0, -0.4472135954999579, 0.4472135954999579, 1.0}};
    static const double tw[3][4] = {{1.0, 1.0},
                                    {0.33333333333333333, 1.33333333333333333, 0.33333333333333333},
                                    {0.16666666666666667, 0.83333333333333333, 0.83333333333333333,
                                     0.16666666666666667}};
    if (p <= 3) {
        for (int i = 0; i <= p; i++) { x[i] = tx[p - 1][i]; w[i] = tw[p - 1][i]; }
        return;
    }
    long double xs[16];
    xs[0] = -1; xs[p] = 1;
    for (int j = 1; j < p; j++) {
        long double t = -cosl(3.14159265358979323846264338327950288L * j / p);
        for (int it = 0; it < 100; it++) {
            long double A, dA, B, dB;
            legendre_pair(p + 1, t, &A, &dA);
            legendre_pair(p - 1, t, &B, &dB);
            long double step = (A - B) / (dA - dB);
            t -= step;
            if (fabsl(step) < 1e-19L) break;
        }
        xs[j] = t;
    }
    for (int j = 0; j <= p; j++) {
        long double L, dL;
        legendre_pair(p, xs[j], &L, &dL);
        x[j] = (double)xs[j];
        w[j] = (double)(2.0L / ((long double)p * (p + 1) * L * L));
    }
    for (int j = 0; j <= p / 2; j++) {   // enforce the exact mirror symmetry
        double a = 0.5 * (x[p - j] - x[j]), b = 0.5 * (w[j] + w[p - j]);
        x[j] = -a; x[p - j] = a; w[j] = b; w[p - j] = b;
    }
    if (p % 2 == 0) x[p / 2] = 0.0;
}

// Barycentric D_ik = l_k'(x_i) with the negative-sum diagonal (:160-179).
static void diff_matrix(int n, const double* x, double* D) {
    double b[16];
    for (int k = 0; k < n; k++) {
        b[k] = 1.0;
        for (int m = 0; m < n; m++) if (m != k) b[k] /= (x[k] - x[m]);
    }
    for (int i = 0; i < n; i++) {
        double diag = 0.0;
        for (int k = 0; k < n; k++) {
            if (k == i) continue;
            D[i * n + k] = (b[k] / b[i]) / (x[i] - x[k]);
            diag -= D[i * n + k];
        }
        D[i * n + i] = diag;
    }
}

// ── spectral-difference operators (fr_lobatto_2d.cu:76-166, :840-867) ──────
// Flux nodes {-1, Gauss-Legendre nodes, +1}: the reference's literal GL
// nodes for p <= 3 (:84-89), else the roots of L_p by Newton in long double
// from Chebyshev guesses.  Isf[k][j] = l_j(xf_k) (solution-node Lagrange
// basis at the flux nodes, :99-118).  Dfs[i][k] = l^f_k'(x_i), the flux-node
// basis differentiated at the solution nodes (what :17-24 defines), written
// as sum_{m != k} 1/(xf_k - xf_m) prod_{l != k,m} (x_i - xf_l)/(xf_k - xf_l),
// which has no 1/(x_i - xf_m) term: the end solution nodes coincide with
// flux nodes, where the reference's construction (:120-166) divides by zero.
static void sd_flux_nodes(int p, double* xf) {
    static const double gl[3][3] = {{0.0}, {-0.5773502691896258, 0.5773502691896258},
                                    {-0.7745966692414834, 0.0, 0.7745966692414834}};
    xf[0] = -1.0;
    xf[p + 1] = 1.0;
    if (p <= 3) {
        for (int j = 0; j < p; j++) xf[1 + j] = gl[p - 1][j];
        return;
    }
    for (int j = 0; j < p; j++) {
        long double t = -cosl(3.14159265358979323846264338327950288L * (j + 0.75L) / (p + 0.5L));
        for (int it = 0; it < 100; it++) {
            long double L, dL;
            legendre_pair(p, t, &L, &dL);
            long double step = L / dL;
            t -= step;
            if (fabsl(step) < 1e-19L) break;
        }
        xf[1 + j] = (double)t;
    }
    for (int j = 0; j < p / 2; j++) {   // exact mirror symmetry
        double a = 0.5 * (xf[p - j] - xf[1 + j]);
        xf[1 + j] = -a; xf[p - j] = a;
    }
    if (p % 2 == 1) xf[1 + p / 2] = 0.0;
}

static void sd_tables(int N, const double* xs, const double* xf, double* Isf, double* Dfs) {
    const int NF = N + 1;
    for (int k = 0; k < NF; k++)
        for (int j = 0; j < N; j++) {
            long double v = 1;
            for (int m = 0; m < N; m++)
                if (m != j) v *= ((long double)xf[k] - xs[m]) / ((long double)xs[j] - xs[m]);
            Isf[k * N + j] = (double)v;
        }
    for (int i = 0; i < N; i++)
        for (int k = 0; k < NF; k++) {
            long double d = 0;
            for (int m = 0; m < NF; m++) {
                if (m == k) continue;
                long double t = 1 / ((long double)xf[k] - xf[m]);
                for (int l = 0; l < NF; l++)
                    if (l != k && l != m) t *= ((long double)xs[i] - xf[l]) / ((long double)xf[k] - xf[l]);
                d += t;
            }
            Dfs[i * NF + k] = (double)d;
        }
}

// ── small kernels ──────────────────────────────────────────────────────────
struct ICArgs {
    int kind, K, N;
    double x0, y0, L, h, gamma;
    double rho_hi, p_hi, eps, xc, yc, u0, v0;
    double sod_width;     // circular Sod front width in units of h
    unsigned long long seed;
    double amp;
    const double* xi;
};

__device__ __forceinline__ unsigned long long splitmix64(unsigned long long z) {
    // SplitMix64 finaliser (Steele, Lea, Flood 2014)
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// Isentropic vortex (q1_dg_euler_2d.cu:2166-2188) at (x, y): strength eps,
// centre (xc, yc), background (u0, v0).  images = 0: the single vortex (the
// IC); images = 1: the velocity and temperature perturbations summed over the
// 3x3 periodic images of the centre (period L), a periodic field.  Evaluated
// without FMA contraction, like the IC.
__device__ __forceinline__ void vortex_state(double x, double y, const ICArgs& a, int images, double& r,
                                             double& u, double& v, double& p) {
#pragma clang fp contract(off)
    const double g = a.gamma, pi = 3.14159265358979323846;
    double T;
    if (!images) {
        const double dx = x - a.xc, dy = y - a.yc, r2 = dx * dx + dy * dy;
        const double f = a.eps / (2.0 * pi) * exp(0.5 * (1.0 - r2));
        u = a.u0 - f * dy;
        v = a.v0 + f * dx;
        T = 1.0 + (-(g - 1.0) * a.eps * a.eps / (8.0 * g * pi * pi) * exp(1.0 - r2));
    } else {
        double du = 0.0, dv = 0.0, dT = 0.0;
        for (int iy = -1; iy <= 1; iy++)
            for (int ix = -1; ix <= 1; ix++) {
                const double dx = x - (a.xc + ix * a.L), dy = y - (a.yc + iy * a.L), r2 = dx * dx + dy * dy;
                const double f = a.eps / (2.0 * pi) * exp(0.5 * (1.0 - r2));
                du += -f * dy;
                dv += f * dx;
                dT += -(g - 1.0) * a.eps * a.eps / (8.0 * g * pi * pi) * exp(1.0 - r2);
            }
        u = a.u0 + du;
        v = a.v0 + dv;
        T = 1.0 + dT;
    }
    if (T < 1e-6) T = 1e-6;
    r = pow(T, 1.0 / (g - 1.0));
    p = pow(T, g / (g - 1.0));
}

// IC at the LGL nodes (ic_kernel :692-717; vortex q1_dg_euler_2d.cu:2166-2188;
// KH step layer with counter-based noise keyed by (seed, node)).  Coordinates
// and the branch tests are evaluated without FMA contraction so that a node on
// an element face lands on the same side of a discontinuity as on the host.
__global__ void ic_kernel(double* Q, ICArgs a) {
#pragma clang fp contract(off)
    const int N = a.N, N2 = N * N, NEL = a.K * a.K;
    const size_t S = (size_t)NEL * N2;
    const double g = a.gamma, pi = 3.14159265358979323846;
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < S; c += (size_t)blockDim.x * gridDim.x) {
        const int el = (int)(c / N2), nd = (int)(c - (size_t)el * N2);
        const int ej = el / a.K, ei = el - ej * a.K, jn = nd / N, in = nd - jn * N;
        const double x = a.x0 + a.h * ((double)ei + 0.5 * (a.xi[in] + 1.0));
        const double y = a.y0 + a.h * ((double)ej + 0.5 * (a.xi[jn] + 1.0));
        double r, u, v, p;
        if (a.kind == DG_IC_CIRCULAR_SOD) {
            const double cx = a.x0 + 0.5 * a.L, cy = a.y0 + 0.5 * a.L;
            const double rr = sqrt((x - cx) * (x - cx) + (y - cy) * (y - cy));
            const double phi = 0.5 * (1.0 + tanh((0.25 * a.L - rr) / (a.sod_width * a.h)));
            r = 0.125 + (a.rho_hi - 0.125) * phi; u = 0.0; v = 0.0; p = 0.1 + (a.p_hi - 0.1) * phi;
        } else if (a.kind == DG_IC_VORTEX) {
            vortex_state(x, y, a, 0, r, u, v, p);
        } else {
            const unsigned long long z = splitmix64(a.seed + ((unsigned long long)c + 1ull) * 0x9e3779b97f4a7c15ull);
            const double uni = (double)(z >> 11) * (1.0 / 9007199254740992.0);
            const bool inner = fabs(y) < 0.5;
            r = inner ? 2.0 : 1.0;
            u = inner ? 0.5 : -0.5;
            v = 0.1 * sin(2.0 * pi * x) + a.amp * (2.0 * uni - 1.0);
            p = 2.5;
        }
        Q[c] = r;
        Q[S + c] = r * u;
        Q[2 * S + c] = r * v;
        Q[3 * S + c] = p / (g - 1.0) + 0.5 * r * (u * u + v * v);
    }
}

// Unfused Shu-Osher stage + sanitize_cons (rk3_s1/s2/s3, :533-576).
__global__ void rk3_kernel(double* Uo, const double* U0, const double* Up, const double* L, double dt, int stage,
                           size_t S, double g) {
    for (size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x; n < S; n += (size_t)blockDim.x * gridDim.x) {
        double Uc[4];
        for (int q = 0; q < 4; q++) {
            const size_t i = q * S + n;
            if (stage == 1) Uc[q] = U0[i] + dt * L[i];
            else if (stage == 2) Uc[q] = 0.75 * U0[i] + 0.25 * (Up[i] + dt * L[i]);
            else Uc[q] = (1.0 / 3.0) * U0[i] + (2.0 / 3.0) * (Up[i] + dt * L[i]);
        }
        Cons64::sanitize_cons(Uc, g);
        for (int q = 0; q < 4; q++) Uo[q * S + n] = Uc[q];
    }
}

// Max wave speed of a state, for the first step of the fused stepping
// (into the partial slots `parts`, as the stage-3 kernels do).
__global__ void lambda_kernel(const double* Q, size_t S, double g, int sd, unsigned long long* parts) {
    __shared__ double sred[BLOCK / 64];
    double m = 0.0;
    for (size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x; n < S; n += (size_t)blockDim.x * gridDim.x) {
        double W[4], ri;
        const double U[4] = {Q[n], Q[S + n], Q[2 * S + n], Q[3 * S + n]};
        Cons64::to_prim(U, W, g, ri);
        m = nanmax(m, sd ? Cons64::wave_speed_sd(W, g, ri) : Cons64::wave_speed(W, g, ri));
    }
    block_max_atomic(m, &parts[blockIdx.x & (LAM_PARTS - 1)], sred);
}

// lam_bits[s] = max of slot s's partial maxima (host read-back after stepping)
__device__ __forceinline__ void fold_lambda(DevState* st) {
    for (int s = 0; s < 2; s++) {
        unsigned long long m = 0ull;
        for (int k = 0; k < LAM_PARTS; k++) m = st->lam_part[s][k] > m ? st->lam_part[s][k] : m;
        st->lam_bits[s] = m;
    }
}

__global__ void fold_lambda_kernel(DevState* st) { fold_lambda(st); }

// max of an array (reduce_max/gpu_max, :646-659) into one bit-pattern slot
__global__ void reduce_max_kernel(const double* v, int n, unsigned long long* dst) {
    __shared__ double sred[BLOCK / 64];
    double m = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += blockDim.x * gridDim.x) m = nanmax(m, v[i]);
    block_max_atomic(m, dst, sred);
}

__device__ __forceinline__ void clear_lambda(DevState* st) {
    st->lam_bits[0] = 0ull;
    st->lam_bits[1] = 0ull;
    for (int k = 0; k < LAM_PARTS; k++) st->lam_part[0][k] = st->lam_part[1][k] = 0ull;
}

__global__ void state_reset_kernel(DevState* st, double t0) {
    clear_lambda(st);
    st->t = t0;
    st->dt = 0.0;
    st->status = ST_OK;
    st->skip = 0;
    st->steps = 0;
}

__global__ void slot_reset_kernel(unsigned long long* slot) { *slot = 0ull; }

// nodal positivity repair of the SD sibling, in place (fr_lobatto_2d.cu:589-615)
__global__ void repair_kernel(double* Q, size_t S, double g, double eps_rho, double eps_p) {
    const Gas64 gs{g, g - 1.0, 1e-14 / (g - 1.0)};
    for (size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x; n < S; n += (size_t)blockDim.x * gridDim.x) {
        double U[4] = {Q[n], Q[S + n], Q[2 * S + n], Q[3 * S + n]};
        Cons64::repair(U, gs, eps_rho, eps_p);
        for (int q = 0; q < 4; q++) Q[q * S + n] = U[q];
    }
}

// GLL weights of one degree, passed by value (no device table needed)
struct GllW {
    double w[16];
};

// GLL-weighted integrals of rho, rho u, rho v, E and -rho s/(g-1); one
// partial per block into part[block*5 + q] (totals_finalize_kernel sums them).
__global__ void totals_kernel(const double* Q, GllW w, int N, size_t S, double jw, double g, double* part) {
    __shared__ double sred[5][BLOCK / 64];
    double acc[5] = {0, 0, 0, 0, 0};
    const int N2 = N * N;
    for (size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x; n < S; n += (size_t)blockDim.x * gridDim.x) {
        const int nd = (int)(n % N2), jn = nd / N, in = nd - jn * N;
        const double wij = jw * w.w[in] * w.w[jn];
        double U[4] = {Q[n], Q[S + n], Q[2 * S + n], Q[3 * S + n]};
        for (int q = 0; q < 4; q++) acc[q] += wij * U[q];
        const double p = (g - 1.0) * (U[3] - 0.5 * (U[1] * U[1] + U[2] * U[2]) / U[0]);
        const double s = log(p) - g * log(U[0]);
        acc[4] += wij * (-U[0] * s / (g - 1.0));
    }
    for (int q = 0; q < 5; q++) {
        double v = acc[q];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        acc[q] = v;
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0)
        for (int q = 0; q < 5; q++) sred[q][wid] = acc[q];
    __syncthreads();
    if (threadIdx.x == 0)
        for (int q = 0; q < 5; q++) {
            double s = 0.0;
            for (int i = 0; i < (int)(blockDim.x >> 6); i++) s += sred[q][i];
            part[(size_t)blockIdx.x * 5 + q] = s;
        }
}

// Sum of the nb per-block partials of totals_kernel into out[0..4], on one
// workgroup of BLOCK threads in a fixed order (thread t takes blocks t,
// t + BLOCK, ...; then fixed shuffle and wave orders): deterministic, and no
// host round trip.  Reads only the nb partials the last totals launch wrote.
__global__ void totals_finalize_kernel(const double* part, int nb, double* out) {
    __shared__ double sred[5][BLOCK / 64];
    double acc[5] = {0, 0, 0, 0, 0};
    for (int b = threadIdx.x; b < nb; b += blockDim.x)
        for (int q = 0; q < 5; q++) acc[q] += part[(size_t)b * 5 + q];
    for (int q = 0; q < 5; q++) {
        double v = acc[q];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        acc[q] = v;
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0)
        for (int q = 0; q < 5; q++) sred[q][wid] = acc[q];
    __syncthreads();
    if (threadIdx.x < 5) {
        double s = 0.0;
        for (int i = 0; i < (int)(blockDim.x >> 6); i++) s += sred[threadIdx.x][i];
        out[threadIdx.x] = s;
    }
}

// Exact vortex advected for time t by the background velocity: the vortex
// field at the node coordinates shifted by -(u0, v0) t and wrapped into the
// periodic domain [x0, x0+L) x [y0, y0+L) (the vortex is a steady solution
// in the frame moving with the background).
__global__ void vortex_exact_kernel(double* Q, ICArgs a, double t, int images) {
#pragma clang fp contract(off)
    const int N = a.N, N2 = N * N, NEL = a.K * a.K;
    const size_t S = (size_t)NEL * N2;
    const double g = a.gamma;
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < S; c += (size_t)blockDim.x * gridDim.x) {
        const int el = (int)(c / N2), nd = (int)(c - (size_t)el * N2);
        const int ej = el / a.K, ei = el - ej * a.K, jn = nd / N, in = nd - jn * N;
        double x = a.x0 + a.h * ((double)ei + 0.5 * (a.xi[in] + 1.0));
        double y = a.y0 + a.h * ((double)ej + 0.5 * (a.xi[jn] + 1.0));
        if (t != 0.0) {
            x -= a.u0 * t;
            y -= a.v0 * t;
            x = a.x0 + (x - a.x0) - a.L * floor((x - a.x0) / a.L);
            y = a.y0 + (y - a.y0) - a.L * floor((y - a.y0) / a.L);
        }
        double r, u, v, p;
        vortex_state(x, y, a, images, r, u, v, p);
        Q[c] = r;
        Q[S + c] = r * u;
        Q[2 * S + c] = r * v;
        Q[3 * S + c] = p / (g - 1.0) + 0.5 * r * (u * u + v * v);
    }
}

// Per-block partials of the GLL-weighted norms of Q - R: sums of w d_q^2 (4),
// max |d_q| (4), sum of w R_q^2 over q, max |R_q| over q.
constexpr int NORM_PARTS = 10;
__global__ void error_norms_kernel(const double* Q, const double* R, const double* w, int N, size_t S, double jw,
                                   double* part) {
    __shared__ double sred[NORM_PARTS][BLOCK / 64];
    double acc[NORM_PARTS] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    const int N2 = N * N;
    for (size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x; n < S; n += (size_t)blockDim.x * gridDim.x) {
        const int nd = (int)(n % N2), jn = nd / N, in = nd - jn * N;
        const double wij = jw * w[in] * w[jn];
        for (int q = 0; q < 4; q++) {
            const double r = R[q * S + n], d = Q[q * S + n] - r;
            acc[q] += wij * d * d;
            acc[4 + q] = nanmax(acc[4 + q], fabs(d));
            acc[8] += wij * r * r;
            acc[9] = nanmax(acc[9], fabs(r));
        }
    }
    for (int k = 0; k < NORM_PARTS; k++) {
        double v = acc[k];
        for (int o = 32; o > 0; o >>= 1) {
            const double x = __shfl_xor(v, o, 64);
            v = (k < 4 || k == 8) ? v + x : nanmax(v, x);
        }
        acc[k] = v;
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0)
        for (int k = 0; k < NORM_PARTS; k++) sred[k][wid] = acc[k];
    __syncthreads();
    if (threadIdx.x == 0)
        for (int k = 0; k < NORM_PARTS; k++) {
            double v = sred[k][0];
            for (int i = 1; i < (int)(blockDim.x >> 6); i++)
                v = (k < 4 || k == 8) ? v + sred[k][i] : nanmax(v, sred[k][i]);
            part[(size_t)blockIdx.x * NORM_PARTS + k] = v;
        }
}

// Element averages of rho and p for frame output (downsample_rho_p,
// :826-844): plain node means, rho floored at 1e-14, negative pressures
// counted as 0; nodes summed in storage order.  One thread per element.
__global__ void frame_fields_kernel(const double* Q, int NEL, int N2, double g, double* rho_el, double* p_el) {
    const size_t S = (size_t)NEL * N2;
    for (int el = blockIdx.x * blockDim.x + threadIdx.x; el < NEL; el += blockDim.x * gridDim.x) {
        double rs = 0.0, ps = 0.0;
        for (int n = 0; n < N2; n++) {
            const size_t i = (size_t)el * N2 + n;
            double rho = Q[i];
            const double mx = Q[S + i], my = Q[2 * S + i], E = Q[3 * S + i];
            if (rho < 1e-14) rho = 1e-14;
            const double pr = (g - 1.0) * (E - 0.5 * (mx * mx + my * my) / rho);
            rs += rho;
            ps += pr > 0.0 ? pr : 0.0;
        }
        rho_el[el] = rs / N2;
        p_el[el] = ps / N2;
    }
}

// ── dispatch tables ───────────────────────────────────────────────────────
static StageLauncher stage_launcher(int p) {
    static const StageLauncher t[] = {nullptr, launch_stage_p1, launch_stage_p2, launch_stage_p3,
                                      launch_stage_p4, launch_stage_p5, launch_stage_p6, launch_stage_p7};
    return t[p];
}
static LimitLauncher limit_launcher(int p) {
    static const LimitLauncher t[] = {nullptr, launch_limit_p1, launch_limit_p2, launch_limit_p3,
                                      launch_limit_p4, launch_limit_p5, launch_limit_p6, launch_limit_p7};
    return t[p];
}
static BatchDegreeLauncher batch_degree_launcher(int p) {
    static const BatchDegreeLauncher t[] = {nullptr, launch_batch_degree_p1, launch_batch_degree_p2,
                                            launch_batch_degree_p3, launch_batch_degree_p4, launch_batch_degree_p5,
                                            launch_batch_degree_p6, launch_batch_degree_p7};
    return t[p];
}
static SensorLauncher sensor_launcher(int p) {
    static const SensorLauncher t[] = {nullptr, launch_sensor_p1, launch_sensor_p2, launch_sensor_p3,
                                       launch_sensor_p4, launch_sensor_p5, launch_sensor_p6, launch_sensor_p7};
    return t[p];
}

static int grid_for(size_t n, int block) {
    size_t g = (n + block - 1) / block;
    return (int)std::min<size_t>(std::max<size_t>(g, 1), 8192);
}

static int max_elems_per_block(int N) { return BLOCK / (2 * N); }
// Batched launches: the most elements whose lines fit in a workgroup of
// `lanes` threads (every degree's tile then fills about the same workgroup).
static int batch_elems_per_block(int N, int lanes) { return std::max(1, lanes / (2 * N)); }

// Default schedule: the largest tile that still gives >= 2048 workgroups
// (8 per CU) and at least one full wave of line threads.
static int default_E(int N, int NEL) {
    const int Emax = max_elems_per_block(N);
    const int Emin = std::min(Emax, (64 + 2 * N - 1) / (2 * N));
    int E = Emax;
    while (E > Emin && (NEL + E - 1) / E < 2048) E = std::max(Emin, E / 2);
    return E;
}

// The MFMA form of the AV exists for the p = 3 FP32 shock-capturing kernels.
// The matrix-core forms: the AV operator of the p = 3 mixed shock-capturing
// kernels, and the volume term of the p = 3 mixed collocation kernel.
static bool mfma_capable(const dg_config& f) {
    return f.p == 3 && f.precision == DG_MIXED && f.scheme == DG_SCHEME_DGSEM &&
           (f.av || f.volume_flux == DG_VOL_COLLOCATION);
}

// ── stage launch ───────────────────────────────────────────────────────────
struct StageCall {
    int mode, stage;
    const double *Up, *U0;
    double* out;
    double* lam_el;
    double ca, cb, cfl, t_end;
    int check_blowup, slot_cur, slot_next;
};

static int launch_stage(dg_ctx* c, const StageCall& s, int E, int xcd, hipStream_t stream, int mfma = -1) {
    const dg_config& f = c->cfg;
    const bool fp32 = f.precision == DG_MIXED;
    const bool sd = f.scheme == DG_SCHEME_SD;
    const bool lam = s.mode == MODE_RHS && s.lam_el != nullptr;
    const bool lim = s.mode == MODE_STAGE && f.limiter && !sd;   // Zhang-Shu scratch
    const bool full = !sd && (f.av || lim || lam);
    const bool mf = (mfma < 0 ? c->mfma : mfma != 0) && mfma_capable(f);
    // the SD line reads the FP64 stage input from LDS in both modes
    LdsLayout L = lds_layout(c->N, E, fp32 ? 4 : 8, s.mode == MODE_STAGE || sd, f.av != 0, lam, lim, mf);
    KArgs a;
    memset(&a, 0, sizeof(a));
    a.Up = s.Up; a.U0 = s.U0; a.out = s.out; a.lam_el = s.lam_el;
    a.tab = fp32 ? c->d_tab32 : c->d_tab64;
    a.w64 = (const double*)c->d_tab64 + c->N2;
    a.st = c->d_st;
    a.P = f.p; a.K = f.K; a.NEL = c->NEL; a.E = E; a.xcd = xcd;
    a.mode = s.mode; a.stage = s.stage; a.vol = sd ? VOL_SD : f.volume_flux; a.surf = f.surface_flux; a.bc = f.bc;
    a.av = f.av; a.limiter = f.limiter; a.check_blowup = s.check_blowup;
    a.slot_cur = s.slot_cur; a.slot_next = s.slot_next;
    a.h = c->h; a.gamma = f.gamma; a.ca = s.ca; a.cb = s.cb; a.cfl = s.cfl; a.t_end = s.t_end;
    a.av_cav = f.av_cav; a.av_k = f.av_ksensor; a.eps_rho = f.eps_rho; a.eps_p = f.eps_p;
    a.off_U = L.off_U; a.off_av = L.off_av; a.off_lam = L.off_lam; a.off_el = L.off_el;
    set_derived(a);
    const int grid = (c->NEL + E - 1) / E;
    const int block = std::min(BLOCK, (E * 2 * c->N + 63) / 64 * 64);
    HIPCK(c, stage_launcher(f.p)(a, fp32, full, mf, grid, block, (size_t)L.total, stream));
    return DG_OK;
}

// ── batched-instance helpers ──────────────────────────────────────────────
// args[3*i] is instance i's stage-1 argument block (its state pointer,
// sizes and device step state).
__global__ void batch_reset_kernel(const KArgs* args, int n, const double* t0) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    DevState* st = args[3 * i].st;
    clear_lambda(st);
    st->t = t0 ? t0[i] : 0.0;
    st->dt = 0.0;
    st->status = ST_OK;
    st->skip = 0;
    st->steps = 0;
}

// wave speed of each instance's state, one tile of E elements per visit
__global__ void batch_lambda_kernel(const KArgs* args, const int2* tiles, int ntiles) {
    __shared__ double sred[BLOCK / 64];
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int2 tl = tiles[t];
        if (tl.x < 0) continue;   // padding of the XCD-binned list
        const KArgs& a = args[(size_t)tl.x * 3];
        const int N2 = (a.P + 1) * (a.P + 1);
        const size_t S = (size_t)a.NEL * N2;
        const int el0 = tl.y * a.E, nE = min(a.E, a.NEL - el0);
        double m = 0.0;
        for (int s = threadIdx.x; s < nE * N2; s += blockDim.x) {
            double W[4];
            load_prim64(a.Up, S, (size_t)el0 * N2 + s, a.gamma, W);
            m = nanmax(m, Cons64::wave_speed(W, a.gamma));
        }
        __syncthreads();   // sred reuse across tiles
        block_max_atomic(m, &a.st->lam_part[0][blockIdx.x & (LAM_PARTS - 1)], sred);
    }
}

__global__ void batch_gather_kernel(const KArgs* args, int n, DevState* out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        fold_lambda(args[3 * i].st);
        out[i] = *args[3 * i].st;
    }
}

// ── exported API ───────────────────────────────────────────────────────────
extern "C" {

void dg_config_default(dg_config* f) {
    if (!f) return;
    memset(f, 0, sizeof(*f));
    f->K = 50; f->p = 3; f->x0 = 0.0; f->y0 = 0.0; f->L = 1.0; f->gamma = 1.4;
    f->volume_flux = DG_VOL_CHANDRASHEKAR; f->surface_flux = DG_SURF_HLLC; f->bc = DG_BC_WALL;
    f->av = 1; f->limiter = 1; f->precision = DG_FP64;
    f->av_cav = 0.25; f->av_ksensor = 0.1; f->eps_rho = 1e-6; f->eps_p = 1e-6;
    f->elems_per_block = 0; f->xcd_remap = -1;
    f->scheme = DG_SCHEME_DGSEM;
    f->mfma = -1;
}

void dg_ic_default(dg_ic_params* ic, int kind) {
    if (!ic) return;
    memset(ic, 0, sizeof(*ic));
    ic->kind = kind;
    ic->rho_hi = 11.0; ic->p_hi = 10.0;
    ic->vortex_eps = 5.0; ic->xc = 5.0; ic->yc = 5.0; ic->u0 = 1.0; ic->v0 = 1.0;
    ic->seed = 1234; ic->noise_amp = 1e-3;
}

const char* dg_status_string(int s) {
    switch (s) {
        case DG_OK: return "ok";
        case DG_ERR_INVALID_ARG: return "invalid argument";
        case DG_ERR_HIP: return "HIP runtime error";
        case DG_ERR_NONFINITE: return "non-finite or non-positive max wave speed";
        case DG_ERR_DT_UNDERFLOW: return "dt underflow";
        case DG_ERR_BLOWUP: return "blow-up (max wave speed > 1e6)";
        default: return "unknown status";
    }
}

const char* dg_last_error(const dg_ctx* c) { return c ? c->err : "null context"; }

long long dg_state_size(const dg_ctx* c) { return c ? 4LL * c->NEL * c->N2 : 0; }
int dg_steps_taken(const dg_ctx* c) { return c ? c->last_steps : 0; }
long long dg_num_elements(const dg_ctx* c) { return c ? c->NEL : 0; }

int dg_operators(const dg_ctx* c, double* x, double* w, double* D) {
    if (!c) return DG_ERR_INVALID_ARG;
    for (int i = 0; i < c->N; i++) {
        if (x) x[i] = c->x[i];
        if (w) w[i] = c->w[i];
    }
    if (D) memcpy(D, c->D, sizeof(double) * c->N2);
    return DG_OK;
}

int dg_sd_operators(const dg_ctx* c, double* xf, double* Isf, double* Dfs) {
    if (!c) return DG_ERR_INVALID_ARG;
    if (c->cfg.scheme != DG_SCHEME_SD) return DG_ERR_INVALID_ARG;
    const int NF = c->N + 1;
    if (xf) memcpy(xf, c->xf, sizeof(double) * NF);
    if (Isf) memcpy(Isf, c->Isf, sizeof(double) * NF * c->N);
    if (Dfs) memcpy(Dfs, c->Dfs, sizeof(double) * c->N * NF);
    return DG_OK;
}

int dg_destroy(dg_ctx* c) {
    if (!c) return DG_OK;
    hipFree(c->d_tab64); hipFree(c->d_tab32); hipFree(c->d_U1); hipFree(c->d_U2);
    hipFree(c->d_st); hipFree(c->d_red); hipFree(c->d_part); hipFree(c->d_xi);
    if (c->gexec) hipGraphExecDestroy(c->gexec);
    if (c->cap) hipStreamDestroy(c->cap);
    if (c->ev_in) hipEventDestroy(c->ev_in);
    if (c->ev_out) hipEventDestroy(c->ev_out);
    delete c;
    return DG_OK;
}

int dg_create(const dg_config* cfg, void* stream, dg_ctx** out) {
    if (!cfg || !out) return DG_ERR_INVALID_ARG;
    *out = nullptr;
    if (cfg->K < 1 || cfg->p < 1 || cfg->p > PMAX_GPU || !(cfg->L > 0) || !(cfg->gamma > 1.0) ||
        cfg->volume_flux < 0 || cfg->volume_flux > 2 || cfg->surface_flux < 0 || cfg->surface_flux > 1 ||
        cfg->bc < 0 || cfg->bc > 1 || cfg->precision < 0 || cfg->precision > 1 ||
        cfg->scheme < 0 || cfg->scheme > 1 || (cfg->scheme == DG_SCHEME_SD && cfg->av) ||
        cfg->mfma < -1 || cfg->mfma > 1 ||
        (long long)cfg->K * cfg->K > (1LL << 30) / ((cfg->p + 1) * (cfg->p + 1)))
        return DG_ERR_INVALID_ARG;
    dg_ctx* c = new dg_ctx();
    c->cfg = *cfg;
    c->stream = (hipStream_t)stream;
    c->N = cfg->p + 1; c->N2 = c->N * c->N; c->NEL = cfg->K * cfg->K;
    c->h = cfg->L / cfg->K;
    lgl_nodes(cfg->p, c->x, c->w);
    diff_matrix(c->N, c->x, c->D);
    sd_flux_nodes(cfg->p, c->xf);
    sd_tables(c->N, c->x, c->xf, c->Isf, c->Dfs);

    const int Emax = max_elems_per_block(c->N);
    c->E = cfg->elems_per_block > 0 ? std::min(cfg->elems_per_block, Emax) : default_E(c->N, c->NEL);
    c->xcd = cfg->xcd_remap >= 0 ? (cfg->xcd_remap != 0) : 1;
    c->mfma = cfg->mfma > 0 && mfma_capable(*cfg);

    // tables: D[N2], w[N], winv[N], Dv[N2], Lav[N2], Isf[(N+1)N], Dsd[N(N+1)].
    // Dv is the volume matrix, D scaled by the volume term's factor -2/h
    // (collocation) or -4/h (the split form's 2 D F#), so that the line loop
    // needs no scaling pass; Lav = W^-1 D^T W D is the artificial-viscosity
    // operator of one line (the reference's two passes, D then W^-1 D^T W, in
    // one matrix); Isf and Dsd = -(2/h) Dfs are the spectral-difference
    // sibling's interpolation and scaled flux-derivative matrices
    const int NF = c->N + 1;
    const size_t off_sd = 3 * c->N2 + 2 * c->N;
    std::vector<double> t64(off_sd + 2 * NF * c->N);
    std::vector<float> t32(t64.size());
    const double vscale = (cfg->volume_flux == DG_VOL_COLLOCATION ? -1.0 : -2.0) * (2.0 / c->h);
    for (int i = 0; i < c->N2; i++) t64[i] = c->D[i];
    for (int i = 0; i < c->N; i++) { t64[c->N2 + i] = c->w[i]; t64[c->N2 + c->N + i] = 1.0 / c->w[i]; }
    for (int i = 0; i < c->N2; i++) t64[c->N2 + 2 * c->N + i] = vscale * c->D[i];
    for (int i = 0; i < c->N; i++)
        for (int m = 0; m < c->N; m++) {
            double sum = 0.0;
            for (int k = 0; k < c->N; k++) sum += c->w[k] * c->D[k * c->N + i] * c->D[k * c->N + m];
            t64[2 * c->N2 + 2 * c->N + i * c->N + m] = sum / c->w[i];
        }
    for (int i = 0; i < NF * c->N; i++) {
        t64[off_sd + i] = c->Isf[i];
        t64[off_sd + NF * c->N + i] = -(2.0 / c->h) * c->Dfs[i];
    }
    for (size_t i = 0; i < t64.size(); i++) t32[i] = (float)t64[i];

    const size_t S = (size_t)c->NEL * c->N2;
    c->part_blocks = grid_for(S, BLOCK);
#define CK_CREATE(x)                                                                           \
    do {                                                                                       \
        hipError_t _e = (x);                                                                   \
        if (_e != hipSuccess) { dg_destroy(c); return DG_ERR_HIP; }                           \
    } while (0)
    CK_CREATE(hipMalloc(&c->d_tab64, t64.size() * sizeof(double)));
    CK_CREATE(hipMalloc(&c->d_tab32, t32.size() * sizeof(float)));
    CK_CREATE(hipMalloc(&c->d_U1, 4 * S * sizeof(double)));
    CK_CREATE(hipMalloc(&c->d_U2, 4 * S * sizeof(double)));
    CK_CREATE(hipMalloc(&c->d_st, sizeof(DevState)));
    CK_CREATE(hipMalloc(&c->d_red, sizeof(unsigned long long)));
    CK_CREATE(hipMalloc(&c->d_part, (size_t)c->part_blocks * NORM_PARTS * sizeof(double)));
    CK_CREATE(hipMalloc(&c->d_xi, 16 * sizeof(double)));
    CK_CREATE(hipMemcpyAsync(c->d_xi, c->x, c->N * sizeof(double), hipMemcpyHostToDevice, c->stream));
    CK_CREATE(hipMemcpyAsync(c->d_tab64, t64.data(), t64.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    CK_CREATE(hipMemcpyAsync(c->d_tab32, t32.data(), t32.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
    CK_CREATE(hipStreamCreateWithFlags(&c->cap, hipStreamNonBlocking));
    CK_CREATE(hipEventCreateWithFlags(&c->ev_in, hipEventDisableTiming));
    CK_CREATE(hipEventCreateWithFlags(&c->ev_out, hipEventDisableTiming));
    CK_CREATE(hipStreamSynchronize(c->stream));
#undef CK_CREATE
    *out = c;
    return DG_OK;
}

static ICArgs ic_args(const dg_ctx* c, const dg_ic_params* ic) {
    ICArgs a;
    a.kind = ic->kind; a.K = c->cfg.K; a.N = c->N;
    a.x0 = c->cfg.x0; a.y0 = c->cfg.y0; a.L = c->cfg.L; a.h = c->h; a.gamma = c->cfg.gamma;
    a.rho_hi = ic->rho_hi; a.p_hi = ic->p_hi; a.eps = ic->vortex_eps; a.xc = ic->xc; a.yc = ic->yc;
    a.u0 = ic->u0; a.v0 = ic->v0; a.seed = ic->seed; a.amp = ic->noise_amp;
    // front width: 0.5h (dgsem_lobatto_2d.cu:710), 2h (fr_lobatto_2d.cu:648)
    a.sod_width = c->cfg.scheme == DG_SCHEME_SD ? 2.0 : 0.5;
    a.xi = c->d_xi;
    return a;
}

int dg_init_state(dg_ctx* c, const dg_ic_params* ic, double* Q) {
    if (!c || !ic || !Q) return fail(c, DG_ERR_INVALID_ARG, "dg_init_state: null argument");
    if (ic->kind < 0 || ic->kind > 2) return fail(c, DG_ERR_INVALID_ARG, "dg_init_state: unknown IC %d", ic->kind);
    const ICArgs a = ic_args(c, ic);
    const size_t S = (size_t)c->NEL * c->N2;
    hipLaunchKernelGGL(ic_kernel, dim3(grid_for(S, BLOCK)), dim3(BLOCK), 0, c->stream, Q, a);
    HIPCK(c, hipGetLastError());
    // the SD driver applies no limiter to its IC (fr_lobatto_2d.cu:922-937)
    if (c->cfg.limiter && c->cfg.scheme != DG_SCHEME_SD) return dg_limit(c, Q);
    return DG_OK;
}

int dg_limit(dg_ctx* c, double* Q) {
    if (!c || !Q) return fail(c, DG_ERR_INVALID_ARG, "dg_limit: null argument");
    if (c->cfg.scheme == DG_SCHEME_SD) {
        const size_t S = (size_t)c->NEL * c->N2;
        hipLaunchKernelGGL(repair_kernel, dim3(grid_for(S, BLOCK)), dim3(BLOCK), 0, c->stream, Q, S, c->cfg.gamma,
                           c->cfg.eps_rho, c->cfg.eps_p);
        HIPCK(c, hipGetLastError());
        return DG_OK;
    }
    const int E = std::max(1, max_elems_per_block(c->N) / 2);
    HIPCK(c, limit_launcher(c->cfg.p)(Q, (const double*)c->d_tab64 + c->N2, c->NEL, E, c->cfg.gamma,
                                       c->cfg.eps_rho, c->cfg.eps_p, c->stream));
    return DG_OK;
}

int dg_compute_rhs(dg_ctx* c, const double* Q, double* RHS, double* lam_el) {
    if (!c || !Q || !RHS) return fail(c, DG_ERR_INVALID_ARG, "dg_compute_rhs: null argument");
    if (!aligned16(Q)) return fail(c, DG_ERR_INVALID_ARG, "dg_compute_rhs: Q is not 16-byte aligned");
    StageCall s{};
    s.mode = MODE_RHS; s.stage = 0; s.Up = Q; s.U0 = Q; s.out = RHS; s.lam_el = lam_el;
    return launch_stage(c, s, c->E, c->xcd, c->stream);
}

int dg_rk3_stage(dg_ctx* c, int stage, double* Uout, const double* U0, const double* Uprev, const double* L,
                 double dt) {
    if (!c || !Uout || !U0 || !L || (stage != 1 && !Uprev) || stage < 1 || stage > 3)
        return fail(c, DG_ERR_INVALID_ARG, "dg_rk3_stage: bad argument");
    const size_t S = (size_t)c->NEL * c->N2;
    hipLaunchKernelGGL(rk3_kernel, dim3(grid_for(S, BLOCK)), dim3(BLOCK), 0, c->stream, Uout, U0,
                       stage == 1 ? U0 : Uprev, L, dt, stage, S, c->cfg.gamma);
    HIPCK(c, hipGetLastError());
    if (c->cfg.limiter) return dg_limit(c, Uout);
    return DG_OK;
}

int dg_cfl_dt(dg_ctx* c, const double* lam_el, double cfl, double* lam_max, double* dt) {
    if (!c || !lam_el) return fail(c, DG_ERR_INVALID_ARG, "dg_cfl_dt: null argument");
    hipLaunchKernelGGL(slot_reset_kernel, dim3(1), dim3(1), 0, c->stream, c->d_red);
    hipLaunchKernelGGL(reduce_max_kernel, dim3(grid_for(c->NEL, BLOCK)), dim3(BLOCK), 0, c->stream, lam_el,
                       c->NEL, c->d_red);
    HIPCK(c, hipGetLastError());
    unsigned long long bits = 0;
    HIPCK(c, hipMemcpyAsync(&bits, c->d_red, sizeof(bits), hipMemcpyDeviceToHost, c->stream));
    HIPCK(c, hipStreamSynchronize(c->stream));
    double m;
    memcpy(&m, &bits, sizeof(m));
    if (lam_max) *lam_max = m;
    if (!(m > 0.0)) return fail(c, DG_ERR_NONFINITE, "max wave speed %g", m);
    if (dt) *dt = cfl * c->h / (m * c->N);
    return DG_OK;
}

// Launch the three fused stages of one step; `n` is the step index in the call.
static int advance_step(dg_ctx* c, double* Q, int n, double cfl, double t_end, int E, int xcd, hipStream_t st) {
    const int cur = n & 1, nxt = cur ^ 1;
    StageCall s{};
    s.mode = MODE_STAGE; s.cfl = cfl; s.t_end = t_end; s.slot_cur = cur; s.slot_next = nxt;
    s.stage = 1; s.Up = Q; s.U0 = Q; s.out = c->d_U1; s.ca = 0.0; s.cb = 1.0; s.check_blowup = n > 0;
    int rc = launch_stage(c, s, E, xcd, st);
    if (rc) return rc;
    s.stage = 2; s.Up = c->d_U1; s.U0 = Q; s.out = c->d_U2; s.ca = 0.75; s.cb = 0.25;
    if ((rc = launch_stage(c, s, E, xcd, st))) return rc;
    s.stage = 3; s.Up = c->d_U2; s.U0 = Q; s.out = Q; s.ca = 1.0 / 3.0; s.cb = 2.0 / 3.0;
    return launch_stage(c, s, E, xcd, st);
}

static int advance_impl(dg_ctx* c, double* Q, int nsteps, double cfl, double t_end, double* t_inout,
                        double* dt_last, int E, int xcd, bool graph) {
    const size_t S = (size_t)c->NEL * c->N2;
    const double t0 = t_inout ? *t_inout : 0.0;
    hipLaunchKernelGGL(state_reset_kernel, dim3(1), dim3(1), 0, c->stream, c->d_st, t0);
    hipLaunchKernelGGL(lambda_kernel, dim3(grid_for(S, BLOCK)), dim3(BLOCK), 0, c->stream, Q, S, c->cfg.gamma,
                       (int)(c->cfg.scheme == DG_SCHEME_SD), c->d_st->lam_part[0]);
    HIPCK(c, hipGetLastError());
    int rc = DG_OK, n = 0;
    if (nsteps > 0 && (rc = advance_step(c, Q, n++, cfl, t_end, E, xcd, c->stream))) return rc;
    if (graph && nsteps >= 3) {
        // steps 1, 2, ... replay a graph of two steps (odd slot parity first)
        const bool same = c->gexec && c->gkey.Q == Q && c->gkey.cfl == cfl && c->gkey.t_end == t_end &&
                          c->gkey.E == E && c->gkey.xcd == xcd && c->gkey.mfma == c->mfma;
        if (!same) {
            if (c->gexec) { hipGraphExecDestroy(c->gexec); c->gexec = nullptr; }
            hipGraph_t g;
            HIPCK(c, hipStreamBeginCapture(c->cap, hipStreamCaptureModeThreadLocal));
            int r1 = advance_step(c, Q, 1, cfl, t_end, E, xcd, c->cap);
            int r2 = r1 ? r1 : advance_step(c, Q, 2, cfl, t_end, E, xcd, c->cap);
            hipError_t ec = hipStreamEndCapture(c->cap, &g);
            if (r2) {
                if (ec == hipSuccess) hipGraphDestroy(g);   // a partial capture is never instantiated
                return r2;
            }
            HIPCK(c, ec);
            hipError_t ei = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
            hipGraphDestroy(g);
            HIPCK(c, ei);
            c->gkey = {Q, cfl, t_end, E, xcd, c->mfma};
        }
        HIPCK(c, hipEventRecord(c->ev_in, c->stream));
        HIPCK(c, hipStreamWaitEvent(c->cap, c->ev_in, 0));
        for (; n + 1 < nsteps; n += 2) HIPCK(c, hipGraphLaunch(c->gexec, c->cap));
        HIPCK(c, hipEventRecord(c->ev_out, c->cap));
        HIPCK(c, hipStreamWaitEvent(c->stream, c->ev_out, 0));
    }
    for (; n < nsteps; n++)
        if ((rc = advance_step(c, Q, n, cfl, t_end, E, xcd, c->stream))) return rc;
    hipLaunchKernelGGL(fold_lambda_kernel, dim3(1), dim3(1), 0, c->stream, c->d_st);
    HIPCK(c, hipGetLastError());
    DevState h;
    HIPCK(c, hipMemcpyAsync(&h, c->d_st, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    HIPCK(c, hipStreamSynchronize(c->stream));
    if (t_inout) *t_inout = h.t;
    if (dt_last) *dt_last = h.dt;
    c->last_steps = h.steps;
    if (h.status != ST_OK) return fail(c, h.status, "%s at t=%.9g", dg_status_string(h.status), h.t);
    if (h.steps > 0) {
        // blow-up check of the last step taken (:1052-1057).  The steps taken
        // are the first h.steps of the call (after t_end every step is a
        // no-op); step n writes its stage-3 wave speed into slot (n & 1) ^ 1,
        // so the last one, n = h.steps - 1, wrote slot h.steps & 1, and the
        // no-op steps after it leave that slot alone.  The next step's stage
        // 1 makes the same check, but a no-op step skips it.
        double lam;
        const unsigned long long b = h.lam_bits[h.steps & 1];
        memcpy(&lam, &b, sizeof(lam));
        if (!(lam > 0.0) || lam > 1e6) return fail(c, DG_ERR_BLOWUP, "blow-up: max wave speed %g at t=%.9g", lam, h.t);
    }
    return DG_OK;
}

int dg_advance(dg_ctx* c, double* Q, int nsteps, double cfl, double t_end, double* t_inout, double* dt_last) {
    if (!c || !Q || nsteps < 0 || !(cfl > 0)) return fail(c, DG_ERR_INVALID_ARG, "dg_advance: bad argument");
    if (!aligned16(Q)) return fail(c, DG_ERR_INVALID_ARG, "dg_advance: Q is not 16-byte aligned");
    return advance_impl(c, Q, nsteps, cfl, t_end, t_inout, dt_last, c->E, c->xcd, c->use_graph);
}

int dg_set_graph(dg_ctx* c, int use_graph) {
    if (!c) return DG_ERR_INVALID_ARG;
    c->use_graph = use_graph != 0;
    return DG_OK;
}

int dg_step(dg_ctx* c, double* Q, double cfl, double* dt_used) {
    return dg_advance(c, Q, 1, cfl, 0.0, nullptr, dt_used);
}

int dg_sensor(dg_ctx* c, const double* Q, double* sensor_el) {
    if (!c || !Q || !sensor_el) return fail(c, DG_ERR_INVALID_ARG, "dg_sensor: null argument");
    HIPCK(c, sensor_launcher(c->cfg.p)(Q, sensor_el, (const double*)c->d_tab64, c->cfg.K, c->h, c->cfg.gamma,
                                        c->cfg.av_ksensor, c->stream));
    return DG_OK;
}

// The totals of state Q on `stream`: totals_kernel's per-block partials in
// part[0 .. 5*blocks), their sum in out_dev[0..4] (device memory).
static hipError_t launch_totals(const double* Q, int p, int K, double L, double gamma, double* part,
                                double* out_dev, hipStream_t stream) {
    const int N = p + 1;
    const size_t S = (size_t)K * K * N * N;
    const double h = L / K;
    GllW w;
    double x[16];
    lgl_nodes(p, x, w.w);
    const int blocks = grid_for(S, BLOCK);
    hipLaunchKernelGGL(totals_kernel, dim3(blocks), dim3(BLOCK), 0, stream, Q, w, N, S, 0.25 * h * h, gamma, part);
    hipLaunchKernelGGL(totals_finalize_kernel, dim3(1), dim3(BLOCK), 0, stream, (const double*)part, blocks, out_dev);
    return hipGetLastError();
}

int dg_totals(dg_ctx* c, const double* Q, double out[5]) {
    if (!c || !Q || !out) return fail(c, DG_ERR_INVALID_ARG, "dg_totals: null argument");
    // partials in d_part[0, 5*part_blocks), the sums after them
    double* sums = c->d_part + (size_t)c->part_blocks * 5;
    HIPCK(c, launch_totals(Q, c->cfg.p, c->cfg.K, c->cfg.L, c->cfg.gamma, c->d_part, sums, c->stream));
    HIPCK(c, hipMemcpyAsync(out, sums, 5 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCK(c, hipStreamSynchronize(c->stream));
    return DG_OK;
}

// ── checkpoint / restart (no counterpart in the reference) ──────────────────
// File: a 128-byte header, then the FP64 state (4*K*K*(p+1)^2 doubles in the
// ABI's layout), both in the host's byte order.  The header records the
// discretisation the state belongs to, the time and the step count; loading
// requires the same discretisation as the context (the volume / surface
// flux, boundaries, shock capturing and precision may differ: they act on
// the state, they do not define it).
namespace {
struct CheckpointHeader {
    char magic[8];                 // "DGSEMCK" + NUL
    uint32_t version, header_bytes;   // 1, 128
    int32_t K, p, scheme, pad0;
    double x0, y0, L, gamma;
    double t;
    int64_t steps, ndoubles;
    char pad[128 - 8 - 8 - 16 - 40 - 16];
};
static_assert(sizeof(CheckpointHeader) == 128, "checkpoint header is 128 bytes");
const char kCheckpointMagic[8] = {'D', 'G', 'S', 'E', 'M', 'C', 'K', 0};
}  // namespace

int dg_save_state(dg_ctx* c, const double* Q, double t, long long steps, const char* path) {
    if (!c || !Q || !path) return fail(c, DG_ERR_INVALID_ARG, "dg_save_state: null argument");
    CheckpointHeader h;
    memset(&h, 0, sizeof(h));
    memcpy(h.magic, kCheckpointMagic, 8);
    h.version = 1;
    h.header_bytes = sizeof(h);
    h.K = c->cfg.K; h.p = c->cfg.p; h.scheme = c->cfg.scheme;
    h.x0 = c->cfg.x0; h.y0 = c->cfg.y0; h.L = c->cfg.L; h.gamma = c->cfg.gamma;
    h.t = t;
    h.steps = steps;
    h.ndoubles = dg_state_size(c);
    std::vector<double> buf((size_t)h.ndoubles);
    HIPCK(c, hipMemcpyAsync(buf.data(), Q, buf.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCK(c, hipStreamSynchronize(c->stream));
    FILE* f = fopen(path, "wb");
    if (!f) return fail(c, DG_ERR_INVALID_ARG, "dg_save_state: cannot open %s", path);
    const bool ok = fwrite(&h, sizeof(h), 1, f) == 1 && fwrite(buf.data(), sizeof(double), buf.size(), f) == buf.size();
    const bool closed = fclose(f) == 0;
    if (!ok || !closed) return fail(c, DG_ERR_INVALID_ARG, "dg_save_state: cannot write %s", path);
    return DG_OK;
}

int dg_load_state(dg_ctx* c, double* Q, double* t, long long* steps, const char* path) {
    if (!c || !Q || !path) return fail(c, DG_ERR_INVALID_ARG, "dg_load_state: null argument");
    FILE* f = fopen(path, "rb");
    if (!f) return fail(c, DG_ERR_INVALID_ARG, "dg_load_state: cannot open %s", path);
    CheckpointHeader h;
    std::vector<double> buf;
    const char* why = nullptr;
    if (fread(&h, sizeof(h), 1, f) != 1) why = "short header";
    else if (memcmp(h.magic, kCheckpointMagic, 8) || h.version != 1 || h.header_bytes != sizeof(h)) why = "not a checkpoint";
    else if (h.K != c->cfg.K || h.p != c->cfg.p || h.scheme != c->cfg.scheme || h.x0 != c->cfg.x0 ||
             h.y0 != c->cfg.y0 || h.L != c->cfg.L || h.gamma != c->cfg.gamma || h.ndoubles != dg_state_size(c))
        why = "another discretisation";
    else {
        buf.resize((size_t)h.ndoubles);
        if (fread(buf.data(), sizeof(double), buf.size(), f) != buf.size()) why = "short state";
        else if (fgetc(f) != EOF) why = "trailing bytes";
    }
    fclose(f);
    if (why) return fail(c, DG_ERR_INVALID_ARG, "dg_load_state: %s: %s", path, why);
    HIPCK(c, hipMemcpyAsync(Q, buf.data(), buf.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCK(c, hipStreamSynchronize(c->stream));   // the host buffer goes out of scope
    if (t) *t = h.t;
    if (steps) *steps = h.steps;
    return DG_OK;
}

// ── totals without allocation (dg_totals_noalloc) ──────────────────────────
// The per-block partials live in `scratch`, one fixed array per GPU (each GPU
// holds its own copy until the process ends), sized for the most blocks a
// totals launch uses (grid_for caps the grid at NOALLOC_BLOCKS).  In host code
// `scratch` names a host-side stand-in, not the device array: the device
// address comes from hipGetSymbolAddress, once per device with that device
// current, and only that address is passed on.  Every call on a GPU shares
// its array: calls that overlap on one GPU race.  Nothing in it has to start
// at zero (a call reads only the partials its own totals launch wrote).
constexpr int NOALLOC_BLOCKS = 8192;
constexpr int NOALLOC_DEVICES = 64;
static __device__ __align__(256) double scratch[NOALLOC_BLOCKS * 5];
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[NOALLOC_DEVICES] = {};

// device address of the current device's scratch array (looked up once)
static hipError_t scratch_ptr(double** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= NOALLOC_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[dev] = (double*)ptr;
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}

int dg_totals_noalloc(const dg_config* cfg, const double* Q, double* out_dev, void* stream) {
    if (!cfg || !Q || !out_dev) return DG_ERR_INVALID_ARG;
    if (cfg->K < 1 || cfg->p < 1 || cfg->p > PMAX_GPU || !(cfg->L > 0) || !(cfg->gamma > 1.0) ||
        (long long)cfg->K * cfg->K > (1LL << 30) / ((cfg->p + 1) * (cfg->p + 1)))
        return DG_ERR_INVALID_ARG;   // the same limits as dg_create: at most 2^30 nodes
    const size_t S = (size_t)cfg->K * cfg->K * (cfg->p + 1) * (cfg->p + 1);
    if (grid_for(S, BLOCK) > NOALLOC_BLOCKS) return DG_ERR_INVALID_ARG;   // more partials than `scratch` holds
    double* part = nullptr;
    if (scratch_ptr(&part) != hipSuccess) return DG_ERR_HIP;
    if (launch_totals(Q, cfg->p, cfg->K, cfg->L, cfg->gamma, part, out_dev, (hipStream_t)stream) != hipSuccess)
        return DG_ERR_HIP;
    return DG_OK;
}

int dg_vortex_exact(dg_ctx* c, const dg_ic_params* ic, double t, int images, double* Q) {
    if (!c || !ic || !Q) return fail(c, DG_ERR_INVALID_ARG, "dg_vortex_exact: null argument");
    if (ic->kind != DG_IC_VORTEX) return fail(c, DG_ERR_INVALID_ARG, "dg_vortex_exact: IC kind %d is not the vortex", ic->kind);
    if (!(t >= 0.0) || (images != 0 && images != 1)) return fail(c, DG_ERR_INVALID_ARG, "dg_vortex_exact: bad t or images");
    const ICArgs a = ic_args(c, ic);
    const size_t S = (size_t)c->NEL * c->N2;
    hipLaunchKernelGGL(vortex_exact_kernel, dim3(grid_for(S, BLOCK)), dim3(BLOCK), 0, c->stream, Q, a, t, images);
    HIPCK(c, hipGetLastError());
    return DG_OK;
}

int dg_error_norms(dg_ctx* c, const double* Q, const double* Qref, double out[10]) {
    if (!c || !Q || !Qref || !out) return fail(c, DG_ERR_INVALID_ARG, "dg_error_norms: null argument");
    const size_t S = (size_t)c->NEL * c->N2;
    hipLaunchKernelGGL(error_norms_kernel, dim3(c->part_blocks), dim3(BLOCK), 0, c->stream, Q, Qref,
                       (const double*)c->d_tab64 + c->N2, c->N, S, 0.25 * c->h * c->h, c->d_part);
    HIPCK(c, hipGetLastError());
    std::vector<double> part((size_t)c->part_blocks * NORM_PARTS);
    HIPCK(c, hipMemcpyAsync(part.data(), c->d_part, part.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCK(c, hipStreamSynchronize(c->stream));
    double tot[NORM_PARTS];
    for (int k = 0; k < NORM_PARTS; k++) {
        double v = part[k];
        for (int b = 1; b < c->part_blocks; b++) {
            const double x = part[(size_t)b * NORM_PARTS + k];
            if (k < 4 || k == 8) v += x;
            else v = (v != v || x != x) ? v + x : std::max(v, x);
        }
        tot[k] = v;
    }
    double sum = 0.0, mx = 0.0;
    for (int q = 0; q < 4; q++) {
        out[q] = std::sqrt(tot[q]);
        out[4 + q] = tot[4 + q];
        sum += tot[q];
        mx = (mx != mx || tot[4 + q] != tot[4 + q]) ? mx + tot[4 + q] : std::max(mx, tot[4 + q]);
    }
    out[8] = std::sqrt(sum) / std::sqrt(tot[8]);
    out[9] = mx / tot[9];
    return DG_OK;
}

int dg_frame_fields(dg_ctx* c, const double* Q, double* rho_el, double* p_el) {
    if (!c || !Q || !rho_el || !p_el) return fail(c, DG_ERR_INVALID_ARG, "dg_frame_fields: null argument");
    hipLaunchKernelGGL(frame_fields_kernel, dim3(grid_for(c->NEL, 256)), dim3(256), 0, c->stream, Q, c->NEL, c->N2,
                       c->cfg.gamma, rho_el, p_el);
    HIPCK(c, hipGetLastError());
    return DG_OK;
}

// Schedule tuner: run a few fused steps on the ctx's own stage buffers for
// each candidate (elements per block x XCD remap) and keep the fastest.
int dg_autotune(dg_ctx* c, int* elems_per_block, int* xcd_remap) {
    if (!c) return DG_ERR_INVALID_ARG;
    const size_t S = (size_t)c->NEL * c->N2;
    double* Q = nullptr;
    HIPCK(c, hipMalloc(&Q, 4 * S * sizeof(double)));
    dg_ic_params ic;
    dg_ic_default(&ic, DG_IC_VORTEX);
    ic.xc = c->cfg.x0 + 0.5 * c->cfg.L; ic.yc = c->cfg.y0 + 0.5 * c->cfg.L;
    int rc = dg_init_state(c, &ic, Q);
    const int Emax = max_elems_per_block(c->N);
    const int Emin = std::max(1, (64 + 2 * c->N - 1) / (2 * c->N) / 2);
    std::vector<int> Es;
    for (int E = Emax; E >= Emin; E /= 2) Es.push_back(E);
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    float best = 1e30f, bestXcd = 1e30f;      // fastest overall, fastest with the XCD order
    int bestE = c->E, bestX = c->xcd, xcdE = 0;
    bool bestM = c->mfma, xcdM = false;
    // the two forms of the AV contraction (VALU in the line phase, MFMA) when
    // the configuration leaves the choice to the tuner
    const int nm = (c->cfg.mfma < 0 && mfma_capable(c->cfg)) ? 2 : 1;
    // Every candidate runs from the same smooth state with a few small steps
    // (CFL 0.1 is stable for every degree), so numerics never interfere.  The
    // candidates are timed round-robin, four rounds of 8-step runs (24
    // launches each), and a candidate's time is its best round: schedules
    // 0.7-2 % apart (vortex 512^2 p = 7: 8 elements per workgroup against 16
    // or 4) are otherwise ranked by noise and clock drift, and one candidate
    // after another would see the drift as a difference.
    struct Cand { int m, x, E; float t; };
    std::vector<Cand> cands;
    for (int mi = 0; mi < nm; mi++)
        for (int x = 0; x < 2; x++)
            for (int E : Es) cands.push_back({mi, x, E, 1e30f});
    for (int round = -1; round < 4 && rc == DG_OK; round++)
        for (Cand& cd : cands) {
            if (nm == 2) c->mfma = cd.m != 0;
            double t = 0.0;
            if ((rc = dg_init_state(c, &ic, Q))) break;
            if (round < 0) {   // warm-up (LDS attribute, caches)
                if ((rc = advance_impl(c, Q, 1, 0.1, 0.0, &t, nullptr, cd.E, cd.x, false))) break;
                continue;
            }
            hipEventRecord(e0, c->stream);
            rc = advance_impl(c, Q, 8, 0.1, 0.0, &t, nullptr, cd.E, cd.x, false);
            hipEventRecord(e1, c->stream);
            hipEventSynchronize(e1);
            if (rc) break;
            float ms = 0.f;
            hipEventElapsedTime(&ms, e0, e1);
            cd.t = std::min(cd.t, ms);
        }
    for (const Cand& cd : cands) {
        const bool m = nm == 2 ? cd.m != 0 : c->mfma;
        if (cd.t < best) { best = cd.t; bestE = cd.E; bestX = cd.x; bestM = m; }
        if (cd.x == 1 && cd.t < bestXcd) { bestXcd = cd.t; xcdE = cd.E; xcdM = m; }
    }
    // The XCD order reads fewer bytes (a tile's face neighbours in its own
    // XCD's L2: vortex 512^2 p = 7, 1.004x the algorithmic traffic with it,
    // 1.05-1.19x without), so it is kept when it runs within 1 % of the
    // fastest schedule (the kernels are VALU-limited: the two orders then
    // differ by noise)
    if (bestX == 0 && xcdE > 0 && bestXcd <= 1.01f * best) { bestE = xcdE; bestX = 1; bestM = xcdM; }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipFree(Q);
    c->mfma = bestM;
    if (rc) return rc;
    c->E = bestE;
    c->xcd = bestX;
    if (elems_per_block) *elems_per_block = bestE;
    if (xcd_remap) *xcd_remap = bestX;
    return DG_OK;
}

// ── batched instances ─────────────────────────────────────────────────────
}  // extern "C"

struct dg_batch {
    int n = 0;
    hipStream_t stream = nullptr;
    hipStream_t cap = nullptr;          // capture / replay stream (default priority)
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    std::vector<dg_ctx*> ctx;
    dg_config common;
    std::vector<KArgs> h_args;          // 3 per instance
    KArgs* d_args = nullptr;
    int2* d_tiles = nullptr;
    DevState* d_gather = nullptr;
    double* d_t0 = nullptr;
    int ntiles = 0, lanes = BATCH_BLOCK, tiles_cap = 0;
    size_t lds = 0;
    bool use_graph = true;
    bool persistent = false;            // persistent form of the stage launches
    bool xcd_bins = false;              // tile list in per-XCD bins (batch_tiles); by degree: XCD-aware order
    int pgrid = 0;                      // its grid: resident workgroups (0 = unavailable)
    // by-degree form: one launch per degree present per stage (batch_tiles)
    struct DegreeRun {
        int p, first, count, block;     // degree, its range of the tile list, workgroup size
        size_t lds;
    };
    bool by_degree = false;
    std::vector<DegreeRun> runs;
    // AV and / or the limiter on (every instance alike): the shock-capturing
    // stage kernels, which run in the by-degree form only
    bool full = false;
    // spectral-difference instances (every instance alike): the SD stage
    // kernels, in the by-degree form only
    bool sd = false;
    int deg_lanes[PMAX_GPU + 1] = {0};  // by-degree form: lanes per workgroup of degree p (0 = lanes)
    hipGraphExec_t gexec = nullptr;
    char err[512] = {0};
};

static int bfail(dg_batch* b, int code, const char* fmt, ...) {
    if (b) {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(b->err, sizeof(b->err), fmt, ap);
        va_end(ap);
    }
    return code;
}
#define BHIPCK(b, x)                                                                          \
    do {                                                                                      \
        hipError_t _e = (x);                                                                  \
        if (_e != hipSuccess)                                                                 \
            return bfail((b), DG_ERR_HIP, "%s:%d: %s", __FILE__, __LINE__, hipGetErrorString(_e)); \
    } while (0)

static hipError_t batch_launch_stage(dg_batch* b, int stage, StepCtl sc) {
    const bool fp32 = b->common.precision == DG_MIXED;
    if (b->by_degree) {
        for (const dg_batch::DegreeRun& r : b->runs) {
            hipError_t e = batch_degree_launcher(r.p)(b->d_args, b->d_tiles + r.first, r.count, stage, sc, fp32,
                                                      b->full, b->sd ? (int)VOL_SD : b->common.volume_flux,
                                                      r.block, r.lds,
                                                      b->xcd_bins ? 1 : 0, b->cap);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    }
    const int grid = b->persistent ? b->pgrid : 0;
    return fp32 ? launch_batch_stage_float(b->d_args, b->d_tiles, b->ntiles, stage, sc, b->common.volume_flux,
                                           b->lanes, b->lds, grid, b->cap)
                : launch_batch_stage_double(b->d_args, b->d_tiles, b->ntiles, stage, sc, b->common.volume_flux,
                                            b->lanes, b->lds, grid, b->cap);
}

static hipError_t batch_step(dg_batch* b, StepCtl sc) {
    hipError_t e;
    for (int stage = 1; stage <= 3; stage++)
        if ((e = batch_launch_stage(b, stage, sc)) != hipSuccess) return e;
    return hipSuccess;
}

// lanes per workgroup of an instance of degree p: its own in the by-degree
// form when set, else the common size (the all-degree kernel has one size)
static int batch_lanes_of(const dg_batch* b, int p) {
    return b->by_degree && b->deg_lanes[p] ? b->deg_lanes[p] : b->lanes;
}

// whether the batched kernels keep the stage input in LDS: the all-degree
// kernel by precision, the by-degree form always (the single-solve schedule)
// Batches that only the by-degree form runs: shock capturing, or the SD
// scheme (the all-degree kernel carries neither).
static bool batch_degree_only(const dg_batch* b) { return b->full || b->sd; }

static bool batch_up_lds(const dg_batch* b) {
    if (b->by_degree) return true;
    return b->common.precision == DG_MIXED ? BatchOcc<float>::up_lds : BatchOcc<double>::up_lds;
}

// Fill the three per-stage argument blocks of every instance for states Qs.
static void batch_bind(dg_batch* b, double* const* Qs, const double* cfl, double t_end) {
    for (int i = 0; i < b->n; i++) {
        dg_ctx* c = b->ctx[i];
        const dg_config& f = c->cfg;
        const bool fp32 = f.precision == DG_MIXED;
        const int E = batch_elems_per_block(c->N, batch_lanes_of(b, c->cfg.p));
        // (the SD kernel's nodal repair needs no limiter scratch)
        LdsLayout L = lds_layout(c->N, E, fp32 ? 4 : 8, batch_up_lds(b), f.av != 0, false, !b->sd && f.limiter != 0);
        for (int s = 1; s <= 3; s++) {
            KArgs a;
            memset(&a, 0, sizeof(a));
            a.Up = s == 1 ? Qs[i] : (s == 2 ? c->d_U1 : c->d_U2);
            a.U0 = Qs[i];
            a.out = s == 1 ? c->d_U1 : (s == 2 ? c->d_U2 : Qs[i]);
            a.lam_el = nullptr;
            a.tab = fp32 ? c->d_tab32 : c->d_tab64;
            a.w64 = (const double*)c->d_tab64 + c->N2;
            a.st = c->d_st;
            a.P = f.p; a.K = f.K; a.NEL = c->NEL; a.E = E; a.xcd = 0;
            a.mode = MODE_STAGE; a.stage = s; a.vol = b->sd ? (int)VOL_SD : f.volume_flux;
            a.surf = f.surface_flux; a.bc = f.bc;
            a.av = f.av; a.limiter = f.limiter;
            a.h = c->h; a.gamma = f.gamma;
            a.ca = s == 1 ? 0.0 : (s == 2 ? 0.75 : 1.0 / 3.0);
            a.cb = s == 1 ? 1.0 : (s == 2 ? 0.25 : 2.0 / 3.0);
            a.cfl = cfl[i]; a.t_end = t_end;
            a.av_cav = f.av_cav; a.av_k = f.av_ksensor; a.eps_rho = f.eps_rho; a.eps_p = f.eps_p;
            a.off_U = L.off_U; a.off_av = L.off_av; a.off_lam = L.off_lam; a.off_el = L.off_el;
            set_derived(a);
            b->h_args[(size_t)i * 3 + (s - 1)] = a;
        }
    }
}

// (Re)build the tile list for the current lanes per workgroup: instance i
// is cut into tiles of batch_elems_per_block(N_i, lanes) elements, cost of a
// tile ~ E (p+1)^3 (pairs per line x lines); the LDS size is the largest
// instance's.  Two orders:
//  - global (xcd_bins off): every tile, heaviest first (the longest tasks
//    start first, the short ones fill the tail);
//  - XCD bins: the instances bin-packed over the 8 XCDs by cost (largest
//    first, each into the least loaded bin), each bin's tiles heaviest first
//    with an instance's tiles contiguous, and the bins interleaved (position
//    8k + x holds bin x's k-th tile, shorter bins padded with empty entries,
//    instance -1).  The hardware deals workgroups round-robin over the XCDs,
//    so XCD x runs bin x in order: the tiles of an instance, which read each
//    other's face nodes, share an L2 and run close in time.
// The by-degree form lists the instances of each degree together, highest
// degree first, each instance's tiles contiguous and in order; each degree's
// range is one launch (b->runs).
static hipError_t batch_tiles(dg_batch* b) {
    std::vector<int2> tiles;
    size_t lds = 0;
    std::vector<double> tile_cost(b->n);
    std::vector<int> ntl(b->n);
    const bool fp32 = b->common.precision == DG_MIXED;
    for (int i = 0; i < b->n; i++) {
        dg_ctx* c = b->ctx[i];
        const int E = batch_elems_per_block(c->N, batch_lanes_of(b, c->cfg.p));
        ntl[i] = (c->NEL + E - 1) / E;
        tile_cost[i] = (double)E * c->N * c->N * c->N;
        LdsLayout L = lds_layout(c->N, E, fp32 ? 4 : 8, batch_up_lds(b), c->cfg.av != 0, false, c->cfg.limiter != 0);
        lds = std::max(lds, (size_t)L.total);
    }
    // instances by tile cost, heaviest first (stable: ties keep their order)
    std::vector<int> order(b->n);
    for (int i = 0; i < b->n; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return tile_cost[x] > tile_cost[y]; });
    b->runs.clear();
    if (b->by_degree) {
        for (int p = PMAX_GPU; p >= 1; p--) {
            dg_batch::DegreeRun r{p, (int)tiles.size(), 0, 0, 0};
            for (int i = 0; i < b->n; i++) {
                dg_ctx* c = b->ctx[i];
                if (c->cfg.p != p) continue;
                for (int t = 0; t < ntl[i]; t++) tiles.push_back(make_int2(i, t));
                const int E = batch_elems_per_block(c->N, batch_lanes_of(b, c->cfg.p));
                r.block = std::min(BLOCK, (E * 2 * c->N + 63) / 64 * 64);
                r.lds = (size_t)lds_layout(c->N, E, fp32 ? 4 : 8, true, c->cfg.av != 0, false,
                                           !b->sd && c->cfg.limiter != 0).total;
            }
            r.count = (int)tiles.size() - r.first;
            if (r.count > 0) b->runs.push_back(r);
        }
    } else if (!b->xcd_bins) {
        for (int i : order)
            for (int t = 0; t < ntl[i]; t++) tiles.push_back(make_int2(i, t));
    } else {
        std::vector<int> by_total(b->n);
        for (int i = 0; i < b->n; i++) by_total[i] = i;
        std::stable_sort(by_total.begin(), by_total.end(), [&](int x, int y) {
            return tile_cost[x] * ntl[x] > tile_cost[y] * ntl[y];
        });
        double load[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        std::vector<int> bin_of(b->n);
        for (int i : by_total) {
            const int x = (int)(std::min_element(load, load + 8) - load);
            bin_of[i] = x;
            load[x] += tile_cost[i] * ntl[i];
        }
        std::vector<int2> bins[8];
        for (int i : order)
            for (int t = 0; t < ntl[i]; t++) bins[bin_of[i]].push_back(make_int2(i, t));
        size_t len = 0;
        for (int x = 0; x < 8; x++) len = std::max(len, bins[x].size());
        tiles.assign(8 * len, make_int2(-1, 0));
        for (int x = 0; x < 8; x++)
            for (size_t k = 0; k < bins[x].size(); k++) tiles[8 * k + x] = bins[x][k];
    }
    if ((int)tiles.size() > b->tiles_cap) {
        hipFree(b->d_tiles);
        b->d_tiles = nullptr;
        hipError_t e = hipMalloc(&b->d_tiles, tiles.size() * sizeof(int2));
        if (e != hipSuccess) return e;
        b->tiles_cap = (int)tiles.size();
    }
    b->ntiles = (int)tiles.size();
    b->lds = lds;
    // the persistent form's grid: the workgroups resident on the chip at
    // once, a multiple of 8 (a block then walks its own XCD's bin)
    {
        const bool fp32 = b->common.precision == DG_MIXED;
        int per_cu = 0, ncu = 0, dev = 0;
        hipError_t e = fp32 ? batch_persistent_per_cu_float(b->common.volume_flux, b->lanes, lds, &per_cu)
                            : batch_persistent_per_cu_double(b->common.volume_flux, b->lanes, lds, &per_cu);
        if (e == hipSuccess) e = hipGetDevice(&dev);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
        if (e != hipSuccess) return e;
        int g = std::min(b->ntiles, std::max(0, per_cu) * ncu);
        b->pgrid = g >= 8 ? g / 8 * 8 : g;
    }
    hipError_t e = hipMemcpyAsync(b->d_tiles, tiles.data(), tiles.size() * sizeof(int2), hipMemcpyHostToDevice,
                                  b->stream);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(b->stream);   // the host vector goes out of scope
}

static int batch_status(dg_batch* b, int nsteps, double* t_inout, int* status_out) {
    std::vector<DevState> h(b->n);
    hipLaunchKernelGGL(batch_gather_kernel, dim3((b->n + 63) / 64), dim3(64), 0, b->stream, b->d_args, b->n,
                       b->d_gather);
    BHIPCK(b, hipGetLastError());
    BHIPCK(b, hipMemcpyAsync(h.data(), b->d_gather, b->n * sizeof(DevState), hipMemcpyDeviceToHost, b->stream));
    BHIPCK(b, hipStreamSynchronize(b->stream));
    int first = DG_OK;
    for (int i = 0; i < b->n; i++) {
        int st = h[i].status;
        if (st == ST_OK && h[i].steps > 0) {
            // the last step taken wrote slot steps & 1 (see advance_impl)
            double lam;
            memcpy(&lam, &h[i].lam_bits[h[i].steps & 1], sizeof(lam));
            if (!(lam > 0.0) || lam > 1e6) st = DG_ERR_BLOWUP;
        }
        if (t_inout) t_inout[i] = h[i].t;
        if (status_out) status_out[i] = st;
        if (st != DG_OK && first == DG_OK) {
            first = st;
            bfail(b, st, "instance %d: %s at t=%.9g", i, dg_status_string(st), h[i].t);
        }
    }
    return first;
}

extern "C" {

int dg_batch_create(const dg_config* cfgs, int n, void* stream, dg_batch** out) {
    if (!cfgs || n < 1 || !out) return DG_ERR_INVALID_ARG;
    *out = nullptr;
    const dg_config& c0 = cfgs[0];
    for (int i = 0; i < n; i++) {
        const dg_config& f = cfgs[i];
        if (f.volume_flux != c0.volume_flux || f.surface_flux != c0.surface_flux || f.bc != c0.bc ||
            f.precision != c0.precision || (f.av != 0) != (c0.av != 0) || (f.limiter != 0) != (c0.limiter != 0) ||
            f.scheme != c0.scheme)
            return DG_ERR_INVALID_ARG;
    }
    dg_batch* b = new dg_batch();
    b->n = n;
    b->sd = c0.scheme == DG_SCHEME_SD;
    b->full = !b->sd && (c0.av || c0.limiter);
    b->by_degree = batch_degree_only(b);
    b->stream = (hipStream_t)stream;
    b->common = c0;
    b->ctx.assign(n, nullptr);
    for (int i = 0; i < n; i++) {
        int rc = dg_create(&cfgs[i], stream, &b->ctx[i]);
        if (rc) { dg_batch_destroy(b); return rc; }
    }
    b->h_args.resize((size_t)3 * n);
#define BCK(x)                                                                 \
    do {                                                                       \
        if ((x) != hipSuccess) { dg_batch_destroy(b); return DG_ERR_HIP; }    \
    } while (0)
    BCK(hipMalloc(&b->d_args, (size_t)3 * n * sizeof(KArgs)));
    BCK(hipMalloc(&b->d_gather, (size_t)n * sizeof(DevState)));
    BCK(hipMalloc(&b->d_t0, (size_t)n * sizeof(double)));
    BCK(hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking));
    BCK(hipEventCreateWithFlags(&b->ev_in, hipEventDisableTiming));
    BCK(hipEventCreateWithFlags(&b->ev_out, hipEventDisableTiming));
    BCK(batch_tiles(b));
    BCK(hipStreamSynchronize(b->stream));
#undef BCK
    *out = b;
    return DG_OK;
}

int dg_batch_destroy(dg_batch* b) {
    if (!b) return DG_OK;
    if (b->gexec) hipGraphExecDestroy(b->gexec);
    for (dg_ctx* c : b->ctx) dg_destroy(c);
    hipFree(b->d_args); hipFree(b->d_tiles); hipFree(b->d_gather); hipFree(b->d_t0);
    if (b->cap) hipStreamDestroy(b->cap);
    if (b->ev_in) hipEventDestroy(b->ev_in);
    if (b->ev_out) hipEventDestroy(b->ev_out);
    delete b;
    return DG_OK;
}

const char* dg_batch_last_error(const dg_batch* b) { return b ? b->err : "null batch"; }

long long dg_batch_state_size(const dg_batch* b, int i) {
    return (b && i >= 0 && i < b->n) ? dg_state_size(b->ctx[i]) : 0;
}

int dg_batch_save_state(dg_batch* b, int i, const double* Q, double t, long long steps, const char* path) {
    if (!b) return DG_ERR_INVALID_ARG;
    if (i < 0 || i >= b->n) return bfail(b, DG_ERR_INVALID_ARG, "dg_batch_save_state: instance %d of %d", i, b->n);
    const int rc = dg_save_state(b->ctx[i], Q, t, steps, path);
    if (rc != DG_OK) bfail(b, rc, "instance %d: %s", i, dg_last_error(b->ctx[i]));
    return rc;
}

int dg_batch_load_state(dg_batch* b, int i, double* Q, double* t, long long* steps, const char* path) {
    if (!b) return DG_ERR_INVALID_ARG;
    if (i < 0 || i >= b->n) return bfail(b, DG_ERR_INVALID_ARG, "dg_batch_load_state: instance %d of %d", i, b->n);
    const int rc = dg_load_state(b->ctx[i], Q, t, steps, path);
    if (rc != DG_OK) bfail(b, rc, "instance %d: %s", i, dg_last_error(b->ctx[i]));
    return rc;
}

int dg_batch_init_state(dg_batch* b, const dg_ic_params* ics, double* const* Qs) {
    if (!b || !ics || !Qs) return bfail(b, DG_ERR_INVALID_ARG, "dg_batch_init_state: null argument");
    for (int i = 0; i < b->n; i++) {
        int rc = dg_init_state(b->ctx[i], &ics[i], Qs[i]);
        if (rc) return bfail(b, rc, "instance %d: %s", i, b->ctx[i]->err);
    }
    return DG_OK;
}

int dg_batch_set_options(dg_batch* b, int use_graph, int lanes) {
    if (!b) return DG_ERR_INVALID_ARG;
    if (lanes == 0) lanes = BATCH_BLOCK;
    if (lanes != 64 && lanes != 128 && lanes != 256)
        return bfail(b, DG_ERR_INVALID_ARG, "dg_batch_set_options: lanes %d not 64, 128 or 256", lanes);
    b->use_graph = use_graph != 0;
    if (lanes != b->lanes) {
        b->lanes = lanes;
        BHIPCK(b, batch_tiles(b));
    }
    if (b->gexec) { hipGraphExecDestroy(b->gexec); b->gexec = nullptr; }   // tiles are baked into the graph
    return DG_OK;
}

int dg_batch_set_xcd_bins(dg_batch* b, int xcd_bins);

int dg_batch_set_persistent(dg_batch* b, int persistent) {
    if (!b) return DG_ERR_INVALID_ARG;
    if (persistent && batch_degree_only(b))
        return bfail(b, DG_ERR_INVALID_ARG,
                     "dg_batch_set_persistent: AV / limiter and the SD scheme run in the by-degree form only");
    if (persistent && b->by_degree) {   // the persistent form is the all-degree kernel's
        b->by_degree = false;
        BHIPCK(b, batch_tiles(b));
        if (b->gexec) { hipGraphExecDestroy(b->gexec); b->gexec = nullptr; }
    }
    if (persistent && b->pgrid < 1)
        return bfail(b, DG_ERR_INVALID_ARG, "dg_batch_set_persistent: no resident workgroup at this size");
    if ((persistent != 0) != b->persistent && b->gexec) { hipGraphExecDestroy(b->gexec); b->gexec = nullptr; }
    b->persistent = persistent != 0;
    return DG_OK;
}

int dg_batch_advance(dg_batch* b, double* const* Qs, int nsteps, const double* cfl, double t_end, double* t_inout,
                     int* status_out) {
    if (!b || !Qs || !cfl || nsteps < 0) return bfail(b, DG_ERR_INVALID_ARG, "dg_batch_advance: bad argument");
    for (int i = 0; i < b->n; i++) {
        if (!Qs[i] || !(cfl[i] > 0)) return bfail(b, DG_ERR_INVALID_ARG, "dg_batch_advance: instance %d", i);
        if (!aligned16(Qs[i]))
            return bfail(b, DG_ERR_INVALID_ARG, "dg_batch_advance: instance %d state is not 16-byte aligned", i);
    }
    batch_bind(b, Qs, cfl, t_end);
    BHIPCK(b, hipMemcpyAsync(b->d_args, b->h_args.data(), b->h_args.size() * sizeof(KArgs), hipMemcpyHostToDevice,
                             b->stream));
    if (t_inout) BHIPCK(b, hipMemcpyAsync(b->d_t0, t_inout, b->n * sizeof(double), hipMemcpyHostToDevice, b->stream));
    hipLaunchKernelGGL(batch_reset_kernel, dim3((b->n + 63) / 64), dim3(64), 0, b->stream, b->d_args, b->n,
                       t_inout ? b->d_t0 : nullptr);
    hipLaunchKernelGGL(batch_lambda_kernel, dim3(std::min(b->ntiles, 8192)), dim3(BLOCK), 0, b->stream, b->d_args,
                       b->d_tiles, b->ntiles);
    BHIPCK(b, hipGetLastError());
    // stage launches run on the capture stream, ordered after the work above
    BHIPCK(b, hipEventRecord(b->ev_in, b->stream));
    BHIPCK(b, hipStreamWaitEvent(b->cap, b->ev_in, 0));
    if (nsteps > 0) BHIPCK(b, batch_step(b, StepCtl{0, 1, 0}));
    int n = 1;
    if (b->use_graph && nsteps >= 3) {
        if (!b->gexec) {   // two steps (odd, even) per replay; arguments live in d_args
            hipGraph_t g;
            BHIPCK(b, hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
            hipError_t e1 = batch_step(b, StepCtl{1, 0, 1});
            hipError_t e2 = batch_step(b, StepCtl{0, 1, 1});
            hipError_t ec = hipStreamEndCapture(b->cap, &g);
            if (ec == hipSuccess && (e1 != hipSuccess || e2 != hipSuccess)) hipGraphDestroy(g);
            BHIPCK(b, e1); BHIPCK(b, e2); BHIPCK(b, ec);
            hipError_t ei = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
            hipGraphDestroy(g);
            BHIPCK(b, ei);
        }
        for (; n + 1 < nsteps; n += 2) BHIPCK(b, hipGraphLaunch(b->gexec, b->cap));
    }
    for (; n < nsteps; n++) BHIPCK(b, batch_step(b, StepCtl{n & 1, (n & 1) ^ 1, 1}));
    BHIPCK(b, hipEventRecord(b->ev_out, b->cap));
    BHIPCK(b, hipStreamWaitEvent(b->stream, b->ev_out, 0));
    return batch_status(b, nsteps, t_inout, status_out);
}

int dg_batch_advance_each(dg_batch* b, double* const* Qs, int nsteps, const double* cfl, double t_end,
                          double* t_inout, int* status_out) {
    if (!b || !Qs || !cfl || nsteps < 0) return bfail(b, DG_ERR_INVALID_ARG, "dg_batch_advance_each: bad argument");
    int first = DG_OK;
    for (int i = 0; i < b->n; i++) {
        double t = t_inout ? t_inout[i] : 0.0;
        int rc = dg_advance(b->ctx[i], Qs[i], nsteps, cfl[i], t_end, &t, nullptr);
        if (t_inout) t_inout[i] = t;
        if (status_out) status_out[i] = rc;
        if (rc && first == DG_OK) { first = rc; bfail(b, rc, "instance %d: %s", i, b->ctx[i]->err); }
    }
    return first;
}

int dg_batch_autotune(dg_batch* b, double* const* Qs, const double* cfl, int* lanes) {
    if (!b || !Qs || !cfl) return DG_ERR_INVALID_ARG;
    const int cands[] = {64, 128, 256};
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    // Every option change drops the cached graph: 3 untimed steps rebuild it
    // (one eager step, instantiation, one replay) and warm the caches, then 6
    // steps are timed (3 replays with the graph on); the states are advanced
    // by 9 steps per timing at cfl[i]
    auto time_steps = [&](float* ms) -> int {
        int r = dg_batch_advance(b, Qs, 3, cfl, 0.0, nullptr, nullptr);
        if (r) return r;
        hipEventRecord(e0, b->stream);
        if ((r = dg_batch_advance(b, Qs, 6, cfl, 0.0, nullptr, nullptr))) return r;
        hipEventRecord(e1, b->stream);
        hipEventSynchronize(e1);
        *ms = 0.f;
        hipEventElapsedTime(ms, e0, e1);
        return DG_OK;
    };
    int rc = DG_OK;
    for (int p = 1; p <= PMAX_GPU && !rc; p++) rc = dg_batch_set_degree_lanes(b, p, 0);
    // each size with both orders in the three launch forms (the all-degree
    // kernel with one workgroup per tile or persistent, one launch per degree),
    // timed round-robin over three rounds, each keeping its fastest run (a
    // slow stretch of the device then costs every candidate alike)
    struct Cand { int lanes, form, bins; float t; };
    std::vector<Cand> cl;
    for (int cand = 0; cand < 18; cand++) {
        const int form = (cand % 6) >> 1;
        if (batch_degree_only(b) && form != 2) continue;   // shock capturing, SD: by degree only
        cl.push_back({cands[cand / 6], form, cand & 1, 1e30f});
    }
    for (int round = 0; round < 3 && !rc; round++)
        for (Cand& cd : cl) {
            if ((rc = dg_batch_set_options(b, b->use_graph, cd.lanes))) break;
            if ((rc = dg_batch_set_by_degree(b, cd.form == 2))) break;
            if ((rc = dg_batch_set_xcd_bins(b, cd.bins))) break;
            if (cd.form == 1 && b->pgrid < 1) continue;   // no resident grid at this size
            if ((rc = dg_batch_set_persistent(b, cd.form == 1))) break;
            float ms;
            if ((rc = time_steps(&ms))) break;
            cd.t = std::min(cd.t, ms);
        }
    float best = 1e30f;
    int best_lanes = b->lanes;
    bool best_pers = b->persistent, best_deg = b->by_degree, best_bins = b->xcd_bins;
    for (const Cand& cd : cl)
        if (cd.t < best) {
            best = cd.t; best_lanes = cd.lanes; best_pers = cd.form == 1; best_deg = cd.form == 2;
            best_bins = cd.bins != 0;
        }
    int rc2 = dg_batch_set_options(b, b->use_graph, best_lanes);
    if (!rc2) rc2 = dg_batch_set_by_degree(b, best_deg);
    if (!rc2) rc2 = dg_batch_set_xcd_bins(b, best_bins);
    if (!rc2) rc2 = dg_batch_set_persistent(b, best_pers);
    // by-degree form: each degree's launch is independent of the others, so
    // its workgroup size is refined on its own, the other degrees fixed; a
    // size is kept only if the whole step runs at least 1 % faster with it
    if (!rc && !rc2 && best_deg) {
        std::vector<int> degs;
        for (const dg_batch::DegreeRun& r : b->runs) degs.push_back(r.p);
        float ref;
        rc = time_steps(&ref);   // the kept configuration on the states as they are now
        for (size_t k = 0; k < degs.size() && !rc; k++) {
            const int p = degs[k];
            int keep = 0;
            for (int l : cands) {
                if (l == best_lanes) continue;
                float ms;
                if ((rc = dg_batch_set_degree_lanes(b, p, l)) || (rc = time_steps(&ms))) break;
                if (ms < 0.99f * ref) { ref = ms; keep = l; }
            }
            if (!rc) rc = dg_batch_set_degree_lanes(b, p, keep);
        }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (lanes) *lanes = b->lanes;
    return rc ? rc : rc2;
}

int dg_batch_get_options(const dg_batch* b, int* use_graph, int* lanes, int* persistent, int* xcd_bins) {
    if (!b) return DG_ERR_INVALID_ARG;
    if (use_graph) *use_graph = b->use_graph;
    if (lanes) *lanes = b->lanes;
    if (persistent) *persistent = b->persistent;
    if (xcd_bins) *xcd_bins = b->xcd_bins;
    return DG_OK;
}

int dg_batch_set_by_degree(dg_batch* b, int by_degree) {
    if (!b) return DG_ERR_INVALID_ARG;
    if (!by_degree && batch_degree_only(b))
        return bfail(b, DG_ERR_INVALID_ARG,
                     "dg_batch_set_by_degree: AV / limiter and the SD scheme run in the by-degree form only");
    if ((by_degree != 0) == b->by_degree) return DG_OK;
    b->by_degree = by_degree != 0;
    if (b->by_degree) b->persistent = false;
    BHIPCK(b, batch_tiles(b));
    if (b->gexec) { hipGraphExecDestroy(b->gexec); b->gexec = nullptr; }   // the launches are baked into the graph
    return DG_OK;
}

int dg_batch_set_degree_lanes(dg_batch* b, int p, int lanes) {
    if (!b) return DG_ERR_INVALID_ARG;
    if (p < 1 || p > PMAX_GPU || (lanes != 0 && lanes != 64 && lanes != 128 && lanes != 256))
        return bfail(b, DG_ERR_INVALID_ARG, "dg_batch_set_degree_lanes: degree %d, lanes %d", p, lanes);
    if (b->deg_lanes[p] == lanes) return DG_OK;
    b->deg_lanes[p] = lanes;
    if (b->by_degree) {
        BHIPCK(b, batch_tiles(b));
        if (b->gexec) { hipGraphExecDestroy(b->gexec); b->gexec = nullptr; }
    }
    return DG_OK;
}

int dg_batch_get_degree_lanes(const dg_batch* b, int p, int* lanes) {
    if (!b || !lanes || p < 1 || p > PMAX_GPU) return DG_ERR_INVALID_ARG;
    *lanes = b->deg_lanes[p] ? b->deg_lanes[p] : b->lanes;
    return DG_OK;
}

int dg_batch_get_by_degree(const dg_batch* b, int* by_degree) {
    if (!b || !by_degree) return DG_ERR_INVALID_ARG;
    *by_degree = b->by_degree;
    return DG_OK;
}

int dg_batch_set_xcd_bins(dg_batch* b, int xcd_bins) {
    if (!b) return DG_ERR_INVALID_ARG;
    if ((xcd_bins != 0) == b->xcd_bins) return DG_OK;
    b->xcd_bins = xcd_bins != 0;
    BHIPCK(b, batch_tiles(b));
    if (b->gexec) { hipGraphExecDestroy(b->gexec); b->gexec = nullptr; }   // the list is baked into the graph
    return DG_OK;
}

}  // extern "C"
