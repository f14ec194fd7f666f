/*
 * oracle_selftest.cpp -- drives every entry point of the CPU oracle
 * (dgsem_oracle.cpp) once per configuration, for the sanitizer build
 * (oracle/Makefile, target `sanitize`: AddressSanitizer + UndefinedBehavior-
 * Sanitizer, every finding fatal).  TEST INFRASTRUCTURE ONLY, like the oracle.
 *
 * Configurations: degrees 1..7, both schemes (DGSEM, spectral difference),
 * every volume flux (Chandrashekar, Ranocha, collocation) and surface flux
 * (HLLC, Rusanov), walls and periodic boundaries, AV and the limiter on and
 * off, and the three initial conditions.  Each runs the initial condition,
 * one right-hand side, the three separate RK stages with the limiter, the
 * sensor, the frame fields, the totals and a few fused steps, on small grids
 * (K = 1, 2 and 5: single-element domains hit every boundary branch).
 * Besides the sanitizers' own checks it requires finite outputs and, for
 * periodic DGSEM runs without the limiter, mass conserved to 1e-11.
 * Exit status 0 = clean.
 */
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../include/dgsem.h"

extern "C" {
int oracle_create(const dg_config* cfg, void** out);
void oracle_destroy(void* ctx);
int oracle_set_threads(int n);
void oracle_lgl(int p, double* x, double* w);
void oracle_build_D(int p, double* D);
void oracle_sd_operators(void* ctx, double* xf, double* Isf, double* Dfs);
double oracle_logmean(double a, double b);
void oracle_volume_flux(int kind, double gamma, const double* WL, const double* WR, int dir, double* F);
void oracle_surface_flux(int kind, double gamma, const double* WL, const double* WR, double nx, double ny,
                         double* F);
void oracle_cons_to_prim(double gamma, const double* U, double* W);
void oracle_physical_flux(double gamma, const double* W, double nx, double ny, double* F);
void oracle_vortex_exact(double gamma, double eps, double xc, double yc, double u0, double v0, double x, double y,
                         double* W);
double oracle_kh_noise(unsigned long long seed, long long node, double amp);
int oracle_init_state(void* ctx, const dg_ic_params* ic, double* Q);
int oracle_compute_rhs(void* ctx, const double* Q, double* RHS, double* lam);
int oracle_rk3_stage(void* ctx, int s, double* Uout, const double* U0, const double* Up, const double* L,
                     double dt);
int oracle_limit(void* ctx, double* Q);
double oracle_max_lambda(void* ctx, const double* lam);
int oracle_advance(void* ctx, double* Q, int nsteps, double cfl, double t_end, double* t_inout, double* dt_last);
int oracle_sensor(void* ctx, const double* Q, double* sensor);
int oracle_frame_fields(void* ctx, const double* Q, double* rho_el, double* p_el);
int oracle_totals(void* ctx, const double* Q, double* out);
}

static int failures = 0;

static void fail(const char* what, const dg_config& c, int ic) {
    std::fprintf(stderr, "FAIL %s: K=%d p=%d scheme=%d vol=%d surf=%d bc=%d av=%d lim=%d ic=%d\n", what, c.K, c.p,
                 c.scheme, c.volume_flux, c.surface_flux, c.bc, c.av, c.limiter, ic);
    failures++;
}

static bool finite(const std::vector<double>& v) {
    for (double x : v)
        if (!std::isfinite(x)) return false;
    return true;
}

static dg_config base_config(int K, int p) {
    dg_config c;
    std::memset(&c, 0, sizeof(c));
    c.K = K; c.p = p; c.L = 1.0; c.gamma = 1.4;
    c.precision = DG_FP64; c.av_cav = 0.25; c.av_ksensor = 0.1; c.eps_rho = 1e-6; c.eps_p = 1e-6;
    c.xcd_remap = -1; c.mfma = -1;
    return c;
}

static dg_ic_params ic_of(int kind, const dg_config& c) {
    dg_ic_params ic;
    std::memset(&ic, 0, sizeof(ic));
    ic.kind = kind;
    ic.rho_hi = 11.0; ic.p_hi = 10.0;
    ic.vortex_eps = 5.0; ic.xc = c.x0 + 0.5 * c.L; ic.yc = c.y0 + 0.5 * c.L; ic.u0 = 1.0; ic.v0 = 1.0;
    ic.seed = 1234; ic.noise_amp = 1e-3;
    return ic;
}

static void run_case(dg_config c, int ic_kind) {
    void* ctx = nullptr;
    if (oracle_create(&c, &ctx) != DG_OK) { fail("create", c, ic_kind); return; }
    const int N = c.p + 1, NEL = c.K * c.K;
    const size_t n = (size_t)4 * NEL * N * N;
    std::vector<double> Q(n), R(n), U1(n), U2(n), lam(NEL), sens(NEL), rho_el(NEL), p_el(NEL), tot(8, 0.0);
    const dg_ic_params ic = ic_of(ic_kind, c);
    if (oracle_init_state(ctx, &ic, Q.data()) != DG_OK || !finite(Q)) fail("init_state", c, ic_kind);
    if (oracle_compute_rhs(ctx, Q.data(), R.data(), lam.data()) != DG_OK || !finite(R) || !finite(lam))
        fail("compute_rhs", c, ic_kind);
    const double lmax = oracle_max_lambda(ctx, lam.data());
    if (!(lmax > 0.0)) fail("max_lambda", c, ic_kind);
    const double dt = 0.2 * (c.L / c.K) / (lmax * N);
    oracle_rk3_stage(ctx, 1, U1.data(), Q.data(), nullptr, R.data(), dt);
    oracle_compute_rhs(ctx, U1.data(), R.data(), lam.data());
    oracle_rk3_stage(ctx, 2, U2.data(), Q.data(), U1.data(), R.data(), dt);
    oracle_compute_rhs(ctx, U2.data(), R.data(), lam.data());
    std::vector<double> Q3(Q);
    oracle_rk3_stage(ctx, 3, Q3.data(), Q.data(), U2.data(), R.data(), dt);
    if (!finite(U1) || !finite(U2) || !finite(Q3)) fail("rk3_stage", c, ic_kind);
    oracle_limit(ctx, Q3.data());
    if (!finite(Q3)) fail("limit", c, ic_kind);
    if (c.scheme == DG_SCHEME_DGSEM && (oracle_sensor(ctx, Q.data(), sens.data()) != DG_OK || !finite(sens)))
        fail("sensor", c, ic_kind);
    if (oracle_frame_fields(ctx, Q.data(), rho_el.data(), p_el.data()) != DG_OK || !finite(rho_el) ||
        !finite(p_el))
        fail("frame_fields", c, ic_kind);
    double m0[8] = {0}, m1[8] = {0};
    oracle_totals(ctx, Q.data(), m0);
    double t = 0.0, dt_last = 0.0;
    const int rc = oracle_advance(ctx, Q.data(), 3, 0.2, 0.0, &t, &dt_last);
    if (rc != DG_OK || !finite(Q) || !(t > 0.0)) fail("advance", c, ic_kind);
    oracle_totals(ctx, Q.data(), m1);
    if (c.bc == DG_BC_PERIODIC && c.scheme == DG_SCHEME_DGSEM && !c.limiter &&
        std::fabs(m1[0] - m0[0]) > 1e-11 * std::fabs(m0[0]))
        fail("mass conservation", c, ic_kind);
    if (c.scheme == DG_SCHEME_SD) {
        std::vector<double> xf(N + 1), Isf((size_t)(N + 1) * N), Dfs((size_t)N * (N + 1));
        oracle_sd_operators(ctx, xf.data(), Isf.data(), Dfs.data());
        if (!finite(xf) || !finite(Isf) || !finite(Dfs)) fail("sd_operators", c, ic_kind);
    }
    oracle_destroy(ctx);
}

int main() {
    oracle_set_threads(2);
    // pointwise entry points
    for (int p = 1; p <= 7; p++) {
        std::vector<double> x(p + 1), w(p + 1), D((size_t)(p + 1) * (p + 1));
        oracle_lgl(p, x.data(), w.data());
        oracle_build_D(p, D.data());
        if (!finite(x) || !finite(w) || !finite(D)) { std::fprintf(stderr, "FAIL lgl/D p=%d\n", p); failures++; }
    }
    const double WL[4] = {1.0, 0.3, -0.2, 1.0}, WR[4] = {0.125, -0.1, 0.4, 0.1}, U[4] = {1.0, 0.5, 0.1, 2.5};
    double F[4], W[4];
    for (int k = 0; k < 3; k++)
        for (int d = 0; d < 2; d++) oracle_volume_flux(k, 1.4, WL, WR, d, F);
    for (int k = 0; k < 2; k++) oracle_surface_flux(k, 1.4, WL, WR, 1.0, 0.0, F);
    oracle_cons_to_prim(1.4, U, W);
    oracle_physical_flux(1.4, W, 0.0, 1.0, F);
    oracle_vortex_exact(1.4, 5.0, 5.0, 5.0, 1.0, 1.0, 4.0, 6.0, W);
    if (!std::isfinite(oracle_logmean(1.0, 1.0 + 1e-9)) || !std::isfinite(oracle_logmean(3.0, 0.5)) ||
        !std::isfinite(oracle_kh_noise(1234, 77, 1e-3))) {
        std::fprintf(stderr, "FAIL pointwise\n");
        failures++;
    }
    int cases = 0;
    for (int p = 1; p <= 7; p++)
        for (int K : {1, 2, 5}) {
            // DGSEM: every flux pair, both boundaries, shock capturing off / on
            for (int vol = 0; vol < 3; vol++)
                for (int surf = 0; surf < 2; surf++)
                    for (int bc = 0; bc < 2; bc++)
                        for (int sc = 0; sc < 2; sc++) {
                            if (K == 5 && (vol != 1 || surf != 1) && sc == 0) continue;   // keep the run short
                            dg_config c = base_config(K, p);
                            c.volume_flux = vol; c.surface_flux = surf; c.bc = bc; c.av = sc; c.limiter = sc;
                            for (int ic = 0; ic < 3; ic++) { run_case(c, ic); cases++; }
                        }
            // spectral difference: the reference's driver configuration
            for (int bc = 0; bc < 2; bc++) {
                dg_config c = base_config(K, p);
                c.scheme = DG_SCHEME_SD; c.surface_flux = DG_SURF_HLLC; c.bc = bc; c.limiter = 1;
                for (int ic = 0; ic < 3; ic++) { run_case(c, ic); cases++; }
            }
        }
    std::printf("oracle selftest: %d configurations, %d failure(s)\n", cases, failures);
    return failures ? 1 : 0;
}
