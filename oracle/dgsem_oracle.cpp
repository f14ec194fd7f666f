/*
 * dgsem_oracle.cpp — CPU FP64 oracle for the 2-D Euler DGSEM step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it is loaded by tests/,
 * by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, never by the
 * product library (fvstuff_amd/), which must not link, import or call it.
 *
 * It restates, function for function and in the reference's operation order,
 * the per-node update of KLozes/fvStuff dgsem_lobatto_2d.cu with `float`
 * replaced by `double` (the 1e-14 floors, 1e-5 velocity cut-off and UMAX clamp
 * are kept).  Each function cites the reference lines it follows.  Added
 * switches, absent from the reference and pinned only by closed forms
 * (see the tests): the Ranocha volume flux, the Rusanov surface flux, periodic
 * boundaries, LGL data for p > 3 and the vortex / Kelvin-Helmholtz ICs.
 * The spectral-difference sibling (DG_SCHEME_SD) restates fr_lobatto_2d.cu
 * the same way, except for its derivative matrix's end rows (see build_sd).
 *
 * Parity pins (tests/test_oracle_*.py):
 *   lgl / build_D .......... reference tables (:145-157), run_fsbp.py literals,
 *                            closed-form invariants (sum w = 2, exactness,
 *                            D x^k = k x^(k-1), SBP, D00 = -p(p+1)/4)
 *   logmean ................ long-double closed form on both branches
 *   ec fluxes .............. consistency, symmetry, Tadmor's entropy identity
 *   hllc / rusanov ......... consistency, supersonic upwinding, brute force
 *   rhs .................... free stream, discrete conservation, entropy
 *                            conservation with an EC surface, convergence order
 *   rk3 .................... exactness on cubic-in-time ODEs
 *   limiter ................ mean preservation, identity on admissible data
 *   sensor / AV ............ closed forms on a linear compression u = a x
 *                            (theta per element; the AV term through the SBP
 *                            identity W^-1 D^T W D U = W^-1 B U' - U'',
 *                            tests/test_oracle_av.py), free stream
 *   ICs .................... (tests/test_oracle_ic.py) vortex closed form at
 *                            the centre and isentropy; Kelvin-Helmholtz closed
 *                            form at every node (nodes on y = +-1/2 belong to
 *                            the outer layer, the noise moves v only); the
 *                            circular Sod front width h/2 recovered from the
 *                            density (:702-716), plateaus, symmetries; the
 *                            p = 7 vortex's CFL limit (blow-up at 0.5)
 *   SD sibling ............. (fr_lobatto_2d.cu, tests/test_oracle_sd.py) flux
 *                            nodes = Gauss-Legendre roots and the reference's
 *                            literals, Isf / Dfs exact on polynomials of degree
 *                            p / p+1, w^T Dfs telescopes, free stream,
 *                            discrete conservation, closed wall box, order
 *                            p+1 on the vortex, the nodal repair, the IC width
 *
 * Build: oracle/Makefile (g++ -O2 -ffp-contract=off -fopenmp -shared).
 */
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <algorithm>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "../include/dgsem.h"

#define OR_API extern "C" __attribute__((visibility("default")))

namespace {

constexpr int NVAR = 4;
constexpr int PMAX = 15;
constexpr double UMAX = 1e4;  /* dgsem_lobatto_2d.cu:346 */

struct Ctx {
    dg_config cfg;
    int K, p, NN, NN2, NEL;
    double h;
    double xi[PMAX + 1], w[PMAX + 1], winv[PMAX + 1];
    double D[PMAX + 1][PMAX + 1];
    /* spectral-difference sibling (fr_lobatto_2d.cu): NF = p + 2 flux nodes */
    int NF;
    double xf[PMAX + 2], Isf[PMAX + 2][PMAX + 1], Dfs[PMAX + 1][PMAX + 2];
    std::vector<double> Q0, Q1, Q2, RHS, lam;
};

/* ── L1: operators ───────────────────────────────────────────────────────
 * LGL nodes / GLL weights for p = 1..3 are the reference's literal tables
 * (dgsem_lobatto_2d.cu:145-157).  For p > 3 (absent from the reference) the
 * nodes are +-1 and the roots of L_{p+1} - L_{p-1} (= const*(1-x^2) L_p'),
 * found by Newton from Chebyshev-Gauss-Lobatto guesses, and
 * w_j = 2 / (p (p+1) L_p(x_j)^2). */
static const double ref_lgl_xi[3][4] = {
    {-1.0, 1.0},
    {-1.0, 0.0, 1.0},
    {-1.0, -0.4472135954999579, 0.4472135954999579, 1.0},
};
static const double ref_gll_w[3][4] = {
    {1.0, 1.0},
    {0.33333333333333333, 1.33333333333333333, 0.33333333333333333},
    {0.16666666666666667, 0.83333333333333333, 0.83333333333333333, 0.16666666666666667},
};

static void legendre(int n, long double x, long double* L, long double* dL)
{
    /* Bonnet recurrence for L_n and its derivative */
    if (n == 0) { *L = 1; *dL = 0; return; }
    long double Lm1 = 1, L0 = x, dLm1 = 0, dL0 = 1;
    for (int k = 2; k <= n; k++) {
        long double Lk = ((2 * k - 1) * x * L0 - (k - 1) * Lm1) / k;
        long double dLk = dLm1 + (2 * k - 1) * L0;
        Lm1 = L0; L0 = Lk; dLm1 = dL0; dL0 = dLk;
    }
    *L = L0; *dL = dL0;
}

static void lgl(int p, double* x, double* w)
{
    if (p >= 1 && p <= 3) {
        for (int i = 0; i <= p; i++) { x[i] = ref_lgl_xi[p - 1][i]; w[i] = ref_gll_w[p - 1][i]; }
        return;
    }
    const long double PI = 3.14159265358979323846264338327950288L;
    std::vector<long double> xs(p + 1);
    xs[0] = -1; xs[p] = 1;
    for (int j = 1; j < p; j++) {
        long double t = -cosl(PI * j / p);
        for (int it = 0; it < 100; it++) {
            long double Lp1, dLp1, Lm1, dLm1;
            legendre(p + 1, t, &Lp1, &dLp1);
            legendre(p - 1, t, &Lm1, &dLm1);
            long double q = Lp1 - Lm1, dq = dLp1 - dLm1;
            long double dt = q / dq;
            t -= dt;
            if (fabsl(dt) < 1e-19L) break;
        }
        xs[j] = t;
    }
    for (int j = 0; j <= p; j++) {
        long double L, dL;
        legendre(p, xs[j], &L, &dL);
        x[j] = (double)xs[j];
        w[j] = (double)(2.0L / ((long double)p * (p + 1) * L * L));
    }
    /* exact symmetry */
    for (int j = 0; j <= p / 2; j++) {
        double a = 0.5 * (x[p - j] - x[j]);
        x[j] = -a; x[p - j] = a;
        double b = 0.5 * (w[j] + w[p - j]);
        w[j] = b; w[p - j] = b;
    }
    if (p % 2 == 0) x[p / 2] = 0.0;
}

/* Barycentric differentiation matrix with negative-sum diagonal,
 * dgsem_lobatto_2d.cu:160-179. */
static void build_D(int n, const double* x, double D[PMAX + 1][PMAX + 1])
{
    double b[PMAX + 1];
    for (int k = 0; k < n; k++) {
        b[k] = 1.0;
        for (int m = 0; m < n; m++)
            if (m != k) b[k] /= (x[k] - x[m]);
    }
    for (int i = 0; i < n; i++) {
        double diag = 0.0;
        for (int k = 0; k < n; k++) {
            if (k != i) {
                D[i][k] = (b[k] / b[i]) / (x[i] - x[k]);
                diag -= D[i][k];
            }
        }
        D[i][i] = diag;
    }
}

/* ── SD operators (fr_lobatto_2d.cu:76-166, :840-867) ────────────────────
 * Flux nodes xf = {-1, Gauss-Legendre nodes, +1}, p + 2 of them: the
 * reference's literal GL nodes for p <= 3 (:84-89); for p > 3 the roots of
 * L_p, Newton from Chebyshev guesses in long double.
 * Isf[k][j] = l_j(xf_k), the Lagrange basis of the solution nodes at the
 * flux nodes (:99-118).
 * Dfs[i][k] = l^f_k'(xs_i), the derivative of the flux nodes' Lagrange basis
 * at the solution nodes, the quantity the reference defines (:17-24,
 * :118-119).  Its construction (:120-166) is followed where xs_i is no flux
 * node: l^f_k(xs_i) * sum_{m != k} 1/(xs_i - xf_m).  It is not followed
 * where a solution node coincides with a flux node (the end nodes): there
 * the reference multiplies l^f_k(xs_i) = 0 by a sum holding 1/0 (NaN) and
 * gives the coincident entry the sign opposite to the derivative.  Here:
 * xs_i = xf_k: sum_{m != k} 1/(xf_k - xf_m); xs_i = xf_m, m != k:
 * prod_{l != k,m} (xf_m - xf_l) / prod_{l != k} (xf_k - xf_l).
 */
static const double ref_gl_interior[3][3] = {   /* fr_lobatto_2d.cu:84-89 */
    {0.0},
    {-0.5773502691896258, 0.5773502691896258},
    {-0.7745966692414834, 0.0, 0.7745966692414834},
};

static void gauss_legendre(int n, double* x)
{
    if (n >= 1 && n <= 3) {
        for (int i = 0; i < n; i++) x[i] = ref_gl_interior[n - 1][i];
        return;
    }
    const long double PI = 3.14159265358979323846264338327950288L;
    for (int j = 0; j < n; j++) {
        long double t = -cosl(PI * (j + 0.75L) / (n + 0.5L));
        for (int it = 0; it < 100; it++) {
            long double L, dL;
            legendre(n, t, &L, &dL);
            long double dt = L / dL;
            t -= dt;
            if (fabsl(dt) < 1e-19L) break;
        }
        x[j] = (double)t;
    }
    for (int j = 0; j < n / 2; j++) {   /* exact symmetry */
        double a = 0.5 * (x[n - 1 - j] - x[j]);
        x[j] = -a; x[n - 1 - j] = a;
    }
    if (n % 2 == 1) x[n / 2] = 0.0;
}

static double lagrange_val(int n, const double* xs, int j, double x)   /* :99-107 */
{
    double v = 1.0;
    for (int k = 0; k < n; k++)
        if (k != j) v *= (x - xs[k]) / (xs[j] - xs[k]);
    return v;
}

static void build_sd(Ctx& c)
{
    const int ns = c.NN, nf = c.NN + 1;
    c.NF = nf;
    c.xf[0] = -1.0;
    gauss_legendre(c.p, c.xf + 1);
    c.xf[nf - 1] = 1.0;
    for (int k = 0; k < nf; k++)                                     /* :109-118 */
        for (int j = 0; j < ns; j++) c.Isf[k][j] = lagrange_val(ns, c.xi, j, c.xf[k]);
    for (int i = 0; i < ns; i++) {
        int hit = -1;   /* flux node that coincides with xs_i */
        for (int m = 0; m < nf; m++)
            if (fabs(c.xi[i] - c.xf[m]) < 1e-14) hit = m;
        for (int k = 0; k < nf; k++) {
            double d = 0.0;
            if (hit < 0) {
                for (int m = 0; m < nf; m++)
                    if (m != k) d += 1.0 / (c.xi[i] - c.xf[m]);
                d *= lagrange_val(nf, c.xf, k, c.xi[i]);
            } else if (hit == k) {
                for (int m = 0; m < nf; m++)
                    if (m != k) d += 1.0 / (c.xf[k] - c.xf[m]);
            } else {
                double num = 1.0, den = 1.0;
                for (int l = 0; l < nf; l++) {
                    if (l == k) continue;
                    den *= c.xf[k] - c.xf[l];
                    if (l != hit) num *= c.xf[hit] - c.xf[l];
                }
                d = num / den;
            }
            c.Dfs[i][k] = d;
        }
    }
}

/* ── L2: pointwise physics (dgsem_lobatto_2d.cu:184-355), FP64 ─────────── */
struct Phys {
    double g;
    double sound_speed(double p, double rho) const {          /* :184-188 */
        double cs2 = g * p / rho;
        return sqrt(cs2 > 1e-14 ? cs2 : 1e-14);
    }
    void p2c(double r, double u, double v, double p, double cv[4]) const {  /* :190-196 */
        cv[0] = r; cv[1] = r * u; cv[2] = r * v;
        cv[3] = p / (g - 1.0) + 0.5 * r * (u * u + v * v);
    }
    void c2p(const double cv[4], double wp[4]) const {        /* :198-204 */
        wp[0] = cv[0];
        wp[1] = cv[1] / cv[0];
        wp[2] = cv[2] / cv[0];
        wp[3] = (g - 1.0) * (cv[3] - 0.5 * (cv[1] * cv[1] + cv[2] * cv[2]) / cv[0]);
    }
    double pressure_from_cons(const double U[4]) const {      /* :208-213 */
        double rho = std::max(U[0], 1e-14);
        double ke = 0.5 * (U[1] * U[1] + U[2] * U[2]) / rho;
        return (g - 1.0) * (U[3] - ke);
    }
    void sanitize_cons(double U[4]) const {                   /* :215-221 */
        U[0] = std::max(U[0], 1e-14);
        double p = pressure_from_cons(U);
        if (p < 1e-14)
            U[3] = 1e-14 / (g - 1.0) + 0.5 * (U[1] * U[1] + U[2] * U[2]) / U[0];
    }
    static void sanitize_prim(double W[4]) {                  /* :347-355 */
        W[0] = std::max(W[0], 1e-14);
        W[3] = std::max(W[3], 1e-14);
        if (W[0] < 1e-5) { W[1] = 0.0; W[2] = 0.0; return; }
        W[1] = std::max(std::min(W[1], UMAX), -UMAX);
        W[2] = std::max(std::min(W[2], UMAX), -UMAX);
    }
    void cons_to_prim_sane(const double U[4], double W[4]) const {  /* :223-229 */
        double Us[4] = {U[0], U[1], U[2], U[3]};
        sanitize_cons(Us);
        c2p(Us, W);
        sanitize_prim(W);
    }
    void euler_flux_n(double rho, double u, double v, double p,
                      double nx, double ny, double F[4]) const {  /* :232-241 */
        double un = u * nx + v * ny;
        double E = p / (g - 1.0) + 0.5 * rho * (u * u + v * v);
        F[0] = rho * un;
        F[1] = rho * un * u + p * nx;
        F[2] = rho * un * v + p * ny;
        F[3] = (E + p) * un;
    }
    static double logmean(double aL, double aR) {            /* :244-253 */
        double d = aL / aR;
        double f = (d - 1.0) / (d + 1.0);
        double u2 = f * f;
        double FF = (u2 < 1e-4)
                  ? (1.0 + u2 * (1.0 / 3.0 + u2 * (1.0 / 5.0 + u2 * (1.0 / 7.0))))
                  : (log(d) / (2.0 * f));
        return (aL + aR) / (2.0 * FF);
    }
    /* Chandrashekar EC flux, :259-285 */
    void ec_chandrashekar(const double WL[4], const double WR[4], int dir, double F[4]) const {
        double rL = WL[0], uL = WL[1], vL = WL[2], pL = WL[3];
        double rR = WR[0], uR = WR[1], vR = WR[2], pR = WR[3];
        double bL = 0.5 * rL / pL, bR = 0.5 * rR / pR;
        double r_ln = logmean(rL, rR);
        double b_ln = logmean(bL, bR);
        double r_av = 0.5 * (rL + rR);
        double b_av = 0.5 * (bL + bR);
        double u_av = 0.5 * (uL + uR);
        double v_av = 0.5 * (vL + vR);
        double p_hat = 0.5 * r_av / b_av;
        double vel2 = 0.5 * (uL * uL + vL * vL) + 0.5 * (uR * uR + vR * vR);
        double e_int = 0.5 * (1.0 / ((g - 1.0) * b_ln) - vel2);
        if (dir == 0) {
            double f1 = r_ln * u_av, f2 = f1 * u_av + p_hat, f3 = f1 * v_av;
            F[0] = f1; F[1] = f2; F[2] = f3; F[3] = f1 * e_int + f2 * u_av + f3 * v_av;
        } else {
            double g1 = r_ln * v_av, g2 = g1 * u_av, g3 = g1 * v_av + p_hat;
            F[0] = g1; F[1] = g2; F[2] = g3; F[3] = g1 * e_int + g2 * u_av + g3 * v_av;
        }
    }
    /* Ranocha EC + KEP flux (absent from the reference; closed form of
     * Ranocha 2018): rho_ln, (rho/p)_ln log means, arithmetic means of u, v, p,
     * E-flux f_rho*(1/2 (uL uR + vL vR) + 1/((g-1)(rho/p)_ln)) + 1/2(pL unR + pR unL). */
    void ec_ranocha(const double WL[4], const double WR[4], int dir, double F[4]) const {
        double rL = WL[0], uL = WL[1], vL = WL[2], pL = WL[3];
        double rR = WR[0], uR = WR[1], vR = WR[2], pR = WR[3];
        double r_ln = logmean(rL, rR);
        double z_ln = logmean(rL / pL, rR / pR);
        double u_av = 0.5 * (uL + uR);
        double v_av = 0.5 * (vL + vR);
        double p_av = 0.5 * (pL + pR);
        double v2 = 0.5 * (uL * uR + vL * vR);
        double eint = v2 + 1.0 / ((g - 1.0) * z_ln);
        if (dir == 0) {
            double f1 = r_ln * u_av;
            F[0] = f1; F[1] = f1 * u_av + p_av; F[2] = f1 * v_av;
            F[3] = f1 * eint + 0.5 * (pL * uR + pR * uL);
        } else {
            double g1 = r_ln * v_av;
            F[0] = g1; F[1] = g1 * u_av; F[2] = g1 * v_av + p_av;
            F[3] = g1 * eint + 0.5 * (pL * vR + pR * vL);
        }
    }
    /* HLLC in the rotated normal frame, :288-335 */
    void hllc_n(const double WL[4], const double WR[4], double nx, double ny, double F[4]) const {
        double rL = WL[0], uL = WL[1], vL = WL[2], pL = WL[3];
        double rR = WR[0], uR = WR[1], vR = WR[2], pR = WR[3];
        double unL = uL * nx + vL * ny, unR = uR * nx + vR * ny;
        double utL = -uL * ny + vL * nx, utR = -uR * ny + vR * nx;
        double cL = sound_speed(pL, rL), cR = sound_speed(pR, rR);
        double EL = pL / (g - 1.0) + 0.5 * rL * (uL * uL + vL * vL);
        double ER = pR / (g - 1.0) + 0.5 * rR * (uR * uR + vR * vR);
        double FnL[4] = {rL * unL, rL * unL * uL + pL * nx, rL * unL * vL + pL * ny, (EL + pL) * unL};
        double FnR[4] = {rR * unR, rR * unR * uR + pR * nx, rR * unR * vR + pR * ny, (ER + pR) * unR};
        double SL = std::min(unL - cL, unR - cR);
        double SR = std::max(unL + cL, unR + cR);
        if (SL >= 0.0) {
            for (int q = 0; q < 4; q++) F[q] = FnL[q];
        } else if (SR <= 0.0) {
            for (int q = 0; q < 4; q++) F[q] = FnR[q];
        } else {
            double dL = rL * (SL - unL), dR = rR * (SR - unR);
            double Ss = (pR - pL + dL * unL - dR * unR) / (dL - dR);
            bool lft = (Ss >= 0.0);
            double rK = lft ? rL : rR, unK = lft ? unL : unR;
            double utK = lft ? utL : utR, pK = lft ? pL : pR;
            double EK = lft ? EL : ER, SK = lft ? SL : SR;
            double fact = rK * (SK - unK) / (SK - Ss);
            double Us[4];
            Us[0] = fact;
            Us[1] = fact * Ss;
            Us[2] = fact * utK;
            Us[3] = fact * (EK / rK + (Ss - unK) * (Ss + pK / (rK * (SK - unK))));
            double FK[4] = {rK * unK, rK * unK * unK + pK, rK * unK * utK, (EK + pK) * unK};
            double UcK[4] = {rK, rK * unK, rK * utK, EK};
            double Fl[4];
            for (int q = 0; q < 4; q++) Fl[q] = FK[q] + SK * (Us[q] - UcK[q]);
            F[0] = Fl[0];
            F[1] = Fl[1] * nx - Fl[2] * ny;
            F[2] = Fl[1] * ny + Fl[2] * nx;
            F[3] = Fl[3];
        }
    }
    /* Rusanov / local Lax-Friedrichs (absent from the reference; closed form):
     * 1/2 (f_L + f_R).n - 1/2 lam (U_R - U_L), lam = max(|un|+c) of both sides. */
    void rusanov_n(const double WL[4], const double WR[4], double nx, double ny, double F[4]) const {
        double FL[4], FR[4], UL[4], UR[4];
        euler_flux_n(WL[0], WL[1], WL[2], WL[3], nx, ny, FL);
        euler_flux_n(WR[0], WR[1], WR[2], WR[3], nx, ny, FR);
        p2c(WL[0], WL[1], WL[2], WL[3], UL);
        p2c(WR[0], WR[1], WR[2], WR[3], UR);
        double cL = sound_speed(WL[3], WL[0]), cR = sound_speed(WR[3], WR[0]);
        double lL = fabs(WL[1] * nx + WL[2] * ny) + cL;
        double lR = fabs(WR[1] * nx + WR[2] * ny) + cR;
        double lam = std::max(lL, lR);
        for (int q = 0; q < 4; q++) F[q] = 0.5 * (FL[q] + FR[q]) - 0.5 * lam * (UR[q] - UL[q]);
    }
    static void wall_state(const double W[4], int dir, double Wg[4]) {  /* :339-344 */
        Wg[0] = W[0]; Wg[1] = W[1]; Wg[2] = W[2]; Wg[3] = W[3];
        if (dir == 0) Wg[1] = -W[1]; else Wg[2] = -W[2];
    }
};

inline size_t sidx(const Ctx& c, int var, int el, int jn, int in) {   /* :126 */
    return (size_t)var * c.NEL * c.NN2 + (size_t)el * c.NN2 + (size_t)jn * c.NN + in;
}

static void surface_flux(const Ctx& c, const Phys& ph, const double WL[4], const double WR[4],
                         double nx, double ny, double F[4])
{
    if (c.cfg.surface_flux == DG_SURF_RUSANOV) ph.rusanov_n(WL, WR, nx, ny, F);
    else ph.hllc_n(WL, WR, nx, ny, F);
}

static void volume_flux(const Ctx& c, const Phys& ph, const double WL[4], const double WR[4],
                        int dir, double F[4])
{
    if (c.cfg.volume_flux == DG_VOL_RANOCHA) ph.ec_ranocha(WL, WR, dir, F);
    else ph.ec_chandrashekar(WL, WR, dir, F);
}

/* Load neighbour node (jn,in) of element nb and convert to sanitized primitives. */
static void nb_prim(const Ctx& c, const Phys& ph, const double* Q, int nb, int jn, int in, double W[4])
{
    double U[4] = {Q[sidx(c, 0, nb, jn, in)], Q[sidx(c, 1, nb, jn, in)],
                   Q[sidx(c, 2, nb, jn, in)], Q[sidx(c, 3, nb, jn, in)]};
    ph.cons_to_prim_sane(U, W);
}

/* ── L3: compute_dg_rhs, dgsem_lobatto_2d.cu:365-530, one element ───────── */
static void rhs_element(const Ctx& c, const double* Q, double* RHS, double* lam, int el)
{
    const Phys ph{c.cfg.gamma};
    const int N = c.K, nn = c.NN, nn2 = c.NN2;
    const double h = c.h;
    const double jac = 2.0 / h;
    const double len_p = h / (2 * c.p + 1);
    const double dx2inv = 1.0 / (len_p * len_p);
    const int ej = el / N, ei = el % N;
    const bool periodic = (c.cfg.bc == DG_BC_PERIODIC);

    /* Phase 1: prims, sanitized cons, physical fluxes (:396-409) */
    std::vector<double> sW(4 * nn2), sUc(4 * nn2), sA(4 * nn2), sB(4 * nn2);
    for (int nd = 0; nd < nn2; nd++) {
        int jn = nd / nn, in = nd % nn;
        double U[4] = {Q[sidx(c, 0, el, jn, in)], Q[sidx(c, 1, el, jn, in)],
                       Q[sidx(c, 2, el, jn, in)], Q[sidx(c, 3, el, jn, in)]};
        double Wt[4]; ph.cons_to_prim_sane(U, Wt);
        double Ut[4]; ph.p2c(Wt[0], Wt[1], Wt[2], Wt[3], Ut);
        double F[4], G[4];
        ph.euler_flux_n(Wt[0], Wt[1], Wt[2], Wt[3], 1.0, 0.0, F);
        ph.euler_flux_n(Wt[0], Wt[1], Wt[2], Wt[3], 0.0, 1.0, G);
        for (int q = 0; q < 4; q++) {
            sW[q * nn2 + nd] = Wt[q]; sUc[q * nn2 + nd] = Ut[q];
            sA[q * nn2 + nd] = F[q]; sB[q * nn2 + nd] = G[q];
        }
    }
    auto Wat = [&](int nd, double W[4]) { for (int q = 0; q < 4; q++) W[q] = sW[q * nn2 + nd]; };

    std::vector<double> R(4 * nn2), theta(nn2), lamn(nn2);
    for (int nd = 0; nd < nn2; nd++) {
        int jn = nd / nn, in = nd % nn;
        double Rn[4] = {0, 0, 0, 0};
        /* Phase 2: volume term (:417-439) */
        if (c.cfg.volume_flux != DG_VOL_COLLOCATION) {
            double Wi[4]; Wat(nd, Wi);
            double ax[4] = {0, 0, 0, 0}, ay[4] = {0, 0, 0, 0};
            for (int k = 0; k < nn; k++) {
                double Wx[4], Wy[4]; Wat(jn * nn + k, Wx); Wat(k * nn + in, Wy);
                double Fs[4], Gs[4];
                volume_flux(c, ph, Wi, Wx, 0, Fs);
                volume_flux(c, ph, Wi, Wy, 1, Gs);
                double dix = c.D[in][k], djy = c.D[jn][k];
                for (int q = 0; q < 4; q++) { ax[q] += dix * Fs[q]; ay[q] += djy * Gs[q]; }
            }
            for (int q = 0; q < 4; q++) Rn[q] = -jac * (2.0 * ax[q] + 2.0 * ay[q]);
        } else {
            for (int q = 0; q < 4; q++) {
                double dfx = 0.0, dgy = 0.0;
                for (int k = 0; k < nn; k++) {
                    dfx += c.D[in][k] * sA[q * nn2 + jn * nn + k];
                    dgy += c.D[jn][k] * sB[q * nn2 + k * nn + in];
                }
                Rn[q] = -jac * (dfx + dgy);
            }
        }
        /* Phase 3: surface lift (:441-481) */
        if (in == 0) {
            double WRi[4]; Wat(jn * nn + 0, WRi);
            double WLn[4];
            if (ei > 0) nb_prim(c, ph, Q, ej * N + (ei - 1), jn, nn - 1, WLn);
            else if (periodic) nb_prim(c, ph, Q, ej * N + (N - 1), jn, nn - 1, WLn);
            else Phys::wall_state(WRi, 0, WLn);
            double fL[4]; surface_flux(c, ph, WLn, WRi, 1.0, 0.0, fL);
            for (int q = 0; q < 4; q++) Rn[q] += jac * c.winv[0] * (fL[q] - sA[q * nn2 + jn * nn + 0]);
        }
        if (in == nn - 1) {
            double WLi[4]; Wat(jn * nn + nn - 1, WLi);
            double WRn[4];
            if (ei < N - 1) nb_prim(c, ph, Q, ej * N + (ei + 1), jn, 0, WRn);
            else if (periodic) nb_prim(c, ph, Q, ej * N + 0, jn, 0, WRn);
            else Phys::wall_state(WLi, 0, WRn);
            double fR[4]; surface_flux(c, ph, WLi, WRn, 1.0, 0.0, fR);
            for (int q = 0; q < 4; q++) Rn[q] -= jac * c.winv[nn - 1] * (fR[q] - sA[q * nn2 + jn * nn + nn - 1]);
        }
        if (jn == 0) {
            double WTi[4]; Wat(in, WTi);
            double WBn[4];
            if (ej > 0) nb_prim(c, ph, Q, (ej - 1) * N + ei, nn - 1, in, WBn);
            else if (periodic) nb_prim(c, ph, Q, (N - 1) * N + ei, nn - 1, in, WBn);
            else Phys::wall_state(WTi, 1, WBn);
            double gB[4]; surface_flux(c, ph, WBn, WTi, 0.0, 1.0, gB);
            for (int q = 0; q < 4; q++) Rn[q] += jac * c.winv[0] * (gB[q] - sB[q * nn2 + in]);
        }
        if (jn == nn - 1) {
            double WBi[4]; Wat((nn - 1) * nn + in, WBi);
            double WTn[4];
            if (ej < N - 1) nb_prim(c, ph, Q, (ej + 1) * N + ei, 0, in, WTn);
            else if (periodic) nb_prim(c, ph, Q, 0 * N + ei, 0, in, WTn);
            else Phys::wall_state(WBi, 1, WTn);
            double gT[4]; surface_flux(c, ph, WBi, WTn, 0.0, 1.0, gT);
            for (int q = 0; q < 4; q++) Rn[q] -= jac * c.winv[nn - 1] * (gT[q] - sB[q * nn2 + (nn - 1) * nn + in]);
        }
        /* Phase 3.5: wave speed + sensor (:483-492) */
        double rho = std::max(sW[0 * nn2 + nd], 1e-12), c2 = ph.g * sW[3 * nn2 + nd] / rho;
        double cs = sqrt(std::max(c2, 1e-14));
        lamn[nd] = fabs(sW[1 * nn2 + nd]) + fabs(sW[2 * nn2 + nd]) + cs;
        theta[nd] = 0.0;
        if (c.cfg.av) {
            double du = 0.0, dv = 0.0;
            for (int k = 0; k < nn; k++) {
                du += c.D[in][k] * sW[1 * nn2 + jn * nn + k];
                dv += c.D[jn][k] * sW[2 * nn2 + k * nn + in];
            }
            double divu = std::min(jac * (du + dv), 0.0), du2 = divu * divu;
            theta[nd] = du2 / (du2 + c.cfg.av_ksensor * c2 * dx2inv + 1e-30);
        }
        for (int q = 0; q < 4; q++) R[q * nn2 + nd] = Rn[q];
    }
    /* per-element max (:495-499) */
    double lam_e = 0.0, theta_e = 0.0;
    for (int k = 0; k < nn2; k++) { theta_e = std::max(theta_e, theta[k]); lam_e = std::max(lam_e, lamn[k]); }

    if (c.cfg.av) {   /* :501-521 */
        double nu_e = c.cfg.av_cav * theta_e * len_p * lam_e;
        for (int nd = 0; nd < nn2; nd++) {
            int jn = nd / nn, in = nd % nn;
            double gx[4] = {0, 0, 0, 0}, gy[4] = {0, 0, 0, 0};
            for (int k = 0; k < nn; k++) for (int q = 0; q < 4; q++) {
                gx[q] += c.D[in][k] * sUc[q * nn2 + jn * nn + k];
                gy[q] += c.D[jn][k] * sUc[q * nn2 + k * nn + in];
            }
            for (int q = 0; q < 4; q++) { sA[q * nn2 + nd] = nu_e * jac * gx[q]; sB[q * nn2 + nd] = nu_e * jac * gy[q]; }
        }
        for (int nd = 0; nd < nn2; nd++) {
            int jn = nd / nn, in = nd % nn;
            for (int q = 0; q < 4; q++) {
                double sx = 0.0, sy = 0.0;
                for (int k = 0; k < nn; k++) {
                    sx += c.w[k] * c.D[k][in] * sA[q * nn2 + jn * nn + k];
                    sy += c.w[k] * c.D[k][jn] * sB[q * nn2 + k * nn + in];
                }
                R[q * nn2 + nd] -= jac * (c.winv[in] * sx + c.winv[jn] * sy);
            }
        }
    }
    for (int nd = 0; nd < nn2; nd++) {
        int jn = nd / nn, in = nd % nn;
        for (int q = 0; q < 4; q++) RHS[sidx(c, q, el, jn, in)] = R[q * nn2 + nd];
    }
    if (lam) lam[el] = lam_e;
}

/* ── compute_sd_rhs, fr_lobatto_2d.cu:332-525, one element ──────────────
 * Per grid line: the conservative state interpolated to the interior flux
 * nodes, converted (cons_to_prim_sane) and its physical flux evaluated; the
 * Riemann flux on the two end flux nodes between the element's own
 * sanitized end-node prims and the neighbour's; dF/dxi at the solution nodes
 * through Dfs, subtracted with the factor 2/h.  x-lines, then y-lines.  The
 * element wave speed is max |u| + |v| + c over the sanitized prims. */
static void sd_rhs_element(const Ctx& c, const double* Q, double* RHS, double* lam, int el)
{
    const Phys ph{c.cfg.gamma};
    const int N = c.K, nn = c.NN, nf = c.NF, nn2 = c.NN2;
    const double jac = 2.0 / c.h;
    const int ej = el / N, ei = el % N;
    const bool periodic = (c.cfg.bc == DG_BC_PERIODIC);

    std::vector<double> Qc(4 * nn2), Qe(4 * nn2), R(4 * nn2, 0.0);   /* :349-365 */
    for (int n = 0; n < nn2; n++) {
        double U[4], W[4];
        for (int q = 0; q < 4; q++) U[q] = Q[sidx(c, q, el, n / nn, n % nn)];
        ph.cons_to_prim_sane(U, W);
        for (int q = 0; q < 4; q++) { Qc[q * nn2 + n] = U[q]; Qe[q * nn2 + n] = W[q]; }
    }

    for (int dir = 0; dir < 2; dir++) {
        const double nx = dir == 0 ? 1.0 : 0.0, ny = dir == 0 ? 0.0 : 1.0;
        for (int line = 0; line < nn; line++) {
            /* node k of the line: x-lines (:367-438) run along in, y-lines (:440-511) along jn */
            auto node = [&](int k) { return dir == 0 ? line * nn + k : k * nn + line; };
            std::vector<double> Ff(4 * nf);
            for (int k = 1; k < nf - 1; k++) {                   /* :370-390 */
                double Uk[4], Wk[4], F4[4];
                for (int q = 0; q < 4; q++) {
                    double v = 0.0;
                    for (int j = 0; j < nn; j++) v += c.Isf[k][j] * Qc[q * nn2 + node(j)];
                    Uk[q] = v;
                }
                ph.cons_to_prim_sane(Uk, Wk);
                ph.euler_flux_n(Wk[0], Wk[1], Wk[2], Wk[3], nx, ny, F4);
                for (int q = 0; q < 4; q++) Ff[q * nf + k] = F4[q];
            }
            double Wm[4], Wp[4];                                    /* :392-397 */
            for (int q = 0; q < 4; q++) { Wm[q] = Qe[q * nn2 + node(0)]; Wp[q] = Qe[q * nn2 + node(nn - 1)]; }
            Phys::sanitize_prim(Wm); Phys::sanitize_prim(Wp);
            const int pos = dir == 0 ? ei : ej;
            {   /* low face (:399-413, :472-486) */
                double Wn[4], F[4];
                if (pos > 0 || periodic) {
                    int pn = pos > 0 ? pos - 1 : N - 1;
                    if (dir == 0) nb_prim(c, ph, Q, ej * N + pn, line, nn - 1, Wn);
                    else nb_prim(c, ph, Q, pn * N + ei, nn - 1, line, Wn);
                } else {
                    Phys::wall_state(Wm, dir, Wn);
                }
                surface_flux(c, ph, Wn, Wm, nx, ny, F);
                for (int q = 0; q < 4; q++) Ff[q * nf + 0] = F[q];
            }
            {   /* high face (:414-428, :487-501) */
                double Wn[4], F[4];
                if (pos < N - 1 || periodic) {
                    int pn = pos < N - 1 ? pos + 1 : 0;
                    if (dir == 0) nb_prim(c, ph, Q, ej * N + pn, line, 0, Wn);
                    else nb_prim(c, ph, Q, pn * N + ei, 0, line, Wn);
                } else {
                    Phys::wall_state(Wp, dir, Wn);
                }
                surface_flux(c, ph, Wp, Wn, nx, ny, F);
                for (int q = 0; q < 4; q++) Ff[q * nf + nf - 1] = F[q];
            }
            for (int i = 0; i < nn; i++)                           /* :430-437 */
                for (int q = 0; q < 4; q++) {
                    double dF = 0.0;
                    for (int k = 0; k < nf; k++) dF += c.Dfs[i][k] * Ff[q * nf + k];
                    R[q * nn2 + node(i)] -= jac * dF;
                }
        }
    }

    double lam_el = 0.0;                                            /* :513-522 */
    for (int n = 0; n < nn2; n++) {
        double rho = Qe[n], u = Qe[nn2 + n], v = Qe[2 * nn2 + n], p = Qe[3 * nn2 + n];
        lam_el = std::max(lam_el, fabs(u) + fabs(v) + ph.sound_speed(p, rho));
        for (int q = 0; q < 4; q++) RHS[sidx(c, q, el, n / nn, n % nn)] = R[q * nn2 + n];
    }
    if (lam) lam[el] = lam_el;
}

static void compute_rhs(const Ctx& c, const double* Q, double* RHS, double* lam)
{
    const bool sd = c.cfg.scheme == DG_SCHEME_SD;
#pragma omp parallel for schedule(static)
    for (int el = 0; el < c.NEL; el++) {
        if (sd) sd_rhs_element(c, Q, RHS, lam, el);
        else rhs_element(c, Q, RHS, lam, el);
    }
}

/* ── nodal positivity repair of the SD sibling, fr_lobatto_2d.cu:589-615 ── */
static void sd_repair(const Ctx& c, double* Q)
{
    const Phys ph{c.cfg.gamma};
    const double eps_rho = c.cfg.eps_rho, eps_p = c.cfg.eps_p;
    const size_t S = (size_t)c.NEL * c.NN2;
#pragma omp parallel for schedule(static)
    for (long long n = 0; n < (long long)S; n++) {
        double U[4];
        for (int q = 0; q < 4; q++) U[q] = Q[q * S + n];
        if (U[0] < eps_rho) { U[0] = eps_rho; U[1] = 0.0; U[2] = 0.0; }
        double p = ph.pressure_from_cons(U);
        if (p < eps_p) U[3] = eps_p / (ph.g - 1.0) + 0.5 * (U[1] * U[1] + U[2] * U[2]) / U[0];
        for (int q = 0; q < 4; q++) Q[q * S + n] = U[q];
    }
}

/* ── SSP-RK3 stages, :533-576 ─────────────────────────────────────────── */
static void rk3_stage(const Ctx& c, int s, double* Uout, const double* U0, const double* Up,
                      const double* L, double dt)
{
    const Phys ph{c.cfg.gamma};
    const size_t ntot = (size_t)c.NEL * c.NN2, str = ntot;
#pragma omp parallel for schedule(static)
    for (long long n = 0; n < (long long)ntot; n++) {
        double Uc[4];
        for (int q = 0; q < 4; q++) {
            size_t ix = q * str + n;
            if (s == 1) Uc[q] = U0[ix] + dt * L[ix];
            else if (s == 2) Uc[q] = 0.75 * U0[ix] + 0.25 * (Up[ix] + dt * L[ix]);
            else Uc[q] = (1.0 / 3.0) * U0[ix] + (2.0 / 3.0) * (Up[ix] + dt * L[ix]);
        }
        ph.sanitize_cons(Uc);
        for (int q = 0; q < 4; q++) Uout[q * str + n] = Uc[q];
    }
}

/* ── Zhang-Shu positivity limiter, :584-643 ───────────────────────────── */
static void zs_limiter(const Ctx& c, double* Q)
{
    const Phys ph{c.cfg.gamma};
    const int nn = c.NN, nn2 = c.NN2;
    const double eps_rho = c.cfg.eps_rho, eps_p = c.cfg.eps_p;
#pragma omp parallel for schedule(static)
    for (int el = 0; el < c.NEL; el++) {
        double Ubar[4] = {0, 0, 0, 0};
        for (int jn = 0; jn < nn; jn++) for (int in = 0; in < nn; in++) {
            double wij = 0.25 * c.w[in] * c.w[jn];
            for (int q = 0; q < 4; q++) Ubar[q] += wij * Q[sidx(c, q, el, jn, in)];
        }
        Ubar[0] = std::max(Ubar[0], eps_rho);
        {
            double p = ph.pressure_from_cons(Ubar);
            if (p < eps_p)
                Ubar[3] = eps_p / (ph.g - 1.0) + 0.5 * (Ubar[1] * Ubar[1] + Ubar[2] * Ubar[2]) / Ubar[0];
        }
        double rho_min = 1e30;
        for (int n = 0; n < nn2; n++) rho_min = std::min(rho_min, Q[sidx(c, 0, el, n / nn, n % nn)]);
        double t_rho = 1.0;
        if (rho_min < eps_rho) t_rho = (Ubar[0] - eps_rho) / std::max(Ubar[0] - rho_min, 1e-30);
        t_rho = std::max(0.0, std::min(1.0, t_rho));
        double theta = t_rho;
        for (int n = 0; n < nn2; n++) {
            double U[4];
            for (int q = 0; q < 4; q++) U[q] = Q[sidx(c, q, el, n / nn, n % nn)];
            double Ud[4];
            for (int q = 0; q < 4; q++) Ud[q] = Ubar[q] + t_rho * (U[q] - Ubar[q]);
            double p = ph.pressure_from_cons(Ud);
            if (p < eps_p) {
                double lo = 0.0, hi = t_rho;
                for (int it = 0; it < 20; it++) {
                    double tm = 0.5 * (lo + hi);
                    double Um[4];
                    for (int q = 0; q < 4; q++) Um[q] = Ubar[q] + tm * (U[q] - Ubar[q]);
                    if (ph.pressure_from_cons(Um) >= eps_p) lo = tm; else hi = tm;
                }
                theta = std::min(theta, lo);
            }
        }
        for (int jn = 0; jn < nn; jn++) for (int in = 0; in < nn; in++)
            for (int q = 0; q < 4; q++) {
                double u = Q[sidx(c, q, el, jn, in)];
                Q[sidx(c, q, el, jn, in)] = Ubar[q] + theta * (u - Ubar[q]);
            }
    }
}

/* the scheme's positivity step: Zhang-Shu (DGSEM) or the nodal repair (SD) */
static void limiter(const Ctx& c, double* Q)
{
    if (c.cfg.scheme == DG_SCHEME_SD) sd_repair(c, Q);
    else zs_limiter(c, Q);
}

static double max_lambda(const Ctx& c, const double* lam)
{
    double v = 0.0;   /* reduce_max starts from 0, :647-648 */
    for (int i = 0; i < c.NEL; i++) v = std::max(v, lam[i]) ;
    /* std::max drops NaN like fmaxf; propagate it so the caller can stop */
    for (int i = 0; i < c.NEL; i++) if (std::isnan(lam[i])) return lam[i];
    return v;
}

/* ── ICs ──────────────────────────────────────────────────────────────── */
static inline double node_x(const Ctx& c, int ei, int in) {
    /* x0 + h*(ei + (xi+1)/2): evaluated without contraction on both sides */
    return c.cfg.x0 + c.h * ((double)ei + 0.5 * (c.xi[in] + 1.0));
}
static inline double node_y(const Ctx& c, int ej, int jn) {
    return c.cfg.y0 + c.h * ((double)ej + 0.5 * (c.xi[jn] + 1.0));
}
static inline uint64_t splitmix64(uint64_t z) {
    /* SplitMix64 finaliser (Steele, Lea, Flood 2014) */
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

}  // namespace

/* ═════════════════════════════ exported API ═════════════════════════════ */

OR_API int oracle_create(const dg_config* cfg, void** out)
{
    if (!cfg || !out || cfg->K < 1 || cfg->p < 1 || cfg->p > PMAX || !(cfg->L > 0)) return DG_ERR_INVALID_ARG;
    Ctx* c = new Ctx();
    c->cfg = *cfg;
    c->K = cfg->K; c->p = cfg->p; c->NN = cfg->p + 1; c->NN2 = c->NN * c->NN; c->NEL = cfg->K * cfg->K;
    c->h = cfg->L / cfg->K;
    lgl(c->p, c->xi, c->w);
    for (int i = 0; i < c->NN; i++) c->winv[i] = 1.0 / c->w[i];
    build_D(c->NN, c->xi, c->D);
    build_sd(*c);
    size_t n = (size_t)4 * c->NEL * c->NN2;
    c->Q0.resize(n); c->Q1.resize(n); c->Q2.resize(n); c->RHS.resize(n); c->lam.resize(c->NEL);
    *out = c;
    return DG_OK;
}

OR_API void oracle_destroy(void* ctx) { delete (Ctx*)ctx; }

OR_API int oracle_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n; return 1;
#endif
}

OR_API void oracle_lgl(int p, double* x, double* w) { lgl(p, x, w); }

/* SD operators of a context: xf[NF], Isf[NF][NN], Dfs[NN][NF] (row-major) */
OR_API void oracle_sd_operators(void* ctx, double* xf, double* Isf, double* Dfs)
{
    const Ctx& c = *(Ctx*)ctx;
    for (int k = 0; k < c.NF; k++) {
        if (xf) xf[k] = c.xf[k];
        for (int j = 0; j < c.NN; j++) if (Isf) Isf[k * c.NN + j] = c.Isf[k][j];
    }
    for (int i = 0; i < c.NN; i++)
        for (int k = 0; k < c.NF; k++) if (Dfs) Dfs[i * c.NF + k] = c.Dfs[i][k];
}

OR_API void oracle_build_D(int p, double* Dout)
{
    double x[PMAX + 1], w[PMAX + 1], D[PMAX + 1][PMAX + 1];
    lgl(p, x, w);
    build_D(p + 1, x, D);
    for (int i = 0; i <= p; i++) for (int k = 0; k <= p; k++) Dout[i * (p + 1) + k] = D[i][k];
}

OR_API double oracle_logmean(double a, double b) { return Phys::logmean(a, b); }

OR_API void oracle_volume_flux(int kind, double gamma, const double* WL, const double* WR, int dir, double* F)
{
    Phys ph{gamma};
    if (kind == DG_VOL_RANOCHA) ph.ec_ranocha(WL, WR, dir, F);
    else ph.ec_chandrashekar(WL, WR, dir, F);
}

OR_API void oracle_surface_flux(int kind, double gamma, const double* WL, const double* WR,
                                double nx, double ny, double* F)
{
    Phys ph{gamma};
    if (kind == DG_SURF_RUSANOV) ph.rusanov_n(WL, WR, nx, ny, F);
    else ph.hllc_n(WL, WR, nx, ny, F);
}

OR_API void oracle_cons_to_prim(double gamma, const double* U, double* W)
{
    Phys ph{gamma};
    ph.cons_to_prim_sane(U, W);
}

OR_API void oracle_physical_flux(double gamma, const double* W, double nx, double ny, double* F)
{
    Phys ph{gamma};
    ph.euler_flux_n(W[0], W[1], W[2], W[3], nx, ny, F);
}

/* Isentropic vortex closed form, q1_dg_euler_2d.cu:2166-2188 (centre and
 * strength as parameters). W = (rho, u, v, p). */
OR_API void oracle_vortex_exact(double gamma, double eps, double xc, double yc, double u0, double v0,
                                double x, double y, double* W)
{
    const double pi = 3.14159265358979323846;
    double dx = x - xc, dy = y - yc;
    double r2 = dx * dx + dy * dy;
    double f = eps / (2.0 * pi) * exp(0.5 * (1.0 - r2));
    double u = u0 - f * dy;
    double v = v0 + f * dx;
    double dT = -(gamma - 1.0) * eps * eps / (8.0 * gamma * pi * pi) * exp(1.0 - r2);
    double T = 1.0 + dT;
    if (T < 1e-6) T = 1e-6;
    W[0] = pow(T, 1.0 / (gamma - 1.0));
    W[1] = u; W[2] = v;
    W[3] = pow(T, gamma / (gamma - 1.0));
}

OR_API double oracle_kh_noise(unsigned long long seed, long long node, double amp)
{
    uint64_t z = splitmix64((uint64_t)seed + ((uint64_t)node + 1ULL) * 0x9e3779b97f4a7c15ULL);
    double u = (double)(z >> 11) * (1.0 / 9007199254740992.0);
    return amp * (2.0 * u - 1.0);
}

OR_API int oracle_init_state(void* ctx, const dg_ic_params* ic, double* Q)
{
    Ctx& c = *(Ctx*)ctx;
    Phys ph{c.cfg.gamma};
    const int N = c.K, nn = c.NN;
    for (int el = 0; el < c.NEL; el++) {
        int ej = el / N, ei = el % N;
        for (int jn = 0; jn < nn; jn++) for (int in = 0; in < nn; in++) {
            double x = node_x(c, ei, in), y = node_y(c, ej, jn);
            double W[4];
            if (ic->kind == DG_IC_CIRCULAR_SOD) {       /* :692-717 */
                double cx = c.cfg.x0 + 0.5 * c.cfg.L, cy = c.cfg.y0 + 0.5 * c.cfg.L;
                double r = sqrt((x - cx) * (x - cx) + (y - cy) * (y - cy));
                /* front width: 0.5h (dgsem_lobatto_2d.cu:710), 2h (fr_lobatto_2d.cu:648) */
                double delta = (c.cfg.scheme == DG_SCHEME_SD ? 2.0 : 0.5) * c.h;
                double phi = 0.5 * (1.0 + tanh((0.25 * c.cfg.L - r) / delta));
                W[0] = 0.125 + (ic->rho_hi - 0.125) * phi; W[1] = 0.0; W[2] = 0.0;
                W[3] = 0.1 + (ic->p_hi - 0.1) * phi;
            } else if (ic->kind == DG_IC_VORTEX) {
                oracle_vortex_exact(ph.g, ic->vortex_eps, ic->xc, ic->yc, ic->u0, ic->v0, x, y, W);
            } else {                                  /* Kelvin-Helmholtz step */
                const double pi = 3.14159265358979323846;
                long long node = (long long)el * c.NN2 + jn * nn + in;
                bool inner = fabs(y) < 0.5;
                W[0] = inner ? 2.0 : 1.0;
                W[1] = inner ? 0.5 : -0.5;
                W[2] = 0.1 * sin(2.0 * pi * x) + oracle_kh_noise(ic->seed, node, ic->noise_amp);
                W[3] = 2.5;
            }
            double U[4]; ph.p2c(W[0], W[1], W[2], W[3], U);
            for (int q = 0; q < 4; q++) Q[sidx(c, q, el, jn, in)] = U[q];
        }
    }
    /* the SD driver applies no limiter to its IC (fr_lobatto_2d.cu:922-937) */
    if (c.cfg.limiter && c.cfg.scheme != DG_SCHEME_SD) limiter(c, Q);
    return DG_OK;
}

OR_API int oracle_compute_rhs(void* ctx, const double* Q, double* RHS, double* lam)
{
    compute_rhs(*(Ctx*)ctx, Q, RHS, lam);
    return DG_OK;
}

OR_API int oracle_rk3_stage(void* ctx, int s, double* Uout, const double* U0, const double* Up,
                            const double* L, double dt)
{
    Ctx& c = *(Ctx*)ctx;
    rk3_stage(c, s, Uout, U0, Up, L, dt);
    if (c.cfg.limiter) limiter(c, Uout);
    return DG_OK;
}

OR_API int oracle_limit(void* ctx, double* Q) { limiter(*(Ctx*)ctx, Q); return DG_OK; }

OR_API double oracle_max_lambda(void* ctx, const double* lam) { return max_lambda(*(Ctx*)ctx, lam); }

/* The reference loop, :1022-1058, for `nsteps` steps (t_end <= 0: no clip). */
OR_API int oracle_advance(void* ctx, double* Q, int nsteps, double cfl, double t_end,
                          double* t_inout, double* dt_last)
{
    Ctx& c = *(Ctx*)ctx;
    size_t n = (size_t)4 * c.NEL * c.NN2;
    double t = t_inout ? *t_inout : 0.0;
    for (int s = 0; s < nsteps; s++) {
        if (t_end > 0.0 && t >= t_end - 1e-12 * std::max(1.0, fabs(t_end))) break;
        compute_rhs(c, Q, c.RHS.data(), c.lam.data());
        double lam_max = max_lambda(c, c.lam.data());
        if (!(lam_max > 0.0)) return DG_ERR_NONFINITE;
        double dt = cfl * c.h / (lam_max * c.NN);
        if (t_end > 0.0 && t + dt > t_end) dt = t_end - t;
        if (dt < 1e-14) return DG_ERR_DT_UNDERFLOW;
        std::memcpy(c.Q0.data(), Q, n * sizeof(double));
        rk3_stage(c, 1, c.Q1.data(), c.Q0.data(), nullptr, c.RHS.data(), dt);
        if (c.cfg.limiter) limiter(c, c.Q1.data());
        compute_rhs(c, c.Q1.data(), c.RHS.data(), c.lam.data());
        rk3_stage(c, 2, c.Q2.data(), c.Q0.data(), c.Q1.data(), c.RHS.data(), dt);
        if (c.cfg.limiter) limiter(c, c.Q2.data());
        compute_rhs(c, c.Q2.data(), c.RHS.data(), c.lam.data());
        rk3_stage(c, 3, Q, c.Q0.data(), c.Q2.data(), c.RHS.data(), dt);
        if (c.cfg.limiter) limiter(c, Q);
        lam_max = max_lambda(c, c.lam.data());
        if (!(lam_max > 0.0) || lam_max > 1e6) return DG_ERR_BLOWUP;
        t += dt;
        if (t_inout) *t_inout = t;
        if (dt_last) *dt_last = dt;
    }
    return DG_OK;
}

/* Diagnostic sensor, :662-689 (divergence not clipped, unlike the RHS). */
OR_API int oracle_sensor(void* ctx, const double* Q, double* sensor)
{
    Ctx& c = *(Ctx*)ctx;
    Phys ph{c.cfg.gamma};
    const int nn = c.NN, nn2 = c.NN2;
    const double jac = 2.0 / c.h, len_p = c.h / (2 * c.p + 1), dx2inv = 1.0 / (len_p * len_p);
    std::vector<double> u(nn2), v(nn2), c2(nn2);
    for (int el = 0; el < c.NEL; el++) {
        for (int n = 0; n < nn2; n++) {
            double U[4] = {Q[sidx(c, 0, el, n / nn, n % nn)], Q[sidx(c, 1, el, n / nn, n % nn)],
                           Q[sidx(c, 2, el, n / nn, n % nn)], Q[sidx(c, 3, el, n / nn, n % nn)]};
            double W[4]; ph.cons_to_prim_sane(U, W);
            u[n] = W[1]; v[n] = W[2]; c2[n] = ph.g * W[3] / std::max(W[0], 1e-12);
        }
        double theta_e = 0.0;
        for (int jn = 0; jn < nn; jn++) for (int in = 0; in < nn; in++) {
            double du = 0.0, dv = 0.0;
            for (int k = 0; k < nn; k++) { du += c.D[in][k] * u[jn * nn + k]; dv += c.D[jn][k] * v[k * nn + in]; }
            double divu = jac * (du + dv), du2 = divu * divu;
            double th = du2 / (du2 + c.cfg.av_ksensor * c2[jn * nn + in] * dx2inv + 1e-30);
            theta_e = std::max(theta_e, th);
        }
        sensor[el] = theta_e;
    }
    return DG_OK;
}

/* Frame fields, downsample_rho_p (dgsem_lobatto_2d.cu:826-844) in FP64:
 * per element the plain mean over its nodes (storage order) of rho, floored
 * at 1e-14, and of the pressure clipped below at 0. */
OR_API int oracle_frame_fields(void* ctx, const double* Q, double* rho_el, double* p_el)
{
    Ctx& c = *(Ctx*)ctx;
    const double g = c.cfg.gamma;
    const int nn = c.NN;
    for (int el = 0; el < c.NEL; el++) {
        double rs = 0.0, ps = 0.0;
        for (int n = 0; n < c.NN2; n++) {
            double rho = Q[sidx(c, 0, el, n / nn, n % nn)];
            double mx = Q[sidx(c, 1, el, n / nn, n % nn)];
            double my = Q[sidx(c, 2, el, n / nn, n % nn)];
            double E = Q[sidx(c, 3, el, n / nn, n % nn)];
            if (rho < 1e-14) rho = 1e-14;
            double p = (g - 1.0) * (E - 0.5 * (mx * mx + my * my) / rho);
            rs += rho;
            ps += (p > 0.0) ? p : 0.0;
        }
        rho_el[el] = rs / c.NN2;
        p_el[el] = ps / c.NN2;
    }
    return DG_OK;
}

/* GLL totals (pattern of q1_dg_euler_2d.cu:3004-3016, on the nodal state):
 * out[0..3] conserved integrals, out[4] entropy integral -rho s/(g-1). */
OR_API int oracle_totals(void* ctx, const double* Q, double* out)
{
    Ctx& c = *(Ctx*)ctx;
    const double g = c.cfg.gamma, jw = 0.25 * c.h * c.h;
    const int nn = c.NN;
    double acc[5] = {0, 0, 0, 0, 0};
    for (int el = 0; el < c.NEL; el++)
        for (int jn = 0; jn < nn; jn++) for (int in = 0; in < nn; in++) {
            double wij = jw * c.w[in] * c.w[jn];
            double U[4];
            for (int q = 0; q < 4; q++) { U[q] = Q[sidx(c, q, el, jn, in)]; acc[q] += wij * U[q]; }
            double p = (g - 1.0) * (U[3] - 0.5 * (U[1] * U[1] + U[2] * U[2]) / U[0]);
            double s = log(p) - g * log(U[0]);
            acc[4] += wij * (-U[0] * s / (g - 1.0));
        }
    for (int q = 0; q < 5; q++) out[q] = acc[q];
    return DG_OK;
}
