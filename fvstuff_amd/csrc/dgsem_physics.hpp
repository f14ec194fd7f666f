// Pointwise Euler physics on the device, templated on the compute type.
//
// Follows dgsem_lobatto_2d.cu:184-355 (sound speed, prim/cons conversion with
// the 1e-14 floors, the physical flux, the logarithmic mean, the Chandrashekar
// EC flux, HLLC, the wall ghost state) plus two closed-form additions that the
// reference does not contain: the Ranocha EC/KEP flux and the Rusanov flux.
//
// Real = double: FP64 throughout.  Quotients are IEEE divisions, or, with
// NEWTON, a * (1/b) with the reciprocal from the hardware estimate
// (v_rcp_f64) refined by two Newton steps (within about 1 ulp for the
// normal-range operands here: densities and pressures are floored at 1e-14;
// about half the instructions of the IEEE sequence, but more live registers
// at a time); logs and square roots are OCML's FP64 routines.
// Real = float : the mixed form.  Reciprocal, log2 and sqrt use the gfx950
// transcendental unit (v_rcp_f32, v_log_f32, v_sqrt_f32, about 1 ulp).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dg {

template <typename Real, bool NEWTON = false> struct M;
__device__ __forceinline__ double rcp_newton(double x);

// 1/x for x in the normal range: the hardware reciprocal estimate plus two
// Newton steps (within about 1 ulp of 1/x; no denormal or overflow scaling,
// which the IEEE division sequence adds).
__device__ __forceinline__ double rcp_newton(double x) {
    double y = __builtin_amdgcn_rcp(x);
    double e = fma(-x, y, 1.0);
    y = fma(y, e, y);
    e = fma(-x, y, 1.0);
    return fma(y, e, y);
}

template <bool NEWTON> struct M<double, NEWTON> {
    static __device__ __forceinline__ double rcp(double x) { return NEWTON ? rcp_newton(x) : 1.0 / x; }
    static __device__ __forceinline__ double div(double a, double b) { return NEWTON ? a * rcp_newton(b) : a / b; }
    static __device__ __forceinline__ double ln(double x) { return log(x); }
    static __device__ __forceinline__ double sqrt_(double x) { return sqrt(x); }
    static __device__ __forceinline__ double fmax_(double a, double b) { return fmax(a, b); }
    static __device__ __forceinline__ double fmin_(double a, double b) { return fmin(a, b); }
    static __device__ __forceinline__ double abs_(double a) { return fabs(a); }
    // below this squared ratio (a-b)/(a+b) the log mean uses its series
    static constexpr double series_thr = 1e-4;
    // a quotient is a product with a reciprocal (which the compiler may
    // compute ahead of a branch unless its operand is kept in place)
    static constexpr bool rcp_div = NEWTON;
};

template <bool NEWTON> struct M<float, NEWTON> {
    static __device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
    static __device__ __forceinline__ float div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
    static __device__ __forceinline__ float ln(float x) {
        return __builtin_amdgcn_logf(x) * 0.693147180559945309f;
    }
    static __device__ __forceinline__ float sqrt_(float x) { return __builtin_amdgcn_sqrtf(x); }
    static __device__ __forceinline__ float fmax_(float a, float b) { return fmaxf(a, b); }
    static __device__ __forceinline__ float fmin_(float a, float b) { return fminf(a, b); }
    static __device__ __forceinline__ float abs_(float a) { return fabsf(a); }
    // below 5e-3 the series truncated after its u^2 term is accurate to 2e-8
    // (the dropped terms are 44u^3/945 and u^3/7), under half an FP32 ulp
    static constexpr float series_thr = 5e-3f;
    static constexpr bool rcp_div = true;
};

// x through an empty asm (ON): no instruction and the same value, but the
// compiler never executes an inline asm speculatively, so arithmetic on the
// result stays inside the branch it is written in.  In the log-mean branches
// below, where a quotient is a product with a reciprocal, 1/R.r and 1/R.z were
// otherwise computed for every pair, also on the path that skips the branch
// (measured: KH 1024^2 p = 3, 0.2 % faster in FP32, 0.7 % in FP64; an IEEE
// FP64 division is not moved, and the asm there only perturbs the schedule).
// Not volatile: a side-effecting asm would keep the kernels' uniform table
// reads from being scalar loads.
template <bool ON, typename T>
__device__ __forceinline__ T in_place(T x) {
    if (ON) asm("" : "+v"(x));
    return x;
}

// Primitive node state held in registers / LDS: rho, u, v, p and z = rho/p.
template <typename Real> struct Prim {
    Real r, u, v, p, z;
};

// Gas constants of the FP64 state side: gamma, gamma - 1 and the energy
// floor 1e-14/(gamma - 1) (a kernel can take them precomputed, or derive them
// with `of`; both give the same bits).
struct Gas64 {
    double g, gm1, efloor;
    static __device__ __forceinline__ Gas64 of(double g) { return Gas64{g, g - 1.0, 1e-14 / (g - 1.0)}; }
};

// ── FP64 state-side helpers (sanitize + prim conversion), :208-229, :347-355 ──
struct Cons64 {
    // 1/x for x in the normal range used here ([1e-14, 1e30])
    static __device__ __forceinline__ double rcp(double x) { return rcp_newton(x); }
    static __device__ __forceinline__ double pressure_from_cons(const double U[4], const Gas64& gs) {
        double rho = fmax(U[0], 1e-14);
        double ke = 0.5 * (U[1] * U[1] + U[2] * U[2]) * rcp(rho);
        return gs.gm1 * (U[3] - ke);
    }
    // returns the kinetic energy it computed: pressure_from_cons of the
    // result is gm1 * (U[3] - ke), bit for bit (the same rho, the same ke)
    static __device__ __forceinline__ double sanitize_cons(double U[4], const Gas64& gs) {
        U[0] = fmax(U[0], 1e-14);
        double ke = 0.5 * (U[1] * U[1] + U[2] * U[2]) * rcp(U[0]);
        if (gs.gm1 * (U[3] - ke) < 1e-14) U[3] = gs.efloor + ke;
        return ke;
    }
    // cons_to_prim_sane: sanitize, convert, floor rho/p, clamp |u|,|v| <= 1e4
    // (one reciprocal of rho serves u, v and the kinetic energy); also
    // returns ri = 1/W[0]
    static __device__ __forceinline__ void to_prim(const double U[4], double W[4], const Gas64& gs, double& ri) {
        const double r = fmax(U[0], 1e-14);
        ri = rcp(r);
        const double ke = 0.5 * (U[1] * U[1] + U[2] * U[2]) * ri;
        double E = U[3];
        if (gs.gm1 * (E - ke) < 1e-14) E = gs.efloor + ke;
        W[0] = r;
        W[1] = U[1] * ri;
        W[2] = U[2] * ri;
        W[3] = fmax(gs.gm1 * (E - ke), 1e-14);
        if (W[0] < 1e-5) { W[1] = 0.0; W[2] = 0.0; return; }
        W[1] = fmax(fmin(W[1], 1e4), -1e4);
        W[2] = fmax(fmin(W[2], 1e4), -1e4);
    }
    static __device__ __forceinline__ void to_prim(const double U[4], double W[4], const Gas64& gs) {
        double ri;
        to_prim(U, W, gs, ri);
    }
    // the same with gamma only
    static __device__ __forceinline__ double pressure_from_cons(const double U[4], double g) {
        return pressure_from_cons(U, Gas64::of(g));
    }
    static __device__ __forceinline__ void sanitize_cons(double U[4], double g) { sanitize_cons(U, Gas64::of(g)); }
    static __device__ __forceinline__ void to_prim(const double U[4], double W[4], double g) {
        to_prim(U, W, Gas64::of(g));
    }
    static __device__ __forceinline__ void to_prim(const double U[4], double W[4], double g, double& ri) {
        to_prim(U, W, Gas64::of(g), ri);
    }
    // max wave speed |u|+|v|+c of sanitized primitives (:483-486)
    static __device__ __forceinline__ double wave_speed(const double W[4], double g) {
        double c2 = g * W[3] * rcp(fmax(W[0], 1e-12));
        return fabs(W[1]) + fabs(W[2]) + sqrt(fmax(c2, 1e-14));
    }
    // the same with ri = 1/W[0] from to_prim (bit-identical: the reciprocal
    // of max(rho, 1e-12) is ri whenever rho >= 1e-12)
    static __device__ __forceinline__ double wave_speed(const double W[4], double g, double ri) {
        double c2 = g * W[3] * (W[0] >= 1e-12 ? ri : rcp(1e-12));
        return fabs(W[1]) + fabs(W[2]) + sqrt(fmax(c2, 1e-14));
    }
    // the spectral-difference sibling's element wave speed: |u|+|v| plus
    // sound_speed(p, rho) of the sanitized prims, without the 1e-12 density
    // floor (fr_lobatto_2d.cu:168-172, :515-518); ri = 1/W[0] from to_prim
    static __device__ __forceinline__ double wave_speed_sd(const double W[4], double g, double ri) {
        return fabs(W[1]) + fabs(W[2]) + sqrt(fmax(g * W[3] * ri, 1e-14));
    }
    // nodal positivity repair of the SD sibling (fr_lobatto_2d.cu:589-615):
    // rho < eps_rho -> rho = eps_rho, zero momentum; then p < eps_p -> E reset
    static __device__ __forceinline__ void repair(double U[4], const Gas64& gs, double eps_rho, double eps_p) {
        if (U[0] < eps_rho) { U[0] = eps_rho; U[1] = 0.0; U[2] = 0.0; }
        const double m2 = 0.5 * (U[1] * U[1] + U[2] * U[2]);
        if (gs.gm1 * (U[3] - m2 / fmax(U[0], 1e-14)) < eps_p) U[3] = eps_p / gs.gm1 + m2 / U[0];
    }
};

template <typename Real>
__device__ __forceinline__ Prim<Real> make_prim(const double W[4]) {
    Prim<Real> w;
    w.r = (Real)W[0]; w.u = (Real)W[1]; w.v = (Real)W[2]; w.p = (Real)W[3];
    if (sizeof(Real) == 8) w.z = (Real)(W[0] * Cons64::rcp(W[3]));
    else w.z = w.r * M<Real>::rcp(w.p);     // FP32 form: rho/p in FP32
    return w;
}

// ── compute-precision physics ─────────────────────────────────────────────
template <typename Real, bool NEWTON = false> struct Phys {
    using Mx = M<Real, NEWTON>;
    Real g, gm1, inv_gm1;

    __device__ __forceinline__ Real sound_speed(Real p, Real rho) const {   // :184-188
        Real cs2 = Mx::div(g * p, rho);
        return Mx::sqrt_(cs2 > (Real)1e-14 ? cs2 : (Real)1e-14);
    }
    // cons_to_prim_sane in the compute precision (:223-229): the state the
    // spectral-difference line interpolates to a flux node (z is not set;
    // only the physical flux uses the result)
    __device__ __forceinline__ Prim<Real> to_prim_sane(const Real U[4]) const {
        Prim<Real> w;
        w.r = Mx::fmax_(U[0], (Real)1e-14);
        const Real ri = Mx::rcp(w.r);
        const Real ke = (Real)0.5 * (U[1] * U[1] + U[2] * U[2]) * ri;
        Real E = U[3];
        if (gm1 * (E - ke) < (Real)1e-14) E = (Real)1e-14 * inv_gm1 + ke;
        w.u = U[1] * ri;
        w.v = U[2] * ri;
        w.p = Mx::fmax_(gm1 * (E - ke), (Real)1e-14);
        w.z = (Real)0;
        if (w.r < (Real)1e-5) {
            w.u = (Real)0; w.v = (Real)0;
        } else {
            w.u = Mx::fmax_(Mx::fmin_(w.u, (Real)1e4), (Real)-1e4);
            w.v = Mx::fmax_(Mx::fmin_(w.v, (Real)1e4), (Real)-1e4);
        }
        return w;
    }
    // cons_to_prim_sane of an FP64 node state in the compute precision, with
    // z = rho/p: the state rounded once to Real (the precision the reference
    // holds it in), then sanitized and converted in Real
    __device__ __forceinline__ Prim<Real> node_prim(const double Ud[4]) const {
        const Real U[4] = {(Real)Ud[0], (Real)Ud[1], (Real)Ud[2], (Real)Ud[3]};
        Prim<Real> w = to_prim_sane(U);
        w.z = w.r * Mx::rcp(w.p);
        return w;
    }
    // node wave speed |u|+|v|+c of sanitized prims in the compute precision:
    // the DGSEM form floors rho at 1e-12 (:483-486), the spectral-difference
    // form takes sound_speed(p, rho) as it is (fr_lobatto_2d.cu:515-518)
    template <bool SD = false>
    __device__ __forceinline__ Real node_speed(const Prim<Real>& w) const {
        const Real c = SD ? sound_speed(w.p, w.r)
                          : Mx::sqrt_(Mx::fmax_(Mx::div(g * w.p, Mx::fmax_(w.r, (Real)1e-12)), (Real)1e-14));
        return Mx::abs_(w.u) + Mx::abs_(w.v) + c;
    }
    // physical flux along x (dir 0) or y (dir 1), :232-241
    template <int DIR>
    __device__ __forceinline__ void flux(const Prim<Real>& w, Real F[4]) const {
        Real un = DIR == 0 ? w.u : w.v;
        Real E = w.p * inv_gm1 + (Real)0.5 * w.r * (w.u * w.u + w.v * w.v);
        F[0] = w.r * un;
        F[1] = w.r * un * w.u + (DIR == 0 ? w.p : (Real)0);
        F[2] = w.r * un * w.v + (DIR == 1 ? w.p : (Real)0);
        F[3] = (E + w.p) * un;
    }

    // Log mean L(a,b) = (a-b)/ln(a/b) and its inverse.  With f = (a-b)/(a+b),
    // ln(a/b) = 2 atanh(f), so L = (a+b)/2 * f/atanh(f); for small f the
    // series atanh(f)/f = 1 + f^2/3 + f^4/5 + f^6/7 (the reference's form,
    // :244-253) and its reciprocal 1 - f^2/3 - 4f^4/45 - 44f^6/945 are used
    // (in FP32 without their f^6 terms, below a smaller threshold).  The log
    // branch keeps the ratio form ln(a/b), well conditioned in FP32.
    // Log means of one pair: the series form first (no log, one reciprocal
    // each), replaced by the ratio-form log only where the series does not
    // apply.  That branch is skipped by the whole wave on smooth data.
    template <bool UNIFORM>
    __device__ __forceinline__ static void pair_means(Real aL, Real aR, Real bL, Real bR, Real& a_ln,
                                                      Real& inv_b_ln) {
        Real sa = aL + aR, sb = bL + bR;
        Real ra = Mx::rcp(sa), rb = Mx::rcp(sb);
        Real fa = (aL - aR) * ra, ua = fa * fa;
        Real fb = (bL - bR) * rb, ub = fb * fb;
        if (sizeof(Real) == 4) {
            a_ln = (Real)0.5 * sa * ((Real)1 - ua * ((Real)(1.0 / 3.0) + ua * (Real)(4.0 / 45.0)));
            inv_b_ln = (Real)2 * rb * ((Real)1 + ub * ((Real)(1.0 / 3.0) + ub * (Real)(1.0 / 5.0)));
        } else {
            a_ln = (Real)0.5 * sa *
                   ((Real)1 - ua * ((Real)(1.0 / 3.0) + ua * ((Real)(4.0 / 45.0) + ua * (Real)(44.0 / 945.0))));
            inv_b_ln = (Real)2 * rb *
                   ((Real)1 + ub * ((Real)(1.0 / 3.0) + ub * ((Real)(1.0 / 5.0) + ub * (Real)(1.0 / 7.0))));
        }
        const bool la = ua >= (Real)Mx::series_thr, lb = ub >= (Real)Mx::series_thr;
        if (UNIFORM) {
            // wave-uniform branch, skipped by the whole wave on smooth data
            // (per-lane branches are if-converted: every pair then evaluates
            // both logs); each lane selects its own value
            if (__any(la || lb)) {
                const Real a_ex = Mx::div(aL - aR, Mx::ln(Mx::div(aL, in_place<Mx::rcp_div>(aR))));
                const Real b_ex = Mx::div(Mx::ln(Mx::div(bL, in_place<Mx::rcp_div>(bR))), bL - bR);
                a_ln = la ? a_ex : a_ln;
                inv_b_ln = lb ? b_ex : inv_b_ln;
            }
        } else {
            if (la) a_ln = Mx::div(aL - aR, Mx::ln(Mx::div(aL, aR)));
            if (lb) inv_b_ln = Mx::div(Mx::ln(Mx::div(bL, bR)), bL - bR);
        }
    }

    // Ranocha flux on half-scaled node states h = (rho, u, v, p, rho/p)/2, the
    // form the line loop uses: (a+b)/2 = aL/2 + aR/2 needs no multiply, one
    // reciprocal of the product (rho_L+rho_R)(z_L+z_R) serves both log means,
    // and c[] = 1/(g-1) * (1, 1/3, 1/5, 1/7) folds the energy factor into the
    // series of the inverse log mean of z = rho/p.  The ratio-form log replaces
    // the series per lane where the series does not apply; the result does not
    // depend on how lines share waves.
    __device__ __forceinline__ static void ranocha_half(const Prim<Real>& L, const Prim<Real>& R, const Real c[4],
                                                        Real inv_gm1, Real F[4]) {
        const Real sr = L.r + R.r, sz = L.z + R.z;          // (rho_L+rho_R)/2, (z_L+z_R)/2
        // the one quotient of the smooth path: the Newton-refined reciprocal
        // in FP64 at every degree (the log branch keeps the policy's division)
        const Real rr = sizeof(Real) == 8 ? (Real)rcp_newton((double)(sr * sz)) : Mx::rcp(sr * sz);
        const Real ir = sz * rr, iz = sr * rr;              // 2/(rho_L+rho_R), 2/(z_L+z_R)
        const Real dr = L.r - R.r, dz = L.z - R.z;
        const Real fr = dr * ir, ur = fr * fr;
        const Real fz = dz * iz, uz = fz * fz;
        Real r_ln, izg;                                     // rho_ln, 1/((g-1) z_ln)
        if (sizeof(Real) == 4) {
            r_ln = sr * ((Real)1 - ur * ((Real)(1.0 / 3.0) + ur * (Real)(4.0 / 45.0)));
            izg = iz * (c[0] + uz * (c[1] + uz * c[2]));
        } else {
            r_ln = sr * ((Real)1 - ur * ((Real)(1.0 / 3.0) + ur * ((Real)(4.0 / 45.0) + ur * (Real)(44.0 / 945.0))));
            izg = iz * (c[0] + uz * (c[1] + uz * (c[2] + uz * c[3])));
        }
        // wave-uniform branch (a divergent one makes the register allocator
        // keep both sides' values of every unrolled pair live); each lane then
        // selects its own value.  Marked unlikely: the log blocks are placed
        // out of line, so smooth pairs fall through instead of jumping over
        // them (vortex 512^2 p = 7: mixed 0.35 %, FP64 1.5 % faster; the KH
        // kernels' Chandrashekar flux, whose front pairs take the log branch
        // more often, was neutral and keeps the plain branch)
        const bool lr = ur >= (Real)Mx::series_thr, lz = uz >= (Real)Mx::series_thr;
        if (__builtin_expect(__any(lr || lz), 0)) {
            // dr is (rho_L-rho_R)/2
            const Real r_ex = Mx::div((Real)2 * dr, Mx::ln(Mx::div(L.r, in_place<Mx::rcp_div>(R.r))));
            const Real z_ex = inv_gm1 * Mx::div(Mx::ln(Mx::div(L.z, in_place<Mx::rcp_div>(R.z))), (Real)2 * dz);
            r_ln = lr ? r_ex : r_ln;
            izg = lz ? z_ex : izg;
        }
        const Real u_av = L.u + R.u, v_av = L.v + R.v, p_av = L.p + R.p;
        const Real f1 = r_ln * u_av;
        const Real ka = L.u * R.u + L.v * R.v;              // (uL uR + vL vR)/4
        const Real kb = L.p * R.u + R.p * L.u;              // (pL uR + pR uL)/4
        F[0] = f1;
        F[1] = f1 * u_av + p_av;
        F[2] = f1 * v_av;
        F[3] = (Real)2 * (f1 * ka + kb) + f1 * izg;
    }

    // Chandrashekar EC flux (:259-285) along x, on half-scaled node states
    // h = (rho, u, v, p, rho/p)/2 like ranocha_half: beta = rho/(2p) = h.z,
    // the arithmetic means are sums of halves, and pair_means on the halved
    // densities gives rho_ln/2 (both of its forms are homogeneous of degree
    // one).  Scaling by 2 is exact, so this is the flux of the full states.
    // UNIFORM: wave-uniform log branch (lines of up to 6 nodes; at 7 and 8
    // nodes the branch form needs more registers than the kernels have)
    template <bool UNIFORM = true>
    __device__ __forceinline__ void chandrashekar_half(const Prim<Real>& L, const Prim<Real>& R, Real F[4]) const {
        Real rh_ln, inv_b_ln;
        pair_means<UNIFORM>(L.r, R.r, L.z, R.z, rh_ln, inv_b_ln);
        // p_hat = rho_avg/(2 beta_avg) = (rL+rR)/(2 (bL+bR))
        const Real p_hat = Mx::div(L.r + R.r, L.z + R.z);
        const Real u_av = L.u + R.u, v_av = L.v + R.v;
        // vel2 = (uL^2+vL^2)/2 + (uR^2+vR^2)/2 = 2 (sum of the halves' squares)
        const Real hv2 = (L.u * L.u + L.v * L.v) + (R.u * R.u + R.v * R.v);
        const Real e_int = (Real)0.5 * inv_gm1 * inv_b_ln - hv2;
        const Real f1 = (Real)2 * rh_ln * u_av, f2 = f1 * u_av + p_hat, f3 = f1 * v_av;
        F[0] = f1; F[1] = f2; F[2] = f3; F[3] = f1 * e_int + f2 * u_av + f3 * v_av;
    }

    // HLLC in the rotated frame, :288-335.  Normal (1,0) for DIR 0, (0,1) for DIR 1.
    __device__ __forceinline__ void hllc(int dir, const Prim<Real>& L, const Prim<Real>& R, Real F[4]) const {
        Real nx = dir == 0 ? (Real)1 : (Real)0, ny = dir == 0 ? (Real)0 : (Real)1;
        Real unL = L.u * nx + L.v * ny, unR = R.u * nx + R.v * ny;
        Real utL = -L.u * ny + L.v * nx, utR = -R.u * ny + R.v * nx;
        Real cL = sound_speed(L.p, L.r), cR = sound_speed(R.p, R.r);
        Real EL = L.p * inv_gm1 + (Real)0.5 * L.r * (L.u * L.u + L.v * L.v);
        Real ER = R.p * inv_gm1 + (Real)0.5 * R.r * (R.u * R.u + R.v * R.v);
        Real SL = Mx::fmin_(unL - cL, unR - cR);
        Real SR = Mx::fmax_(unL + cL, unR + cR);
        if (SL >= (Real)0) {
            F[0] = L.r * unL; F[1] = L.r * unL * L.u + L.p * nx; F[2] = L.r * unL * L.v + L.p * ny;
            F[3] = (EL + L.p) * unL;
        } else if (SR <= (Real)0) {
            F[0] = R.r * unR; F[1] = R.r * unR * R.u + R.p * nx; F[2] = R.r * unR * R.v + R.p * ny;
            F[3] = (ER + R.p) * unR;
        } else {
            Real dL = L.r * (SL - unL), dR = R.r * (SR - unR);
            Real Ss = Mx::div(R.p - L.p + dL * unL - dR * unR, dL - dR);
            bool lft = Ss >= (Real)0;
            Real rK = lft ? L.r : R.r, unK = lft ? unL : unR, utK = lft ? utL : utR;
            Real pK = lft ? L.p : R.p, EK = lft ? EL : ER, SK = lft ? SL : SR;
            Real fact = Mx::div(rK * (SK - unK), SK - Ss);
            Real Us0 = fact, Us1 = fact * Ss, Us2 = fact * utK;
            Real Us3 = fact * (Mx::div(EK, rK) + (Ss - unK) * (Ss + Mx::div(pK, rK * (SK - unK))));
            Real Fl0 = rK * unK + SK * (Us0 - rK);
            Real Fl1 = rK * unK * unK + pK + SK * (Us1 - rK * unK);
            Real Fl2 = rK * unK * utK + SK * (Us2 - rK * utK);
            Real Fl3 = (EK + pK) * unK + SK * (Us3 - EK);
            F[0] = Fl0;
            F[1] = Fl1 * nx - Fl2 * ny;
            F[2] = Fl1 * ny + Fl2 * nx;
            F[3] = Fl3;
        }
    }

    // Rusanov flux along x, F*(L,R) = 1/2 (f_L + f_R) - lam/2 (U_R - U_L) with
    // lam = max(|u| + c), seen from one side of a face: `own` is the left
    // state (own_left) or the right one, n the other.  Returns F*(L,R) (JUMP
    // false) or F*(L,R) - f(own) (JUMP true), written as
    //   1/2 (f_own + f_nb) - s lam/2 (U_nb - U_own)   or
    //   1/2 (f_nb - f_own) - s lam/2 (U_nb - U_own),   s = +1 (own left), -1,
    // with no left/right selection of the states.  Both sides of a face compute F* with the same
    // bits (the sums are commutative and U_nb - U_own changes sign exactly).
    template <bool JUMP>
    __device__ __forceinline__ void rusanov_side(const Prim<Real>& o, const Prim<Real>& n, bool own_left,
                                                 Real d[4]) const {
        const Real Eo = o.p * inv_gm1 + (Real)0.5 * o.r * (o.u * o.u + o.v * o.v);
        const Real En = n.p * inv_gm1 + (Real)0.5 * n.r * (n.u * n.u + n.v * n.v);
        const Real mo = o.r * o.u, mn = n.r * n.u;        // mass flux = x-momentum
        const Real lam = Mx::fmax_(Mx::abs_(o.u) + sound_speed(o.p, o.r), Mx::abs_(n.u) + sound_speed(n.p, n.r));
        const Real hl = own_left ? (Real)0.5 * lam : (Real)-0.5 * lam;
        const Real fo[4] = {mo, mo * o.u + o.p, mo * o.v, (Eo + o.p) * o.u};
        const Real fn[4] = {mn, mn * n.u + n.p, mn * n.v, (En + n.p) * n.u};
        const Real du[4] = {n.r - o.r, mn - mo, n.r * n.v - o.r * o.v, En - Eo};
#pragma unroll
        for (int q = 0; q < 4; q++)
            d[q] = (Real)0.5 * (JUMP ? fn[q] - fo[q] : fo[q] + fn[q]) - hl * du[q];
    }
};

// wall_state, :339-344: reflect the normal velocity
template <typename Real>
__device__ __forceinline__ Prim<Real> wall_ghost(const Prim<Real>& w, int dir) {
    Prim<Real> g = w;
    if (dir == 0) g.u = -w.u; else g.v = -w.v;
    return g;
}

}  // namespace dg
