// DGSEM stage kernels for gfx950.
//
// One workgroup owns E consecutive elements (a tile of one element row when
// K >= E).  The work is split in three phases that hand data over through LDS:
//
//  A  node phase  - each thread loads whole nodes (4 coalesced FP64 streams,
//                   the reference's var-major SoA, dgsem_lobatto_2d.cu:126),
//                   applies cons_to_prim_sane (FP64 kernels: in FP64; mixed
//                   kernels: in FP32 on the node rounded once, as the
//                   reference holds it) and stores (rho, u, v, p, rho/p) in
//                   the compute precision.
//  B  line phase  - one thread per grid line of an element (p+1 x-lines and
//                   p+1 y-lines).  The thread keeps the line's p+1 states in
//                   registers and evaluates every two-point volume flux of the
//                   line once: F#(U_i,U_k) is symmetric, so each pair feeds
//                   D_ik F# to node i and D_ki F# to node k (p(p+1)/2 flux
//                   evaluations per line instead of the reference's (p+1)^2
//                   per line, :417-430).  y-lines work on states with u and v
//                   swapped so that every lane runs the same x-direction code
//                   (no divergence between x- and y-lines).  The thread also
//                   evaluates the Riemann flux on the two faces its line ends
//                   on (:441-481) and the line's share of the artificial
//                   viscosity (:501-521).
//  C  node phase  - RHS per node = x-line part + y-line part; either written
//                   out (RHS mode) or fused into the SSP-RK3 Shu-Osher update
//                   (:533-576) with sanitize_cons, the Zhang-Shu limiter
//                   (:584-643, FP64) and the wave-speed reduction for the next
//                   step's dt (:483-486, :646-659).
//
// The spectral-difference sibling of fr_lobatto_2d.cu (VOL_SD) shares the
// three phases: its line thread interpolates the line's conservative states
// to the interior flux nodes instead of evaluating two-point fluxes
// (sd_line), and its update applies the nodal positivity repair.
#pragma once
#include "dgsem_physics.hpp"

namespace dg {

enum { MODE_RHS = 0, MODE_STAGE = 1 };
// VOL_SD: the spectral-difference sibling (fr_lobatto_2d.cu), whose line
// operator replaces the volume term
enum { VOL_CHANDRASHEKAR = 0, VOL_RANOCHA = 1, VOL_COLLOCATION = 2, VOL_SD = 3 };
enum { SURF_HLLC = 0, SURF_RUSANOV = 1 };
enum { BC_WALL = 0, BC_PERIODIC = 1 };
enum { ST_OK = 0, ST_NONFINITE = 3, ST_DT_UNDERFLOW = 4, ST_BLOWUP = 5 };

constexpr int BLOCK = 256;

// Device-resident step state for the fused stepping (no host round trip).
// The stage-3 wave-speed maximum is gathered in LAM_PARTS partial slots
// (workgroup w updates part w mod LAM_PARTS), so no single address takes one
// atomic per workgroup; the next stage 1 folds the parts.
constexpr int LAM_PARTS = 64;

struct DevState {
    unsigned long long lam_bits[2];  // max wave speed as ordered bit pattern (>= 0), folded
    unsigned long long lam_part[2][LAM_PARTS];   // per-slot partial maxima
    double t;                        // simulation time
    double dt;                       // dt of the current step
    int status;                      // first numerical failure (ST_*)
    int skip;                        // current step is a no-op (t >= t_end)
    int steps;                       // steps taken since the last reset
    int pad;
};

struct KArgs {
    const double* Up;    // state whose RHS is evaluated
    const double* U0;    // U^n for the Shu-Osher combination
    double* out;         // RHS (MODE_RHS) or new stage state (MODE_STAGE)
    double* lam_el;      // MODE_RHS: per-element max wave speed (nullable)
    const void* tab;     // compute-precision tables: D[N*N], w[N], winv[N]
    const double* w64;   // FP64 GLL weights (limiter)
    DevState* st;
    int P, K, NEL, E, xcd;
    int zero;            // always 0: see batch_stage_persistent
    int mode, stage, vol, surf, bc, av, limiter, slot_cur, slot_next;
    int check_blowup;    // stage 1: the wave speed in slot_cur came from a step of this call
    double h, gamma, ca, cb, cfl, t_end, av_cav, av_k, eps_rho, eps_p;
    // derived constants, computed once on the host (set_derived): 2/h,
    // g-1, 1/(g-1), 1/K, the energy floor 1e-14/(g-1), the AV length
    // h/(2p+1) and its inverse square; FP32 copies for the mixed form
    double jac, gm1, inv_gm1, inv_K, efloor, len_p, dx2inv;
    float jac_f, g_f, gm1_f, inv_gm1_f, len_p_f, dx2inv_f;
    int off_U, off_av, off_lam, off_el;   // byte offsets of the LDS regions
};

// Read-only operator table (D, weights, Lav, SD matrices) read through the
// constant address space: a uniform index then compiles to a scalar load.
// Read as plain global memory, every table read after an `asm volatile`
// (the accumulator pins, which the compiler must assume may write memory)
// or a global store became a per-lane vector load (FP64 p = 7: 56 per line
// thread in the volume loop, 120 with the AV).  Divergent indices (the MFMA
// form's Lav rows) still compile to vector loads.  The tables are written
// once at set-up and never by a kernel.
template <typename Real>
struct CTab {
    const Real* p;
    __device__ __forceinline__ Real operator[](int i) const {
#ifdef __HIP_DEVICE_COMPILE__
        return ((const __attribute__((address_space(4))) Real*)p)[i];
#else
        return p[i];
#endif
    }
    __device__ __forceinline__ CTab operator+(int o) const { return CTab{p + o}; }
};

// The same table read as plain global memory (vector loads after a pin or a
// store).  The FP64 shock-capturing kernels at p <= 5 read the AV operators
// (D for the velocity divergence, Lav) so: as scalar loads, next to the
// volume matrix, their doubles overflow the scalar register file and the
// compiler spills scalars into vector-register lanes (p = 3: 312 v_readlane
// instead of 10; KH 2048^2 p = 3 FP64 1.4 % faster this way).  At p = 7 the
// scalar loads are faster (1.8 %); FP32 compiles to scalar loads either way.
template <typename Real>
struct GTab {
    const Real* p;
    __device__ __forceinline__ Real operator[](int i) const { return p[i]; }
    __device__ __forceinline__ GTab operator+(int o) const { return GTab{p + o}; }
};

// Host side: fill the derived constants of an argument block whose h, gamma,
// K and P are set (the same IEEE divisions the device would do).
inline void set_derived(KArgs& a) {
    a.jac = 2.0 / a.h;
    a.gm1 = a.gamma - 1.0;
    a.inv_gm1 = 1.0 / (a.gamma - 1.0);
    a.inv_K = 1.0 / (double)a.K;
    a.efloor = 1e-14 / (a.gamma - 1.0);
    a.jac_f = (float)a.jac;
    a.g_f = (float)a.gamma;
    a.gm1_f = (float)a.gm1;
    a.inv_gm1_f = (float)a.inv_gm1;
    a.len_p = a.h / (2 * a.P + 1);
    a.dx2inv = 1.0 / (a.len_p * a.len_p);
    a.len_p_f = (float)a.len_p;
    a.dx2inv_f = 1.0f / (a.len_p_f * a.len_p_f);
}

// a derived constant in the compute precision (no conversion for FP32)
template <typename Real>
__device__ __forceinline__ Real kconst(double d, float f) {
    if (sizeof(Real) == 4) return (Real)f;
    return (Real)d;
}

// Global FP64 stream access at a 32-bit byte offset from a uniform base held
// in scalar registers (global_load/store with a scalar base and a 32-bit
// vector offset); the callers keep the offset within one tile of one
// variable.  `sbase` materialises the base in SGPRs through an empty asm (no
// instruction), so that the address arithmetic is not re-associated into
// 64-bit vector adds.
#define DG_GLOBAL __attribute__((address_space(1)))
typedef DG_GLOBAL const double* gcptr;
typedef DG_GLOBAL double* gptr;
__device__ __forceinline__ gcptr sbase(const double* p) {
    gcptr g = (gcptr)p;
    asm("" : "+s"(g));
    return g;
}
__device__ __forceinline__ gptr sbase(double* p) {
    gptr g = (gptr)p;
    asm("" : "+s"(g));
    return g;
}
__device__ __forceinline__ double ld_at(gcptr base, unsigned byte_off) {
    return *(gcptr)((DG_GLOBAL const char*)base + byte_off);
}
// Stage outputs and U^n are streamed (non-temporal: read or written once per
// launch, no reuse in L2 before it is evicted; measured 2 % faster).  The
// stage input stays cached: other tiles read its face nodes.
__device__ __forceinline__ void st_at(gptr base, unsigned byte_off, double v) {
    __builtin_nontemporal_store(v, (gptr)((DG_GLOBAL char*)base + byte_off));
}
__device__ __forceinline__ double ld_stream(gcptr base, unsigned byte_off) {
    return __builtin_nontemporal_load((gcptr)((DG_GLOBAL const char*)base + byte_off));
}
// the same for two adjacent FP64 values (16-byte aligned offsets)
typedef double dbl2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ dbl2 ld2_at(gcptr base, unsigned byte_off) {
    return *(DG_GLOBAL const dbl2*)((DG_GLOBAL const char*)base + byte_off);
}
__device__ __forceinline__ dbl2 ld2_stream(gcptr base, unsigned byte_off) {
    return __builtin_nontemporal_load((DG_GLOBAL const dbl2*)((DG_GLOBAL const char*)base + byte_off));
}
__device__ __forceinline__ void st2_at(gptr base, unsigned byte_off, dbl2 v) {
    __builtin_nontemporal_store(v, (DG_GLOBAL dbl2*)((DG_GLOBAL char*)base + byte_off));
}

// (row, column) of element el of a K x K grid without an integer division:
// the FP64 quotient is within one of the true one for el < 2^31.
__device__ __forceinline__ void elem_rc(int el, int K, double inv_K, int& ej, int& ei) {
    int q = (int)((double)el * inv_K);
    int r = el - q * K;
    if (r < 0) { q -= 1; r += K; }
    if (r >= K) { q += 1; r -= K; }
    ej = q; ei = r;
}

// Per-launch step control of the fused stepping (ping-pong wave-speed slots).
struct StepCtl {
    int slot_cur, slot_next, check_blowup;
};

__device__ __forceinline__ unsigned long long dbits(double x) {
    return (unsigned long long)__double_as_longlong(x);
}
__device__ __forceinline__ double bitsd(unsigned long long b) {
    return __longlong_as_double((long long)b);
}

// XCD-aware tile order: the hardware deals workgroups round-robin over the 8
// XCDs; give each XCD a contiguous range of tiles so that the y-neighbours a
// tile reads (K/E tiles away) sit in the same XCD's L2.
__device__ __forceinline__ int tile_of_block(int b, int G, int xcd) {
    if (!xcd) return b;
    int q = G >> 3, r = G & 7;
    int x = b & 7, i = b >> 3;
    return x * q + (x < r ? x : r) + i;
}

// max that propagates NaN (fmax would drop it)
__device__ __forceinline__ double nanmax(double a, double b) {
    return (a != a || b != b) ? a + b : fmax(a, b);
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = nanmax(v, __shfl_xor(v, o, 64));
    return v;
}

// Wave-wide unsigned max through DPP (no LDS round trips): row shifts give
// lane 15 of every row its row's maximum, two row broadcasts carry it to lane
// 63.  Lanes whose DPP source lies outside the row keep their own value.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ unsigned long long umax_dpp_step(unsigned long long v) {
    const unsigned lo = (unsigned)v, hi = (unsigned)(v >> 32);
    const unsigned olo = (unsigned)__builtin_amdgcn_update_dpp((int)lo, (int)lo, CTRL, ROW_MASK, 0xf, false);
    const unsigned ohi = (unsigned)__builtin_amdgcn_update_dpp((int)hi, (int)hi, CTRL, ROW_MASK, 0xf, false);
    const unsigned long long o = ((unsigned long long)ohi << 32) | olo;
    return o > v ? o : v;
}
// the maximum over the 64 lanes, valid in lane 63
__device__ __forceinline__ unsigned long long wave_umax_to_last(unsigned long long v) {
    v = umax_dpp_step<0x111, 0xf>(v);   // row_shr:1
    v = umax_dpp_step<0x112, 0xf>(v);   // row_shr:2
    v = umax_dpp_step<0x114, 0xf>(v);   // row_shr:4
    v = umax_dpp_step<0x118, 0xf>(v);   // row_shr:8
    v = umax_dpp_step<0x142, 0xa>(v);   // row_bcast:15 into rows 1 and 3
    v = umax_dpp_step<0x143, 0xc>(v);   // row_bcast:31 into rows 2 and 3
    return v;
}
// the same on 32-bit values (one register per step)
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ unsigned umax32_dpp_step(unsigned v) {
    const unsigned o = (unsigned)__builtin_amdgcn_update_dpp((int)v, (int)v, CTRL, ROW_MASK, 0xf, false);
    return o > v ? o : v;
}
// 32-bit unsigned maximum over the 64 lanes, in every lane
__device__ __forceinline__ unsigned wave_umax32(unsigned v) {
    v = umax32_dpp_step<0x111, 0xf>(v);   // row_shr:1
    v = umax32_dpp_step<0x112, 0xf>(v);   // row_shr:2
    v = umax32_dpp_step<0x114, 0xf>(v);   // row_shr:4
    v = umax32_dpp_step<0x118, 0xf>(v);   // row_shr:8
    v = umax32_dpp_step<0x142, 0xa>(v);   // row_bcast:15 into rows 1 and 3
    v = umax32_dpp_step<0x143, 0xc>(v);   // row_bcast:31 into rows 2 and 3
    return (unsigned)__builtin_amdgcn_readlane((int)v, 63);
}
// the maximum over the 64 lanes of 64-bit unsigned values, in every lane, as
// two 32-bit passes: the high words, then the low words of the lanes that
// hold the highest high word
__device__ __forceinline__ unsigned long long wave_umax(unsigned long long v) {
    const unsigned hi = wave_umax32((unsigned)(v >> 32));
    const unsigned lo = wave_umax32((unsigned)(v >> 32) == hi ? (unsigned)v : 0u);
    return ((unsigned long long)hi << 32) | lo;
}

// Element maximum into LDS for values of one element held by G consecutive
// lanes (one node per lane, G = N2; or a node pair per lane, G = N2/2): when
// the group is 4, 8, 16 or 32 lanes or the whole wave, it is reduced with
// DPP (row shifts within a row of 16, then a row broadcast for 32) and one
// lane per element issues the LDS atomic (nodes of one element would
// otherwise serialise on the same address); other group sizes use one
// atomic per lane.  All lanes of the wave must take part; `live` is uniform
// over an element's group.
template <int G>
__device__ __forceinline__ void elem_max_lds(unsigned long long* dst, unsigned long long v, bool live) {
    const int lane = threadIdx.x & 63;
    if (G == 64) {
        v = wave_umax_to_last(v);
        if (live && lane == 63) atomicMax(dst, v);
    } else if (G == 4 || G == 8 || G == 16 || G == 32) {
        v = umax_dpp_step<0x111, 0xf>(v);                    // row_shr:1
        v = umax_dpp_step<0x112, 0xf>(v);                    // row_shr:2
        if (G >= 8) v = umax_dpp_step<0x114, 0xf>(v);        // row_shr:4
        if (G >= 16) v = umax_dpp_step<0x118, 0xf>(v);       // row_shr:8
        if (G == 32) v = umax_dpp_step<0x142, 0xa>(v);       // row_bcast:15 into rows 1 and 3
        if (live && (lane & (G - 1)) == G - 1) atomicMax(dst, v);
    } else if (live) {
        atomicMax(dst, v);
    }
}

// The same for a non-negative FP32 value (or a NaN, of either sign): its bit
// pattern orders like the value as a 32-bit unsigned, so the DPP steps move
// one register instead of two; the LDS atomic takes the FP64 bits of the
// element maximum (FP32 -> FP64 is exact and keeps the order).
__device__ __forceinline__ float fmax_bits(float a, float b) {
    return __uint_as_float(max(__float_as_uint(a), __float_as_uint(b)));
}
template <int G>
__device__ __forceinline__ void elem_max_lds(unsigned long long* dst, float f, bool live) {
    const int lane = threadIdx.x & 63;
    unsigned v = __float_as_uint(f);
    if (G == 64 || G == 4 || G == 8 || G == 16 || G == 32) {
        v = umax32_dpp_step<0x111, 0xf>(v);                    // row_shr:1
        v = umax32_dpp_step<0x112, 0xf>(v);                    // row_shr:2
        if (G >= 8) v = umax32_dpp_step<0x114, 0xf>(v);        // row_shr:4
        if (G >= 16) v = umax32_dpp_step<0x118, 0xf>(v);       // row_shr:8
        if (G >= 32) v = umax32_dpp_step<0x142, 0xa>(v);       // row_bcast:15 into rows 1 and 3
        if (G == 64) v = umax32_dpp_step<0x143, 0xc>(v);       // row_bcast:31 into rows 2 and 3
        if (live && (lane & (G - 1)) == G - 1) atomicMax(dst, dbits((double)__uint_as_float(v)));
    } else if (live) {
        atomicMax(dst, dbits((double)f));
    }
}

// Block-wide max of v (FP64) -> one atomicMax on the bit pattern.
__device__ __forceinline__ void block_max_atomic(double v, unsigned long long* dst, double* sred) {
    v = wave_max(v);
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) sred[wid] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = sred[0];
        for (int i = 1; i < (int)(blockDim.x >> 6); i++) m = nanmax(m, sred[i]);
        // a positive NaN orders above every finite value as an unsigned bit pattern
        if (m != m) m = __longlong_as_double(0x7ff8000000000000ll);
        atomicMax(dst, dbits(m));
    }
}

// ── Zhang-Shu limiter on E elements whose stage values sit in LDS (FP64) ──
// sU[e][q][nd]; on return the limited values are in sU.  :584-643.
template <int N>
__device__ void limit_block(double* sU, const double* w64, int nE, double g, double eps_rho, double eps_p,
                            double* sBar, double* sTrho, unsigned long long* sTheta) {
    constexpr int N2 = N * N;
    const int tid = threadIdx.x, nthr = blockDim.x;
    __syncthreads();
    if (tid < nE) {
        const int e = tid;
        const double* U = sU + (size_t)e * 4 * N2;
        double Ubar[4] = {0, 0, 0, 0};
        for (int jn = 0; jn < N; jn++)
            for (int in = 0; in < N; in++) {
                double wij = 0.25 * w64[in] * w64[jn];
                for (int q = 0; q < 4; q++) Ubar[q] += wij * U[q * N2 + jn * N + in];
            }
        Ubar[0] = fmax(Ubar[0], eps_rho);
        {
            double p = Cons64::pressure_from_cons(Ubar, g);
            if (p < eps_p) Ubar[3] = eps_p / (g - 1.0) + 0.5 * (Ubar[1] * Ubar[1] + Ubar[2] * Ubar[2]) / Ubar[0];
        }
        double rho_min = 1e30;
        for (int n = 0; n < N2; n++) rho_min = fmin(rho_min, U[n]);
        double t_rho = 1.0;
        if (rho_min < eps_rho) t_rho = (Ubar[0] - eps_rho) / fmax(Ubar[0] - rho_min, 1e-30);
        t_rho = fmax(0.0, fmin(1.0, t_rho));
        for (int q = 0; q < 4; q++) sBar[e * 4 + q] = Ubar[q];
        sTrho[e] = t_rho;
        sTheta[e] = dbits(t_rho);
    }
    __syncthreads();
    for (int s = tid; s < nE * N2; s += nthr) {
        const int e = s / N2, nd = s - e * N2;
        const double* U = sU + (size_t)e * 4 * N2;
        double t_rho = sTrho[e];
        double Ub[4], Un[4], Ud[4];
        for (int q = 0; q < 4; q++) { Ub[q] = sBar[e * 4 + q]; Un[q] = U[q * N2 + nd]; }
        for (int q = 0; q < 4; q++) Ud[q] = Ub[q] + t_rho * (Un[q] - Ub[q]);
        if (Cons64::pressure_from_cons(Ud, g) < eps_p) {
            double lo = 0.0, hi = t_rho;
            for (int it = 0; it < 20; it++) {
                double tm = 0.5 * (lo + hi), Um[4];
                for (int q = 0; q < 4; q++) Um[q] = Ub[q] + tm * (Un[q] - Ub[q]);
                if (Cons64::pressure_from_cons(Um, g) >= eps_p) lo = tm; else hi = tm;
            }
            atomicMin(&sTheta[e], dbits(lo));   // theta in [0,1]: bit order = value order
        }
    }
    __syncthreads();
    for (int s = tid; s < nE * N2; s += nthr) {
        const int e = s / N2, nd = s - e * N2;
        double* U = sU + (size_t)e * 4 * N2;
        double th = bitsd(sTheta[e]);
        for (int q = 0; q < 4; q++) {
            double ub = sBar[e * 4 + q];
            U[q * N2 + nd] = ub + th * (U[q * N2 + nd] - ub);
        }
    }
    __syncthreads();
}

// Load node `nd` of element `el` (4 FP64 streams) and convert.
__device__ __forceinline__ void load_prim64(const double* __restrict__ Q, size_t S, size_t off, double g,
                                            double W[4]) {
    double U[4] = {Q[off], Q[S + off], Q[2 * S + off], Q[3 * S + off]};
    Cons64::to_prim(U, W, g);
}

// LDS holds node states scaled by 1/2 (exact in binary floating point); the
// full state is recovered exactly by doubling.
template <typename Real>
__device__ __forceinline__ Prim<Real> doubled(const Prim<Real>& h) {
    Prim<Real> w;
    w.r = (Real)2 * h.r; w.u = (Real)2 * h.u; w.v = (Real)2 * h.v; w.p = (Real)2 * h.p; w.z = (Real)2 * h.z;
    return w;
}

template <typename Real>
__device__ __forceinline__ void rot(Prim<Real>& w, bool swap) {
    Real u = w.u, v = w.v;
    w.u = swap ? v : u;
    w.v = swap ? u : v;
}

// PIN: materialise a pair's two accumulator rows where the pair is
// evaluated (an empty asm that reads and writes them, no instruction; the
// arithmetic and its results are unchanged).  Without it the compiler sinks
// every pair's flux and accumulation past the later pairs' log-branches into
// one block at the end of the line, keeping all pairs' intermediate values
// live at once: in FP64 at p = 7 that needs the whole register file (1 wave
// per SIMD) instead of about 220 registers (2 waves).  The FP32 single-solve kernel
// keeps the sunk schedule, whose wide independent block runs faster at its
// occupancy (measured); the batched kernel pins (see batch_stage_kernel).
template <int N, bool PIN, typename Real>
__device__ __forceinline__ void pin_pair(Real (&acc)[N][4], int i, int k) {
    if (!PIN) return;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        asm volatile("" : "+v"(acc[i][q]));
        asm volatile("" : "+v"(acc[k][q]));
    }
}

// ── symmetric pair loop of one line: F#(U_i,U_k) once per i < k ──────────
// Dv is the volume matrix (D scaled by the volume term's factor).  acc[0] and
// acc[N-1] hold their first terms already; node k of 0 < k < N-1 receives its
// first term from the pair (0, k), which initialises it.
template <int N, typename Real, int VOL, bool PIN, typename PH>
__device__ __forceinline__ void line_pairs(const PH& ph, const Prim<Real> (&hw)[N], CTab<Real> Dv,
                                           Real (&acc)[N][4]) {
    Real c[4] = {0, 0, 0, 0};
    if (VOL == VOL_RANOCHA) {
        c[0] = ph.inv_gm1; c[1] = ph.inv_gm1 * (Real)(1.0 / 3.0);
        c[2] = ph.inv_gm1 * (Real)(1.0 / 5.0); c[3] = ph.inv_gm1 * (Real)(1.0 / 7.0);
    }
#pragma unroll
    for (int i = 0; i < N; i++) {
#pragma unroll
        for (int k = i + 1; k < N; k++) {
            Real F[4];
            if (VOL == VOL_RANOCHA) PH::ranocha_half(hw[i], hw[k], c, ph.inv_gm1, F);
            else ph.template chandrashekar_half<(N <= 6)>(hw[i], hw[k], F);
            const Real dik = Dv[i * N + k], dki = Dv[k * N + i];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                acc[i][q] += dik * F[q];
                if (i == 0 && k < N - 1) acc[k][q] = dki * F[q];
                else acc[k][q] += dki * F[q];
            }
            pin_pair<N, PIN>(acc, i, k);
        }
    }
}

// ── volume contribution of one line plus its two face terms:
//    acc[i] = sum_k Dv_ik F#(U_i,U_k) + [i = 0] fm + [i = N-1] fp ──────────
template <int N, typename Real, int VOL, bool PIN, typename PH>
__device__ __forceinline__ void line_volume(const PH& ph, const Prim<Real> (&hw)[N], CTab<Real> Dv,
                                            const Real (&fm)[4], const Real (&fp)[4], Real (&acc)[N][4]) {
    if (VOL == VOL_COLLOCATION) {
        // linear contraction sum_k D_ik f(U_k) (:431-439)
#pragma unroll
        for (int k = 0; k < N; k++) {
            Real f[4];
            ph.template flux<0>(doubled(hw[k]), f);
#pragma unroll
            for (int i = 0; i < N; i++) {
                const Real dik = Dv[i * N + k];
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    if (k > 0) acc[i][q] += dik * f[q];
                    else if (i == 0) acc[i][q] = dik * f[q] + fm[q];
                    else if (i == N - 1) acc[i][q] = dik * f[q] + fp[q];
                    else acc[i][q] = dik * f[q];
                }
            }
        }
        return;
    }
    // diagonal: F#(U_i,U_i) = f(U_i) (consistency).  On LGL nodes D_ii = 0
    // except D_00 = -p(p+1)/4 = -1/(2 w_0) and D_pp = 1/(2 w_p), so the
    // split form's 2 D_00 f(U_0) equals the -f(U_0)/w_0 of the strong-form
    // surface term with the opposite sign: the two cancel exactly, and the
    // face terms fm, fp carry the Riemann flux alone (weak form at the end
    // nodes; the interior diagonal, ~1e-16 from the negative-sum
    // construction, is dropped).
#pragma unroll
    for (int q = 0; q < 4; q++) { acc[0][q] = fm[q]; acc[N - 1][q] = fp[q]; }
    line_pairs<N, Real, VOL, PIN>(ph, hw, Dv, acc);
}

// ── spectral-difference line (fr_lobatto_2d.cu:367-438): the conservative
//    states of the line's p+1 solution nodes (FP64 in LDS, eU[q*N2 + node])
//    interpolated to the p interior flux nodes, converted (cons_to_prim_sane)
//    and their physical flux evaluated; the Riemann fluxes fm, fp sit on the
//    two end flux nodes; the flux polynomial's derivative at the solution
//    nodes through Dsd = -(2/h) Dfs:
//    acc[i] = Dsd[i][0] fm + Dsd[i][N] fp + sum_{m=1}^{N-1} Dsd[i][m] f(Isf[m] U)
//    y-lines read the momenta swapped and run the x-direction code. ──
template <int N, typename Real, typename PH>
__device__ __forceinline__ void sd_line(const PH& ph, const double* eU, int lbase, int lstride, bool ydir,
                                        CTab<Real> Isf, CTab<Real> Dsd, const Real (&fm)[4],
                                        const Real (&fp)[4], Real (&acc)[N][4]) {
    constexpr int N2 = N * N, NF = N + 1;
    const int q1 = ydir ? 2 : 1, q2 = ydir ? 1 : 2;
    Real uc[N][4];
#pragma unroll
    for (int k = 0; k < N; k++) {
        const double* u = eU + lbase + k * lstride;
        uc[k][0] = (Real)u[0]; uc[k][1] = (Real)u[q1 * N2]; uc[k][2] = (Real)u[q2 * N2]; uc[k][3] = (Real)u[3 * N2];
    }
#pragma unroll
    for (int i = 0; i < N; i++)
#pragma unroll
        for (int q = 0; q < 4; q++) acc[i][q] = Dsd[i * NF] * fm[q] + Dsd[i * NF + N] * fp[q];
#pragma unroll
    for (int m = 1; m < N; m++) {
        Real uf[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            Real t = 0;
#pragma unroll
            for (int j = 0; j < N; j++) t += Isf[m * N + j] * uc[j][q];
            uf[q] = t;
        }
        Real F[4];
        ph.template flux<0>(ph.to_prim_sane(uf), F);
#pragma unroll
        for (int i = 0; i < N; i++)
#pragma unroll
            for (int q = 0; q < 4; q++) acc[i][q] += Dsd[i * NF + m] * F[q];
    }
}

// Register budget: a line keeps its p+1 node states and 4(p+1) accumulators
// live ((p+1)*9 values per lane, twice the registers in FP64).  The floor on
// waves per SIMD lets FP64 at p >= 5 use the full 512-register file (VGPRs +
// AGPRs) instead of spilling to scratch.
template <int P, typename Real, int VOL, bool FULL> struct Occ { static constexpr int waves = 3; };
template <int P, int VOL, bool FULL> struct Occ<P, double, VOL, FULL> {
    // FP64 at p >= 6 (p = 5 with the Chandrashekar flux): the line state and
    // accumulators plus one pair's temporaries exceed 256 registers, so use
    // the whole file rather than spill
    // the same for the spectral-difference line (its p+1 conservative states
    // and 4(p+1) accumulators) from p = 5 on
    static constexpr int waves =
        VOL == VOL_SD ? (P >= 5 ? 1 : 2) : (P >= 6 || (P == 5 && VOL == VOL_CHANDRASHEKAR)) ? 1 : 2;
};

// One tile (E consecutive elements) of one stage.  `lead` marks the one tile
// per launch that updates the device step state.  Every early exit is taken
// by the whole block (decided by thread 0, broadcast through LDS).
//
// UP_LDS: the FP64 stage input is kept in LDS from the node phase to the
// update (required by the limiter); otherwise it is loaded again for the
// update, which takes more than half of the LDS of a workgroup away (the
// batched kernel, whose large workgroups are LDS-limited, runs faster so; the
// single-solve kernels run slower, measured).
// U0_EARLY: U^n is loaded with the stage input at the start (see below).
// MFMA_AV (p = 3, FP32): the tile's linear tensor-product contraction on the
// matrix cores instead of in the line phase: in the shock-capturing kernels
// the artificial viscosity's (see the AV pre-pass below), in the lean
// collocation kernel the volume term D f (see the collocation pre-pass).
template <int P, typename Real, bool FULL, int VOL, bool PIN = (sizeof(Real) == 8), bool UP_LDS = true,
          bool U0_EARLY = !FULL, bool PAIR_OK = true, bool MFMA_AV = false>
__device__ __forceinline__ void stage_tile(const KArgs& a, const StepCtl sc, int tile, bool lead, int tid) {
    static_assert(UP_LDS || !FULL, "the limiter needs the stage values in LDS");
    static_assert(!MFMA_AV || (P == 3 && sizeof(Real) == 4 && VOL != VOL_SD && (FULL || VOL == VOL_COLLOCATION)),
                  "the MFMA forms are the p = 3 FP32 shock-capturing and collocation kernels'");
    constexpr bool MFMA_COL = MFMA_AV && !FULL && VOL == VOL_COLLOCATION;
    constexpr bool SD = VOL == VOL_SD;
    static_assert(UP_LDS || !SD, "the SD line reads the stage input from LDS");
    constexpr int N = P + 1, N2 = N * N;
    // FP64 quotients through the Newton-refined reciprocal where the kernel
    // has registers to spare (p <= 5: KH 1024^2 p = 3 FP64 8 % faster, SD
    // 1.5 %); at p >= 6 its extra live values cost the FP64 kernels a wave
    // per SIMD (p = 7: 270 registers, 40 % slower), so they divide exactly
    constexpr bool NEWTON = sizeof(Real) == 8 && P <= 5;
    using Mx = M<Real, NEWTON>;
    // the mixed kernels convert their nodes to primitives in FP32 (the FP64
    // state rounded once) and take the node wave speeds of the fused stages
    // (the AV's element maxima, stage 3's for the next dt) from those prims
    // in FP32, the reference's precision for both (:396-409, :483-486); the
    // maxima are FP64, and lam_el of RHS calls stays FP64 in every form
    constexpr bool PRIM32 = sizeof(Real) == 4;
    extern __shared__ __align__(16) unsigned char smem[];
    // dt and the go flag go from thread 0 to the block through the first
    // bytes of the prims region, which the node phase writes only after every
    // wave has read them (multi-wave blocks wait at a barrier; one wave reads
    // before it writes)
    double* s_dt = (double*)smem;
    int* s_go = (int*)(smem + 8);
    const int nthr = blockDim.x;
    const int E = a.E;
    const int el0 = tile * E;
    const int nE = min(E, a.NEL - el0);
    const size_t S = (size_t)a.NEL * N2;
    const double g = a.gamma;
    const Gas64 gs{g, a.gm1, a.efloor};

    const CTab<Real> D{(const Real*)a.tab};
    using AvTab = typename std::conditional<(sizeof(Real) == 8 && P <= 5), GTab<Real>, CTab<Real>>::type;
    const AvTab Dav{(const Real*)a.tab};              // the AV's operators (GTab)
    const CTab<Real> winv = D + (N * N + N);
    const CTab<Real> Dv = D + (N * N + 2 * N);      // volume matrix
    // the tile's first node in every FP64 stream (variable q at + q*S)
    const size_t tile_off = (size_t)el0 * N2;

    const Phys<Real, NEWTON> ph{kconst<Real>(g, a.g_f), kconst<Real>(a.gm1, a.gm1_f),
                                kconst<Real>(a.inv_gm1, a.inv_gm1_f)};

    // LDS regions
    Real* sW = (Real*)smem;                          // [E][N2][5] prims; reused as sR [E][4][N2]
    Real* sR = sW;
    double* sU = (double*)(smem + a.off_U);         // [E][4][N2] FP64 stage input, then stage values
    Real* sDu = (Real*)(smem + a.off_av);           // [E][N2] AV: du/dx
    Real* sDv = sDu + E * N2;                       // [E][N2] AV: dv/dy
    Real* sNu = sDv + E * N2;                       // [E]     AV: element viscosity
    Real* sUc = sNu + E;                            // [E][4][N2] MFMA_AV: conserved nodes, then the AV
    // [2][E][4][N2] MFMA collocation: the nodes' x- and y-fluxes; the x plane
    // then holds the volume term D f
    Real* sMf = (Real*)(smem + a.off_av);
    unsigned long long* sLamE = (unsigned long long*)(smem + a.off_lam);   // [E] max wave speed (bits)
    unsigned long long* sThE = sLamE + E;           // [E] max sensor (bits of the double)
    double* sEl = (double*)(smem + a.off_el);       // per-element scratch
    double* sBar = sEl;                              // [E][4]
    double* sTrho = sEl + 4 * E;                     // [E]
    unsigned long long* sTheta = (unsigned long long*)(sEl + 5 * E);  // [E]

    // SD: no AV; its lean kernel gathers the element wave speed for lam_el
    const bool need_lam = (FULL || SD) && (a.av || (a.mode == MODE_RHS && a.lam_el != nullptr));

    // line owned by this thread in phase B: element le, x-line (u, v as
    // stored) or y-line (u <-> v swapped); node k of the line is lbase + k*lstride
    const int nlines = E * 2 * N;
    const bool is_line = tid < nlines;
    const int le = is_line ? tid / (2 * N) : 0;
    const int ll = is_line ? tid - le * 2 * N : 0;
    const bool ydir = ll >= N;
    const int lidx = ydir ? ll - N : ll;
    const int lbase = ydir ? lidx : lidx * N;
    const int lstride = ydir ? N : 1;
    const bool line_live = is_line && le < nE;
    const Real jac = kconst<Real>(a.jac, a.jac_f);

    // Face neighbours of the line's two end nodes: inside the tile they are
    // read from LDS in phase B; outside it their FP64 state is loaded here,
    // with the tile's own nodes, so that no global load waits inside the
    // line phase.  hk: 0 = in tile, 1 = load from hn[], 2 = wall ghost.
    int hk[2] = {0, 0}, hnd[2] = {0, 0}, hel[2] = {0, 0};
    double Hn[2][4];
    if (line_live) {
        int ej, ei;
        elem_rc(el0 + le, a.K, a.inv_K, ej, ei);
        const int cpos = ydir ? ej : ei;
#pragma unroll
        for (int f = 0; f < 2; f++) {
            const bool edge = f ? (cpos == a.K - 1) : (cpos == 0);
            int c2 = f ? cpos + 1 : cpos - 1;
            if (edge) c2 = f ? 0 : a.K - 1;
            hel[f] = ydir ? c2 * a.K + ei : ej * a.K + c2;
            hnd[f] = lbase + (f ? 0 : N - 1) * lstride;
            hk[f] = (edge && a.bc == BC_WALL) ? 2 : (hel[f] >= el0 && hel[f] < el0 + nE) ? 0 : 1;
        }
    }
    // Every lane issues every global load of the start (neighbours, stage
    // input, U^n) as straight-line code: vector-memory counters retire in
    // order, and a load that some path may skip makes the compiler's waits
    // cover every load issued after the one needed (measured: 8 % faster
    // batched FP32 kernel, +-1 % elsewhere).  Lanes with nothing to load read
    // a valid stand-in address (the tile's first node).
#pragma unroll
    for (int f = 0; f < 2; f++) {
        const size_t o = hk[f] == 1 ? (size_t)hel[f] * N2 + hnd[f] : tile_off;
#pragma unroll
        for (int q = 0; q < 4; q++) Hn[f][q] = a.Up[q * S + o];
    }

    // ── A: load nodes, prims (FP64 or FP32, see PRIM32), store compute-precision prims ──
    // A thread owns at most ceil((p+1)/2) node slots (E(p+1)^2 nodes over at
    // least 2E(p+1) threads), slot(j) for j < JMAX, here and in the update
    // phase.  All of its global loads are issued before any conversion, so
    // each wave keeps many requests in flight.
    constexpr int JMAX = (N + 1) / 2;
    // PAIR (an even node count per element and an even slot count, p = 3
    // and 7): a thread owns its slots in pairs of adjacent nodes of one
    // element, slot(2m) = 2(tid + m*blockDim) and slot(2m)+1, and moves each
    // pair with one 16-byte load or store (the state pointers are 16-byte
    // aligned).  An element's nodes then span N2/2 lanes: the element wave
    // speed folds a lane's two nodes first and reduces over N2/2 lanes.
    constexpr bool PAIR = PAIR_OK && N2 % 2 == 0 && JMAX % 2 == 0;
    auto slot = [&](int j) { return PAIR ? 2 * (tid + (j >> 1) * nthr) + (j & 1) : tid + j * nthr; };
    const bool staged = a.mode == MODE_STAGE;
    double Ul[JMAX][4], U0l[JMAX][4], Upl[JMAX][4];
    gcptr upq[4];
#pragma unroll
    for (int q = 0; q < 4; q++) upq[q] = sbase(a.Up + tile_off + q * S);
    if (PAIR) {
#pragma unroll
        for (int j = 0; j < JMAX; j += 2) {
            const int s = slot(j);
            const unsigned sl = s < nE * N2 ? s : 0;     // pairs past the tile read nodes 0, 1
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const dbl2 v = ld2_at(upq[q], sl * 8u);
                Ul[j][q] = v.x; Ul[j + 1][q] = v.y;
            }
        }
    } else {
#pragma unroll
        for (int j = 0; j < JMAX; j++) {
            const int s = slot(j);
            const unsigned sl = s < nE * N2 ? s : 0;     // slots past the tile read node 0
#pragma unroll
            for (int q = 0; q < 4; q++) Ul[j][q] = ld_at(upq[q], sl * 8u);
        }
    }
    // U^n for the update (stages 2, 3).  With U0_EARLY (the lean kernels) it
    // is issued with the stage input, so that the update never waits for
    // memory (32 more registers, within their occupancy target; measured 2-4 %
    // faster); the shock-capturing kernels, whose AV and limiter code needs
    // the registers, and the FP32 batched kernel, whose occupancy they would
    // lower, issue it after the line phase.
    auto load_u0 = [&](const double* src) {
        gcptr u0q[4];
#pragma unroll
        for (int q = 0; q < 4; q++) u0q[q] = sbase(src + tile_off + q * S);
        if (PAIR) {
#pragma unroll
            for (int j = 0; j < JMAX; j += 2) {
                const int s = slot(j);
                const unsigned sl = s < nE * N2 ? s : 0;
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const dbl2 v = ld2_stream(u0q[q], sl * 8u);
                    U0l[j][q] = v.x; U0l[j + 1][q] = v.y;
                }
            }
        } else {
#pragma unroll
            for (int j = 0; j < JMAX; j++) {
                const int s = slot(j);
                const unsigned sl = s < nE * N2 ? s : 0;
#pragma unroll
                for (int q = 0; q < 4; q++) U0l[j][q] = ld_stream(u0q[q], sl * 8u);
            }
        }
    };
    // Issued in every stage (without U^n, in stage 1 and RHS mode, the stage
    // input is read a second time, from L2, and not used), behind a
    // scheduling barrier that keeps them after the stage input's loads, so
    // that a wait for the stage input never covers them.
    __builtin_amdgcn_sched_barrier(0);
    if (U0_EARLY) load_u0(staged && a.ca != 0.0 ? a.U0 : a.Up);
    // ── fused dt (device side, :1026-1035) ──
    // Decided after the tile's loads are issued (they are in bounds either
    // way), so the step-state read and the loads overlap.
    double dt = 0.0;
    if (a.mode == MODE_STAGE) {
        unsigned long long lam_b = 0ull;
        if (a.stage == 1 && tid < 64) {
            // max over the partial slots (non-negative doubles and NaN order
            // as their unsigned bit patterns, NaN above every number)
            lam_b = wave_umax(a.st->lam_part[sc.slot_cur][tid & (LAM_PARTS - 1)]);
        }
        if (tid == 0) {
            DevState* st = a.st;
            int go = 0;
            double d = 0.0;
            if (st->status == ST_OK) {
                if (a.stage == 1) {
                    const double lam = bitsd(lam_b);
                    const double t = st->t;
                    int status = ST_OK, skip = 0;
                    if (a.t_end > 0.0 && t >= a.t_end - 1e-12 * fmax(1.0, fabs(a.t_end))) {
                        skip = 1;
                    } else if (!(lam > 0.0)) {
                        // NaN / non-positive speed at step start (:1028); after a
                        // step of this call it is that step's blow-up (:1054)
                        status = sc.check_blowup ? ST_BLOWUP : ST_NONFINITE;
                    } else if (sc.check_blowup && lam > 1e6) {
                        status = ST_BLOWUP;
                    } else {
                        d = a.cfl * a.h / (lam * (double)N);
                        if (a.t_end > 0.0 && t + d > a.t_end) d = a.t_end - t;
                        if (d < 1e-14) status = ST_DT_UNDERFLOW;
                    }
                    go = status == ST_OK && !skip;
                    if (lead) {
                        st->dt = go ? d : 0.0;
                        st->skip = skip;
                        if (status != ST_OK) st->status = status;
                    }
                } else if (!st->skip) {
                    d = st->dt;
                    go = 1;
                    if (a.stage == 2 && lead) {
                        st->t += d;
                        st->steps += 1;
                        for (int k = 0; k < LAM_PARTS; k++) st->lam_part[sc.slot_next][k] = 0ull;
                    }
                }
            }
            *s_go = go;
            *s_dt = d;
        }
        __syncthreads();
        const int go = *s_go;
        dt = *s_dt;
        if (nthr > 64) __syncthreads();
        if (!go) return;
    }

    if ((FULL || SD) && (need_lam || a.av)) {
        // element maxima gathered with LDS atomics on the ordered bit patterns
        // of non-negative doubles (order-independent, so deterministic)
        if (tid < E) { sLamE[tid] = 0ull; sThE[tid] = 0ull; }
        __syncthreads();
    }
    unsigned long long lam_pair = 0ull;   // PAIR: wave speed of the pair's first node
    float lam_pair32 = 0.f;
    Real c2s[JMAX];                       // AV: c^2 of the thread's nodes, for the sensor
#pragma unroll
    for (int j = 0; j < JMAX; j++) {
        const int s = slot(j);
        if (s >= E * N2) continue;
        const int e = s / N2, nd = s - e * N2;
        double W[4] = {1.0, 0.0, 0.0, 1.0}, ri = 1.0;
        Prim<Real> w = {1, 0, 0, 1, 1};
        if (s < nE * N2) {
            if (UP_LDS && (staged || SD)) {
#pragma unroll
                for (int q = 0; q < 4; q++) sU[e * 4 * N2 + q * N2 + nd] = Ul[j][q];
            }
            if (PRIM32) w = ph.node_prim(Ul[j]);
            else Cons64::to_prim(Ul[j], W, gs, ri);
        }
        if (!PRIM32) w = make_prim<Real>(W);
        if (FULL && !SD) c2s[j] = Mx::div(ph.g * w.p, Mx::fmax_(w.r, (Real)1e-12));
        Real* p = sW + s * 5;
        p[0] = (Real)0.5 * w.r; p[1] = (Real)0.5 * w.u; p[2] = (Real)0.5 * w.v;
        p[3] = (Real)0.5 * w.p; p[4] = (Real)0.5 * w.z;
        if (MFMA_AV && FULL && a.av) {
            // the conserved variables the AV acts on, from the prims as the
            // line phase's VALU form computes them
            Real* u = sUc + e * 4 * N2 + nd;
            u[0] = w.r; u[N2] = w.r * w.u; u[2 * N2] = w.r * w.v;
            u[3 * N2] = w.p * ph.inv_gm1 + (Real)0.5 * w.r * (w.u * w.u + w.v * w.v);
        }
        if (MFMA_COL) {
            // the x- and y-fluxes the collocation pre-pass contracts
            Real fx[4], gy[4];
            ph.template flux<0>(w, fx);
            ph.template flux<1>(w, gy);
            Real* F = sMf + e * 4 * N2 + nd;
#pragma unroll
            for (int q = 0; q < 4; q++) { F[q * N2] = fx[q]; F[4 * E * N2 + q * N2] = gy[q]; }
        }
        if (PRIM32 && !SD && need_lam && staged) {
            // FP32 node speeds, reduced on their 32-bit patterns (they feed
            // the AV only)
            const float sp = (float)ph.node_speed(w);
            if (!PAIR) {
                elem_max_lds<N2>(&sLamE[e], sp, s < nE * N2);
            } else if (j & 1) {
                elem_max_lds<N2 / 2>(&sLamE[e], fmax_bits(sp, lam_pair32), s < nE * N2);
            } else {
                lam_pair32 = sp;
            }
        } else if ((FULL || SD) && need_lam) {
            // lam_el of RHS calls (the dt of separate calls comes from it)
            // stays FP64 in every form
            if (PRIM32) Cons64::to_prim(Ul[j], W, gs, ri);
            const unsigned long long sp =
                dbits(SD ? Cons64::wave_speed_sd(W, g, ri) : Cons64::wave_speed(W, g, ri));
            if (!PAIR) {
                elem_max_lds<N2>(&sLamE[e], sp, s < nE * N2);
            } else if (j & 1) {
                // second node of the pair: fold with the first, reduce the element
                const unsigned long long m = sp > lam_pair ? sp : lam_pair;
                elem_max_lds<N2 / 2>(&sLamE[e], m, s < nE * N2);
            } else {
                lam_pair = sp;
            }
        }
    }
    __syncthreads();

    // ── collocation pre-pass on the matrix cores: the volume term
    //    sum_k D_ik f(U_k) along x plus sum_k D_jk g(U_k) along y (:431-439)
    //    for every node, F_q D^T + D G_q on the 4 x 4 node arrays F_q, G_q of
    //    each element and variable, with the block layout of the AV's MFMA
    //    form below (block b = (element, variable), lane 4b + c; x-term step
    //    m: A[r] = F_q[r][m], B[c] = D[c][m]; y-term: A[r] = D[r][m],
    //    B[c] = G_q[m][c]); the result at node (jn = r, in = c) replaces F_q
    //    and the line phase adds only the face terms.  Each wave reads and
    //    writes the planes of the 8 elements of its own lines (reads first);
    //    elements past the tile hold the stand-in state. ──
    if (MFMA_COL) {
        typedef float f4 __attribute__((ext_vector_type(4)));
        const int l = tid & 63, c = l & 3, q = (l >> 2) & 3;
        float dv[4];
#pragma unroll
        for (int m = 0; m < 4; m++) dv[m] = Dv[c * 4 + m];
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int e = 8 * (tid >> 6) + 4 * h + (l >> 4);
            const bool ok = e < E;
            Real* Fq = sMf + (ok ? e : 0) * 4 * N2 + q * N2;
            const Real* Gq = Fq + 4 * E * N2;
            float fx[4], gy[4];
#pragma unroll
            for (int m = 0; m < 4; m++) { fx[m] = Fq[c * 4 + m]; gy[m] = Gq[m * 4 + c]; }
            f4 d = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int m = 0; m < 4; m++) d = __builtin_amdgcn_mfma_f32_4x4x1f32(fx[m], dv[m], d, 0, 0, 0);
#pragma unroll
            for (int m = 0; m < 4; m++) d = __builtin_amdgcn_mfma_f32_4x4x1f32(dv[m], gy[m], d, 0, 0, 0);
            if (ok) {
#pragma unroll
                for (int r = 0; r < 4; r++) Fq[r * 4 + c] = d[r];
            }
        }
    }

    // ── AV pre-pass: velocity divergence, sensor, element viscosity (:483-521) ──
    if (FULL && !SD && a.av) {
        if (is_line) {
            Real un[N];
#pragma unroll
            for (int k = 0; k < N; k++)
                un[k] = (Real)2 * sW[(le * N2 + lbase + k * lstride) * 5 + (ydir ? 2 : 1)];
            Real* dst = ydir ? sDv : sDu;
#pragma unroll
            for (int i = 0; i < N; i++) {
                Real d = 0;
#pragma unroll
                for (int k = 0; k < N; k++) d += Dav[i * N + k] * un[k];
                dst[le * N2 + lbase + i * lstride] = d;
            }
        }
        __syncthreads();
        // Ducros-type sensor per node, max per element (:488-491)
        const Real len_p = kconst<Real>(a.len_p, a.len_p_f);
        const Real dx2inv = kconst<Real>(a.dx2inv, a.dx2inv_f);
        // the node phase's slots (so c^2 comes from its registers) and its
        // element reduction: every lane of an element's group takes part;
        // slots past the tile write nothing
        float th_pair32 = 0.f;
        unsigned long long th_pair = 0ull;
#pragma unroll
        for (int j = 0; j < JMAX; j++) {
            const int s = slot(j);
            if (s >= E * N2) continue;
            const bool live = s < nE * N2;
            const int e = s / N2;
            // (FP32: quotients through the hardware reciprocal, as in the fluxes)
            Real divu = Mx::fmin_(jac * (sDu[s] + sDv[s]), (Real)0), du2 = divu * divu;
            Real th = Mx::div(du2, du2 + (Real)a.av_k * c2s[j] * dx2inv + (Real)1e-30);
            if (sizeof(Real) == 4) {
                if (!PAIR) {
                    elem_max_lds<N2>(&sThE[e], (float)th, live);
                } else if (j & 1) {
                    elem_max_lds<N2 / 2>(&sThE[e], fmax_bits((float)th, th_pair32), live);
                } else {
                    th_pair32 = (float)th;
                }
            } else {
                const unsigned long long tb = dbits((double)th);
                if (!PAIR) {
                    elem_max_lds<N2>(&sThE[e], tb, live);
                } else if (j & 1) {
                    elem_max_lds<N2 / 2>(&sThE[e], tb > th_pair ? tb : th_pair, live);
                } else {
                    th_pair = tb;
                }
            }
        }
        __syncthreads();
        if (tid < E)
            sNu[tid] = (Real)a.av_cav * (Real)bitsd(sThE[tid]) * len_p * (Real)bitsd(sLamE[tid]);
        __syncthreads();
        if (MFMA_AV) {
            // AV_q = -nu jac^2 (U_q Lav^T + Lav U_q) on the 4 x 4 node array
            // U_q[jn][in] of each element and variable, i.e. Lav along the
            // x-lines plus Lav along the y-lines (:501-521 in one operator,
            // Lav = W^-1 D^T W D), as 16 blocks of 4 x 4 x 1 MFMA per
            // instruction: block b = (element e, variable q), lane 4b + c.
            // The instruction forms D[r][c] += A[r] B[c] from the A values of
            // the block's lanes 4b + r and the B value of lane 4b + c (checked
            // on gfx950).  x-term, step m: A[r] = U_q[r][m], B[c] = Lav[c][m];
            // y-term: A[r] = Lav[r][m], B[c] = U_q[m][c].  Both give the
            // output at node (jn = r, in = c), written back over U_q (each wave
            // owns the 8 elements of its lines, and its reads precede its
            // writes).  Elements past the tile hold the stand-in state.
            typedef float f4 __attribute__((ext_vector_type(4)));
            const CTab<Real> Lav = D + (2 * N * N + 2 * N);
            const int l = tid & 63, c = l & 3, q = (l >> 2) & 3;
            float lav[4];
#pragma unroll
            for (int m = 0; m < 4; m++) lav[m] = Lav[c * 4 + m];
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int e = 8 * (tid >> 6) + 4 * h + (l >> 4);
                const bool ok = e < E;
                Real* U = sUc + (ok ? e : 0) * 4 * N2 + q * N2;
                float ux[4], uy[4];
#pragma unroll
                for (int m = 0; m < 4; m++) { ux[m] = U[c * 4 + m]; uy[m] = U[m * 4 + c]; }
                f4 d = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int m = 0; m < 4; m++) d = __builtin_amdgcn_mfma_f32_4x4x1f32(ux[m], lav[m], d, 0, 0, 0);
#pragma unroll
                for (int m = 0; m < 4; m++) d = __builtin_amdgcn_mfma_f32_4x4x1f32(lav[m], uy[m], d, 0, 0, 0);
                const Real f = -sNu[ok ? e : 0] * jac * jac;
                if (ok) {
#pragma unroll
                    for (int r = 0; r < 4; r++) U[r * 4 + c] = f * d[r];
                }
            }
        }
    }

    // ── B: line phase ──
    Real acc[N][4];     // set by line_volume (lines only; other lanes write nothing)

    if (is_line) {
        const int yd = ydir ? 1 : 0;
        Prim<Real> hw[N];    // half-scaled line states
        const Real* eb = sW + le * N2 * 5;
#pragma unroll
        for (int k = 0; k < N; k++) {
            // y-lines read u and v swapped (the rotation is in the LDS index)
            const Real* p = eb + (lbase + k * lstride) * 5;
            hw[k].r = p[0]; hw[k].u = p[1 + yd]; hw[k].v = p[2 - yd]; hw[k].p = p[3]; hw[k].z = p[4];
        }

        // ── surface terms at both ends of the line, evaluated first: their
        //    neighbour states are ready.  Unrolled: a loop over the two faces
        //    picks each face's inputs and outputs with selects (measured:
        //    unrolled, 1.9-4 % fewer VALU instructions per wave; vortex p = 7
        //    1 % (mixed) / 2 % (FP64), KH p = 3 2 % / 3.5 % faster)
        Real cm[4] = {0, 0, 0, 0}, cp[4] = {0, 0, 0, 0};
        if (line_live) {
#pragma unroll
            for (int f = 0; f < 2; f++) {
                const Prim<Real> own = doubled(f ? hw[N - 1] : hw[0]);
                const int kind = f ? hk[1] : hk[0];
                Prim<Real> nb;
                if (kind == 2) {
                    nb = wall_ghost(own, 0);
                } else {
                    if (kind == 0) {
                        // neighbour staged by this tile: the same FP64 -> prim
                        // conversion, so the face sees bit-identical states
                        const int elnb = f ? hel[1] : hel[0], ndnb = f ? hnd[1] : hnd[0];
                        const Real* pn = sW + ((elnb - el0) * N2 + ndnb) * 5;
                        nb.r = pn[0]; nb.u = pn[1 + yd]; nb.v = pn[2 - yd]; nb.p = pn[3]; nb.z = pn[4];
                        nb = doubled(nb);
                    } else {
                        double U[4], W[4];
#pragma unroll
                        for (int q = 0; q < 4; q++) U[q] = f ? Hn[1][q] : Hn[0][q];
                        if (PRIM32) {
                            nb = ph.node_prim(U);
                        } else {
                            Cons64::to_prim(U, W, gs);
                            nb = make_prim<Real>(W);
                        }
                        rot(nb, ydir);
                    }
                }
                // split forms: F*(face) (the diagonal volume term cancels the
                // -f(own) of the strong-form surface term exactly on LGL
                // nodes); collocation: F*(face) - f(own); SD: F*(face), the
                // flux on the end flux node
                constexpr bool JUMP = VOL == VOL_COLLOCATION;
                Real dj[4];
                if (a.surf == SURF_RUSANOV) {
                    ph.template rusanov_side<JUMP>(own, nb, f != 0, dj);
                } else {
                    const Prim<Real> L = f ? own : nb, R = f ? nb : own;
                    ph.hllc(0, L, R, dj);
                    if (JUMP) {
                        Real fo[4];
                        ph.template flux<0>(own, fo);
#pragma unroll
                        for (int q = 0; q < 4; q++) dj[q] -= fo[q];
                    }
                }
                const Real c = SD ? (Real)1 : (f ? -jac : jac) * winv[f ? N - 1 : 0];
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    Real v = c * dj[q];
                    if (f) cp[q] = v; else cm[q] = v;
                }
            }
        }

        if (SD) {
            const CTab<Real> Isf = D + (3 * N * N + 2 * N);     // [N+1][N]
            const CTab<Real> Dsd = Isf + (N + 1) * N;           // [N][N+1]
            sd_line<N, Real>(ph, sU + le * 4 * N2, lbase, lstride, ydir, Isf, Dsd, cm, cp, acc);
        } else if (MFMA_COL) {
            // the volume term came from the matrix cores: the face terms alone
#pragma unroll
            for (int i = 0; i < N; i++)
#pragma unroll
                for (int q = 0; q < 4; q++) acc[i][q] = i == 0 ? cm[q] : (i == N - 1 ? cp[q] : (Real)0);
        } else {
            line_volume<N, Real, VOL, PIN>(ph, hw, Dv, cm, cp, acc);
        }

        // ── artificial viscosity along the line: -jac W^-1 D^T W (nu jac D Uc)
        //    = -nu jac^2 Lav Uc, Lav = W^-1 D^T W D precomputed (one pass) ──
        if (FULL && !SD && !MFMA_AV && a.av) {
            const Real f = -sNu[le] * jac * jac;
            const AvTab Lav = Dav + (2 * N * N + 2 * N);
            Real uc[N][4];      // conserved variables of the line's nodes (u, v as rotated)
#pragma unroll
            for (int k = 0; k < N; k++) {
                const Prim<Real> w = doubled(hw[k]);
                uc[k][0] = w.r;
                uc[k][1] = w.r * w.u;
                uc[k][2] = w.r * w.v;
                uc[k][3] = w.p * ph.inv_gm1 + (Real)0.5 * w.r * (w.u * w.u + w.v * w.v);
            }
#pragma unroll
            for (int i = 0; i < N; i++)
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    Real s = 0;
#pragma unroll
                    for (int m = 0; m < N; m++) s += Lav[i * N + m] * uc[m][q];
                    acc[i][q] += f * s;
                }
        }
    }
    // U^n (shock-capturing kernels) and, without UP_LDS, the stage input
    // again, issued now so that the loads are in flight across the barriers
    // and the LDS exchange below
    if (!U0_EARLY && staged && a.ca != 0.0) load_u0(a.U0);
    if (!UP_LDS && staged) {
#pragma unroll
        for (int j = 0; j < JMAX; j++) {
            const int s = slot(j);
            if (s < nE * N2) {
#pragma unroll
                for (int q = 0; q < 4; q++) Upl[j][q] = ld_at(upq[q], (unsigned)s * 8u);
            }
        }
    }
    __syncthreads();   // every read of sW is done: reuse it as sR
    if (is_line && !ydir) {
        Real* r = sR + le * 4 * N2;
#pragma unroll
        for (int i = 0; i < N; i++)
#pragma unroll
            for (int q = 0; q < 4; q++) r[q * N2 + lbase + i * lstride] = acc[i][q];
    }
    __syncthreads();
    if (is_line && ydir) {
        // y-line momentum components are in the rotated frame: (v, u)
        Real* r = sR + le * 4 * N2;
        constexpr int qs[4] = {0, 2, 1, 3};
#pragma unroll
        for (int i = 0; i < N; i++)
#pragma unroll
            for (int q = 0; q < 4; q++) r[qs[q] * N2 + lbase + i * lstride] += acc[i][q];
    }
    __syncthreads();

    // ── C: node phase ──
    // R of node nd of element e, variable q (with the matrix-core forms' AV or
    // collocation volume term)
    auto rhs = [&](int e, int nd, int q) -> Real {
        const Real r = sR[e * 4 * N2 + q * N2 + nd];
        if (MFMA_AV && FULL && a.av) return r + sUc[e * 4 * N2 + q * N2 + nd];
        if (MFMA_COL) return r + sMf[e * 4 * N2 + q * N2 + nd];
        return r;
    };
    if (a.mode == MODE_RHS) {
        for (int s = tid; s < nE * N2; s += nthr) {
            const int e = s / N2, nd = s - e * N2;
            const size_t o = (size_t)(el0 + e) * N2 + nd;
#pragma unroll
            for (int q = 0; q < 4; q++) a.out[q * S + o] = (double)rhs(e, nd, q);
        }
        if ((FULL || SD) && a.lam_el != nullptr && tid < nE) a.lam_el[el0 + tid] = bitsd(sLamE[tid]);
        return;
    }

    // MODE_STAGE: U_s = ca*U0 + cb*(Up + dt*R), sanitize, [limit], write, [lambda]
    const bool lim = FULL && !SD && a.limiter;     // Zhang-Shu; SD repairs per node below
    const bool want_lambda = a.stage == 3;
    double lmax = 0.0;
    gptr outq[4];
#pragma unroll
    for (int q = 0; q < 4; q++) outq[q] = sbase(a.out + tile_off + q * S);
    // stage value of slot j (node s of the tile): Shu-Osher combination,
    // sanitize_cons, the SD repair
    // (ke: the kinetic energy sanitize_cons computed, for the limiter test)
    auto stage_value = [&](int j, int s, double (&Us)[4], double& ke) {
        const int e = s / N2, nd = s - e * N2;
        const double* up = sU + e * 4 * N2 + nd;
        double Ui[4];
#pragma unroll
        for (int q = 0; q < 4; q++) Ui[q] = UP_LDS ? up[q * N2] : Upl[j][q];
        if (a.ca != 0.0) {
#pragma unroll
            for (int q = 0; q < 4; q++) Us[q] = a.ca * U0l[j][q] + a.cb * (Ui[q] + dt * (double)rhs(e, nd, q));
        } else {
#pragma unroll
            for (int q = 0; q < 4; q++) Us[q] = Ui[q] + dt * (double)rhs(e, nd, q);
        }
        ke = Cons64::sanitize_cons(Us, gs);
        if (SD && a.limiter) Cons64::repair(Us, gs, a.eps_rho, a.eps_p);
    };
    // a stage value the Zhang-Shu limiter may change: rho < eps_rho, or a
    // pressure within a rounding margin of eps_p (NaN counts as such)
    bool lim_bad = false;
    // (the pressure of a sanitized value from the kinetic energy sanitize_cons
    // computed: pressure_from_cons, bit for bit)
    auto needs_limit = [&](const double (&Us)[4], double ke) {
        return !(Us[0] >= a.eps_rho) || !(gs.gm1 * (Us[3] - ke) >= 2.0 * a.eps_p + 1e-13 * fabs(Us[3]));
    };
    // wave speed of a new node for the next dt (its maximum is taken in FP64);
    // the mixed kernels compute it from the node's FP32 prims, as they do for
    // the flux
    auto speed = [&](const double (&Us)[4]) {
        if (PRIM32) return (double)ph.template node_speed<SD>(ph.node_prim(Us));
        double W[4], ri;
        Cons64::to_prim(Us, W, gs, ri);
        return SD ? Cons64::wave_speed_sd(W, g, ri) : Cons64::wave_speed(W, g, ri);
    };
    // With the limiter on, the mixed kernels keep the stage values in
    // registers, and LDS takes them only on a tile that the limiter may
    // change (see below); the FP64 kernels, with fewer registers to spare
    // (p = 3: 170 instead of 153, a wave per SIMD less), write them to LDS
    // at once.
    constexpr bool LIM_REG = sizeof(Real) == 4;
    double Sv[JMAX][4];
    if (PAIR) {
#pragma unroll
        for (int j = 0; j < JMAX; j += 2) {
            const int s = slot(j);
            if (s >= nE * N2) continue;
            double A[4], B[4];
            double keA, keB;
            stage_value(j, s, A, keA);
            stage_value(j + 1, s + 1, B, keB);
            if (lim) {
                const int e = s / N2, nd = s - e * N2;
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    if (LIM_REG) {
                        Sv[j][q] = A[q]; Sv[j + 1][q] = B[q];
                    } else {
                        sU[e * 4 * N2 + q * N2 + nd] = A[q];
                        sU[e * 4 * N2 + q * N2 + nd + 1] = B[q];
                    }
                }
                lim_bad = lim_bad || needs_limit(A, keA) || needs_limit(B, keB);
            } else {
#pragma unroll
                for (int q = 0; q < 4; q++) st2_at(outq[q], (unsigned)s * 8u, dbl2{A[q], B[q]});
                if (want_lambda) lmax = nanmax(nanmax(lmax, speed(A)), speed(B));
            }
        }
    } else {
#pragma unroll
        for (int j = 0; j < JMAX; j++) {
            const int s = slot(j);
            if (s >= nE * N2) continue;
            double Us[4];
            double ke;
            stage_value(j, s, Us, ke);
            if (lim) {
                const int e = s / N2, nd = s - e * N2;
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    if (LIM_REG) Sv[j][q] = Us[q];
                    else sU[e * 4 * N2 + q * N2 + nd] = Us[q];
                }
                lim_bad = lim_bad || needs_limit(Us, ke);
            } else {
#pragma unroll
                for (int q = 0; q < 4; q++) st_at(outq[q], (unsigned)s * 8u, Us[q]);
                if (want_lambda) lmax = nanmax(lmax, speed(Us));
            }
        }
    }
    if (lim) {
        // With every node at rho >= eps_rho and p >= 2 eps_p (+ a relative
        // margin), the limiter's theta is 1 in every element of the tile and
        // it returns Ubar + (U - Ubar), U up to rounding: such a tile (on
        // smooth and on most shocked data, every tile) skips it, and its
        // stage values go from registers to memory.  Otherwise they pass
        // through LDS, where the limiter acts on whole elements.  The
        // workgroup reduction is also the barrier after every read of the
        // stage input in LDS, which the stage values overwrite.
        const bool bad = __syncthreads_or(lim_bad);
        if (bad) {
            if (LIM_REG) {
#pragma unroll
                for (int j = 0; j < JMAX; j++) {
                    const int s = slot(j);
                    if (s >= nE * N2) continue;
                    const int e = s / N2, nd = s - e * N2;
#pragma unroll
                    for (int q = 0; q < 4; q++) sU[e * 4 * N2 + q * N2 + nd] = Sv[j][q];
                }
            }
            limit_block<N>(sU, a.w64, nE, g, a.eps_rho, a.eps_p, sBar, sTrho, sTheta);   // barriers on entry and exit
        }
        if (bad || !LIM_REG) {
#pragma unroll
            for (int j = 0; j < JMAX; j++) {
                const int s = slot(j);
                if (s >= nE * N2) continue;
                const int e = s / N2, nd = s - e * N2;
#pragma unroll
                for (int q = 0; q < 4; q++) Sv[j][q] = sU[e * 4 * N2 + q * N2 + nd];
            }
        }
        if (PAIR) {
#pragma unroll
            for (int j = 0; j < JMAX; j += 2) {
                const int s = slot(j);
                if (s >= nE * N2) continue;
#pragma unroll
                for (int q = 0; q < 4; q++) st2_at(outq[q], (unsigned)s * 8u, dbl2{Sv[j][q], Sv[j + 1][q]});
                if (want_lambda) lmax = nanmax(nanmax(lmax, speed(Sv[j])), speed(Sv[j + 1]));
            }
        } else {
#pragma unroll
            for (int j = 0; j < JMAX; j++) {
                const int s = slot(j);
                if (s >= nE * N2) continue;
#pragma unroll
                for (int q = 0; q < 4; q++) st_at(outq[q], (unsigned)s * 8u, Sv[j][q]);
                if (want_lambda) lmax = nanmax(lmax, speed(Sv[j]));
            }
        }
    }
    if (want_lambda) {
        // one atomic per wave on the ordered bit pattern (the speeds are
        // non-negative; a NaN, of either sign, orders above every number)
        const unsigned long long m = wave_umax(dbits(lmax));
        if ((tid & 63) == 63) atomicMax(&a.st->lam_part[sc.slot_next][blockIdx.x & (LAM_PARTS - 1)], m);
    }
}

template <int P, typename Real, bool FULL, int VOL, bool MFMA_AV = false>
__global__ __launch_bounds__(BLOCK, (Occ<P, Real, VOL, FULL>::waves)) void stage_kernel(KArgs a) {
    const StepCtl sc{a.slot_cur, a.slot_next, a.check_blowup};
    stage_tile<P, Real, FULL, VOL, (sizeof(Real) == 8), true, !FULL, true, MFMA_AV>(
        a, sc, tile_of_block(blockIdx.x, gridDim.x, a.xcd), blockIdx.x == 0, (int)threadIdx.x);
}

// Batched stage: many independent solves (own K, p, domain, dt) in one
// launch.  tiles[] lists (instance, tile) pairs, heaviest first, one
// workgroup per tile (a grid-stride loop over the list lets LLVM hoist the
// loop-invariant work of all seven degree bodies out of the loop, which
// costs ~70 registers and spills).  args[3*i + stage-1] holds instance i's
// stage arguments (its own device step state and wave-speed slots).  The
// kernel carries every degree's body; with the accumulator pin its register
// budget fits 4 waves/SIMD in FP32 (2 in FP64).  In FP32 the stage input is
// loaded again for the update instead of being kept in LDS (UP_LDS false):
// the LDS of its up-to-256-lane workgroups would otherwise limit the
// occupancy (11 % faster); in FP64 registers limit it first and keeping the
// input in LDS is faster, as is loading U^n at the start (5 %; in FP32 that
// costs occupancy and 13 %, measured).
template <typename Real> struct BatchOcc {
    static constexpr int waves = sizeof(Real) == 4 ? 2 : 1;
    static constexpr bool up_lds = sizeof(Real) == 8;     // stage input kept in LDS (FP64 only)
    static constexpr bool u0_early = sizeof(Real) == 8;   // U^n loaded at the start (FP64 only)
    // node slots in pairs with 16-byte accesses (see stage_tile): FP64 1.4 %
    // faster, FP32 2.4 % slower (measured, same box and ensemble)
    static constexpr bool pair = sizeof(Real) == 8;
};

// One tile of the batched stage: the instance's degree selects the body.
template <typename Real, bool FULL, int VOL>
__device__ __forceinline__ void batch_tile(const KArgs& a, const StepCtl sc, int tile, bool lead, int tid) {
    constexpr bool UL = BatchOcc<Real>::up_lds, UE = BatchOcc<Real>::u0_early, PR = BatchOcc<Real>::pair;
    switch (a.P) {
        case 1: stage_tile<1, Real, FULL, VOL, true, UL, UE, PR>(a, sc, tile, lead, tid); break;
        case 2: stage_tile<2, Real, FULL, VOL, true, UL, UE, PR>(a, sc, tile, lead, tid); break;
        case 3: stage_tile<3, Real, FULL, VOL, true, UL, UE, PR>(a, sc, tile, lead, tid); break;
        case 4: stage_tile<4, Real, FULL, VOL, true, UL, UE, PR>(a, sc, tile, lead, tid); break;
        case 5: stage_tile<5, Real, FULL, VOL, true, UL, UE, PR>(a, sc, tile, lead, tid); break;
        case 6: stage_tile<6, Real, FULL, VOL, true, UL, UE, PR>(a, sc, tile, lead, tid); break;
        default: stage_tile<7, Real, FULL, VOL, true, UL, UE, PR>(a, sc, tile, lead, tid); break;
    }
}

// One workgroup per list entry; entries with instance -1 pad the XCD-binned
// list (see batch_tiles) and do nothing.
template <typename Real, bool FULL, int VOL>
__global__ __launch_bounds__(BLOCK, (BatchOcc<Real>::waves)) void batch_stage_kernel(
    const KArgs* __restrict__ args, const int2* __restrict__ tiles, int ntiles, int stage, StepCtl sc) {
    const int t = blockIdx.x;
    if (t >= ntiles) return;
    const int2 tl = tiles[t];
    if (tl.x < 0) return;
    const KArgs& a = args[(size_t)tl.x * 3 + (stage - 1)];
    batch_tile<Real, FULL, VOL>(a, sc, tl.y, tl.y == 0, (int)threadIdx.x);
}

// Persistent form of the batched stage: a grid of as many workgroups as are
// resident on the chip at once walks the tile list (heaviest first) with a
// grid stride.  The thread index is taken through the tile's argument block
// (tid ^ a.zero, a.zero = 0): otherwise the compiler hoists the index
// arithmetic of all seven degree bodies out of the loop and keeps it live
// across it (256 registers and spills).  The barrier at the end of an
// iteration keeps the next tile's LDS writes behind this tile's reads.
// (With a grid that is a multiple of 8, a block walks only list positions
// of its own XCD's bin.)
template <typename Real, bool FULL, int VOL>
__global__ __launch_bounds__(BLOCK, (BatchOcc<Real>::waves)) void batch_stage_persistent(
    const KArgs* __restrict__ args, const int2* __restrict__ tiles, int ntiles, int stage, StepCtl sc) {
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int2 tl = tiles[t];
        if (tl.x < 0) continue;
        const KArgs& a = args[(size_t)tl.x * 3 + (stage - 1)];
        batch_tile<Real, FULL, VOL>(a, sc, tl.y, tl.y == 0, (int)threadIdx.x ^ a.zero);
        __syncthreads();
    }
}

// Host-side LDS layout shared by the launcher and the kernel.
struct LdsLayout {
    int off_U, off_av, off_lam, off_el, total;
};

inline int align_up(int x, int a) { return (x + a - 1) / a * a; }

// `stage`: the FP64 input state is staged in LDS (sU) for the update phase;
// `lim`: the limiter's per-element scratch.
// `mfma`: the matrix-core forms' planes: with av the AV's conserved-node
// planes [E][4][N2] after its other arrays; without it the collocation
// form's flux planes [2][E][4][N2].
inline LdsLayout lds_layout(int N, int E, int real_bytes, bool stage, bool av, bool lam, bool lim,
                            bool mfma = false) {
    const int N2 = N * N;
    LdsLayout L;
    int off = align_up(E * N2 * 5 * real_bytes, 16);
    L.off_U = off;
    if (stage) off += align_up(E * 4 * N2 * 8, 16);
    L.off_av = off;
    if (av) off += align_up((2 * E * N2 + E + (mfma ? 4 * E * N2 : 0)) * real_bytes, 16);
    else if (mfma) off += align_up(8 * E * N2 * real_bytes, 16);
    L.off_lam = off;
    if (lam || av) off += align_up(2 * E * 8, 16);
    L.off_el = off;
    if (lim) off += align_up(6 * E * 8, 16);
    L.total = off;
    return L;
}

}  // namespace dg
