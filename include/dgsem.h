/*
 * dgsem.h — C ABI of the MI355X-native 2-D compressible-Euler DGSEM step.
 *
 * The reference (KLozes/fvStuff, dgsem_lobatto_2d.cu) has no library interface:
 * its "API" is the set of __global__ kernels plus the __constant__ operator
 * tables that main() drives.  Each entry point below replaces one of them
 * (file:line cited per function).  Everything here is plain C: device buffers
 * are passed as raw pointers (hipMalloc'd by the caller), sizes as ints.
 *
 * State layout (identical to the reference's SIDX macro, dgsem_lobatto_2d.cu:126):
 *   Q[var*NEL*NN2 + el*NN2 + jn*NN + in],  var in {rho, rho*u, rho*v, E},
 *   el = ej*K + ei,  NN = p+1,  NN2 = NN*NN,  NEL = K*K.
 * The state is always FP64.  DG_MIXED computes the right-hand side in FP32 and
 * keeps the stage accumulation, reductions and positivity checks in FP64.
 * The state passed to dg_compute_rhs, dg_advance / dg_step and
 * dg_batch_advance must be 16-byte aligned (hipMalloc'd memory is; the stage
 * kernels move node pairs with 16-byte accesses): a misaligned state is
 * rejected with DG_ERR_INVALID_ARG.
 *
 * Errors: every function returns DG_OK (0) or a positive status code; nothing
 * calls exit() (the reference's CK macro exits, dgsem_lobatto_2d.cu:129-136).
 * dg_last_error() gives a human-readable message for the last failure on a ctx.
 */
#ifndef DGSEM_H
#define DGSEM_H

#ifdef __cplusplus
extern "C" {
#endif

/* ── status codes ───────────────────────────────────────────────────────── */
#define DG_OK                 0
#define DG_ERR_INVALID_ARG    1  /* bad config / null pointer / unsupported p      */
#define DG_ERR_HIP            2  /* a HIP runtime call failed                      */
#define DG_ERR_NONFINITE      3  /* max wave speed not > 0 (NaN state), ref :1028  */
#define DG_ERR_DT_UNDERFLOW   4  /* dt < 1e-14, ref :1032-1035                     */
#define DG_ERR_BLOWUP         5  /* max wave speed > 1e6, ref :1054-1057           */

/* ── switches ───────────────────────────────────────────────────────────── */
enum dg_volume_flux {
    DG_VOL_CHANDRASHEKAR = 0,  /* EC_FLUX=1 of the reference, :259-285, :417-430 */
    DG_VOL_RANOCHA       = 1,  /* entropy-conservative + kinetic-energy-preserving */
    DG_VOL_COLLOCATION   = 2   /* EC_FLUX=0 of the reference, :431-439           */
};
enum dg_surface_flux {
    DG_SURF_HLLC    = 0,       /* hllc_n, :288-335                               */
    DG_SURF_RUSANOV = 1        /* local Lax-Friedrichs                           */
};
enum dg_bc {
    DG_BC_WALL     = 0,        /* wall_state reflecting ghost, :339-344          */
    DG_BC_PERIODIC = 1
};
enum dg_precision {
    DG_FP64  = 0,
    DG_MIXED = 1               /* FP32 volume/surface, FP64 state and reductions */
};
/* Spatial scheme.  DG_SCHEME_SD is the sibling solver of fr_lobatto_2d.cu:
 * spectral difference on the same LGL solution nodes and state layout, with
 * p + 2 flux nodes per direction {-1, Gauss-Legendre nodes, +1}; the
 * conservative state is interpolated to the interior flux nodes, the
 * physical flux is evaluated there, the Riemann flux (surface_flux) on the
 * two end nodes, and the flux polynomial is differentiated at the solution
 * nodes (compute_sd_rhs, fr_lobatto_2d.cu:332-525).  It has no volume flux
 * and no artificial viscosity (av must be 0); its `limiter` is the
 * reference's nodal positivity repair (fr_lobatto_2d.cu:589-615) instead of
 * the Zhang-Shu limiter.  Batched SD solves (every instance SD) run in the
 * by-degree form only. */
enum dg_scheme {
    DG_SCHEME_DGSEM = 0,       /* dgsem_lobatto_2d.cu                            */
    DG_SCHEME_SD    = 1        /* fr_lobatto_2d.cu                               */
};
enum dg_ic_kind {
    DG_IC_CIRCULAR_SOD      = 0,  /* ic_kernel, :692-717                          */
    DG_IC_VORTEX            = 1,  /* isentropic vortex, q1_dg_euler_2d.cu:2166-2188 */
    DG_IC_KELVIN_HELMHOLTZ  = 2   /* step shear layer + seeded noise              */
};

/* Solver configuration.  dg_config_default() fills the reference's defaults
 * (p=3, walls, Chandrashekar + HLLC, AV on, limiter on, [0,1]^2, FP64). */
typedef struct dg_config {
    int    K;             /* elements per side (NEL = K*K)                       */
    int    p;             /* polynomial degree, 1..7                             */
    double x0, y0;        /* lower-left corner of the square domain              */
    double L;             /* side length; h = L/K (reference: L=1, :922)         */
    double gamma;         /* ratio of specific heats (reference GAMMA_V=1.4)     */
    int    volume_flux;   /* enum dg_volume_flux                                 */
    int    surface_flux;  /* enum dg_surface_flux                                */
    int    bc;            /* enum dg_bc                                          */
    int    av;            /* Ducros-sensor artificial viscosity on/off (AV_ON)   */
    int    limiter;       /* Zhang-Shu limiter after each stage (SD: nodal repair) */
    int    precision;     /* enum dg_precision                                   */
    double av_cav;        /* AV_CAV (0.25)                                       */
    double av_ksensor;    /* AV_KSENSOR (0.1)                                    */
    double eps_rho;       /* limiter floors (1e-6, :1007)                        */
    double eps_p;
    int    elems_per_block; /* tuning: elements per workgroup, 0 = heuristic     */
    int    xcd_remap;       /* tuning: XCD-aware block->tile map, -1 = default on */
    int    scheme;          /* enum dg_scheme (DG_SCHEME_DGSEM by default)      */
    int    mfma;            /* tuning: the tensor-product contraction on the
                               matrix cores (p = 3, DG_MIXED): the artificial
                               viscosity's (av on), or the collocation volume
                               term's (DG_VOL_COLLOCATION, av off); -1 =
                               dg_autotune decides (off until then), 0 = off,
                               1 = on                                            */
} dg_config;

/* Initial-condition parameters (unused fields are ignored). */
typedef struct dg_ic_params {
    int    kind;          /* enum dg_ic_kind                                     */
    /* circular Sod: inside (rho_hi, p_hi), outside (0.125, 0.1), radius L/4     */
    double rho_hi, p_hi;  /* RHO_HI=11, P_HI=10                                  */
    /* vortex: strength eps, centre (xc, yc), background velocity (u0, v0)       */
    double vortex_eps, xc, yc, u0, v0;
    /* Kelvin-Helmholtz: seed and amplitude of the uniform velocity noise        */
    unsigned long long seed;
    double noise_amp;
} dg_ic_params;

typedef struct dg_ctx dg_ctx;

void        dg_config_default(dg_config* cfg);
void        dg_ic_default(dg_ic_params* ic, int kind);

/* Build LGL nodes/weights, the differentiation matrix and the workspace
 * (replaces build_D + cudaMemcpyToSymbol, dgsem_lobatto_2d.cu:160-179,924-940,
 * and the scratch cudaMallocs of :957-974).  `stream` is a hipStream_t (NULL =
 * the null stream).  Every later call on the ctx is ordered on that stream. */
int         dg_create(const dg_config* cfg, void* stream, dg_ctx** out);
int         dg_destroy(dg_ctx* ctx);
const char* dg_last_error(const dg_ctx* ctx);
const char* dg_status_string(int status);

/* Sizes: number of doubles in one state vector (4*NEL*NN2) and NEL. */
long long   dg_state_size(const dg_ctx* ctx);
long long   dg_num_elements(const dg_ctx* ctx);
/* Copy the host-side operator tables: x[NN], w[NN], D[NN*NN] (row-major D_ik). */
int         dg_operators(const dg_ctx* ctx, double* x, double* w, double* D);

/* DG_SCHEME_SD only: the flux nodes xf[p+2], the interpolation matrix
 * Isf[(p+2)*(p+1)] (row-major, Isf[k][j] = l_j(xf_k), the Lagrange basis of
 * the solution nodes at the flux nodes) and the derivative matrix
 * Dfs[(p+1)*(p+2)] (row-major, Dfs[i][k] = derivative of the flux nodes'
 * Lagrange basis l^f_k at solution node x_i).  Replaces build_Isf /
 * build_Dfs + cudaMemcpyToSymbol, fr_lobatto_2d.cu:99-166,840-867.  Any
 * pointer may be NULL. */
int         dg_sd_operators(const dg_ctx* ctx, double* xf, double* Isf, double* Dfs);

/* Evaluate the IC at the LGL nodes into device state Q, then apply the
 * positivity limiter if it is enabled (ic_kernel + positivity_limiter,
 * dgsem_lobatto_2d.cu:692-717,1006-1007).  DG_SCHEME_SD: the circular Sod
 * front has the width 2h of fr_lobatto_2d.cu:633-657 (0.5h for DGSEM) and no
 * limiter pass follows (fr_lobatto_2d.cu:922-937). */
int         dg_init_state(dg_ctx* ctx, const dg_ic_params* ic, double* Q);

/* du/dt of device state Q into RHS (FP64, same layout) and the per-element
 * max wave speed lam_el[NEL] (may be NULL).  compute_dg_rhs, :365-530;
 * DG_SCHEME_SD: compute_sd_rhs, fr_lobatto_2d.cu:332-525. */
int         dg_compute_rhs(dg_ctx* ctx, const double* Q, double* RHS, double* lam_el);

/* One SSP-RK3 Shu-Osher stage followed by sanitize_cons and (if enabled) the
 * positivity limiter (rk3_s1/s2/s3 + positivity_limiter, :533-576,584-643):
 *   stage 1: Uout = U0 + dt*L
 *   stage 2: Uout = 3/4*U0 + 1/4*(Uprev + dt*L)
 *   stage 3: Uout = 1/3*U0 + 2/3*(Uprev + dt*L)
 * Uprev is ignored for stage 1.  Uout may alias U0 (stage 3 of the reference
 * writes Q while Q0 holds the copy).  DG_SCHEME_SD: rk3_fr_s1/s2/s3 +
 * positivity_limiter, fr_lobatto_2d.cu:542-615 (the same stages, the nodal
 * repair). */
int         dg_rk3_stage(dg_ctx* ctx, int stage, double* Uout, const double* U0,
                         const double* Uprev, const double* L, double dt);

/* Zhang-Shu positivity limiter in place (positivity_limiter, :584-643);
 * DG_SCHEME_SD: the nodal repair (rho < eps_rho: rho = eps_rho and zero
 * momentum; then p < eps_p: E reset to p = eps_p), fr_lobatto_2d.cu:589-615. */
int         dg_limit(dg_ctx* ctx, double* Q);

/* Max of lam_el[NEL] (device) -> *lam_max (host) and the CFL step
 * dt = cfl*h/(lam_max*(p+1))  (reduce_max/gpu_max + :1029).  Either output
 * pointer may be NULL.  Returns DG_ERR_NONFINITE if lam_max is not > 0. */
int         dg_cfl_dt(dg_ctx* ctx, const double* lam_el, double cfl,
                      double* lam_max, double* dt);

/* Fused time stepping: `nsteps` SSP-RK3 steps of device state Q in place.
 * Each stage is one kernel that computes the right-hand side and applies the
 * Shu-Osher update (and the limiter) without writing the RHS to memory; dt is
 * computed on the device from the wave speed of the state at the start of the
 * step, dt = min(cfl*h/(lam_max*(p+1)), t_end - t) (the reference's loop,
 * :1022-1058, with frame clipping replaced by one end time; pass t_end <= 0
 * for no clipping).  No host synchronisation happens until the end.  Steps
 * after t_end is reached are still launched and end early on the device
 * (each still streams the state once), so size nsteps from the last dt when
 * advancing to an end time.
 * t_inout: simulation time in/out (may be NULL); dt_last: last dt (may be NULL).
 * Returns the first numerical failure.  NONFINITE (at the first step of the
 * call) and DT_UNDERFLOW are found when a step starts, before it changes
 * anything: Q and *t_inout are those of the last good step.  BLOWUP (max
 * wave speed not > 0 or > 1e6 after a step, :1052-1057) is found after the
 * step that blew up has written Q: Q is that step's result and *t_inout
 * includes its dt (the reference's loop breaks before adding it, :1054-1058).
 * The last step taken is checked even when the steps after it are no-ops.
 * DG_SCHEME_SD: the same loop (fr_lobatto_2d.cu:937-990). */
int         dg_advance(dg_ctx* ctx, double* Q, int nsteps, double cfl, double t_end,
                       double* t_inout, double* dt_last);

/* Steps actually taken by the last dg_advance / dg_step on this ctx (steps
 * past t_end are launched as no-ops and not counted). */
int         dg_steps_taken(const dg_ctx* ctx);

/* Replay the fused loop's stage launches from a captured HIP graph (two steps
 * per replay; default on).  The graph is re-captured when Q, cfl, t_end or
 * the schedule change.  Results are identical either way. */
int         dg_set_graph(dg_ctx* ctx, int use_graph);

/* Same as dg_advance with nsteps = 1 (one full step of the reference loop). */
int         dg_step(dg_ctx* ctx, double* Q, double cfl, double* dt_used);

/* Per-element shock sensor theta_e (compute_sensor, :662-689), device out. */
int         dg_sensor(dg_ctx* ctx, const double* Q, double* sensor_el);

/* GLL-quadrature totals on the device, copied to host:
 * out[0..3] = integrals of rho, rho*u, rho*v, E;  out[4] = integral of the
 * entropy -rho*s/(gamma-1), s = ln(p*rho^-gamma). */
int         dg_totals(dg_ctx* ctx, const double* Q, double out[5]);

/* dg_totals without a context and without any device allocation, for
 * callers that cannot allocate (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  No counterpart in the reference.
 * The same five integrals as dg_totals, from the same kernels (bit-identical
 * results), for the state Q of configuration cfg (K, p, L and gamma are
 * used); written to DEVICE memory out_dev[0..4] on `stream` (a hipStream_t,
 * NULL = null stream), asynchronously: no host synchronisation, no memset.
 * The per-block partial sums go to a fixed array of the library (8192 blocks
 * x 5 doubles = 320 KiB per GPU, held until the process ends); calls that
 * overlap on one GPU race on it.  Returns DG_ERR_INVALID_ARG for a null
 * pointer, an invalid configuration or more than 2^30 nodes (dg_create's
 * limits), DG_ERR_HIP if a launch fails. */
int         dg_totals_noalloc(const dg_config* cfg, const double* Q, double* out_dev, void* stream);

/* Exact isentropic vortex (vortex_paper_exact, q1_dg_euler_2d.cu:2166-2188,
 * the reference solution of its L2 error kernel :2247-2294) advected for
 * time t >= 0 by the background velocity (u0, v0) on the periodic domain:
 * the vortex field at the node coordinates shifted by -(u0, v0) t and wrapped
 * into the domain, written to Q (SIDX).  images = 0: the single vortex (at
 * t = 0 identical to dg_init_state); images = 1: velocity and temperature
 * perturbations summed over the 3x3 periodic images of the centre, a
 * periodic field (the single vortex is periodic only to ~1e-5 at the edges).
 * ic->kind must be DG_IC_VORTEX. */
int         dg_vortex_exact(dg_ctx* ctx, const dg_ic_params* ic, double t, int images, double* Q);

/* Norms of Q - Qref with the GLL quadrature of the discretisation, FP64 on
 * the device, deterministic, copied to host (the per-variable L2 errors of
 * compute_l2_err_paper, q1_dg_euler_2d.cu:2247-2294):
 * out[0..3] = sqrt(integral of (Q_q - Qref_q)^2) for rho, rho*u, rho*v, E;
 * out[4..7] = max |Q_q - Qref_q|;  out[8] = relative L2 over all four
 * variables, sqrt(sum_q int d_q^2) / sqrt(sum_q int Qref_q^2);
 * out[9] = max_q max|d_q| / max_q max|Qref_q|.  A NaN propagates. */
int         dg_error_norms(dg_ctx* ctx, const double* Q, const double* Qref, double out[10]);

/* Frame fields (downsample_rho_p, dgsem_lobatto_2d.cu:826-844, the data of
 * the reference's PPM frames and of plot_rt.py's binary frames): per element,
 * the mean over its nodes of rho (floored at 1e-14) and of max(p, 0).
 * rho_el, p_el: device arrays of K*K doubles, element order el = ej*K + ei
 * (row-major, j = 0 at the bottom).  FP64, deterministic. */
int         dg_frame_fields(dg_ctx* ctx, const double* Q, double* rho_el, double* p_el);

/* Tuning: time the candidate schedules (elements per block x XCD remap, and
 * the matrix-core form where the configuration leaves it open) on scratch
 * buffers, round-robin over four rounds of 8-step runs, and keep the fastest
 * for this ctx (the XCD order when it runs within 1 % of the fastest: it
 * reads fewer bytes).  Optional; dg_create picks a heuristic schedule.
 * Writes the chosen values if pointers non-NULL. */
int         dg_autotune(dg_ctx* ctx, int* elems_per_block, int* xcd_remap);

/* ── Checkpoint / restart (no counterpart in the reference, whose frames are
 * lossy PPM images): dg_save_state writes device state Q with its time and
 * step count to `path` (a 128-byte header -- magic "DGSEMCK", version 1, K,
 * p, scheme, x0, y0, L, gamma, t, steps, the number of doubles -- then the
 * FP64 state in the layout above, host byte order); dg_load_state reads such
 * a file into device state Q (t, steps may be NULL).  Loading requires the
 * file's discretisation (K, p, scheme, domain, gamma) to be the context's,
 * else DG_ERR_INVALID_ARG with the reason in dg_last_error.  A run resumed
 * from a checkpoint takes the same steps, bit for bit, as one that continues
 * with the same state and time in a new dg_advance call. */
int         dg_save_state(dg_ctx* ctx, const double* Q, double t, long long steps, const char* path);
int         dg_load_state(dg_ctx* ctx, double* Q, double* t, long long* steps, const char* path);

/* ── Batched independent solves (no counterpart in the reference, which runs
 * one solve per process): ensembles of instances that differ in K, p, domain,
 * IC and CFL but share the flux / boundary / precision switches.  Each stage
 * of all instances is one launch over a tile list (heaviest tiles first, one
 * workgroup per tile, every degree's code in the same kernel), or, in the
 * by-degree form, one launch per degree present (the instances of that
 * degree, each with the single-solve kernel of that degree); each instance
 * keeps its own device-side dt and time.  The stage launches of the time loop
 * are replayed from a captured HIP graph.  AV and the limiter must be on in
 * every instance or in none, and the scheme alike (DG_ERR_INVALID_ARG
 * otherwise); a batch with AV / limiter, or of spectral-difference
 * instances, runs in the by-degree form only, with the single-solve kernel of
 * each degree (shock-capturing, or SD with its nodal repair), and refuses
 * dg_batch_set_by_degree(b, 0) and dg_batch_set_persistent(b, 1).  The AV /
 * limiter parameters (av_cav, av_ksensor, eps_rho, eps_p) may differ between
 * instances. */
typedef struct dg_batch dg_batch;

int         dg_batch_create(const dg_config* cfgs, int n, void* stream, dg_batch** out);
int         dg_batch_destroy(dg_batch* b);
const char* dg_batch_last_error(const dg_batch* b);
/* doubles in instance i's state vector */
long long   dg_batch_state_size(const dg_batch* b, int i);
/* Qs[i]: device state of instance i (caller-owned), ics[i]: its IC */
int         dg_batch_init_state(dg_batch* b, const dg_ic_params* ics, double* const* Qs);
/* Checkpoint / restart of instance i: dg_save_state / dg_load_state with the
 * instance's discretisation (the same file format); a batch resumed from its
 * instances' checkpoints (their times passed back through t_inout) takes the
 * same steps, bit for bit, as one that continues in a new dg_batch_advance
 * call. */
int         dg_batch_save_state(dg_batch* b, int i, const double* Q, double t, long long steps, const char* path);
int         dg_batch_load_state(dg_batch* b, int i, double* Q, double* t, long long* steps, const char* path);
/* nsteps fused SSP-RK3 steps of every instance; cfl[i] per instance; t_inout[i]
 * in/out (may be NULL); status_out[i] per-instance DG_* code (may be NULL).
 * Returns the first instance failure. */
int         dg_batch_advance(dg_batch* b, double* const* Qs, int nsteps, const double* cfl, double t_end,
                             double* t_inout, int* status_out);
/* The same steps issued instance by instance with dg_advance (for comparison). */
int         dg_batch_advance_each(dg_batch* b, double* const* Qs, int nsteps, const double* cfl, double t_end,
                                  double* t_inout, int* status_out);
/* Tuning: use_graph (0/1); lanes per workgroup, 64 / 128 / 256 (0 = 64):
 * instance i's tiles hold max(1, lanes / (2(p_i+1))) elements. */
int         dg_batch_set_options(dg_batch* b, int use_graph, int lanes);
/* Persistent form of the stage launches (0 = one workgroup per tile, the
 * default; 1 = as many workgroups as are resident on the chip walk the tile
 * list with a grid stride).  Results are identical either way. */
int         dg_batch_set_persistent(dg_batch* b, int persistent);
/* Order of the tile list (0 = all tiles heaviest first, the default; 1 =
 * instances bin-packed by cost over the 8 XCDs, each bin heaviest first,
 * the bins interleaved so that each XCD runs one bin and the tiles of an
 * instance share an L2).  Results are identical either way. */
int         dg_batch_set_xcd_bins(dg_batch* b, int xcd_bins);
/* By-degree form (1): each stage is one launch per degree present, over the
 * tiles of the instances of that degree (an instance's tiles contiguous,
 * highest degree first), each launch running the single-solve lean kernel of
 * its degree with its own register budget and schedule; xcd_bins then selects
 * the XCD-aware block order within each launch.  0 = the all-degree kernel
 * (the default without AV / limiter; a batch with them is always by degree).
 * The by-degree and persistent forms exclude each other:
 * setting one clears the other.  Results are identical in every form. */
int         dg_batch_set_by_degree(dg_batch* b, int by_degree);
int         dg_batch_get_by_degree(const dg_batch* b, int* by_degree);
/* By-degree form: lanes per workgroup of the launch of degree p (1..7):
 * 64 / 128 / 256, or 0 to follow dg_batch_set_options (the default); the
 * getter gives the size in effect for p. */
int         dg_batch_set_degree_lanes(dg_batch* b, int p, int lanes);
int         dg_batch_get_degree_lanes(const dg_batch* b, int p, int* lanes);
/* Current options (any pointer may be NULL). */
int         dg_batch_get_options(const dg_batch* b, int* use_graph, int* lanes, int* persistent, int* xcd_bins);
/* Time the three workgroup sizes, each with both orders in the three launch
 * forms (all-degree kernel one workgroup per tile, persistent, by degree; a
 * batch with AV / limiter or the SD scheme: the by-degree form only), on the
 * given states, round-robin over three rounds (each timing: 3 untimed steps
 * that rebuild the graph, then 6 timed steps; a candidate keeps its fastest
 * timing), and keep the fastest; in the by-degree form then refine each
 * degree's workgroup size on its own (kept if the step runs >= 1 % faster).
 * The states are advanced by 9 steps per timing at cfl[i] (at most 18 x 3
 * timings, plus 1 + 2 per degree present).  Writes the common size to *lanes
 * if non-NULL. */
int         dg_batch_autotune(dg_batch* b, double* const* Qs, const double* cfl, int* lanes);

#ifdef __cplusplus
}
#endif
#endif /* DGSEM_H */
