// dgsem_run — a plain C++ host program driving libdgsem.so through the C ABI
// (include/dgsem.h), the way the reference's main() drives its kernels
// (dgsem_lobatto_2d.cu:940-1090): caller-owned device buffers (:966-974), IC +
// limiter (:1006-1007), then per step the RHS and wave speed (:1024-1026), dt
// (:1029-1035), the copy of Q into Q0 (:1037), the three SSP-RK3 stages with
// the limiter (:1040-1051) and the blow-up check (:1052-1057).  The frame-time
// clipping (:1030-1031) is one end time; PPM frames are not written.
//
//   dgsem_run [--case sod|vortex|kh] [--K n] [--p n] [--steps n] [--t-end t]
//             [--cfl c] [--precision fp64|mixed] [--mode reference|fused]
//             [--out file] [--checkpoint file] [--restart file]
//
// --mode reference issues the reference's separate calls (dg_compute_rhs,
// dg_cfl_dt, dg_rk3_stage); --mode fused runs the same steps with
// dg_advance.  Prints one JSON line (steps, t, last dt, status, totals) and,
// with --out, writes the final state (4*K*K*(p+1)^2 doubles, SIDX order).
// --restart starts from a checkpoint (dg_load_state: its state, time and
// step count) instead of the initial condition; --checkpoint writes one at
// the end (dg_save_state).  The reference has neither: it restarts from
// t = 0.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "dgsem.h"

namespace {

struct Options {
    std::string ic_case = "sod", precision = "fp64", mode = "reference", out, checkpoint, restart;
    int K = 50, p = 3, steps = 100;
    double t_end = 0.0, cfl = 0.6;
};

bool parse(int argc, char** argv, Options& o) {
    for (int i = 1; i < argc; i++) {
        const std::string a = argv[i];
        if (i + 1 >= argc) return false;
        const char* v = argv[++i];
        if (a == "--case") o.ic_case = v;
        else if (a == "--K") o.K = atoi(v);
        else if (a == "--p") o.p = atoi(v);
        else if (a == "--steps") o.steps = atoi(v);
        else if (a == "--t-end") o.t_end = atof(v);
        else if (a == "--cfl") o.cfl = atof(v);
        else if (a == "--precision") o.precision = v;
        else if (a == "--mode") o.mode = v;
        else if (a == "--out") o.out = v;
        else if (a == "--checkpoint") o.checkpoint = v;
        else if (a == "--restart") o.restart = v;
        else return false;
    }
    return (o.ic_case == "sod" || o.ic_case == "vortex" || o.ic_case == "kh") &&
           (o.precision == "fp64" || o.precision == "mixed") && (o.mode == "reference" || o.mode == "fused");
}

// The configurations of the three workloads: the reference's own defaults
// (circular Sod, walls, Chandrashekar + HLLC, AV, limiter), the isentropic
// vortex on periodic [0,10]^2 (Ranocha + Rusanov) and the Kelvin-Helmholtz
// layer on periodic [-1,1]^2 with the reference's shock capturing.
void configure(const Options& o, dg_config& cfg, dg_ic_params& ic) {
    dg_config_default(&cfg);
    cfg.K = o.K;
    cfg.p = o.p;
    cfg.precision = o.precision == "mixed" ? DG_MIXED : DG_FP64;
    if (o.ic_case == "sod") {
        dg_ic_default(&ic, DG_IC_CIRCULAR_SOD);
    } else if (o.ic_case == "vortex") {
        cfg.L = 10.0;
        cfg.bc = DG_BC_PERIODIC;
        cfg.volume_flux = DG_VOL_RANOCHA;
        cfg.surface_flux = DG_SURF_RUSANOV;
        cfg.av = 0;
        cfg.limiter = 0;
        dg_ic_default(&ic, DG_IC_VORTEX);
    } else {
        cfg.x0 = -1.0;
        cfg.y0 = -1.0;
        cfg.L = 2.0;
        cfg.bc = DG_BC_PERIODIC;
        dg_ic_default(&ic, DG_IC_KELVIN_HELMHOLTZ);
    }
}

#define HCK(x)                                                                            \
    do {                                                                                  \
        hipError_t e_ = (x);                                                              \
        if (e_ != hipSuccess) {                                                           \
            fprintf(stderr, "%s:%d: %s\n", __FILE__, __LINE__, hipGetErrorString(e_));    \
            return 2;                                                                     \
        }                                                                                 \
    } while (0)

}  // namespace

int main(int argc, char** argv) {
    Options o;
    if (!parse(argc, argv, o)) {
        fprintf(stderr,
                "usage: dgsem_run [--case sod|vortex|kh] [--K n] [--p n] [--steps n] [--t-end t] [--cfl c]\n"
                "                 [--precision fp64|mixed] [--mode reference|fused] [--out file]\n"
                "                 [--checkpoint file] [--restart file]\n");
        return 1;
    }
    dg_config cfg;
    dg_ic_params ic;
    configure(o, cfg, ic);
    dg_ctx* ctx = nullptr;
    int rc = dg_create(&cfg, nullptr, &ctx);
    if (rc) {
        fprintf(stderr, "dg_create: %s\n", dg_status_string(rc));
        return 1;
    }
    const long long n = dg_state_size(ctx), nel = dg_num_elements(ctx);
    const size_t bytes = (size_t)n * sizeof(double);

    // caller-owned buffers, as the reference allocates them (:966-974)
    double *d_Q, *d_Q0, *d_Q1, *d_Q2, *d_RHS, *d_lam;
    HCK(hipMalloc(&d_Q, bytes));
    HCK(hipMalloc(&d_Q0, bytes));
    HCK(hipMalloc(&d_Q1, bytes));
    HCK(hipMalloc(&d_Q2, bytes));
    HCK(hipMalloc(&d_RHS, bytes));
    HCK(hipMalloc(&d_lam, (size_t)nel * sizeof(double)));

    double t = 0.0, dt = 0.0;
    int steps = 0;
    long long steps0 = 0;   // steps of the run a checkpoint was taken from
    if (o.restart.empty()) rc = dg_init_state(ctx, &ic, d_Q);   // ic_kernel + positivity_limiter (:1006-1007)
    else rc = dg_load_state(ctx, d_Q, &t, &steps0, o.restart.c_str());
    if (rc == DG_OK && o.mode == "reference") {
        while (steps < o.steps && (o.t_end <= 0.0 || t < o.t_end)) {   // :1022
            double lam_max;
            if ((rc = dg_compute_rhs(ctx, d_Q, d_RHS, d_lam))) break;              // :1024
            if ((rc = dg_cfl_dt(ctx, d_lam, o.cfl, &lam_max, &dt))) break;         // :1026-1029
            if (o.t_end > 0.0 && t + dt > o.t_end) dt = o.t_end - t;              // :1030-1031
            if (dt < 1e-14) { rc = DG_ERR_DT_UNDERFLOW; break; }                 // :1032-1035
            HCK(hipMemcpy(d_Q0, d_Q, bytes, hipMemcpyDeviceToDevice));             // :1037
            if ((rc = dg_rk3_stage(ctx, 1, d_Q1, d_Q0, nullptr, d_RHS, dt))) break;   // :1040-1041
            if ((rc = dg_compute_rhs(ctx, d_Q1, d_RHS, nullptr))) break;              // :1044
            if ((rc = dg_rk3_stage(ctx, 2, d_Q2, d_Q0, d_Q1, d_RHS, dt))) break;      // :1045-1046
            if ((rc = dg_compute_rhs(ctx, d_Q2, d_RHS, d_lam))) break;                // :1049
            if ((rc = dg_rk3_stage(ctx, 3, d_Q, d_Q0, d_Q2, d_RHS, dt))) break;       // :1050-1051
            // blow-up check on the last RHS's wave speeds (:1052-1057)
            if ((rc = dg_cfl_dt(ctx, d_lam, o.cfl, &lam_max, nullptr)) || lam_max > 1e6) {
                rc = DG_ERR_BLOWUP;
                break;
            }
            t += dt;
            steps++;
        }
    } else if (rc == DG_OK) {
        rc = dg_advance(ctx, d_Q, o.steps, o.cfl, o.t_end, &t, &dt);
        steps = dg_steps_taken(ctx);
    }
    if (rc == DG_OK && !o.checkpoint.empty()) rc = dg_save_state(ctx, d_Q, t, steps0 + steps, o.checkpoint.c_str());
    double tot[5] = {0, 0, 0, 0, 0};
    if (rc == DG_OK) rc = dg_totals(ctx, d_Q, tot);
    if (rc != DG_OK) fprintf(stderr, "status %d (%s): %s\n", rc, dg_status_string(rc), dg_last_error(ctx));
    printf("{\"case\": \"%s\", \"mode\": \"%s\", \"precision\": \"%s\", \"K\": %d, \"p\": %d, \"steps\": %d, "
           "\"total_steps\": %lld, \"t\": %.17g, \"dt\": %.17g, \"status\": %d, "
           "\"totals\": [%.17g, %.17g, %.17g, %.17g, %.17g]}\n",
           o.ic_case.c_str(), o.mode.c_str(), o.precision.c_str(), o.K, o.p, steps, steps0 + steps, t, dt, rc,
           tot[0], tot[1], tot[2], tot[3], tot[4]);
    if (!o.out.empty()) {
        std::vector<double> h((size_t)n);
        HCK(hipMemcpy(h.data(), d_Q, bytes, hipMemcpyDeviceToHost));
        FILE* f = fopen(o.out.c_str(), "wb");
        if (!f || fwrite(h.data(), sizeof(double), h.size(), f) != h.size()) {
            fprintf(stderr, "cannot write %s\n", o.out.c_str());
            rc = rc ? rc : 1;
        }
        if (f) fclose(f);
    }
    hipFree(d_Q); hipFree(d_Q0); hipFree(d_Q1); hipFree(d_Q2); hipFree(d_RHS); hipFree(d_lam);
    dg_destroy(ctx);
    return rc;
}
