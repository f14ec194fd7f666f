#!/usr/bin/env python3
"""Benchmark of the 2-D Euler DGSEM step on one MI355X.

Workload: isentropic vortex (strength 5, centre (5,5), background rho=1,
u=v=1, p=1, gamma=1.4) on periodic [0,10]^2, 512x512 elements, degree 7
(8x8 LGL nodes, 16.8M nodes, 67.1M DOF), Ranocha EC volume flux, Rusanov
faces, SSP-RK3 at CFL 0.25 (dt = CFL*h/(lam*(p+1)); at p=7 the scheme is
linearly unstable above CFL ~0.3, so the 0.5 used at p=3 cannot be kept).  A "step" is one SSP-RK3 step (three fused stage
kernels).  The state is resident in HBM before timing starts.

Prints one JSON line: value = DOF*stage updates/s of the mixed form (FP32
volume/surface, FP64 state, RK accumulation and reductions); the FP64 form,
the mixed/FP64 ratio, the L2/Linf error of each form against the other and
against the exact advected vortex, the roofline of the stage kernel and the
CPU oracle baseline are reported beside it.
"""
import argparse
import json
import os
import sys
import time


ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# HBM bytes per stage launch from rocprofv3 FETCH_SIZE / WRITE_SIZE runs of the
# stage kernel (scaled by a device copy's readings), keyed by (K, p, form):
# (bytes, profile that lists the runs and the scaling, KERNEL_SOURCES hash of
# the build that was profiled).  A figure whose hash differs from the current
# sources' is not reported (the kernel changed since it was measured).
# Keyed (K, p, form, elements per workgroup, XCD order): the tile order sets
# how many face-neighbour reads hit another XCD's L2.
_PROF = "profiles/stage_kernel_p7_vortex512.md"
_SRC = "60a95d92ae0aa40d"
PMC_TRAFFIC = {
    (512, 7, "mixed", 16, 1): (1.43688e9, _PROF, _SRC), (512, 7, "mixed", 16, 0): (1.50031e9, _PROF, _SRC),
    (512, 7, "mixed", 8, 1): (1.43688e9, _PROF, _SRC), (512, 7, "mixed", 8, 0): (1.56743e9, _PROF, _SRC),
    (512, 7, "mixed", 4, 1): (1.43687e9, _PROF, _SRC), (512, 7, "mixed", 4, 0): (1.70164e9, _PROF, _SRC),
    (512, 7, "fp64", 8, 1): (1.43701e9, _PROF, _SRC), (512, 7, "fp64", 8, 0): (1.56755e9, _PROF, _SRC),
    (512, 7, "fp64", 4, 1): (1.43701e9, _PROF, _SRC), (512, 7, "fp64", 4, 0): (1.70177e9, _PROF, _SRC),
}
FP64_PEAK_TFLOPS = 78.6
# the sources the stage kernels are compiled from, and the build script that
# holds the compiler flags
KERNEL_SOURCES = ("fvstuff_amd/csrc/dgsem_kernels.hpp", "fvstuff_amd/csrc/dgsem_physics.hpp",
                  "fvstuff_amd/csrc/dgsem_inst.hip", "fvstuff_amd/csrc/dgsem_launch.hpp",
                  "fvstuff_amd/build.py")
# SSP-RK3 steps of the accuracy comparison: the headline configuration's count
ACCURACY_STEPS = 200


def kernel_sources_hash():
    """First 16 hex digits of the SHA-256 over the stage kernels' sources."""
    import hashlib
    h = hashlib.sha256()
    for rel in KERNEL_SOURCES:
        with open(os.path.join(ROOT, rel), "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--K", type=int, default=512)
    ap.add_argument("--p", type=int, default=7)
    ap.add_argument("--cfl", type=float, default=0.25,
                    help="0.5 is outside the SSP-RK3 stability limit at p=7 (about 0.3)")
    ap.add_argument("--forms", default="fp64,mixed")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the error report")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="vortex: skip the short samples of the other workloads (KH 1024^2 / 2048^2 p=3, "
                         "256 batched vortex instances, 256 batched Sod instances with AV + limiter)")
    ap.add_argument("--extra-steps", type=int, default=30,
                    help="timed steps of each short sample of the other workloads")
    ap.add_argument("--schedule", default="",
                    help="E,X: elements per workgroup, XCD remap (implies --no-autotune)")
    ap.add_argument("--cpu-steps", type=int, default=15,
                    help="SSP-RK3 steps of the CPU oracle baseline (about 10 s on 16 host threads)")
    ap.add_argument("--workload", default="vortex", choices=["vortex", "kh", "sd", "batched", "copy", "colloc"],
                    help="vortex: the headline line; kh: Kelvin-Helmholtz (--K, --p, --t-end); "
                         "sd: spectral-difference sibling on circular Sod (--K, --p, --steps); "
                         "batched: seeded ensemble of vortex instances; copy: device copy probe; "
                         "colloc: the collocation volume term on the vortex (the MFMA A/B, --mfma)")
    ap.add_argument("--t-end", type=float, default=1.5)
    ap.add_argument("--no-shock-capturing", action="store_true",
                    help="kh: switch off the Ducros AV and the Zhang-Shu limiter")
    ap.add_argument("--kh-steps", type=int, default=0,
                    help="kh: run this many steps instead of advancing to --t-end (0 = to --t-end)")
    ap.add_argument("--t-compare", type=float, default=0.01,
                    help="kh: early time of the mixed-vs-FP64 comparison (0 = none)")
    ap.add_argument("--mfma", type=int, default=-1, choices=[-1, 0, 1],
                    help="kh: the AV contraction, colloc: the volume term, on the matrix cores (1), "
                         "on the VALU (0), or tuned (-1)")
    ap.add_argument("--instances", type=int, default=256)
    ap.add_argument("--persistent", type=int, default=-1, choices=[-1, 0, 1],
                    help="batched: persistent stage launches (1), one workgroup per tile (0), tuned (-1)")
    ap.add_argument("--lanes", type=int, default=0, help="batched: lanes per workgroup (0 = tuned)")
    ap.add_argument("--by-degree", type=int, default=-1, choices=[-1, 0, 1],
                    help="batched: one launch per degree with the single-solve kernels (1), the all-degree "
                         "kernel (0), tuned (-1)")
    ap.add_argument("--xcd-bins", type=int, default=-1, choices=[-1, 0, 1],
                    help="batched: tile list in per-XCD bins (1), heaviest first (0), tuned (-1)")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--ensemble", default="vortex", choices=["vortex", "sod"],
                    help="batched: vortex instances, or the circular-Sod setup with AV and limiter")
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def time_form(D, torch, cfg, ic, steps, warmup, cfl, autotune):
    s = D.Solver(cfg)
    tuned = s.autotune() if autotune else None
    Q = s.init_state(ic)
    if warmup:
        s.advance(Q, warmup, cfl)
    Q = s.init_state(ic, Q)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    t, dt = s.advance(Q, steps, cfl)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    return s, Q, ms, t, tuned


def hbm_reference(n, reps=20):
    """One step's HBM byte pattern of the fused stages moved by torch's own
    elementwise kernels, for reference: a copy (stage 1 reads U^n and writes
    U1) and two two-input combinations (stages 2 and 3 read two vectors and
    write one), on FP64 vectors of n values (the state's size), HIP events
    around `reps` such steps: 64 n bytes per step, as the stage kernels'
    algorithmic traffic."""
    import torch
    U = torch.rand(n, dtype=torch.float64, device="cuda")
    U1 = torch.empty_like(U)
    U2 = torch.empty_like(U)

    def step():
        U1.copy_(U)
        torch.add(U, U1, alpha=0.25, out=U2)
        torch.add(U, U2, alpha=2.0 / 3.0, out=U1)

    for _ in range(3):
        step()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        step()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    gbs = 64.0 * n / (ms * 1e-3) / 1e9
    del U, U1, U2
    return {"ms_per_step": ms, "GBs": gbs, "frac": gbs / HBM_PEAK_GBS,
            "method": "torch copy_ + 2 x add(out=) on FP64 vectors of the state's size, 64 B per DOF per step"}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def cpu_share():
    """Host threads this process may use: the OpenMP thread count the
    environment sets (a shared host that gives each GPU a share of its cores
    sets it), else the CPUs in the process's affinity mask."""
    try:
        aff = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        aff = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS", "")
    return (int(env) if env.isdigit() and int(env) > 0 else aff), aff


def cpu_baseline(cfg, ic, steps, cfl, what="vortex", single_steps=1, keep_state=False):
    """The FP64 oracle on the host: `steps` SSP-RK3 steps on all the threads
    of the process's CPU share (the reported baseline) and `single_steps` on
    one thread ("single_thread"); both bounded samples of the same workload.
    With keep_state, also returns the oracle's state after `steps` steps."""
    from oracle import oracle as O
    threads, aff = cpu_share()
    o = O.Oracle(cfg)
    dof = o.size
    rates = {}
    kept = None
    for n, k in ((threads, steps), (1, single_steps)):
        if k <= 0:
            continue
        O.set_threads(n)
        Q = o.init_state(ic)
        t0 = time.perf_counter()
        rc, t, _ = o.advance(Q, k, cfl)
        wall = time.perf_counter() - t0
        rates[n] = (dof * 3 * k / wall, k, wall)
        if keep_state and n == threads:
            kept = (Q, k, t, rc)
    model = cpu_model()
    v, k, wall = rates[threads]
    out = {"value": v, "unit": "DOF*stage/s", "cores": threads, "kind": "port",
           "host_cpus": os.cpu_count(), "affinity_cpus": aff, "cpu_model": model,
           "sample": f"FP64 oracle, {k} SSP-RK3 step(s) of the same {cfg.K}x{cfg.K} p={cfg.p} {what}, "
                     f"{threads} OpenMP threads (the process's CPU share; the host has {os.cpu_count()} CPUs) "
                     f"on {model}, {wall:.1f} s"}
    if 1 in rates:
        v1, k1, w1 = rates[1]
        out["single_thread"] = {"value": v1, "cores": 1, "sample": f"{k1} step(s) on 1 thread, {w1:.1f} s"}
    return (out, kept) if keep_state else out


def run_kh(K, p, forms, cfl=0.5, seed=2024, t_end=None, steps=None, schedule="", autotune=True, mfma=-1,
           shock=True, t_compare=0.01, tag="bench-kh"):
    """Kelvin-Helmholtz step shear layer on periodic [-1,1]^2 (rho 2 / 1,
    u +-0.5, p 2.5, v = 0.1 sin(2 pi x) + seeded noise 1e-3) with the
    reference's shock capturing (Chandrashekar + HLLC, Ducros AV, Zhang-Shu
    limiter), in each form: either to t_end or for a fixed number of steps,
    timed with HIP events (state resident in HBM).  Beside the rate: mass,
    energy and entropy drift, and the mixed form against the FP64 form at the
    early time t_compare (both forms advanced there from the IC, untimed;
    later the roll-up is chaotic and the two forms drift apart)."""
    import torch
    from fvstuff_amd import dgsem as D
    from fvstuff_amd.capi import kh_config, make_ic
    ic = make_ic("kh", seed=seed, noise_amp=1e-3)
    N = p + 1
    dof = 4 * K * K * N * N
    out, early = {}, {}
    for form in forms:
        # without the shock capturing the unresolved step layer loses
        # positivity at t ~ 0.15-0.25 (Ranocha or Chandrashekar, Rusanov or
        # HLLC); with it the layer rolls up to t = 1.5
        sched = {}
        if schedule:
            E_, X_ = (int(v) for v in schedule.split(","))
            sched = dict(elems_per_block=E_, xcd_remap=X_)
        s = D.Solver(kh_config(K=K, p=p, precision=form, av=shock, limiter=shock,
                               volume_flux="chandrashekar", surface_flux="hllc", mfma=mfma, **sched))
        tuned = s.autotune() if (autotune and not schedule) else None
        if t_compare:
            Qc = s.init_state(ic)
            t_c = early["fp64"][1] if (form == "mixed" and "fp64" in early) else t_compare
            tc, _ = s.advance_to(Qc, t_c, cfl)
            early[form] = (s, tc, Qc)
        Q = s.init_state(ic)
        T0 = s.totals(Q)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True); e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        last = [time.perf_counter()]

        def progress(t, n):
            if time.perf_counter() - last[0] > 10.0:
                last[0] = time.perf_counter()
                log(f"[{tag}] {form}: t={t:.4f} steps={n}")

        if steps:
            t, _ = s.advance(Q, steps, cfl)
            taken, launched = s.steps_taken(), steps
        else:
            t, taken = s.advance_to(Q, t_end, cfl, chunk=512, progress=progress)
            launched = s.last_launched
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1)
        T1 = s.totals(Q)
        rate = dof * 3 * taken / (ms * 1e-3)
        # the rate counts the steps taken; the timed region also holds the
        # no-op steps launched past t_end (each still streams the state once)
        out[form] = {"value": rate, "steps": taken, "steps_launched": launched, "t": t, "ms": ms,
                     "us_per_launch": ms * 1e3 / (3 * launched), "schedule": tuned,
                     "hbm_GBs_algorithmic": rate * (64.0 / 3.0) / 1e9,
                     "hbm_frac": rate * (64.0 / 3.0) / 1e9 / HBM_PEAK_GBS,
                     "energy_drift_rel": (T1[3] - T0[3]) / abs(T0[3]),
                     "mass_drift_rel": (T1[0] - T0[0]) / abs(T0[0]),
                     "entropy_change_rel": (T1[4] - T0[4]) / abs(T0[4])}
        log(f"[{tag}] {form}: {out[form]}")
        del Q
    if "fp64" in out and "mixed" in out:
        out["mixed_over_fp64"] = out["mixed"]["value"] / out["fp64"]["value"]
    if "fp64" in early and "mixed" in early:
        s64, t64, Q64 = early["fp64"]
        nrm = s64.error_norms(early["mixed"][2], Q64)
        out["early_compare"] = {"t": t64, "mixed_vs_fp64_rel_l2": nrm["rel_l2"],
                                "mixed_vs_fp64_rel_linf": nrm["rel_linf"]}
    return {"workload": f"Kelvin-Helmholtz {K}x{K} p={p}, "
                        + (f"{steps} steps" if steps else f"to t={t_end}")
                        + f", CFL {cfl}, Chandrashekar + HLLC, AV + limiter {'on' if shock else 'off'}",
            "unit": "DOF*stage/s", "dof": dof, **out}


def bench_kh(args):
    """The Kelvin-Helmholtz workload to --t-end (run_kh); the mixed form
    against the FP64 form at t = 0.01 and at the end time."""
    res = run_kh(args.K, args.p, args.forms.split(","), cfl=args.cfl, seed=args.seed, t_end=args.t_end,
                 steps=args.kh_steps or None,
                 schedule=args.schedule, autotune=not args.no_autotune, mfma=args.mfma,
                 shock=not args.no_shock_capturing, t_compare=args.t_compare)
    print(json.dumps(res))


def bench_sd(args):
    """The spectral-difference sibling (fr_lobatto_2d.cu) on the problem of
    its own driver: circular Sod on [0,1]^2 with reflecting walls, HLLC
    faces, the nodal positivity repair after every stage and CFL 0.1/p, for a
    fixed number of fused SSP-RK3 steps in both forms (state resident in HBM,
    HIP events around the timed steps).  Mass and energy drift over the run
    (the walls carry neither) and the mixed-vs-FP64 difference are reported
    beside the rate; the FP64 oracle runs a bounded sample on the host."""
    import torch
    from fvstuff_amd import dgsem as D
    from fvstuff_amd.capi import make_ic, sd_config
    K, p = args.K, args.p
    N = p + 1
    dof = 4 * K * K * N * N
    cfl = 0.1 / p
    ic = make_ic("sod")
    out, Qs = {}, {}
    sched = {}
    if args.schedule:
        E_, X_ = (int(v) for v in args.schedule.split(","))
        sched = dict(elems_per_block=E_, xcd_remap=X_)
    for form in args.forms.split(","):
        s = D.Solver(sd_config(K=K, p=p, precision=form, **sched))
        tuned = None if (args.no_autotune or args.schedule) else s.autotune()
        Q = s.init_state(ic)
        if args.warmup:
            s.advance(Q, args.warmup, cfl)
        Q = s.init_state(ic, Q)
        T0 = s.totals(Q)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True); e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        t, _ = s.advance(Q, args.steps, cfl)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1)
        T1 = s.totals(Q)
        rate = dof * 3 * args.steps / (ms * 1e-3)
        out[form] = {"value": rate, "ms_per_step": ms / args.steps, "t": t, "schedule": tuned,
                     "hbm_GBs_algorithmic": rate * (64.0 / 3.0) / 1e9,
                     "hbm_frac": rate * (64.0 / 3.0) / 1e9 / HBM_PEAK_GBS,
                     "mass_drift_rel": (T1[0] - T0[0]) / abs(T0[0]),
                     "energy_drift_rel": (T1[3] - T0[3]) / abs(T0[3])}
        Qs[form] = (s, Q)
        log(f"[bench-sd] {form}: {out[form]}")
    if "fp64" in Qs and "mixed" in Qs:
        s64, Q64 = Qs["fp64"]
        nrm = s64.error_norms(Qs["mixed"][1], Q64)
        out["mixed_vs_fp64_rel_l2"] = nrm["rel_l2"]
        out["mixed_vs_fp64_rel_linf"] = nrm["rel_linf"]
        out["mixed_over_fp64"] = out["mixed"]["value"] / out["fp64"]["value"]
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(sd_config(K=K, p=p), ic, args.cpu_steps, cfl,
                                           what="circular Sod (spectral difference)")
    print(json.dumps({"workload": f"spectral-difference circular Sod {K}x{K} p={p}, walls, HLLC, nodal repair, "
                                  f"CFL {cfl:.4g}, {args.steps} steps",
                      "unit": "DOF*stage/s", "dof": dof, **out}))


def run_batched(instances, seed, form, steps, warmup, lanes=0, persistent=-1, xcd_bins=-1, autotune=True,
                by_degree=-1, kind="vortex"):
    """Seeded ensemble stepped in one launch per stage for all instances (or
    one per degree present) vs instance by instance (wall clock around each,
    after a synchronisation).  kind "vortex": isentropic-vortex instances
    (K ~ U{16..128}, p ~ U{3,5,7}, strength ~ U[1,5], Mach ~ U[0.1,0.8]);
    "sod": the reference's circular-Sod setup with AV and the limiter
    (K ~ U{16..128}, p ~ U{3,5,7}, inner density ~ U[2,11], pressure
    ~ U[2,10]), which runs one launch per degree only."""
    import torch
    from fvstuff_amd import dgsem as D
    from fvstuff_amd.capi import ensemble, sod_ensemble
    make = sod_ensemble if kind == "sod" else ensemble
    cfgs, ics, cfls = make(n=instances, seed=seed, precision=form)
    b = D.Batch(cfgs)
    Qs = b.init_states(ics)
    tuned = autotune and not (lanes or persistent >= 0 or xcd_bins >= 0 or by_degree >= 0)
    if tuned:
        b.autotune(Qs, cfls)
    else:
        b.set_options(use_graph=True, lanes=lanes)
        if by_degree >= 0 or kind != "sod":
            b.set_by_degree(max(0, by_degree))
        b.set_xcd_bins(max(0, xcd_bins))
        if persistent > 0:
            b.set_persistent(True)
    _, lanes, persistent, xcd_bins = b.options()
    by_degree = b.by_degree()
    lanes_by_degree = {p: b.degree_lanes(p) for p in sorted({c.p for c in cfgs})} if by_degree else None
    res = {}
    for mode in ("batched", "each"):
        Qs = b.init_states(ics, Qs)
        fn = b.advance if mode == "batched" else b.advance_each
        fn(Qs, warmup, cfls)
        Qs = b.init_states(ics, Qs)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fn(Qs, steps, cfls)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        rate = b.dofs * 3 * steps / wall
        # algorithmic HBM share: 64/3 bytes per DOF*stage (the state streams)
        res[mode] = {"value": rate, "ms_per_step": wall * 1e3 / steps,
                     "hbm_frac": rate * (64.0 / 3.0) / 1e9 / HBM_PEAK_GBS}
        log(f"[bench-batched] {form} {mode}: {res[mode]}")
    what = "circular-Sod instances with AV and limiter" if kind == "sod" else "vortex instances"
    return {"workload": f"{instances} {what} (seed {seed}), {form}", "unit": "DOF*stage/s",
            "dof": b.dofs, "steps": steps, "lanes_per_workgroup": lanes, "persistent": persistent,
            "xcd_bins": xcd_bins, "by_degree": by_degree, "lanes_by_degree": lanes_by_degree, "tuned": tuned,
            "batched": res["batched"], "one_by_one": res["each"],
            "speedup": res["batched"]["value"] / res["each"]["value"]}


def bench_batched(args):
    print(json.dumps(run_batched(args.instances, args.seed, args.forms.split(",")[-1], args.steps, args.warmup,
                                 lanes=args.lanes, persistent=args.persistent, xcd_bins=args.xcd_bins,
                                 autotune=not args.no_autotune, by_degree=args.by_degree,
                                 kind=args.ensemble)))


def bench_colloc(args):
    """The collocation volume term (dgsem_lobatto_2d.cu:431-439, EC_FLUX=0 of
    the reference) on the vortex workload, Rusanov faces: the A/B of its
    contraction on the VALU (--mfma 0) or on the matrix cores (--mfma 1, p = 3
    mixed form; -1: the autotuner decides), timed with HIP events."""
    import torch
    from fvstuff_amd import dgsem as D
    from fvstuff_amd.capi import vortex_config, make_ic
    ic = make_ic("vortex")
    K, p = args.K, args.p
    dof = 4 * K * K * (p + 1) ** 2
    sched = {}
    if args.schedule:
        E_, X_ = (int(v) for v in args.schedule.split(","))
        sched = dict(elems_per_block=E_, xcd_remap=X_)
    out = {}
    for form in args.forms.split(","):
        cfg = vortex_config(K=K, p=p, precision=form, volume_flux="collocation", mfma=args.mfma, **sched)
        s, Q, ms, t, tuned = time_form(D, torch, cfg, ic, args.steps, args.warmup, args.cfl,
                                       not (args.no_autotune or args.schedule))
        rate = dof * 3 * args.steps / (ms * 1e-3)
        out[form] = {"value": rate, "ms_per_step": ms / args.steps, "us_per_launch": ms * 1e3 / (3 * args.steps),
                     "schedule": tuned, "hbm_frac": rate * (64.0 / 3.0) / 1e9 / HBM_PEAK_GBS, "t": t}
        log(f"[bench-colloc] {form}: {out[form]}")
        del Q, s
    print(json.dumps({"workload": f"isentropic vortex {K}x{K} p={p}, collocation volume term + Rusanov, "
                                  f"{args.steps} steps, CFL {args.cfl}, mfma {args.mfma}",
                      "unit": "DOF*stage/s", "dof": dof, **out}))


def bench_copy(args):
    """Device copy of one state vector (known bytes) for counter scaling."""
    import torch
    K, p = args.K, args.p
    n = 4 * K * K * (p + 1) ** 2
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"workload": "copy", "bytes_read": n * 8, "bytes_written": n * 8}))


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is design-only "
                                   "(halo exchange of element faces), no N>1 figure is produced"}))
        sys.exit(2)
    if args.workload == "kh":
        return bench_kh(args)
    if args.workload == "sd":
        return bench_sd(args)
    if args.workload == "batched":
        return bench_batched(args)
    if args.workload == "copy":
        return bench_copy(args)
    if args.workload == "colloc":
        return bench_colloc(args)
    import torch
    from fvstuff_amd import dgsem as D
    from fvstuff_amd.capi import vortex_config, make_ic

    ic = make_ic("vortex")
    K, p = args.K, args.p
    N = p + 1
    dof = 4 * K * K * N * N
    res = {}
    sols = {}
    sched = {}
    if args.schedule:
        E_, X_ = (int(v) for v in args.schedule.split(","))
        sched = dict(elems_per_block=E_, xcd_remap=X_)
    for form in args.forms.split(","):
        cfg = vortex_config(K=K, p=p, precision=form, **sched)
        s, Q, ms, t, tuned = time_form(D, torch, cfg, ic, args.steps, args.warmup, args.cfl,
                                       not (args.no_autotune or args.schedule))
        rate = dof * 3 * args.steps / (ms * 1e-3)
        res[form] = dict(ms=ms, rate=rate, t=t, tuned=tuned)
        sols[form] = (s, Q, cfg)
        log(f"[bench] {form}: {ms / args.steps:.3f} ms/step  {rate:.4e} DOF*stage/s  t={t:.6f} tuned={tuned}")

    # bytes per DOF*stage: stage 1 reads U^n, writes U1 (16 B); stages 2, 3 read
    # U_prev and U^n and write U_s (24 B)  ->  64/3 B averaged over a step
    b_alg = 64.0 / 3.0
    main_form = "mixed" if "mixed" in res else args.forms.split(",")[0]
    ms_launch = res[main_form]["ms"] / (3 * args.steps)
    achieved = b_alg * dof / (ms_launch * 1e-3) / 1e9

    acc = {}
    if not args.no_accuracy and "fp64" in sols and "mixed" in sols:
        # untimed: ACCURACY_STEPS steps of each form from the IC (the
        # configuration's own step count, whatever --steps is), then the
        # mixed form again to exactly the FP64 form's end time
        s64, Q64, cfg64 = sols["fp64"]
        s32, Q32, _ = sols["mixed"]
        Q64 = s64.init_state(ic, Q64)
        t64, _ = s64.advance(Q64, ACCURACY_STEPS, args.cfl)
        Q32 = s32.init_state(ic, Q32)
        s32.advance(Q32, 10 * ACCURACY_STEPS, args.cfl, t_end=t64)
        mf = s64.error_norms(Q32, Q64)
        # exact advected vortex (periodic images summed), on the device
        ex = s64.vortex_exact(ic, t64, images=True)
        e64 = s64.error_norms(Q64, ex)
        e32 = s64.error_norms(Q32, ex)
        acc = {"steps": ACCURACY_STEPS, "t": t64,
               "mixed_vs_fp64_rel_l2": mf["rel_l2"], "mixed_vs_fp64_rel_linf": mf["rel_linf"],
               "fp64_vs_exact_rel_l2": e64["rel_l2"], "mixed_vs_exact_rel_l2": e32["rel_l2"],
               "fp64_vs_exact_rel_linf": e64["rel_linf"], "mixed_vs_exact_rel_linf": e32["rel_linf"],
               "fp64_vs_exact_l2_per_variable": e64["l2"]}
        log(f"[bench] accuracy: {acc}")

    cpu = None
    if not args.no_cpu:
        cfgc = vortex_config(K=K, p=p, precision="fp64")
        cpu, kept = cpu_baseline(cfgc, ic, args.cpu_steps, args.cfl, keep_state=True)
        log(f"[bench] cpu baseline: {cpu}")
        if kept is not None and kept[3] == 0:
            # the full-size GPU states after the oracle's steps, against the oracle
            import numpy as np
            Qo, k, to, _ = kept
            vs = {}
            for form, (sf, Qf, _) in sols.items():
                Qf = sf.init_state(ic, Qf)
                tf, _ = sf.advance(Qf, k, args.cfl)
                ref = torch.from_numpy(Qo).to(Qf.device)
                nrm = sf.error_norms(Qf, ref)
                vs[form] = {"rel_l2": nrm["rel_l2"], "rel_linf": nrm["rel_linf"], "t": tf,
                            "t_rel_diff": abs(tf - to) / to}
                del ref
            acc["vs_oracle"] = {"steps": k, "t_oracle": to, **vs}
            log(f"[bench] vs oracle after {k} steps: {acc['vs_oracle']}")

    # short, bounded samples of the other workloads (after the headline's
    # solvers are released): KH 1024^2 and 2048^2 p = 3 in both forms, the
    # 256-instance vortex batch and the 256-instance batch of the reference's
    # Sod setup (AV + limiter); rates over --extra-steps timed steps
    extra = None
    if not args.no_extra:
        del sols
        extra = {}
        for Kx in (1024, 2048):
            extra[f"kelvin_helmholtz_{Kx}_p3"] = run_kh(Kx, 3, ["fp64", "mixed"], steps=args.extra_steps,
                                                        tag=f"bench-kh{Kx}")
        for form in ("mixed", "fp64"):
            extra[f"batched_256_vortex_{form}"] = run_batched(256, 2024, form, args.extra_steps, 3)
        for form in ("mixed", "fp64"):
            extra[f"batched_256_sod_{form}"] = run_batched(256, 2024, form, args.extra_steps, 3, kind="sod")

    ref = hbm_reference(dof)
    ref["stage_kernels_over_reference"] = ref["ms_per_step"] / (res[main_form]["ms"] / args.steps)
    log(f"[bench] HBM reference: {ref}")

    src_hash = kernel_sources_hash()
    sched_used = res[main_form]["tuned"] or (tuple(int(v) for v in args.schedule.split(",")) if args.schedule
                                              else None)
    traffic = PMC_TRAFFIC.get((K, p, main_form) + tuple(sched_used)) if sched_used else None
    traffic_ok = traffic is not None and traffic[2] == src_hash
    out = {
        "metric": "DOF*stage updates/s, mixed vs FP64 DGSEM at equal L2 error; % HBM peak",
        "value": res[main_form]["rate"],
        "unit": "DOF*stage/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res[main_form]["ms"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 RHS / fp64 state" if main_form == "mixed" else "f64",
        "data": "synthetic (isentropic vortex IC evaluated at the LGL nodes)",
        "config": {"workload": f"isentropic vortex, {K}x{K} elements, p={p} ({N}x{N} LGL nodes), "
                               f"Ranocha EC volume + Rusanov faces, SSP-RK3 CFL {args.cfl}, periodic [0,10]^2",
                   "K": K, "p": p, "dof": dof, "parallelism": "single GPU"},
        "fp64": {"value": res.get("fp64", {}).get("rate"), "ms_per_step": res.get("fp64", {}).get("ms", 0) / args.steps},
        "mixed_over_fp64": (res["mixed"]["rate"] / res["fp64"]["rate"]) if ("mixed" in res and "fp64" in res) else None,
        "accuracy": acc,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic[0] if traffic_ok else None,
                     "traffic_source": (f"{traffic[1]} (kernel sources {traffic[2]}, schedule {sched_used})"
                                        if traffic_ok else
                                        f"not measured for kernel sources {src_hash}, schedule {sched_used}"),
                     "schedule": sched_used,
                     "kernel": f"stage_kernel<{p},{'float' if main_form == 'mixed' else 'double'},lean,Ranocha>",
                     "kernel_sources": src_hash,
                     "bytes_per_dof_stage": b_alg, "avg_launch_ms": ms_launch,
                     "hbm_reference": ref},
        "cpu_baseline": cpu,
        "other_workloads": extra,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
