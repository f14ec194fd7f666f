"""Batched launch: every instance must match its own single-solve
run bit for bit (same per-tile arithmetic), with or without graph replay,
and per-instance dt / end time / status must be honoured."""
import numpy as np
import pytest
import torch

from fvstuff_amd.capi import ensemble, make_config, make_ic, vortex_config

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def D():
    from fvstuff_amd import dgsem
    dgsem.load_library()
    return dgsem


def host(ts):
    torch.cuda.synchronize()
    return [t.cpu().numpy() for t in ts]


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_batch_matches_single_solves(D, precision):
    cfgs, ics, cfls = ensemble(n=12, seed=5, precision=precision, k_range=(4, 20), degrees=(1, 2, 3, 4, 5, 6, 7))
    b = D.Batch(cfgs)
    Qs = b.init_states(ics)
    ts = b.advance(Qs, 7, cfls)
    got = host(Qs)
    for i, (cfg, ic, cfl) in enumerate(zip(cfgs, ics, cfls)):
        s = D.Solver(cfg)
        Q = s.init_state(ic)
        t, _ = s.advance(Q, 7, cfl)
        assert t == ts[i]
        assert np.array_equal(host([Q])[0], got[i]), i


def test_batch_graph_and_each_agree(D):
    cfgs, ics, cfls = ensemble(n=10, seed=11, k_range=(6, 24))
    b = D.Batch(cfgs)
    runs = []
    for mode in ("graph", "eager", "each"):
        Qs = b.init_states(ics)
        if mode == "each":
            ts = b.advance_each(Qs, 9, cfls)
        else:
            b.set_options(use_graph=(mode == "graph"))
            ts = b.advance(Qs, 9, cfls)
        runs.append((ts, host(Qs)))
    for ts, Q in runs[1:]:
        assert ts == runs[0][0]
        for a, c in zip(Q, runs[0][1]):
            assert np.array_equal(a, c)


def test_batch_lanes_agree(D):
    """Tiles of 64, 128 or 256 lanes cut the instances differently; the
    results are the same bit for bit."""
    cfgs, ics, cfls = ensemble(n=8, seed=3, k_range=(5, 17), degrees=(1, 3, 4, 5, 7))
    b = D.Batch(cfgs)
    ref = None
    for lanes in (64, 128, 256):
        b.set_options(use_graph=True, lanes=lanes)
        Qs = b.init_states(ics)
        ts = b.advance(Qs, 6, cfls)
        got = (ts, host(Qs))
        if ref is None:
            ref = got
            continue
        assert got[0] == ref[0]
        for a, c in zip(got[1], ref[1]):
            assert np.array_equal(a, c)
    with pytest.raises(D.DGError):
        b.set_options(lanes=96)


def test_batch_end_time_and_status(D):
    cfgs = [vortex_config(K=8, p=3, precision="mixed"), vortex_config(K=6, p=5, precision="mixed")]
    ics = [make_ic("vortex"), make_ic("vortex")]
    b = D.Batch(cfgs)
    Qs = b.init_states(ics)
    ts = b.advance(Qs, 10000, [0.5, 0.35], t_end=0.2)
    assert ts[0] == pytest.approx(0.2, abs=1e-13) and ts[1] == pytest.approx(0.2, abs=1e-13)
    Qs = b.init_states(ics)
    with pytest.raises(D.DGError) as e:
        b.advance(Qs, 1, [0.5, 1e-20])
    assert e.value.status == 4 and b.last_status == [0, 4]


def test_batch_shock_capturing_rules(D):
    """Shock capturing (AV and / or the limiter) must be alike in every
    instance, and runs in the by-degree form only."""
    with pytest.raises(D.DGError):
        D.Batch([vortex_config(K=4, p=3, av=True), vortex_config(K=4, p=3)])
    with pytest.raises(D.DGError):
        D.Batch([make_config(K=4, p=3), make_config(K=4, p=3, limiter=False)])
    b = D.Batch([make_config(K=4, p=3), make_config(K=5, p=2)])
    assert b.by_degree()
    with pytest.raises(D.DGError, match="by-degree form only"):
        b.set_by_degree(False)
    with pytest.raises(D.DGError, match="by-degree form only"):
        b.set_persistent(True)
    assert b.by_degree() and not b.options()[2]


@pytest.mark.parametrize("by_degree", [False, True], ids=["all-degree", "by-degree"])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("mixed", 1e-6)])
def test_batch_matches_oracle(D, precision, tol, by_degree):
    """Every instance of a 12-instance ensemble (degrees 1..7, K 4..20, seeded
    strength and Mach) after 20 batched steps against the FP64 oracle run
    of the same instance: FP64 within 1e-12 relative L2, mixed within 1e-6.
    The all-degree kernel has its own load schedule (in FP32 the stage input
    is loaded twice instead of kept in LDS, and U^n after the line phase);
    the by-degree form runs one launch per degree with the single-solve body."""
    from oracle.oracle import Oracle, lgl
    from tests.layout import gll_weights_2d, rel_l2
    # seed 4 draws every degree from 1 to 7
    cfgs, ics, cfls = ensemble(n=12, seed=4, precision=precision, k_range=(4, 20), degrees=(1, 2, 3, 4, 5, 6, 7))
    assert sorted({c.p for c in cfgs}) == [1, 2, 3, 4, 5, 6, 7]
    b = D.Batch(cfgs)
    b.set_by_degree(by_degree)
    assert b.by_degree() == by_degree
    Qs = b.init_states(ics)
    ts = b.advance(Qs, 20, cfls)
    got = host(Qs)
    for i, (cfg, ic, cfl) in enumerate(zip(cfgs, ics, cfls)):
        c64 = vortex_config(K=cfg.K, p=cfg.p)
        o = Oracle(c64)
        Q = o.init_state(ic)
        rc, t, _ = o.advance(Q, 20, cfl)
        assert rc == 0
        _, w = lgl(cfg.p)
        err = rel_l2(got[i], Q, gll_weights_2d(w, cfg.K, c64.L / cfg.K))
        print(f"instance {i} K={cfg.K} p={cfg.p}: rel L2 vs oracle {err:.2e}")
        assert err < tol, (i, cfg.K, cfg.p, err)
        assert ts[i] == pytest.approx(t, rel=1e-13 if precision == "fp64" else 1e-6)


def test_batch_blowup_found_when_trailing_steps_are_noops(D):
    """One instance carries an energy spike (E = 1e13 at one node); the
    common end time is that instance's first dt, so its second step and
    every later one are no-ops.  The last step taken must still be checked:
    status 5 for the spiked instance, 0 for the other."""
    cfgs = [vortex_config(K=8, p=3, precision="mixed"), vortex_config(K=8, p=3, precision="mixed")]
    ics = [make_ic("vortex"), make_ic("vortex")]
    b = D.Batch(cfgs)
    Qs = b.init_states(ics)
    Qs[1].view(4, 8, 8, 4, 4)[3, 2, 2, 1, 1] = 1e13
    s = D.Solver(cfgs[1])
    _, lam = s.compute_rhs(Qs[1])
    _, dt = s.cfl_dt(lam, 0.5)
    with pytest.raises(D.DGError) as e:
        b.advance(Qs, 5, [0.5, 0.5], t_end=dt)
    assert e.value.status == 5 and b.last_status == [0, 5]


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_batch_persistent_matches(D, precision):
    """The persistent form of the stage launches (resident workgroups walk
    the tile list) and the XCD-binned list order (padded with empty entries)
    give the same bits as one workgroup per tile in the global order, at
    every workgroup size, with and without graph replay."""
    cfgs, ics, cfls = ensemble(n=9, seed=4, precision=precision, k_range=(4, 20), degrees=(1, 2, 3, 4, 5, 6, 7))
    b = D.Batch(cfgs)
    ref = None
    for lanes in (64, 128, 256):
        for persistent in (False, True):
            for bins in (False, True):
                for use_graph in (True, False):
                    b.set_options(use_graph=use_graph, lanes=lanes)
                    b.set_xcd_bins(bins)
                    b.set_persistent(persistent)
                    assert b.options() == (use_graph, lanes, persistent, bins)
                    Qs = b.init_states(ics)
                    ts = b.advance(Qs, 7, cfls)
                    got = (ts, host(Qs))
                    if ref is None:
                        ref = got
                        continue
                    key = (lanes, persistent, bins, use_graph)
                    assert got[0] == ref[0], key
                    for a, c in zip(got[1], ref[1]):
                        assert np.array_equal(a, c), key


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_batch_by_degree_matches(D, precision):
    """The by-degree form (one launch per degree present, each with the
    single-solve lean kernel of its degree) gives every instance the bits of
    its own single solve, and so of the all-degree kernel, at every workgroup
    size, in both block orders, with and without graph replay; per-instance
    end times are honoured."""
    cfgs, ics, cfls = ensemble(n=11, seed=4, precision=precision, k_range=(4, 20), degrees=(1, 2, 3, 4, 5, 6, 7))
    ref = []
    for cfg, ic, cfl in zip(cfgs, ics, cfls):
        s = D.Solver(cfg)
        Q = s.init_state(ic)
        t, _ = s.advance(Q, 7, cfl)
        ref.append((t, host([Q])[0]))
    b = D.Batch(cfgs)
    for lanes in (64, 128, 256):
        for bins in (False, True):
            for use_graph in (True, False):
                b.set_options(use_graph=use_graph, lanes=lanes)
                b.set_by_degree(True)
                b.set_xcd_bins(bins)
                assert b.by_degree() and b.options() == (use_graph, lanes, False, bins)
                Qs = b.init_states(ics)
                ts = b.advance(Qs, 7, cfls)
                got = host(Qs)
                key = (lanes, bins, use_graph)
                for i, (t, Q) in enumerate(ref):
                    assert ts[i] == t, (key, i)
                    assert np.array_equal(got[i], Q), (key, i)
    # each degree's launch with its own workgroup size
    b.set_options(use_graph=True, lanes=128)
    b.set_by_degree(True)
    for p, lanes in ((1, 256), (3, 64), (5, 256), (7, 64)):
        b.set_degree_lanes(p, lanes)
    assert [b.degree_lanes(p) for p in range(1, 8)] == [256, 128, 64, 128, 256, 128, 64]
    Qs = b.init_states(ics)
    ts = b.advance(Qs, 7, cfls)
    got = host(Qs)
    for i, (t, Q) in enumerate(ref):
        assert ts[i] == t and np.array_equal(got[i], Q), ("per-degree lanes", i)
    for bad in ((0, 64), (8, 64), (3, 96)):
        with pytest.raises(D.DGError):
            b.set_degree_lanes(*bad)
    for p in range(1, 8):
        b.set_degree_lanes(p, 0)
    # a common end time inside the run: every instance stops there
    Qs = b.init_states(ics)
    ts = b.advance(Qs, 100000, cfls, t_end=0.01)
    assert all(t == pytest.approx(0.01, abs=1e-13) for t in ts)


def test_batch_forms_exclude_each_other(D):
    """The by-degree and persistent forms exclude each other: setting one
    clears the other, and each gives the same bits."""
    cfgs, ics, cfls = ensemble(n=6, seed=9, k_range=(4, 12), degrees=(2, 3, 5))
    b = D.Batch(cfgs)
    out = []
    b.set_by_degree(True)
    b.set_persistent(True)
    assert not b.by_degree() and b.options()[2]
    Qs = b.init_states(ics)
    out.append((b.advance(Qs, 5, cfls), host(Qs)))
    b.set_by_degree(True)
    assert b.by_degree() and not b.options()[2]
    Qs = b.init_states(ics)
    out.append((b.advance(Qs, 5, cfls), host(Qs)))
    assert out[0][0] == out[1][0]
    for a, c in zip(out[0][1], out[1][1]):
        assert np.array_equal(a, c)


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_batch_by_degree_tiny_instances(D, precision):
    """Instances of 1, 2 and 3 elements per side at every degree (launches of
    fewer tiles than there are XCDs, single-element periodic domains whose
    faces are their own) in the by-degree form with the XCD-aware order:
    every instance gets its single solve's bits."""
    cfgs = [vortex_config(K=K, p=p, precision=precision) for p in range(1, 8) for K in (1, 2, 3)]
    ics = [make_ic("vortex") for _ in cfgs]
    cfls = [0.2] * len(cfgs)
    b = D.Batch(cfgs)
    b.set_by_degree(True)
    b.set_xcd_bins(True)
    for lanes in (64, 256):
        b.set_options(use_graph=True, lanes=lanes)
        Qs = b.init_states(ics)
        ts = b.advance(Qs, 5, cfls)
        got = host(Qs)
        for i, (cfg, ic) in enumerate(zip(cfgs, ics)):
            s = D.Solver(cfg)
            Q = s.init_state(ic)
            t, _ = s.advance(Q, 5, 0.2)
            assert ts[i] == t, (lanes, cfg.K, cfg.p)
            assert np.array_equal(host([Q])[0], got[i]), (lanes, cfg.K, cfg.p)


def _sod_ensemble(precision, n=10, seed=7):
    """Circular-Sod instances with walls and the reference's shock capturing
    (Chandrashekar + HLLC, Ducros AV, Zhang-Shu limiter), degrees 1..7,
    seeded K, density / pressure ratios and AV strength."""
    rng = np.random.default_rng(seed)
    cfgs, ics = [], []
    for i in range(n):
        p = 1 + i % 7
        K = int(rng.integers(3, 11))
        cfgs.append(make_config(K=K, p=p, precision=precision, av_cav=float(rng.uniform(0.15, 0.35))))
        ics.append(make_ic("sod", rho_hi=float(rng.uniform(2.0, 11.0)), p_hi=float(rng.uniform(2.0, 10.0))))
    return cfgs, ics


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_batch_shock_capturing_matches_single_solves(D, precision):
    """The reference's default configuration (circular Sod, walls, AV,
    limiter) batched over degrees 1..7: every instance gets the bits of its
    own dg_advance at every workgroup size, in both block orders, with and
    without graph replay, with per-degree workgroup sizes and after the
    autotuner (which keeps the by-degree form)."""
    cfgs, ics = _sod_ensemble(precision)
    cfl = 0.3
    ref = []
    for cfg, ic in zip(cfgs, ics):
        s = D.Solver(cfg)
        Q = s.init_state(ic)
        t, _ = s.advance(Q, 8, cfl)
        ref.append((t, host([Q])[0]))
    b = D.Batch(cfgs)
    for lanes in (64, 256):
        for bins in (False, True):
            for use_graph in (True, False):
                b.set_options(use_graph=use_graph, lanes=lanes)
                b.set_xcd_bins(bins)
                Qs = b.init_states(ics)
                ts = b.advance(Qs, 8, [cfl] * len(cfgs))
                got = host(Qs)
                for i, (t, Q) in enumerate(ref):
                    key = (lanes, bins, use_graph, cfgs[i].K, cfgs[i].p)
                    assert ts[i] == t, key
                    assert np.array_equal(got[i], Q), key
    # per-degree workgroup sizes, then the autotuner's schedule
    for step in ("degree_lanes", "autotune"):
        if step == "degree_lanes":
            b.set_options(use_graph=True, lanes=64)
            for p, lanes in ((7, 128), (5, 256), (3, 64), (1, 128)):
                b.set_degree_lanes(p, lanes)
        else:
            Qs = b.init_states(ics)
            b.autotune(Qs, [cfl] * len(cfgs))
            assert b.by_degree()
        Qs = b.init_states(ics)
        ts = b.advance(Qs, 8, [cfl] * len(cfgs))
        got = host(Qs)
        for i, (t, Q) in enumerate(ref):
            key = (step, b.options(), cfgs[i].K, cfgs[i].p, b.degree_lanes(cfgs[i].p))
            assert ts[i] == t, key
            assert np.array_equal(got[i], Q), key


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-11), ("mixed", 1e-6)])
def test_batch_shock_capturing_matches_oracle(D, precision, tol):
    """A batch of circular blasts (pressure ratio 1e6, where the limiter
    changes the state) and reference Sod instances, 20 fused steps, against
    the FP64 oracle run of each instance: FP64 within 1e-11 relative L2,
    mixed within 1e-6 (p <= 3: at p = 7 the FP32 flux error at the blast
    front reaches the low-pressure state, as for single solves)."""
    from oracle.oracle import Oracle, lgl
    from tests.layout import gll_weights_2d, rel_l2
    degrees = (1, 2, 3) if precision == "mixed" else (1, 3, 5, 7)
    cases = []
    for p in degrees:
        K = {1: 12, 2: 10, 3: 8, 5: 6, 7: 6}[p]   # (K = 5, p = 7 Sod blows up at CFL 0.4 in the oracle too)
        cases.append((make_config(K=K, p=p, precision=precision), make_ic("sod", rho_hi=1.0, p_hi=1e6)))
        cases.append((make_config(K=K, p=p, precision=precision), make_ic("sod")))
    cfgs = [c for c, _ in cases]
    ics = [ic for _, ic in cases]
    b = D.Batch(cfgs)
    Qs = b.init_states(ics)
    ts = b.advance(Qs, 20, [0.4] * len(cfgs))
    got = host(Qs)
    for i, (cfg, ic) in enumerate(cases):
        o = Oracle(make_config(K=cfg.K, p=cfg.p))
        Q = o.init_state(ic)
        rc, t, _ = o.advance(Q, 20, 0.4)
        assert rc == 0
        _, w = lgl(cfg.p)
        err = rel_l2(got[i], Q, gll_weights_2d(w, cfg.K, cfg.L / cfg.K))
        print(f"instance {i} K={cfg.K} p={cfg.p} p_hi={ic.p_hi:g}: rel L2 vs oracle {err:.2e}")
        assert err < tol, (i, cfg.K, cfg.p, err)
        assert ts[i] == pytest.approx(t, rel=1e-12 if precision == "fp64" else 1e-6)


def _sd_ensemble(precision, kind):
    """Spectral-difference instances (fr_lobatto_2d.cu), every switch alike
    within a batch: "sod", the reference driver's circular Sod with walls,
    HLLC and the nodal repair at p = 1..4; "vortex", periodic vortices with
    Rusanov faces at p = 1..7."""
    from fvstuff_amd.capi import sd_config
    cfgs, ics, cfls = [], [], []
    if kind == "sod":
        for p, K in ((1, 12), (2, 10), (3, 9), (4, 7)):
            cfgs.append(sd_config(K=K, p=p, precision=precision))
            ics.append(make_ic("sod"))
            cfls.append(0.1 / p)
    else:
        for p, K in ((1, 12), (2, 9), (3, 8), (4, 6), (5, 5), (6, 4), (7, 4)):
            cfgs.append(sd_config(K=K, p=p, precision=precision, L=10.0, bc="periodic", surface_flux="rusanov",
                                  limiter=False))
            ics.append(make_ic("vortex"))
            cfls.append(0.2 if p <= 5 else 0.15)
    return cfgs, ics, cfls


@pytest.mark.parametrize("kind", ["sod", "vortex"])
@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_batch_sd_matches_single_solves(D, precision, kind):
    """Batched spectral-difference solves (by-degree form only) give every
    instance the bits of its own dg_advance at every workgroup size, in both
    block orders, with and without graph replay; the all-degree and
    persistent forms are refused; a batch mixing SD and DGSEM is refused."""
    cfgs, ics, cfls = _sd_ensemble(precision, kind)
    ref = []
    for cfg, ic, cfl in zip(cfgs, ics, cfls):
        s = D.Solver(cfg)
        Q = s.init_state(ic)
        t, _ = s.advance(Q, 8, cfl)
        ref.append((t, host([Q])[0]))
    b = D.Batch(cfgs)
    assert b.by_degree()
    with pytest.raises(D.DGError, match="by-degree form only"):
        b.set_by_degree(False)
    with pytest.raises(D.DGError, match="by-degree form only"):
        b.set_persistent(True)
    for lanes in (64, 256):
        for bins in (False, True):
            for use_graph in (True, False):
                b.set_options(use_graph=use_graph, lanes=lanes)
                b.set_xcd_bins(bins)
                Qs = b.init_states(ics)
                ts = b.advance(Qs, 8, cfls)
                got = host(Qs)
                for i, (t, Q) in enumerate(ref):
                    key = (lanes, bins, use_graph, cfgs[i].K, cfgs[i].p)
                    assert ts[i] == t, key
                    assert np.array_equal(got[i], Q), key
    with pytest.raises(D.DGError):
        D.Batch([cfgs[0], make_config(K=6, p=3, precision=precision)])


@pytest.mark.parametrize("kind", ["sod", "vortex"])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("mixed", 1e-6)])
def test_batch_sd_matches_oracle(D, precision, tol, kind):
    """The SD ensemble after 20 batched steps against the FP64 oracle run of
    each instance: FP64 within 1e-12 relative L2, mixed within 1e-6 (the
    bars of the single-solve SD loops)."""
    from oracle.oracle import Oracle, lgl
    from tests.layout import gll_weights_2d, rel_l2
    cfgs, ics, cfls = _sd_ensemble(precision, kind)
    b = D.Batch(cfgs)
    Qs = b.init_states(ics)
    ts = b.advance(Qs, 20, cfls)
    got = host(Qs)
    for i, (cfg, ic, cfl) in enumerate(zip(cfgs, ics, cfls)):
        c64 = type(cfg).from_buffer_copy(cfg)
        c64.precision = 0
        o = Oracle(c64)
        Q = o.init_state(ic)
        rc, t, _ = o.advance(Q, 20, cfl)
        assert rc == 0
        _, w = lgl(cfg.p)
        err = rel_l2(got[i], Q, gll_weights_2d(w, cfg.K, cfg.L / cfg.K))
        print(f"SD instance {i} K={cfg.K} p={cfg.p}: rel L2 vs oracle {err:.2e}")
        assert err < tol, (i, cfg.K, cfg.p, err)
        assert ts[i] == pytest.approx(t, rel=1e-12 if precision == "fp64" else 1e-6)
