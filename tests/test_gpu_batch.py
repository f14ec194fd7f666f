"""Batched launch: every instance must match its own single-solve
run bit for bit (same per-tile arithmetic), with or without graph replay,
and per-instance dt / end time / status must be honoured."""
import numpy as np
import pytest
import torch

from fvstuff_amd.capi import ensemble, vortex_config, make_ic

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


def test_batch_rejects_av_and_limiter(D):
    cfg = vortex_config(K=4, p=3, av=True)
    with pytest.raises(D.DGError):
        D.Batch([cfg])
