"""The seed-drawn ensembles of the batched workloads (vortex instances and
the reference's Sod setup): ranges, degrees, determinism and the per-degree
stable CFL."""
import math

from fvstuff_amd.capi import ensemble, sod_ensemble


def test_ensemble_ranges_and_determinism():
    cfgs, ics, cfls = ensemble(n=256, seed=2024)
    assert len(cfgs) == len(ics) == len(cfls) == 256
    Ks = [c.K for c in cfgs]
    ps = [c.p for c in cfgs]
    assert min(Ks) >= 16 and max(Ks) <= 128 and len(set(Ks)) > 50
    assert set(ps) == {3, 5, 7}
    for c, ic, cfl in zip(cfgs, ics, cfls):
        assert c.volume_flux == 1 and c.surface_flux == 1 and c.bc == 1 and c.precision == 1
        assert 1.0 <= ic.vortex_eps <= 5.0
        mach = math.hypot(ic.u0, ic.v0) / math.sqrt(1.4)
        assert 0.1 - 1e-12 <= mach <= 0.8 + 1e-12
        assert cfl == (0.4 if c.p == 3 else (0.25 if c.p == 5 else 0.2))
    cfgs2, ics2, _ = ensemble(n=256, seed=2024)
    assert [(c.K, c.p) for c in cfgs] == [(c.K, c.p) for c in cfgs2]
    assert [i.vortex_eps for i in ics] == [i.vortex_eps for i in ics2]
    cfgs3, _, _ = ensemble(n=256, seed=7)
    assert [(c.K, c.p) for c in cfgs3] != [(c.K, c.p) for c in cfgs]


def test_sod_ensemble_is_the_reference_setup():
    cfgs, ics, cfls = sod_ensemble(n=128, seed=2024)
    assert {c.p for c in cfgs} == {3, 5, 7}
    for c, ic, cfl in zip(cfgs, ics, cfls):
        # circular Sod, walls, Chandrashekar + HLLC, AV and limiter on
        assert (c.volume_flux, c.surface_flux, c.bc, c.av, c.limiter) == (0, 0, 0, 1, 1)
        assert (c.x0, c.y0, c.L) == (0.0, 0.0, 1.0) and c.precision == 1
        assert ic.kind == 0 and 2.0 <= ic.rho_hi <= 11.0 and 2.0 <= ic.p_hi <= 10.0
        assert 16 <= c.K <= 128
        assert cfl == (0.5 if c.p == 3 else (0.3 if c.p == 5 else 0.2))
    again = sod_ensemble(n=128, seed=2024)
    assert [(c.K, c.p, i.rho_hi) for c, i in zip(again[0], again[1])] == [(c.K, c.p, i.rho_hi) for c, i in zip(cfgs, ics)]
