"""Solver-level pins of the oracle: the RHS (dgsem_lobatto_2d.cu:365-530), the
SSP-RK3 stages (:533-576), the limiter (:584-643), dt (:1029) and the step
loop (:1022-1058), checked by properties the discretisation must have."""
import math

import numpy as np
import pytest

from fvstuff_amd.capi import make_config, make_ic, vortex_config, kh_config
from oracle.oracle import Oracle, lgl
from tests.layout import as5, node_xy, prim_to_cons, gll_weights_2d, rel_l2, entropy_vars

VOLS = ["chandrashekar", "ranocha", "collocation"]
SURFS = ["hllc", "rusanov"]


def uniform_state(o, r=1.3, u=0.4, v=-0.2, p=0.9):
    n = o.NN
    U = prim_to_cons(*(np.full((o.K, o.K, n, n), a) for a in (r, u, v, p)))
    return np.ascontiguousarray(U.reshape(-1))


@pytest.mark.parametrize("vol", VOLS)
@pytest.mark.parametrize("surf", SURFS)
@pytest.mark.parametrize("p", [1, 3, 4, 7])
def test_free_stream_periodic(vol, surf, p):
    o = Oracle(make_config(K=4, p=p, volume_flux=vol, surface_flux=surf, bc="periodic", av=True, limiter=True))
    R, lam = o.compute_rhs(uniform_state(o))
    assert np.abs(R).max() < 1e-12
    assert np.allclose(lam, 0.6 + math.sqrt(1.4 * 0.9 / 1.3), rtol=1e-14)


@pytest.mark.parametrize("surf", SURFS)
def test_free_stream_wall_at_rest(surf):
    o = Oracle(make_config(K=5, p=3, surface_flux=surf, bc="wall"))
    R, _ = o.compute_rhs(uniform_state(o, u=0.0, v=0.0))
    assert np.abs(R).max() < 1e-12


@pytest.mark.parametrize("vol", VOLS)
@pytest.mark.parametrize("surf", SURFS)
@pytest.mark.parametrize("p", [2, 3, 5])
def test_discrete_conservation_periodic(vol, surf, p):
    cfg = vortex_config(K=6, p=p, volume_flux=vol, surface_flux=surf)
    o = Oracle(cfg)
    Q = o.init_state(make_ic("vortex"))
    R, _ = o.compute_rhs(Q)
    x, w = lgl(p)
    w2 = gll_weights_2d(w, o.K, o.h)
    R5 = as5(R, o.K, p)
    for q in range(4):
        tot = np.sum(w2 * R5[q])
        assert abs(tot) < 1e-12 * np.sum(w2 * np.abs(R5[q])) + 1e-14, (q, tot)


@pytest.mark.parametrize("vol", ["ranocha", "chandrashekar"])
def test_entropy_stable_with_rusanov(vol):
    """EC volume flux + dissipative Rusanov faces: sum_w v(U).dU/dt <= 0,
    while the collocation volume term has no such sign guarantee."""
    cfg = kh_config(K=8, p=3, precision="fp64", volume_flux=vol)
    o = Oracle(cfg)
    Q = o.init_state(make_ic("kh", noise_amp=1e-2))
    R, _ = o.compute_rhs(Q)
    x, w = lgl(3)
    w2 = gll_weights_2d(w, o.K, o.h)
    V = entropy_vars(as5(Q, o.K, 3))
    prod = np.sum(w2[None] * V * as5(R, o.K, 3))
    scale = np.sum(w2[None] * np.abs(V * as5(R, o.K, 3)))
    assert prod < 1e-12 * scale and prod < 0


def test_rk3_stability_polynomial():
    """du/dt = a*u through the three Shu-Osher stages gives u0*(1+z+z^2/2+z^3/6)."""
    o = Oracle(vortex_config(K=3, p=2))
    U0 = o.init_state(make_ic("vortex"))
    a, dt = -0.7, 0.3
    U1 = o.rk3_stage(1, U0, None, a * U0, dt)
    U2 = o.rk3_stage(2, U0, U1, a * U1, dt)
    U3 = o.rk3_stage(3, U0, U2, a * U2, dt)
    z = a * dt
    np.testing.assert_allclose(U3, U0 * (1 + z + z * z / 2 + z ** 3 / 6), rtol=1e-14)


def test_limiter_identity_and_repair():
    o = Oracle(make_config(K=4, p=3, bc="periodic"))
    Q = o.init_state(make_ic("sod"))
    np.testing.assert_allclose(o.limit(Q), Q, rtol=1e-14, atol=1e-14)
    x, w = lgl(3)
    Q5 = as5(Q, 4, 3).copy()
    Q5[0, 1, 2, 1, 1] = -0.05          # negative density at one node
    Q5[3, 2, 2, 0, 0] = 1e-8           # negative pressure at another
    Qb = np.ascontiguousarray(Q5.reshape(-1))
    Ql = as5(o.limit(Qb), 4, 3)
    wij = 0.25 * np.outer(w, w)
    for ej, ei in [(1, 2), (2, 2)]:
        np.testing.assert_allclose((wij * Ql[:, ej, ei]).sum((1, 2)), (wij * Q5[:, ej, ei]).sum((1, 2)), rtol=1e-13)
    r = Ql[0]; p = 0.4 * (Ql[3] - 0.5 * (Ql[1] ** 2 + Ql[2] ** 2) / r)
    assert r.min() >= 1e-6 * (1 - 1e-9) and p.min() >= 1e-6 * (1 - 1e-4)


def test_dt_brute_force_and_step():
    cfg = vortex_config(K=8, p=3)
    o = Oracle(cfg)
    Q = o.init_state(make_ic("vortex"))
    R, lam = o.compute_rhs(Q)
    U = as5(Q, 8, 3)
    r = U[0]; u = U[1] / r; v = U[2] / r; p = 0.4 * (U[3] - 0.5 * r * (u * u + v * v))
    lam_nodes = np.abs(u) + np.abs(v) + np.sqrt(1.4 * p / r)
    np.testing.assert_allclose(lam, lam_nodes.max((2, 3)).reshape(-1), rtol=1e-14)
    assert o.max_lambda(lam) == lam.max()
    rc, t, dt = o.advance(Q.copy(), 1, 0.5)
    assert rc == 0 and dt == pytest.approx(0.5 * o.h / (lam.max() * 4), rel=1e-15) and t == dt


def test_end_time_clip():
    o = Oracle(vortex_config(K=4, p=2))
    Q = o.init_state(make_ic("vortex"))
    rc, t, dt = o.advance(Q, 1000, 0.5, t_end=0.3)
    assert rc == 0 and t == pytest.approx(0.3, abs=1e-13)


def test_failure_codes():
    o = Oracle(make_config(K=4, p=2, bc="periodic", av=False, limiter=False))
    Q = uniform_state(o)
    rc, _, _ = o.advance(Q.copy(), 1, 1e-20)
    assert rc == 4                                 # dt underflow
    Qb = as5(Q, 4, 2).copy(); Qb[3, 1, 1, 1, 1] = 1e13   # c ~ 1e6.5
    rc, _, _ = o.advance(np.ascontiguousarray(Qb.reshape(-1)), 1, 0.5)
    assert rc == 5                                 # blow-up after the step


@pytest.mark.parametrize("p", [2, 3])
def test_vortex_convergence_order(p):
    """Error vs the advected exact vortex after t=0.5 drops at about h^(p+1)."""
    errs = []
    for K in (10, 20):
        cfg = vortex_config(K=K, p=p)
        o = Oracle(cfg)
        Q = o.init_state(make_ic("vortex"))
        rc, t, _ = o.advance(Q, 100000, 0.25, t_end=0.5)
        assert rc == 0
        x, w = lgl(p)
        X, Y = node_xy(K, p, 0.0, 0.0, 10.0, x)
        from oracle import oracle as O
        ex = np.zeros((4,) + X.shape)
        for idx in np.ndindex(X.shape):     # exact solution: IC advected by (1,1)*t, periodic
            W = O.vortex_exact((X[idx] - t) % 10.0, (Y[idx] - t) % 10.0)
            ex[(slice(None),) + idx] = prim_to_cons(*W)
        errs.append(rel_l2(Q, ex, gll_weights_2d(w, K, o.h)))
    rate = math.log2(errs[0] / errs[1])
    assert rate > p + 0.5, (errs, rate)
    assert errs[1] < 2e-3
