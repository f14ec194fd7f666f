"""Closed-form pins of the oracle's initial conditions and of the CFL used
for the degree-7 vortex.

- Kelvin-Helmholtz step shear layer (the workload's own definition, no
  counterpart in the reference): rho = 2, u = 0.5 for |y| < 0.5, else
  rho = 1, u = -0.5; p = 2.5; v = 0.1 sin(2 pi x) + uniform noise in
  [-amp, amp].  Nodes exactly on the element faces y = +-0.5 belong to the
  outer layer (strict inequality).
- Circular Sod of the DGSEM driver (dgsem_lobatto_2d.cu:692-717): the
  density and pressure share one tanh profile phi = (1 + tanh((R - r)/delta))/2
  around the centre (L/2, L/2) with R = L/4 and the front width delta = h/2
  (:707); the spectral-difference driver uses 2h (fr_lobatto_2d.cu:648).
- SSP-RK3 with dt = cfl h/(lam (p+1)) on the degree-7 vortex is unstable at
  CFL 0.5 and stable at 0.25 (the headline configuration's value).
"""
import math

import numpy as np
import pytest

from fvstuff_amd.capi import kh_config, make_config, make_ic, sd_config, vortex_config
from oracle.oracle import Oracle, lgl
from tests.layout import as5, cons_to_prim, node_xy


def prims(o, Q):
    r, u, v, p = cons_to_prim(as5(Q, o.K, o.p))
    return r, u, v, p


def coords(cfg):
    xi, _ = lgl(cfg.p)
    return node_xy(cfg.K, cfg.p, cfg.x0, cfg.y0, cfg.L, xi)


@pytest.mark.parametrize("p", [1, 3, 4])
def test_kh_ic_closed_form(p):
    cfg = kh_config(K=8, p=p, precision="fp64")
    o = Oracle(cfg)
    r, u, v, pr = prims(o, o.init_state(make_ic("kh", noise_amp=0.0)))
    x, y = coords(cfg)
    inner = np.abs(y) < 0.5
    # the faces y = +-0.5 carry nodes (h = 1/4) and they are outer nodes
    on_face = np.abs(np.abs(y) - 0.5) == 0.0
    assert on_face.sum() == 2 * 2 * cfg.K * (p + 1)
    assert not np.any(inner & on_face)
    np.testing.assert_array_equal(r, np.where(inner, 2.0, 1.0))
    np.testing.assert_allclose(u, np.where(inner, 0.5, -0.5), rtol=0, atol=1e-15)
    np.testing.assert_allclose(v, 0.1 * np.sin(2 * np.pi * x), rtol=0, atol=1e-15)
    np.testing.assert_allclose(pr, 2.5, rtol=1e-14)
    # both layers are present in equal area: the inner band is half the domain
    assert 0.3 < inner.mean() < 0.7


def test_kh_ic_noise_moves_v_only():
    cfg = kh_config(K=8, p=3, precision="fp64")
    o = Oracle(cfg)
    amp = 1e-3
    r0, u0, v0, p0 = prims(o, o.init_state(make_ic("kh", noise_amp=0.0)))
    r1, u1, v1, p1 = prims(o, o.init_state(make_ic("kh", noise_amp=amp, seed=77)))
    np.testing.assert_array_equal(r1, r0)
    np.testing.assert_allclose(u1, u0, rtol=0, atol=1e-15)
    np.testing.assert_allclose(p1, p0, rtol=1e-12)
    d = v1 - v0
    assert np.all(np.abs(d) <= amp * (1 + 1e-9))
    # uniform on [-amp, amp]: mean ~0, std amp/sqrt(3), both signs present
    assert abs(d.mean()) < 0.1 * amp and abs(d.std() - amp / math.sqrt(3)) < 0.1 * amp
    assert d.min() < -0.8 * amp and d.max() > 0.8 * amp


def sod_profile(cfg, limiter_off=True):
    o = Oracle(cfg)
    r, u, v, pr = prims(o, o.init_state(make_ic("sod")))
    x, y = coords(cfg)
    rr = np.hypot(x - (cfg.x0 + 0.5 * cfg.L), y - (cfg.y0 + 0.5 * cfg.L))
    return r, u, v, pr, rr


@pytest.mark.parametrize("scheme,width", [("dgsem", 0.5), ("sd", 2.0)])
def test_circular_sod_front_width(scheme, width):
    """The tanh width recovered from the nodes: delta = (R - r)/atanh(2 phi - 1)
    with phi from the density, at nodes where the profile is not saturated.
    It is h/2 for DGSEM (:707) and 2h for the spectral-difference driver."""
    K, p = 16, 3
    if scheme == "dgsem":
        cfg = make_config(K=K, p=p, limiter=False)
    else:
        cfg = sd_config(K=K, p=p, limiter=False)
    h = cfg.L / K
    r, u, v, pr, rr = sod_profile(cfg)
    assert np.all(u == 0.0) and np.all(v == 0.0)
    phi_r = (r - 0.125) / (11.0 - 0.125)
    phi_p = (pr - 0.1) / (10.0 - 0.1)
    np.testing.assert_allclose(phi_p, phi_r, rtol=0, atol=1e-12)
    ok = (phi_r > 0.02) & (phi_r < 0.98) & (np.abs(rr - 0.25 * cfg.L) > 1e-3 * h)
    assert ok.sum() > 20
    delta = (0.25 * cfg.L - rr[ok]) / np.arctanh(2 * phi_r[ok] - 1)
    np.testing.assert_allclose(delta, width * h, rtol=1e-9)
    # far inside / outside: the two plateaus
    assert np.all(np.abs(r[rr < 0.25 * cfg.L - 8 * width * h] - 11.0) < 1e-5)
    assert np.all(np.abs(r[rr > 0.25 * cfg.L + 8 * width * h] - 0.125) < 1e-5)


def test_circular_sod_centre_and_symmetry():
    """Centre (L/2, L/2) on a shifted domain; the field is symmetric under
    x -> -x and x <-> y about the centre (LGL nodes are mirror-symmetric)."""
    K, p = 12, 3
    cfg = make_config(K=K, p=p, x0=2.0, y0=-1.0, L=3.0, limiter=False)
    r, _, _, _, rr = sod_profile(cfg)
    phi = (r - 0.125) / (11.0 - 0.125)
    # the profile is a function of the distance to (3.5, 0.5) alone
    order = np.argsort(rr.ravel())
    assert np.all(np.diff(phi.ravel()[order]) <= 1e-12)
    # mirror in x: element column ei <-> K-1-ei, node in <-> p-in
    np.testing.assert_allclose(r, r[:, ::-1, :, ::-1], rtol=1e-13)
    np.testing.assert_allclose(r, np.transpose(r, (1, 0, 3, 2)), rtol=1e-13)


def test_degree7_vortex_cfl():
    """dt = cfl h/(lam (p+1)) on the strong vortex (eps = 5, background
    (1, 1)) at p = 7: CFL 0.5 blows up (max wave speed > 1e6, status 5)
    within 20 steps; CFL 0.25 keeps the max wave speed within 1 % of its
    initial value over 200 steps.  p = 3 runs at CFL 0.5."""
    ic = make_ic("vortex")
    o = Oracle(vortex_config(K=8, p=7))
    Q = o.init_state(ic)
    rc, t, _ = o.advance(Q, 20, 0.5)
    assert rc == 5
    lam0 = o.compute_rhs(o.init_state(ic))[1].max()
    Q = o.init_state(ic)
    for _ in range(10):
        rc, t, _ = o.advance(Q, 20, 0.25, t=0.0)
        assert rc == 0
        assert abs(o.compute_rhs(Q)[1].max() / lam0 - 1) < 0.01
    o3 = Oracle(vortex_config(K=8, p=3))
    Q = o3.init_state(ic)
    rc, _, _ = o3.advance(Q, 200, 0.5)
    assert rc == 0
