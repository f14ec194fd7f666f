"""Device diagnostics: the exact advected vortex and the GLL-quadrature error
norms (dg_vortex_exact, dg_error_norms), checked against plain numpy
restatements of their definitions and against the IC kernel."""
import math

import numpy as np
import pytest
import torch

from fvstuff_amd.capi import make_ic, vortex_config
from oracle.oracle import lgl
from tests.layout import gll_weights_2d

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def D():
    from fvstuff_amd import dgsem
    dgsem.load_library()
    return dgsem


def cpu(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy().copy()


def node_xy(K, p, L, x0=0.0, y0=0.0):
    """Node coordinates in SIDX order (element-major, x fastest)."""
    x, _ = lgl(p)
    h = L / K
    e = np.arange(K, dtype=np.float64)
    X = x0 + h * (e[None, :, None, None] + 0.5 * (x[None, None, None, :] + 1.0))
    Y = y0 + h * (e[:, None, None, None] + 0.5 * (x[None, None, :, None] + 1.0))
    X = np.broadcast_to(X, (K, K, p + 1, p + 1)).reshape(-1)
    Y = np.broadcast_to(Y, (K, K, p + 1, p + 1)).reshape(-1)
    return X, Y


def vortex_images_np(X, Y, t, L=10.0, eps=5.0, xc=5.0, yc=5.0, u0=1.0, v0=1.0, g=1.4):
    """The advected vortex with its perturbation summed over the 3x3 periodic
    images (numpy restatement of the definition)."""
    xs = np.mod(X - u0 * t, L)
    ys = np.mod(Y - v0 * t, L)
    du = np.zeros_like(X); dv = np.zeros_like(X); dT = np.zeros_like(X)
    for iy in (-1, 0, 1):
        for ix in (-1, 0, 1):
            dx = xs - (xc + ix * L); dy = ys - (yc + iy * L)
            r2 = dx * dx + dy * dy
            f = eps / (2 * math.pi) * np.exp(0.5 * (1 - r2))
            du += -f * dy; dv += f * dx
            dT += -(g - 1) * eps * eps / (8 * g * math.pi ** 2) * np.exp(1 - r2)
    T = 1 + dT
    r = T ** (1 / (g - 1)); pr = T ** (g / (g - 1)); u = u0 + du; v = v0 + dv
    return np.stack([r, r * u, r * v, pr / (g - 1) + 0.5 * r * (u * u + v * v)])


def test_vortex_exact_single_image_at_t0_is_the_ic(D):
    cfg = vortex_config(K=12, p=4)
    s = D.Solver(cfg)
    ic = make_ic("vortex")
    Q0 = cpu(s.init_state(ic))
    Qe = cpu(s.vortex_exact(ic, 0.0, images=False))
    assert np.array_equal(Q0, Qe)


@pytest.mark.parametrize("t", [0.0, 0.37, 7.3])
def test_vortex_exact_images_matches_definition(D, t):
    K, p = 10, 5
    cfg = vortex_config(K=K, p=p)
    s = D.Solver(cfg)
    ic = make_ic("vortex")
    Qe = cpu(s.vortex_exact(ic, t, images=True)).reshape(4, -1)
    X, Y = node_xy(K, p, 10.0)
    ref = vortex_images_np(X, Y, t)
    assert np.max(np.abs(Qe - ref)) <= 1e-13 * np.max(np.abs(ref))


def test_vortex_exact_is_periodic_with_images(D):
    """Nodes on the domain's left and right edges carry the same state."""
    K, p = 8, 3
    N = p + 1
    s = D.Solver(vortex_config(K=K, p=p))
    Qe = cpu(s.vortex_exact(make_ic("vortex"), 0.0, images=True)).reshape(4, K, K, N, N)
    left = Qe[:, :, 0, :, 0]
    right = Qe[:, :, K - 1, :, N - 1]
    assert np.max(np.abs(left - right)) <= 1e-12
    single = cpu(s.vortex_exact(make_ic("vortex"), 0.0, images=False)).reshape(4, K, K, N, N)
    assert np.max(np.abs(single[:, :, 0, :, 0] - single[:, :, K - 1, :, N - 1])) > 1e-9


def test_error_norms_match_quadrature(D):
    K, p = 9, 3
    cfg = vortex_config(K=K, p=p)
    s = D.Solver(cfg)
    rng = np.random.default_rng(3)
    n = 4 * K * K * (p + 1) ** 2
    A = rng.standard_normal(n); B = rng.standard_normal(n) + 2.0
    out = s.error_norms(torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda())
    _, w = lgl(p)
    wq = gll_weights_2d(w, K, 10.0 / K).reshape(-1)
    d = (A - B).reshape(4, -1); Bv = B.reshape(4, -1)
    l2 = np.sqrt(np.sum(wq * d * d, axis=1))
    assert np.allclose(out["l2"], l2, rtol=1e-12, atol=0)
    assert np.array_equal(out["max"], np.max(np.abs(d), axis=1))
    assert out["rel_l2"] == pytest.approx(np.sqrt(np.sum(wq * d * d)) / np.sqrt(np.sum(wq * Bv * Bv)), rel=1e-12)
    assert out["rel_linf"] == pytest.approx(np.max(np.abs(d)) / np.max(np.abs(Bv)), rel=1e-15)
    # a NaN anywhere propagates
    A[17] = np.nan
    out = s.error_norms(torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda())
    assert math.isnan(out["rel_l2"]) and math.isnan(out["rel_linf"]) and math.isnan(out["l2"][0])


def test_error_vs_exact_after_advection(D):
    """FP64 vortex advected for a short time: the error against the exact
    solution is small and equals the norms computed from the host copies."""
    K, p = 16, 3
    cfg = vortex_config(K=K, p=p, precision="fp64")
    s = D.Solver(cfg)
    ic = make_ic("vortex")
    Q = s.init_state(ic)
    t, _ = s.advance(Q, 20, 0.3)
    Qe = s.vortex_exact(ic, t, images=True)
    out = s.error_norms(Q, Qe)
    _, w = lgl(p)
    wq = gll_weights_2d(w, K, 10.0 / K).reshape(-1)
    A = cpu(Q).reshape(4, -1); B = cpu(Qe).reshape(4, -1)
    ref = np.sqrt(np.sum(wq * (A - B) ** 2)) / np.sqrt(np.sum(wq * B * B))
    assert out["rel_l2"] == pytest.approx(ref, rel=1e-10)
    assert 0.0 < out["rel_l2"] < 1e-3


def test_frame_fields_match_oracle(D):
    from fvstuff_amd.capi import make_config
    from oracle.oracle import Oracle
    cfg = make_config(K=12, p=3)          # the reference's default run: circular Sod, walls
    s = D.Solver(cfg)
    ic = make_ic("sod")
    Q = s.init_state(ic)
    s.advance(Q, 10, 0.4)
    rho, p = s.frame_fields(Q)
    o = Oracle(cfg)
    rb, pb = o.frame_fields(cpu(Q))
    assert np.allclose(cpu(rho), rb, rtol=1e-15, atol=0)
    assert np.allclose(cpu(p), pb, rtol=1e-14, atol=0)


def test_frame_writer_files(D, tmp_path):
    from fvstuff_amd import frames
    from fvstuff_amd.capi import make_config
    K = 10
    s = D.Solver(make_config(K=K, p=3))
    Q = s.init_state(make_ic("sod"))
    fw = frames.FrameWriter(s, outdir=str(tmp_path), binary=True)
    rho0, _, _ = fw.write(Q, 0)
    s.advance(Q, 5, 0.4)
    fw.write(Q, 1)
    w, h, pix = frames.read_ppm(str(tmp_path / "dgsem_sod_0001.ppm"))
    assert (w, h) == (2 * K, K)
    assert frames.read_ppm(str(tmp_path / "dgsem_sensor_0000.ppm"))[:2] == (K, K)
    K2, r, _ = frames.read_rt_bin(str(tmp_path / "dgsem_sod_0000.bin"))
    assert K2 == K and np.array_equal(r.reshape(-1), rho0.astype(np.float32))
    # the colour range is fixed by the first frame
    assert fw.rho_range == frames.image_range(rho0)
