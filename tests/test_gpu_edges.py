"""Degenerate grids on the GPU against the FP64 oracle: one element (a
periodic element is its own neighbour on all four faces, a walled one has
walls on all four) and two elements per side, for both schemes, both
boundary kinds and the smallest, the reference's and the largest degree.

Bars as in the parity suite: IC to 2e-15, RHS <= 1e-13 relative L2, wave
speeds to 1e-15, three fused steps <= 1e-12.
"""
import numpy as np
import pytest
import torch

from fvstuff_amd.capi import make_config, make_ic, sd_config, vortex_config
from oracle.oracle import Oracle, lgl
from tests.layout import gll_weights_2d, rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dgsem():
    from fvstuff_amd import dgsem as D
    D.load_library()
    assert torch.cuda.is_available()
    return D


def cpu(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy().copy()


def l2_close(A, B, w2, tol):
    """sqrt(sum w (A-B)^2) <= tol * sqrt(sum w B^2), with an absolute floor
    for a zero reference (a uniform state has a zero right-hand side, which
    both sides evaluate to round-off of the O(1) fluxes times 2/h)."""
    A = np.asarray(A).reshape((4,) + w2.shape); B = np.asarray(B).reshape((4,) + w2.shape)
    err = np.sqrt(np.sum(w2[None] * (A - B) ** 2))
    ref = np.sqrt(np.sum(w2[None] * B ** 2))
    return err <= tol * ref + 1e-12, (err, ref)


def config(scheme, bc, K, p):
    if bc == "wall":       # the reference drivers' own problems: circular Sod in a closed box
        cfg = sd_config(K=K, p=p) if scheme == "sd" else make_config(K=K, p=p)
        return cfg, "sod"
    if scheme == "sd":     # periodic vortex on [0,10]^2
        return sd_config(K=K, p=p, L=10.0, bc="periodic", surface_flux="rusanov", limiter=False), "vortex"
    return vortex_config(K=K, p=p), "vortex"


CASES = [(sch, bc, K, p) for sch in ("dgsem", "sd") for bc in ("wall", "periodic")
         for K in (1, 2) for p in (1, 3, 7)]


@pytest.mark.parametrize("scheme,bc,K,p", CASES, ids=[f"{c[0]}-{c[1]}-K{c[2]}-p{c[3]}" for c in CASES])
def test_tiny_grids(dgsem, scheme, bc, K, p):
    cfg, ic = config(scheme, bc, K, p)
    print(scheme, bc, K, p, flush=True)
    o = Oracle(cfg); s = dgsem.Solver(cfg)
    _, w = lgl(p)
    w2 = gll_weights_2d(w, K, cfg.L / K)
    Q_ref = o.init_state(make_ic(ic))
    Q = s.init_state(make_ic(ic))
    np.testing.assert_allclose(cpu(Q), Q_ref, rtol=2e-15, atol=1e-15)
    R_ref, lam_ref = o.compute_rhs(Q_ref)
    R, lam = s.compute_rhs(torch.from_numpy(Q_ref.copy()).cuda())
    ok, info = l2_close(cpu(R), R_ref, w2, 1e-13)
    assert ok, info
    np.testing.assert_allclose(cpu(lam), lam_ref, rtol=1e-15)
    rc, t_ref, _ = o.advance(Q_ref, 3, 0.1)
    assert rc == 0
    t, _ = s.advance(Q, 3, 0.1)
    assert t == pytest.approx(t_ref, rel=1e-13)
    assert rel_l2(cpu(Q), Q_ref, w2) < 1e-12


def test_misaligned_state_rejected(dgsem):
    """The stage kernels move node pairs with 16-byte accesses: a state
    pointer that is only 8-byte aligned is rejected before any launch."""
    cfg = vortex_config(K=4, p=3)
    s = dgsem.Solver(cfg)
    buf = torch.zeros(s.size + 1, dtype=torch.float64, device="cuda")
    Q = buf[1:]
    assert Q.data_ptr() % 16 == 8
    with pytest.raises(dgsem.DGError) as e:
        s.advance(Q, 1, 0.1)
    assert e.value.status == 1
    with pytest.raises(dgsem.DGError) as e:
        s.compute_rhs(Q)
    assert e.value.status == 1
    b = dgsem.Batch([vortex_config(K=4, p=3), vortex_config(K=5, p=5)])
    buf2 = torch.zeros(b.sizes[1] + 1, dtype=torch.float64, device="cuda")
    Qs = [torch.zeros(b.sizes[0], dtype=torch.float64, device="cuda"), buf2[1:]]
    with pytest.raises(dgsem.DGError) as e:
        b.advance(Qs, 1, [0.1, 0.1])
    assert e.value.status == 1
