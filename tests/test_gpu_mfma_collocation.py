"""The collocation volume term on the matrix cores (p = 3, mixed form).

stage_kernel<3, float, false, VOL_COLLOCATION, true> contracts the nodes'
x- and y-fluxes with the volume matrix as 4x4x1 FP32 MFMA blocks, one per
element and variable (F_q D^T + D G_q on the element's 4 x 4 node arrays),
instead of in the line phase (sum_k D_ik f(U_k), dgsem_lobatto_2d.cu:431-439);
the line phase adds only the face terms.  Bars: one right-hand side within
5e-6 relative L2 of the FP64 oracle and within 10 % of the VALU form's own
error (measured on these states: VALU 3.4e-6 on the vortex, whose
collocation volume term is a difference of fluxes several times the RHS,
1.4e-6 on the Sod state; the matrix-core form the same to 1 %), 20 fused
steps within 1e-6; and within 2e-6 of the VALU form (on the vortex the two
differ by 0.9e-6 to 1.0e-6, each 3.4e-6 from the oracle) but not bit-equal
to it (the blocks sum in another order), which shows that the matrix-core
form ran.  Tiles of 3 elements (a partial wave and a ragged last tile), 8 (one
wave) and 16 (two waves), periodic vortex, and Sod with walls after 8
oracle steps (at rest, the Sod state's fluxes have one nonzero component
per direction and both forms give the same bits).
"""
import numpy as np
import pytest
import torch

from fvstuff_amd.capi import make_config, make_ic, vortex_config
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


def gpu(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def w2(cfg):
    _, w = lgl(cfg.p)
    return gll_weights_2d(w, cfg.K, cfg.L / cfg.K)


def with_fields(cfg, **kw):
    c = type(cfg).from_buffer_copy(cfg)
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def colloc_vortex(K, E, mfma):
    return vortex_config(K=K, p=3, precision="mixed", volume_flux="collocation", elems_per_block=E, mfma=mfma)


def colloc_walls(K, E, mfma):
    # circular Sod with reflecting walls, HLLC faces, no AV and no limiter
    return make_config(K=K, p=3, precision="mixed", volume_flux="collocation", av=False, limiter=False,
                       elems_per_block=E, mfma=mfma)


RHS_CASES = [
    ("vortex-E8", colloc_vortex(8, 8, 1), "vortex"),
    ("vortex-E16-two-waves", colloc_vortex(8, 16, 1), "vortex"),
    ("vortex-E3-ragged", colloc_vortex(7, 3, 1), "vortex"),
    ("walls-hllc-E8", colloc_walls(8, 8, 1), "sod"),
]


@pytest.mark.parametrize("name,cfg,ic", RHS_CASES, ids=[c[0] for c in RHS_CASES])
def test_mfma_collocation_rhs(dgsem, name, cfg, ic):
    o = Oracle(with_fields(cfg, precision=0, mfma=-1, elems_per_block=0))
    Q = o.init_state(make_ic(ic))
    if ic == "sod":
        assert o.advance(Q, 8, 0.3)[0] == 0     # the front moves: every flux component is active
    R_ref, _ = o.compute_rhs(Q)
    # no per-element wave speeds requested: the lean kernel runs
    R, _ = dgsem.Solver(cfg).compute_rhs(gpu(Q), want_lambda=False)
    R_valu, _ = dgsem.Solver(with_fields(cfg, mfma=0)).compute_rhs(gpu(Q), want_lambda=False)
    R, R_valu = cpu(R), cpu(R_valu)
    err = rel_l2(R, R_ref, w2(cfg))
    err_valu = rel_l2(R_valu, R_ref, w2(cfg))
    d = rel_l2(R, R_valu, w2(cfg))
    print(f"{name}: MFMA collocation RHS rel L2 vs oracle {err:.2e} (VALU form {err_valu:.2e}), "
          f"vs the VALU form {d:.2e}")
    assert err < 5e-6
    assert err < 1.1 * err_valu
    assert 0.0 < d < 2e-6


@pytest.mark.parametrize("E", [8, 16])
def test_mfma_collocation_fused_loop(dgsem, E):
    """20 fused steps of the vortex with the matrix-core form, against the
    FP64 oracle and the VALU form."""
    cfg = colloc_vortex(8, E, 1)
    o = Oracle(with_fields(cfg, precision=0, mfma=-1, elems_per_block=0))
    ic = make_ic("vortex")
    Q_ref = o.init_state(ic)
    Q0 = Q_ref.copy()
    rc, t_ref, _ = o.advance(Q_ref, 20, 0.4)
    assert rc == 0
    s = dgsem.Solver(cfg)
    Q = gpu(Q0)
    t, _ = s.advance(Q, 20, 0.4)
    Qv = gpu(Q0)
    tv, _ = dgsem.Solver(with_fields(cfg, mfma=0)).advance(Qv, 20, 0.4)
    err = rel_l2(cpu(Q), Q_ref, w2(cfg))
    d = rel_l2(cpu(Q), cpu(Qv), w2(cfg))
    print(f"E={E}: MFMA collocation state rel L2 vs oracle after 20 steps {err:.2e}, vs the VALU form {d:.2e}")
    assert t == pytest.approx(t_ref, rel=1e-6)
    assert err < 1e-6
    assert d < 1e-6


def test_mfma_collocation_autotune(dgsem):
    """With mfma = -1 the autotuner times both forms and keeps one; the
    result stays within the mixed bars either way."""
    cfg = colloc_vortex(16, 0, -1)
    s = dgsem.Solver(cfg)
    sched = s.autotune()
    o = Oracle(with_fields(cfg, precision=0, elems_per_block=0))
    ic = make_ic("vortex")
    Q_ref = o.init_state(ic)
    Q0 = Q_ref.copy()
    assert o.advance(Q_ref, 10, 0.4)[0] == 0
    Q = gpu(Q0)
    s.advance(Q, 10, 0.4)
    err = rel_l2(cpu(Q), Q_ref, w2(cfg))
    print(f"autotuned schedule {sched}: rel L2 vs oracle after 10 steps {err:.2e}")
    assert err < 1e-6
