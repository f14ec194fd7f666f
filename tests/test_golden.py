"""Committed oracle vectors (tests/golden/dgsem_golden.npz, written by
tests/golden/make_golden.py): the oracle must keep reproducing them, and the
GPU path must match them."""
import os

import numpy as np
import pytest

from tests.golden.make_golden import CASES
from oracle.oracle import Oracle, lgl, diff_matrix
from tests.layout import gll_weights_2d, rel_l2

HERE = os.path.dirname(os.path.abspath(__file__))
G = np.load(os.path.join(HERE, "golden", "dgsem_golden.npz"))


@pytest.mark.parametrize("p", range(1, 8))
def test_oracle_operators_match_golden(p):
    x, w = lgl(p)
    np.testing.assert_array_equal(x, G[f"lgl/x{p}"])
    np.testing.assert_array_equal(w, G[f"lgl/w{p}"])
    np.testing.assert_array_equal(diff_matrix(p), G[f"lgl/D{p}"])


@pytest.mark.parametrize("p", range(1, 8))
def test_oracle_sd_operators_match_golden(p):
    from fvstuff_amd.capi import sd_config
    xf, Isf, Dfs = Oracle(sd_config(K=1, p=p)).sd_operators()
    np.testing.assert_array_equal(xf, G[f"sd/xf{p}"])
    np.testing.assert_array_equal(Isf, G[f"sd/Isf{p}"])
    np.testing.assert_array_equal(Dfs, G[f"sd/Dfs{p}"])


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_matches_golden(name):
    cfg, ic, steps, cfl = CASES[name]
    o = Oracle(cfg)
    Q = o.init_state(ic)
    np.testing.assert_allclose(Q, G[f"{name}/Q0"], rtol=1e-14, atol=1e-15)
    R, lam = o.compute_rhs(G[f"{name}/Q0"].copy())
    np.testing.assert_allclose(R, G[f"{name}/rhs0"], rtol=1e-12, atol=1e-12 * np.abs(R).max())
    np.testing.assert_allclose(lam, G[f"{name}/lam0"], rtol=1e-14)
    Qn = G[f"{name}/Q0"].copy()
    rc, t, dt = o.advance(Qn, steps, cfl)
    assert rc == 0
    x, w = lgl(cfg.p)
    assert rel_l2(Qn, G[f"{name}/Qn"], gll_weights_2d(w, cfg.K, cfg.L / cfg.K)) < 1e-13
    np.testing.assert_allclose([t, dt], G[f"{name}/t_dt"], rtol=1e-13)


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(CASES))
def test_gpu_matches_golden(name):
    import torch
    from fvstuff_amd import dgsem as D
    cfg, ic, steps, cfl = CASES[name]
    s = D.Solver(cfg)
    Q0 = G[f"{name}/Q0"]
    Qi = s.init_state(ic)
    torch.cuda.synchronize()
    np.testing.assert_allclose(Qi.cpu().numpy(), Q0, rtol=2e-15, atol=1e-15)
    Q = torch.from_numpy(Q0.copy()).cuda()
    R, lam = s.compute_rhs(Q)
    torch.cuda.synchronize()
    x, w = lgl(cfg.p)
    w2 = gll_weights_2d(w, cfg.K, cfg.L / cfg.K)
    assert rel_l2(R.cpu().numpy(), G[f"{name}/rhs0"], w2) < 1e-13
    np.testing.assert_allclose(lam.cpu().numpy(), G[f"{name}/lam0"], rtol=1e-15)
    t, dt = s.advance(Q, steps, cfl)
    torch.cuda.synchronize()
    assert rel_l2(Q.cpu().numpy(), G[f"{name}/Qn"], w2) < 1e-12
    np.testing.assert_allclose([t, dt], G[f"{name}/t_dt"], rtol=1e-12)
