"""Failure reporting of the fused loop (dg_advance): the blow-up check of
the reference's loop (dgsem_lobatto_2d.cu:1052-1057) must cover the last
step taken, also when the steps after it are no-ops because t_end was
reached."""
import numpy as np
import pytest
import torch

from fvstuff_amd.capi import make_ic, vortex_config
from tests.layout import as5

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dgsem():
    from fvstuff_amd import dgsem as D
    D.load_library()
    return D


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("use_graph", [True, False])
def test_blowup_found_when_trailing_steps_are_noops(dgsem, precision, use_graph):
    """E = 1e13 at one node: the first step's wave speed (~2e6) is computed
    without the blow-up check (the loop's first step only requires it to be
    positive), and t_end is exactly that step's dt, so the remaining four
    steps are no-ops.  The status must be BLOWUP (5) with t = t_end."""
    cfg = vortex_config(K=8, p=3, precision=precision)
    s = dgsem.Solver(cfg)
    s.set_graph(use_graph)
    Q = s.init_state(make_ic("vortex"))
    Q.view(4, 8, 8, 4, 4)[3, 2, 2, 1, 1] = 1e13
    _, lam = s.compute_rhs(Q)
    lam_max, dt = s.cfl_dt(lam, 0.5)
    assert lam_max > 1e6
    with pytest.raises(dgsem.DGError) as e:
        s.advance(Q, 5, 0.5, t_end=dt)
    assert e.value.status == 5
    assert s.steps_taken() == 1


def test_no_false_blowup_after_end_time(dgsem):
    """A healthy run that reaches t_end early and idles for the rest of its
    steps returns DG_OK with the steps actually taken."""
    cfg = vortex_config(K=8, p=3)
    s = dgsem.Solver(cfg)
    Q = s.init_state(make_ic("vortex"))
    for use_graph in (True, False):
        s.set_graph(use_graph)
        Q = s.init_state(make_ic("vortex"), Q)
        t, dt = s.advance(Q, 50, 0.5, t_end=0.05)
        assert t == pytest.approx(0.05, abs=1e-13)
        assert 0 < s.steps_taken() < 50
