"""dg_totals_noalloc, the allocation-free entry point (no context; the
per-block partial sums go to the library's fixed device array), against
dg_totals through the C ABI: the same kernels, so the same bits, on
consecutive calls whose second state is smaller (fewer partial blocks: the
partials the first call left behind must not leak into the second sum)."""
import ctypes as C

import numpy as np
import pytest
import torch

from fvstuff_amd.capi import DG_ERR_INVALID_ARG, kh_config, make_ic, vortex_config

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dgsem():
    from fvstuff_amd import dgsem as D
    D.load_library()
    return D


def noalloc_totals(lib, cfg, Q, out, stream):
    return lib.dg_totals_noalloc(C.byref(cfg), C.c_void_p(Q.data_ptr()), C.c_void_p(out.data_ptr()),
                                 C.c_void_p(stream.cuda_stream))


def test_noalloc_totals_match_dg_totals(dgsem):
    lib = dgsem.load_library()
    stream = torch.cuda.current_stream()
    out = torch.empty(5, dtype=torch.float64, device="cuda")
    cases = [
        # 4.2M nodes: the grid is at its cap of 8192 blocks
        (vortex_config(K=256, p=7), make_ic("vortex")),
        # then 2,304 nodes: 9 blocks
        (kh_config(K=12, p=3, precision="fp64"), make_ic("kh")),
    ]
    for cfg, ic in cases:
        s = dgsem.Solver(cfg)
        Q = s.init_state(ic)
        ref = np.array(s.totals(Q))
        out.fill_(float("nan"))
        assert noalloc_totals(lib, cfg, Q, out, stream) == 0
        got = out.cpu().numpy()
        print(f"K={cfg.K} p={cfg.p}: {got}")
        assert np.array_equal(got, ref), (got, ref)
        del s, Q


def test_noalloc_totals_rejects_bad_arguments(dgsem):
    lib = dgsem.load_library()
    stream = torch.cuda.current_stream()
    out = torch.empty(5, dtype=torch.float64, device="cuda")
    Q = torch.zeros(4 * 16 * 16, dtype=torch.float64, device="cuda")
    for bad in (vortex_config(K=2, p=9), vortex_config(K=0, p=3), vortex_config(K=2, p=3, gamma=1.0),
                vortex_config(K=8193, p=3)):
        assert noalloc_totals(lib, bad, Q, out, stream) == DG_ERR_INVALID_ARG
    cfg = vortex_config(K=2, p=3)
    assert lib.dg_totals_noalloc(C.byref(cfg), None, C.c_void_p(out.data_ptr()), None) == DG_ERR_INVALID_ARG
    assert lib.dg_totals_noalloc(C.byref(cfg), C.c_void_p(Q.data_ptr()), None, None) == DG_ERR_INVALID_ARG
    assert lib.dg_totals_noalloc(None, C.c_void_p(Q.data_ptr()), C.c_void_p(out.data_ptr()), None) == \
        DG_ERR_INVALID_ARG
