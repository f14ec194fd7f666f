"""The compiled C++ host driver (fvstuff_amd/driver/dgsem_run.cpp), which
calls the C ABI the way the reference's main() calls its kernels
(dgsem_lobatto_2d.cu:1006-1058), against the FP64 oracle: the reference's
own configuration (circular Sod, walls, Chandrashekar + HLLC, AV, limiter)
issued as separate RHS / dt / stage calls and as the fused loop, and the
isentropic vortex."""
import json
import os
import subprocess

import numpy as np
import pytest

from fvstuff_amd.build import DRIVER
from fvstuff_amd.capi import make_config, make_ic, vortex_config
from oracle.oracle import Oracle, lgl
from tests.layout import gll_weights_2d, rel_l2

pytestmark = pytest.mark.gpu


def run_driver(tmp_path, **kw):
    out = tmp_path / f"state_{kw['case']}_{kw['mode']}_{kw['precision']}.bin"
    cmd = [DRIVER]
    for k, v in kw.items():
        cmd += ["--" + k.replace("_", "-"), str(v)]
    cmd += ["--out", str(out)]
    print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    info = json.loads(r.stdout.strip().splitlines()[-1])
    return info, np.fromfile(out, dtype=np.float64)


CASES = [
    ("sod", make_config(K=16, p=3), make_ic("sod"), 20, 0.6),
    ("vortex", vortex_config(K=16, p=3), make_ic("vortex"), 20, 0.5),
]


@pytest.mark.parametrize("mode", ["reference", "fused"])
@pytest.mark.parametrize("case,cfg,ic,steps,cfl", CASES, ids=[c[0] for c in CASES])
def test_driver_matches_oracle(tmp_path, mode, case, cfg, ic, steps, cfl):
    assert os.path.exists(DRIVER), "driver not built (python -m fvstuff_amd.build)"
    o = Oracle(cfg)
    Q_ref = o.init_state(ic)
    T0 = o.totals(Q_ref)
    rc, t_ref, dt_ref = o.advance(Q_ref, steps, cfl)
    assert rc == 0
    _, w = lgl(cfg.p)
    w2 = gll_weights_2d(w, cfg.K, cfg.L / cfg.K)
    for precision, tol in (("fp64", 1e-12), ("mixed", 1e-6)):
        info, Q = run_driver(tmp_path, case=case, K=cfg.K, p=cfg.p, steps=steps, cfl=cfl, precision=precision,
                             mode=mode)
        assert info["status"] == 0 and info["steps"] == steps
        assert info["t"] == pytest.approx(t_ref, rel=1e-12 if precision == "fp64" else 1e-6)
        err = rel_l2(Q, Q_ref, w2)
        print(f"driver {case} {mode} {precision}: rel L2 vs oracle {err:.2e}")
        assert err <= tol
        # mass and energy: walls and periodic boundaries carry none
        for q in (0, 3):
            assert info["totals"][q] == pytest.approx(T0[q], rel=1e-12 if precision == "fp64" else 1e-7)
