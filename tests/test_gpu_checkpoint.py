"""Checkpoint / restart (dg_save_state / dg_load_state): the state comes back
bit for bit with its time and step count, a resumed run takes the same steps
as one that continues with the same state in a new call, and a file of
another discretisation, a truncated file or a file that is no checkpoint is
refused.  The compiled driver's --checkpoint / --restart run the reference's
separate calls across a restart with the bits of an uninterrupted run.
Batched instances checkpoint and restart the same way."""
import json
import os
import subprocess

import numpy as np
import pytest
import torch

from fvstuff_amd.build import DRIVER
from fvstuff_amd.capi import make_config, make_ic, vortex_config

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def D():
    from fvstuff_amd import dgsem
    dgsem.load_library()
    return dgsem


CASES = [
    ("vortex_p5", lambda prec: vortex_config(K=8, p=5, precision=prec), "vortex", 0.3),
    ("sod_p3_av_limiter", lambda prec: make_config(K=8, p=3, precision=prec), "sod", 0.5),
]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("name,mk,ic,cfl", CASES, ids=[c[0] for c in CASES])
def test_restart_equals_continuing(D, tmp_path, precision, name, mk, ic, cfl):
    cfg = mk(precision)
    path = str(tmp_path / f"{name}_{precision}.ck")
    s = D.Solver(cfg)
    Q = s.init_state(make_ic(ic))
    t4, _ = s.advance(Q, 4, cfl)
    s.save_state(path, Q, t4, 4)
    saved = Q.clone()
    tA, _ = s.advance(Q, 6, cfl, t=t4)
    torch.cuda.synchronize()
    assert os.path.getsize(path) == 128 + 8 * Q.numel()
    s2 = D.Solver(cfg)
    Q2, t2, n2 = s2.load_state(path)
    assert (t2, n2) == (t4, 4)
    assert torch.equal(Q2, saved)
    tB, _ = s2.advance(Q2, 6, cfl, t=t2)
    torch.cuda.synchronize()
    assert tB == tA
    assert torch.equal(Q2, Q)


def test_load_refuses_other_files(D, tmp_path):
    cfg = vortex_config(K=6, p=3)
    s = D.Solver(cfg)
    Q = s.init_state(make_ic("vortex"))
    path = tmp_path / "a.ck"
    s.save_state(str(path), Q, 0.25, 7)
    # flux, precision and shock capturing may differ: they act on the state
    other = D.Solver(vortex_config(K=6, p=3, precision="mixed", volume_flux="chandrashekar", surface_flux="hllc"))
    Q2, t, n = other.load_state(str(path))
    assert (t, n) == (0.25, 7) and torch.equal(Q2, Q)
    for bad_cfg in (vortex_config(K=7, p=3), vortex_config(K=6, p=4), make_config(K=6, p=3)):
        with pytest.raises(D.DGError, match="another discretisation"):
            D.Solver(bad_cfg).load_state(str(path))
    raw = path.read_bytes()
    cases = {"short state": raw[:-8], "trailing bytes": raw + b"\0", "not a checkpoint": b"X" + raw[1:],
             "short header": raw[:100]}
    for why, data in cases.items():
        p = tmp_path / "bad.ck"
        p.write_bytes(data)
        with pytest.raises(D.DGError, match=why):
            s.load_state(str(p))
    with pytest.raises(D.DGError, match="cannot open"):
        s.load_state(str(tmp_path / "missing.ck"))


def _drive(*args):
    cmd = [DRIVER, *map(str, args)]
    print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout.strip().splitlines()[-1])


def test_driver_restart(tmp_path):
    """The reference's call sequence (separate RHS, dt and stage calls, Sod
    with walls, AV and the limiter), 10 steps in one run or 4 + 6 across a
    checkpoint: the same state bit for bit, the same time, 10 steps."""
    assert os.path.exists(DRIVER), "driver not built (python -m fvstuff_amd.build)"
    common = ["--case", "sod", "--K", 8, "--p", 3, "--precision", "fp64", "--mode", "reference"]
    a, b, ck = tmp_path / "a.bin", tmp_path / "b.bin", tmp_path / "s.ck"
    full = _drive(*common, "--steps", 10, "--out", a)
    first = _drive(*common, "--steps", 4, "--checkpoint", ck)
    second = _drive(*common, "--steps", 6, "--restart", ck, "--out", b)
    assert full["status"] == first["status"] == second["status"] == 0
    assert first["total_steps"] == 4 and second["total_steps"] == 10 and second["steps"] == 6
    assert second["t"] == full["t"]
    assert np.array_equal(np.fromfile(a, dtype=np.float64), np.fromfile(b, dtype=np.float64))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_restart_equals_continuing(D, tmp_path, precision):
    """A batch of vortex instances (degrees 1..7, their own dt) advanced 6
    steps, checkpointed instance by instance, loaded into a new batch and
    advanced 6 more steps from the loaded times, against the same batch
    continuing with 6 more steps in a new call: the same bits and times.
    An out-of-range instance and a file of another discretisation are
    refused."""
    from fvstuff_amd.capi import ensemble
    cfgs, ics, cfls = ensemble(n=7, seed=4, precision=precision, k_range=(4, 12), degrees=(1, 2, 3, 4, 5, 6, 7))
    b = D.Batch(cfgs)
    Qs = b.init_states(ics)
    t0 = b.advance(Qs, 6, cfls)
    for i, Q in enumerate(Qs):
        b.save_state(i, str(tmp_path / f"inst{i}.ck"), Q, t0[i], 6)
    t_cont = b.advance(Qs, 6, cfls, t=t0)
    cont = [q.cpu().numpy() for q in Qs]
    b2 = D.Batch(cfgs)
    Q2 = b2.new_states()
    t_load = []
    for i in range(len(cfgs)):
        t, n = b2.load_state(i, str(tmp_path / f"inst{i}.ck"), Q2[i])
        assert t == t0[i] and n == 6
        t_load.append(t)
    t_res = b2.advance(Q2, 6, cfls, t=t_load)
    torch.cuda.synchronize()
    assert t_res == t_cont
    for i, q in enumerate(Q2):
        assert np.array_equal(q.cpu().numpy(), cont[i]), i
    with pytest.raises(D.DGError):
        b2.save_state(len(cfgs), str(tmp_path / "x.ck"), Q2[0], 0.0, 0)
    other = next(j for j, c in enumerate(cfgs) if (c.K, c.p) != (cfgs[1].K, cfgs[1].p))
    with pytest.raises(D.DGError):
        b2.load_state(1, str(tmp_path / f"inst{other}.ck"), Q2[1])
