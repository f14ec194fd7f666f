"""The C-ABI library loads and exports every function include/dgsem.h
declares, and the ctypes mirror matches the C struct layout.  No device calls."""
import ctypes as C
import os
import re
import subprocess

import pytest

from fvstuff_amd import capi
from fvstuff_amd.dgsem import LIB_PATH, SIGNATURES, load_library

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "dgsem.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(dg_[a-z_0-9]+)\s*\(", src)))


def test_header_and_wrapper_agree():
    assert declared_functions() == sorted(SIGNATURES)


def test_library_exports_every_symbol():
    assert os.path.exists(LIB_PATH), "libdgsem.so not built (python -m fvstuff_amd.build)"
    lib = load_library()
    for name in declared_functions():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", LIB_PATH], capture_output=True, text=True).stdout
    for name in declared_functions():
        assert re.search(rf"\bT {name}$", out, re.M), name


def test_per_degree_kernels_distinct():
    """Every kernel template compiled in the per-degree objects carries the
    degree among its template arguments, so the link keeps one instantiation
    per degree (a template without it is instantiated with the same symbol in
    all seven objects, and only one of them survives the link)."""
    out = subprocess.run(["nm", "-C", LIB_PATH], capture_output=True, text=True).stdout
    for kern in ("stage_kernel", "batch_degree_kernel", "limit_kernel", "sensor_kernel"):
        degrees = {int(m) for m in re.findall(rf"dg::{kern}<(\d+)\b", out)}
        assert degrees == set(range(1, 8)), (kern, sorted(degrees))


def test_host_only_calls():
    lib = load_library()
    cfg = capi.DGConfig()
    lib.dg_config_default(C.byref(cfg))
    ref = capi.make_config()
    for f, _ in capi.DGConfig._fields_:
        assert getattr(cfg, f) == getattr(ref, f), f
    assert lib.dg_status_string(5).decode().startswith("blow-up")
    ic = capi.DGICParams()
    lib.dg_ic_default(C.byref(ic), 1)
    assert ic.kind == 1 and ic.vortex_eps == 5.0 and ic.xc == 5.0 and ic.u0 == 1.0
    # checkpoint calls reject null arguments before touching a device
    assert lib.dg_save_state(None, None, 0.0, 0, b"x") == capi.DG_ERR_INVALID_ARG
    assert lib.dg_load_state(None, None, None, None, b"x") == capi.DG_ERR_INVALID_ARG
    assert lib.dg_batch_save_state(None, 0, None, 0.0, 0, b"x") == capi.DG_ERR_INVALID_ARG
    assert lib.dg_batch_load_state(None, 0, None, None, None, b"x") == capi.DG_ERR_INVALID_ARG


def test_struct_layout_matches_header(tmp_path):
    prog = tmp_path / "layout.c"
    fields_cfg = [f for f, _ in capi.DGConfig._fields_]
    fields_ic = [f for f, _ in capi.DGICParams._fields_]
    body = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', 'int main(void){']
    body.append('printf("%zu %zu\\n", sizeof(dg_config), sizeof(dg_ic_params));')
    for f in fields_cfg:
        body.append(f'printf("%zu\\n", offsetof(dg_config, {f}));')
    for f in fields_ic:
        body.append(f'printf("%zu\\n", offsetof(dg_ic_params, {f}));')
    body.append("return 0;}")
    prog.write_text("\n".join(body))
    exe = tmp_path / "layout"
    subprocess.check_call(["gcc", "-o", str(exe), str(prog)])
    vals = subprocess.check_output([str(exe)]).decode().split()
    assert int(vals[0]) == C.sizeof(capi.DGConfig) and int(vals[1]) == C.sizeof(capi.DGICParams)
    offs = [int(v) for v in vals[2:]]
    exp = [getattr(capi.DGConfig, f).offset for f in fields_cfg] + [getattr(capi.DGICParams, f).offset for f in fields_ic]
    assert offs == exp


def test_scheme_validation_before_any_device_call():
    """dg_create / dg_batch_create reject invalid configurations (empty grid,
    degree 0, bad domain or gas, AV or an unknown scheme with spectral
    difference, a batch that mixes spectral difference and DGSEM) during
    validation, before any HIP call."""
    lib = load_library()
    ctx = C.c_void_p()
    for bad in (capi.sd_config(K=4, p=3, av=True), capi.make_config(K=4, p=3, scheme=2),
                capi.sd_config(K=4, p=8), capi.make_config(K=0, p=3), capi.make_config(K=4, p=0),
                capi.make_config(K=4, p=3, L=0.0), capi.make_config(K=4, p=3, gamma=1.0),
                capi.make_config(K=4, p=3, mfma=2), capi.make_config(K=4, p=3, mfma=-2)):
        assert lib.dg_create(C.byref(bad), None, C.byref(ctx)) == capi.DG_ERR_INVALID_ARG
        assert not ctx.value
    arr = (capi.DGConfig * 2)(capi.sd_config(K=4, p=3, limiter=False), capi.make_config(K=4, p=3, limiter=False))
    b = C.c_void_p()
    assert lib.dg_batch_create(arr, 2, None, C.byref(b)) == capi.DG_ERR_INVALID_ARG
    assert lib.dg_sd_operators(None, None, None, None) == capi.DG_ERR_INVALID_ARG


def test_driver_built_and_rejects_bad_arguments():
    """The C++ host driver links the library; bad arguments are rejected
    before any device call."""
    from fvstuff_amd.build import DRIVER
    assert os.path.exists(DRIVER), "driver not built (python -m fvstuff_amd.build)"
    r = subprocess.run([DRIVER, "--case", "nope"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "usage" in r.stderr
