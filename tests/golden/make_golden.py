#!/usr/bin/env python3
"""Writes tests/golden/dgsem_golden.npz from the FP64 oracle.

The reference ships no vectors for its DGSEM path (it has no tests); these
fixtures freeze the oracle's answers on small cases so that the GPU tests can
check against committed data and any change to the oracle is caught.
Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from fvstuff_amd.capi import make_config, make_ic, vortex_config, kh_config, sd_config  # noqa: E402
from oracle.oracle import Oracle, lgl, diff_matrix  # noqa: E402

# name -> (config, ic, steps, cfl)
CASES = {
    "vortex_k10_p3": (vortex_config(K=10, p=3), make_ic("vortex"), 50, 0.5),
    "sod_reference_k8_p3": (make_config(K=8, p=3), make_ic("sod"), 20, 0.6),
    "kh_k8_p3": (kh_config(K=8, p=3, precision="fp64"), make_ic("kh"), 20, 0.5),
    "kh_av_limiter_k8_p4": (kh_config(K=8, p=4, precision="fp64", volume_flux="chandrashekar",
                                      surface_flux="hllc", av=True, limiter=True), make_ic("kh"), 20, 0.3),
    "vortex_k4_p7": (vortex_config(K=4, p=7), make_ic("vortex"), 10, 0.25),
    "sod_collocation_k6_p2": (make_config(K=6, p=2, volume_flux="collocation", surface_flux="rusanov"),
                              make_ic("sod"), 10, 0.4),
    # the spectral-difference sibling (fr_lobatto_2d.cu): its driver's Sod and a periodic vortex
    "sd_sod_k8_p3": (sd_config(K=8, p=3), make_ic("sod"), 20, 0.1 / 3),
    "sd_vortex_k5_p5": (sd_config(K=5, p=5, L=10.0, bc="periodic", surface_flux="rusanov", limiter=False),
                        make_ic("vortex"), 10, 0.2),
}


def case_arrays(name):
    cfg, ic, steps, cfl = CASES[name]
    o = Oracle(cfg)
    Q = o.init_state(ic)
    R, lam = o.compute_rhs(Q)
    Q1 = Q.copy()
    rc, t, dt = o.advance(Q1, steps, cfl)
    assert rc == 0, (name, rc)
    return {f"{name}/Q0": Q, f"{name}/rhs0": R, f"{name}/lam0": lam, f"{name}/Qn": Q1,
            f"{name}/t_dt": np.array([t, dt])}


def main():
    out = {}
    for p in range(1, 8):
        x, w = lgl(p)
        out[f"lgl/x{p}"] = x; out[f"lgl/w{p}"] = w; out[f"lgl/D{p}"] = diff_matrix(p)
        xf, Isf, Dfs = Oracle(sd_config(K=1, p=p)).sd_operators()
        out[f"sd/xf{p}"] = xf; out[f"sd/Isf{p}"] = Isf; out[f"sd/Dfs{p}"] = Dfs
    for name in CASES:
        out.update(case_arrays(name))
    path = sys.argv[1] if len(sys.argv) > 1 else os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                                               "dgsem_golden.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
