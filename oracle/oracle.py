"""ctypes binding of the CPU FP64 oracle (liboracle_dgsem.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the fvstuff_amd package.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from fvstuff_amd.capi import DGConfig, DGICParams, make_config, make_ic  # noqa: F401

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_dgsem.so")
_lib = None

_dp = C.POINTER(C.c_double)


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.oracle_create.argtypes = [C.POINTER(DGConfig), C.POINTER(C.c_void_p)]
        L.oracle_destroy.argtypes = [C.c_void_p]
        L.oracle_set_threads.argtypes = [C.c_int]
        L.oracle_lgl.argtypes = [C.c_int, _dp, _dp]
        L.oracle_build_D.argtypes = [C.c_int, _dp]
        L.oracle_logmean.argtypes = [C.c_double, C.c_double]
        L.oracle_logmean.restype = C.c_double
        L.oracle_volume_flux.argtypes = [C.c_int, C.c_double, _dp, _dp, C.c_int, _dp]
        L.oracle_surface_flux.argtypes = [C.c_int, C.c_double, _dp, _dp, C.c_double, C.c_double, _dp]
        L.oracle_cons_to_prim.argtypes = [C.c_double, _dp, _dp]
        L.oracle_physical_flux.argtypes = [C.c_double, _dp, C.c_double, C.c_double, _dp]
        L.oracle_vortex_exact.argtypes = [C.c_double] * 8 + [_dp]
        L.oracle_kh_noise.argtypes = [C.c_ulonglong, C.c_longlong, C.c_double]
        L.oracle_kh_noise.restype = C.c_double
        L.oracle_init_state.argtypes = [C.c_void_p, C.POINTER(DGICParams), _dp]
        L.oracle_compute_rhs.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.oracle_rk3_stage.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp, _dp, C.c_double]
        L.oracle_limit.argtypes = [C.c_void_p, _dp]
        L.oracle_max_lambda.argtypes = [C.c_void_p, _dp]
        L.oracle_max_lambda.restype = C.c_double
        L.oracle_advance.argtypes = [C.c_void_p, _dp, C.c_int, C.c_double, C.c_double, _dp, _dp]
        L.oracle_sensor.argtypes = [C.c_void_p, _dp, _dp]
        L.oracle_totals.argtypes = [C.c_void_p, _dp, _dp]
        L.oracle_frame_fields.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.oracle_sd_operators.argtypes = [C.c_void_p, _dp, _dp, _dp]
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def set_threads(n):
    return lib().oracle_set_threads(int(n))


def lgl(p):
    x = np.zeros(p + 1); w = np.zeros(p + 1)
    lib().oracle_lgl(p, _p(x), _p(w))
    return x, w


def diff_matrix(p):
    D = np.zeros((p + 1) * (p + 1))
    lib().oracle_build_D(p, _p(D))
    return D.reshape(p + 1, p + 1)


def logmean(a, b):
    return lib().oracle_logmean(float(a), float(b))


def volume_flux(kind, WL, WR, direction, gamma=1.4):
    F = np.zeros(4)
    lib().oracle_volume_flux(int(kind), gamma, _p(np.ascontiguousarray(WL, dtype=np.float64)),
                             _p(np.ascontiguousarray(WR, dtype=np.float64)), int(direction), _p(F))
    return F


def surface_flux(kind, WL, WR, nx, ny, gamma=1.4):
    F = np.zeros(4)
    lib().oracle_surface_flux(int(kind), gamma, _p(np.ascontiguousarray(WL, dtype=np.float64)),
                              _p(np.ascontiguousarray(WR, dtype=np.float64)), nx, ny, _p(F))
    return F


def physical_flux(W, nx, ny, gamma=1.4):
    F = np.zeros(4)
    lib().oracle_physical_flux(gamma, _p(np.ascontiguousarray(W, dtype=np.float64)), nx, ny, _p(F))
    return F


def cons_to_prim(U, gamma=1.4):
    W = np.zeros(4)
    lib().oracle_cons_to_prim(gamma, _p(np.ascontiguousarray(U, dtype=np.float64)), _p(W))
    return W


def vortex_exact(x, y, eps=5.0, xc=5.0, yc=5.0, u0=1.0, v0=1.0, gamma=1.4):
    W = np.zeros(4)
    lib().oracle_vortex_exact(gamma, eps, xc, yc, u0, v0, float(x), float(y), _p(W))
    return W


def kh_noise(seed, node, amp):
    return lib().oracle_kh_noise(int(seed), int(node), float(amp))


class Oracle:
    """FP64 CPU restatement of the reference's solver on one dg_config."""

    def __init__(self, cfg):
        self.cfg = cfg
        self._ctx = C.c_void_p()
        rc = lib().oracle_create(C.byref(cfg), C.byref(self._ctx))
        if rc != 0:
            raise ValueError(f"oracle_create failed ({rc})")
        self.K, self.p = cfg.K, cfg.p
        self.NN = cfg.p + 1
        self.NEL = cfg.K * cfg.K
        self.size = 4 * self.NEL * self.NN * self.NN
        self.h = cfg.L / cfg.K

    def __del__(self):
        if getattr(self, "_ctx", None) and self._ctx.value:
            lib().oracle_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def new_state(self):
        return np.zeros(self.size)

    def init_state(self, ic):
        Q = self.new_state()
        lib().oracle_init_state(self._ctx, C.byref(ic), _p(Q))
        return Q

    def compute_rhs(self, Q):
        R = np.zeros(self.size); lam = np.zeros(self.NEL)
        lib().oracle_compute_rhs(self._ctx, _p(Q), _p(R), _p(lam))
        return R, lam

    def rk3_stage(self, stage, U0, Uprev, L, dt):
        out = np.zeros(self.size)
        lib().oracle_rk3_stage(self._ctx, int(stage), _p(out), _p(U0),
                               _p(Uprev if Uprev is not None else U0), _p(L), float(dt))
        return out

    def limit(self, Q):
        Q = Q.copy()
        lib().oracle_limit(self._ctx, _p(Q))
        return Q

    def max_lambda(self, lam):
        return lib().oracle_max_lambda(self._ctx, _p(lam))

    def advance(self, Q, nsteps, cfl, t_end=0.0, t=0.0):
        """Returns (status, t, dt_last); Q is updated in place."""
        tt = C.c_double(t); dt = C.c_double(0.0)
        rc = lib().oracle_advance(self._ctx, _p(Q), int(nsteps), float(cfl), float(t_end),
                                  C.byref(tt), C.byref(dt))
        return rc, tt.value, dt.value

    def sensor(self, Q):
        s = np.zeros(self.NEL)
        lib().oracle_sensor(self._ctx, _p(Q), _p(s))
        return s

    def frame_fields(self, Q):
        rho = np.zeros(self.NEL); p = np.zeros(self.NEL)
        lib().oracle_frame_fields(self._ctx, _p(Q), _p(rho), _p(p))
        return rho, p

    def sd_operators(self):
        """Spectral-difference tables: xf[p+2], Isf[p+2, p+1], Dfs[p+1, p+2]."""
        nf, nn = self.NN + 1, self.NN
        xf = np.zeros(nf); Isf = np.zeros(nf * nn); Dfs = np.zeros(nn * nf)
        lib().oracle_sd_operators(self._ctx, _p(xf), _p(Isf), _p(Dfs))
        return xf, Isf.reshape(nf, nn), Dfs.reshape(nn, nf)

    def totals(self, Q):
        out = np.zeros(5)
        lib().oracle_totals(self._ctx, _p(Q), _p(out))
        return out
