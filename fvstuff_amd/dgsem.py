"""Host-side mirror of the DGSEM step over the C ABI (include/dgsem.h).

The reference (dgsem_lobatto_2d.cu) drives its kernels from main(): build the
operators (:924-940), set the IC (:1006-1007), then per step compute the RHS
and wave speed (:1024-1026), pick dt (:1029), and apply the three SSP-RK3
stages with the positivity limiter (:1040-1051).  `Solver` exposes exactly
those operations (init_state, compute_rhs, cfl_dt, rk3_stage, limit) plus the
fused device-side stepping (advance / step).

Device memory comes from torch (float64 CUDA tensors, var-major layout
Q[var, el, jn, in] flattened); the numerics run in libdgsem.so.  There is no
CPU fallback: a missing library or a missing GPU raises.
"""
import ctypes as C
import math
import os

import torch

from .capi import DGConfig, DGICParams, STATUS_NAMES, make_config, make_ic  # noqa: F401

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdgsem.so")

_lib = None
_dp = C.POINTER(C.c_double)
_vp = C.c_void_p

# name -> (restype, argtypes); every function declared in include/dgsem.h
SIGNATURES = {
    "dg_config_default": (None, [C.POINTER(DGConfig)]),
    "dg_ic_default": (None, [C.POINTER(DGICParams), C.c_int]),
    "dg_create": (C.c_int, [C.POINTER(DGConfig), _vp, C.POINTER(_vp)]),
    "dg_destroy": (C.c_int, [_vp]),
    "dg_last_error": (C.c_char_p, [_vp]),
    "dg_status_string": (C.c_char_p, [C.c_int]),
    "dg_state_size": (C.c_longlong, [_vp]),
    "dg_num_elements": (C.c_longlong, [_vp]),
    "dg_operators": (C.c_int, [_vp, _dp, _dp, _dp]),
    "dg_sd_operators": (C.c_int, [_vp, _dp, _dp, _dp]),
    "dg_init_state": (C.c_int, [_vp, C.POINTER(DGICParams), _vp]),
    "dg_compute_rhs": (C.c_int, [_vp, _vp, _vp, _vp]),
    "dg_rk3_stage": (C.c_int, [_vp, C.c_int, _vp, _vp, _vp, _vp, C.c_double]),
    "dg_limit": (C.c_int, [_vp, _vp]),
    "dg_cfl_dt": (C.c_int, [_vp, _vp, C.c_double, _dp, _dp]),
    "dg_advance": (C.c_int, [_vp, _vp, C.c_int, C.c_double, C.c_double, _dp, _dp]),
    "dg_step": (C.c_int, [_vp, _vp, C.c_double, _dp]),
    "dg_steps_taken": (C.c_int, [_vp]),
    "dg_set_graph": (C.c_int, [_vp, C.c_int]),
    "dg_sensor": (C.c_int, [_vp, _vp, _vp]),
    "dg_totals": (C.c_int, [_vp, _vp, _dp]),
    "dg_save_state": (C.c_int, [_vp, _vp, C.c_double, C.c_longlong, C.c_char_p]),
    "dg_batch_save_state": (C.c_int, [_vp, C.c_int, _vp, C.c_double, C.c_longlong, C.c_char_p]),
    "dg_batch_load_state": (C.c_int, [_vp, C.c_int, _vp, _dp, C.POINTER(C.c_longlong), C.c_char_p]),
    "dg_load_state": (C.c_int, [_vp, _vp, _dp, C.POINTER(C.c_longlong), C.c_char_p]),
    "dg_totals_noalloc": (C.c_int, [C.POINTER(DGConfig), _vp, _vp, _vp]),
    "dg_vortex_exact": (C.c_int, [_vp, C.POINTER(DGICParams), C.c_double, C.c_int, _vp]),
    "dg_error_norms": (C.c_int, [_vp, _vp, _vp, _dp]),
    "dg_frame_fields": (C.c_int, [_vp, _vp, _vp, _vp]),
    "dg_autotune": (C.c_int, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "dg_batch_create": (C.c_int, [C.POINTER(DGConfig), C.c_int, _vp, C.POINTER(_vp)]),
    "dg_batch_destroy": (C.c_int, [_vp]),
    "dg_batch_last_error": (C.c_char_p, [_vp]),
    "dg_batch_state_size": (C.c_longlong, [_vp, C.c_int]),
    "dg_batch_init_state": (C.c_int, [_vp, C.POINTER(DGICParams), C.POINTER(_vp)]),
    "dg_batch_advance": (C.c_int, [_vp, C.POINTER(_vp), C.c_int, _dp, C.c_double, _dp, C.POINTER(C.c_int)]),
    "dg_batch_advance_each": (C.c_int, [_vp, C.POINTER(_vp), C.c_int, _dp, C.c_double, _dp, C.POINTER(C.c_int)]),
    "dg_batch_set_options": (C.c_int, [_vp, C.c_int, C.c_int]),
    "dg_batch_autotune": (C.c_int, [_vp, C.POINTER(_vp), _dp, C.POINTER(C.c_int)]),
    "dg_batch_set_persistent": (C.c_int, [_vp, C.c_int]),
    "dg_batch_set_xcd_bins": (C.c_int, [_vp, C.c_int]),
    "dg_batch_set_by_degree": (C.c_int, [_vp, C.c_int]),
    "dg_batch_get_by_degree": (C.c_int, [_vp, C.POINTER(C.c_int)]),
    "dg_batch_set_degree_lanes": (C.c_int, [_vp, C.c_int, C.c_int]),
    "dg_batch_get_degree_lanes": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_int)]),
    "dg_batch_get_options": (C.c_int, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                       C.POINTER(C.c_int)]),
}


class DGError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


def load_library(path=LIB_PATH):
    """Load libdgsem.so (built by fvstuff_amd.build); raises if it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built: run `python -m fvstuff_amd.build`")
        lib = C.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def _ptr(t):
    if t is None:
        return None
    if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()):
        raise ValueError("expected a contiguous float64 CUDA tensor")
    return C.c_void_p(t.data_ptr())


class Solver:
    """One DGSEM discretisation (dg_ctx) on one GPU."""

    def __init__(self, cfg, device="cuda:0", stream=None):
        if not torch.cuda.is_available():
            raise RuntimeError("fvstuff_amd.dgsem needs a ROCm GPU (torch.cuda.is_available() is False)")
        self.lib = load_library()
        self.cfg = cfg
        self.device = torch.device(device)
        torch.cuda.set_device(self.device)
        self.stream = stream if stream is not None else torch.cuda.current_stream(self.device)
        self._ctx = C.c_void_p()
        rc = self.lib.dg_create(C.byref(cfg), C.c_void_p(self.stream.cuda_stream), C.byref(self._ctx))
        if rc != 0:
            raise DGError(rc, "dg_create failed (bad config or device allocation)")
        self.K, self.p = cfg.K, cfg.p
        self.NN = cfg.p + 1
        self.NEL = cfg.K * cfg.K
        self.size = int(self.lib.dg_state_size(self._ctx))
        self.h = cfg.L / cfg.K

    def close(self):
        if getattr(self, "_ctx", None) is not None and self._ctx.value:
            self.lib.dg_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise DGError(rc, self.lib.dg_last_error(self._ctx).decode())

    # ── buffers ──
    def new_state(self):
        return torch.empty(self.size, dtype=torch.float64, device=self.device)

    def new_elem(self):
        return torch.empty(self.NEL, dtype=torch.float64, device=self.device)

    def operators(self):
        import numpy as np
        x = np.zeros(self.NN); w = np.zeros(self.NN); D = np.zeros(self.NN * self.NN)
        self._check(self.lib.dg_operators(self._ctx, x.ctypes.data_as(_dp), w.ctypes.data_as(_dp),
                                          D.ctypes.data_as(_dp)))
        return x, w, D.reshape(self.NN, self.NN)

    def sd_operators(self):
        """Spectral-difference tables of an SD context: flux nodes xf[p+2],
        Isf[p+2, p+1] and Dfs[p+1, p+2] (fr_lobatto_2d.cu:99-166)."""
        import numpy as np
        nf = self.NN + 1
        xf = np.zeros(nf); Isf = np.zeros(nf * self.NN); Dfs = np.zeros(self.NN * nf)
        self._check(self.lib.dg_sd_operators(self._ctx, xf.ctypes.data_as(_dp), Isf.ctypes.data_as(_dp),
                                             Dfs.ctypes.data_as(_dp)))
        return xf, Isf.reshape(nf, self.NN), Dfs.reshape(self.NN, nf)

    # ── reference operations ──
    def init_state(self, ic, Q=None):
        Q = self.new_state() if Q is None else Q
        self._check(self.lib.dg_init_state(self._ctx, C.byref(ic), _ptr(Q)))
        return Q

    def compute_rhs(self, Q, RHS=None, lam=None, want_lambda=True):
        RHS = self.new_state() if RHS is None else RHS
        if lam is None and want_lambda:
            lam = self.new_elem()
        self._check(self.lib.dg_compute_rhs(self._ctx, _ptr(Q), _ptr(RHS), _ptr(lam)))
        return RHS, lam

    def rk3_stage(self, stage, U0, Uprev, L, dt, out=None):
        out = self.new_state() if out is None else out
        self._check(self.lib.dg_rk3_stage(self._ctx, int(stage), _ptr(out), _ptr(U0), _ptr(Uprev), _ptr(L),
                                          float(dt)))
        return out

    def limit(self, Q):
        self._check(self.lib.dg_limit(self._ctx, _ptr(Q)))
        return Q

    def cfl_dt(self, lam, cfl):
        lm = C.c_double(); dt = C.c_double()
        self._check(self.lib.dg_cfl_dt(self._ctx, _ptr(lam), float(cfl), C.byref(lm), C.byref(dt)))
        return lm.value, dt.value

    # ── fused stepping ──
    def advance(self, Q, nsteps, cfl, t_end=0.0, t=0.0):
        """nsteps fused SSP-RK3 steps in place; returns (t, dt_last)."""
        tt = C.c_double(t); dt = C.c_double(0.0)
        self._check(self.lib.dg_advance(self._ctx, _ptr(Q), int(nsteps), float(cfl), float(t_end),
                                        C.byref(tt), C.byref(dt)))
        return tt.value, dt.value

    def set_graph(self, use_graph):
        self._check(self.lib.dg_set_graph(self._ctx, int(bool(use_graph))))

    def steps_taken(self):
        return int(self.lib.dg_steps_taken(self._ctx))

    def advance_to(self, Q, t_end, cfl, chunk=256, t=0.0, max_steps=10 ** 7, progress=None):
        """Advance until t_end in chunks of fused steps (one host sync per
        chunk); returns (t, steps taken).  progress(t, steps) is called after
        every chunk if given.  self.last_launched holds the steps launched,
        which include the no-op steps of the last chunk past t_end."""
        steps = 0
        self.last_launched = 0
        n = min(chunk, 8)            # a short first chunk measures dt
        while t < t_end - 1e-12 * max(1.0, abs(t_end)) and steps < max_steps:
            t, dt = self.advance(Q, n, cfl, t_end=t_end, t=t)
            self.last_launched += n
            steps += self.steps_taken()
            if progress is not None:
                progress(t, steps)
            # steps past t_end are launched but skipped on the device: size the
            # next chunk from the last dt (dt varies slowly), with a margin
            if dt > 0.0:
                n = int(min(chunk, max(2, math.ceil(1.02 * (t_end - t) / dt) + 2)))
        return t, steps

    def step(self, Q, cfl):
        dt = C.c_double(0.0)
        self._check(self.lib.dg_step(self._ctx, _ptr(Q), float(cfl), C.byref(dt)))
        return dt.value

    def reference_step(self, Q, cfl, scratch=None):
        """One step as the reference's loop issues it (:1024-1051): separate
        RHS, dt, stage and limiter calls.  Returns dt."""
        Q0, Q1, Q2, R, lam = scratch if scratch is not None else (
            self.new_state(), self.new_state(), self.new_state(), self.new_state(), self.new_elem())
        self.compute_rhs(Q, R, lam)
        _, dt = self.cfl_dt(lam, cfl)
        Q0.copy_(Q)
        self.rk3_stage(1, Q0, None, R, dt, out=Q1)
        self.compute_rhs(Q1, R, lam)
        self.rk3_stage(2, Q0, Q1, R, dt, out=Q2)
        self.compute_rhs(Q2, R, lam)
        self.rk3_stage(3, Q0, Q2, R, dt, out=Q)
        return dt

    # ── diagnostics ──
    def sensor(self, Q):
        s = self.new_elem()
        self._check(self.lib.dg_sensor(self._ctx, _ptr(Q), _ptr(s)))
        return s

    def save_state(self, path, Q, t, steps):
        """Checkpoint: state Q (device) with its time and step count."""
        self._check(self.lib.dg_save_state(self._ctx, _ptr(Q), float(t), int(steps), os.fsencode(path)))

    def load_state(self, path, Q=None):
        """Restart: the checkpoint's state into Q (a new state if None);
        returns (Q, t, steps)."""
        Q = self.new_state() if Q is None else Q
        t, n = C.c_double(), C.c_longlong()
        self._check(self.lib.dg_load_state(self._ctx, _ptr(Q), C.byref(t), C.byref(n), os.fsencode(path)))
        return Q, t.value, n.value

    def totals(self, Q):
        out = (C.c_double * 5)()
        self._check(self.lib.dg_totals(self._ctx, _ptr(Q), out))
        return list(out)

    def vortex_exact(self, ic, t, images=True, Q=None):
        """Exact advected vortex at time t (periodic images summed by default)."""
        Q = self.new_state() if Q is None else Q
        self._check(self.lib.dg_vortex_exact(self._ctx, C.byref(ic), float(t), int(bool(images)), _ptr(Q)))
        return Q

    def error_norms(self, Q, Qref):
        """GLL-quadrature norms of Q - Qref: dict with per-variable L2 and max
        errors and the relative L2 / Linf over all variables."""
        out = (C.c_double * 10)()
        self._check(self.lib.dg_error_norms(self._ctx, _ptr(Q), _ptr(Qref), out))
        v = list(out)
        return {"l2": v[0:4], "max": v[4:8], "rel_l2": v[8], "rel_linf": v[9]}

    def frame_fields(self, Q):
        """Per-element mean rho and p (device tensors of K*K), the data of a frame."""
        rho, p = self.new_elem(), self.new_elem()
        self._check(self.lib.dg_frame_fields(self._ctx, _ptr(Q), _ptr(rho), _ptr(p)))
        return rho, p

    def autotune(self):
        E = C.c_int(); X = C.c_int()
        self._check(self.lib.dg_autotune(self._ctx, C.byref(E), C.byref(X)))
        return E.value, X.value


class Batch:
    """Many independent solves (own K, p, domain, IC, CFL) stepped together:
    one launch per stage for all instances (one workgroup per tile), or one per
    degree present (set_by_degree), graph-replayed.  AV / limiter must be on
    in all instances or in none, and the scheme alike; with them, or with the
    spectral-difference scheme, the batch runs by degree only."""

    def __init__(self, cfgs, device="cuda:0", stream=None):
        if not torch.cuda.is_available():
            raise RuntimeError("fvstuff_amd.dgsem needs a ROCm GPU (torch.cuda.is_available() is False)")
        self.lib = load_library()
        self.n = len(cfgs)
        self.cfgs = list(cfgs)
        self.device = torch.device(device)
        torch.cuda.set_device(self.device)
        self.stream = stream if stream is not None else torch.cuda.current_stream(self.device)
        arr = (DGConfig * self.n)(*cfgs)
        self._b = C.c_void_p()
        rc = self.lib.dg_batch_create(arr, self.n, C.c_void_p(self.stream.cuda_stream), C.byref(self._b))
        if rc != 0:
            raise DGError(rc, "dg_batch_create failed (flux / boundary / precision / AV / limiter / scheme switches differ between instances, or allocation)")
        self.sizes = [int(self.lib.dg_batch_state_size(self._b, i)) for i in range(self.n)]
        self.dofs = sum(self.sizes)

    def close(self):
        if getattr(self, "_b", None) is not None and self._b.value:
            self.lib.dg_batch_destroy(self._b)
            self._b = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise DGError(rc, self.lib.dg_batch_last_error(self._b).decode())

    def _ptrs(self, Qs):
        if len(Qs) != self.n:
            raise ValueError("one state per instance")
        return (C.c_void_p * self.n)(*[_ptr(q).value for q in Qs])

    def new_states(self):
        return [torch.empty(n, dtype=torch.float64, device=self.device) for n in self.sizes]

    def init_states(self, ics, Qs=None):
        Qs = self.new_states() if Qs is None else Qs
        arr = (DGICParams * self.n)(*ics)
        self._check(self.lib.dg_batch_init_state(self._b, arr, self._ptrs(Qs)))
        return Qs

    def _advance(self, fn, Qs, nsteps, cfl, t_end, t):
        cfl = [float(cfl)] * self.n if isinstance(cfl, (int, float)) else [float(c) for c in cfl]
        c = (C.c_double * self.n)(*cfl)
        tt = (C.c_double * self.n)(*([0.0] * self.n if t is None else t))
        st = (C.c_int * self.n)()
        rc = fn(self._b, self._ptrs(Qs), int(nsteps), c, float(t_end), tt, st)
        self.last_status = list(st)
        self._check(rc)
        return list(tt)

    def advance(self, Qs, nsteps, cfl, t_end=0.0, t=None):
        """nsteps steps of every instance; returns the per-instance times."""
        return self._advance(self.lib.dg_batch_advance, Qs, nsteps, cfl, t_end, t)

    def advance_each(self, Qs, nsteps, cfl, t_end=0.0, t=None):
        """The same steps, one dg_advance per instance (for comparison)."""
        return self._advance(self.lib.dg_batch_advance_each, Qs, nsteps, cfl, t_end, t)

    def save_state(self, i, path, Q, t, steps):
        """Checkpoint of instance i: its state Q (device), time and step count."""
        self._check(self.lib.dg_batch_save_state(self._b, int(i), _ptr(Q), float(t), int(steps), os.fsencode(path)))

    def load_state(self, i, path, Q):
        """Restart of instance i from its checkpoint into Q; returns (t, steps)."""
        t, n = C.c_double(), C.c_longlong()
        self._check(self.lib.dg_batch_load_state(self._b, int(i), _ptr(Q), C.byref(t), C.byref(n), os.fsencode(path)))
        return t.value, n.value

    def set_options(self, use_graph=True, lanes=0):
        self._check(self.lib.dg_batch_set_options(self._b, int(bool(use_graph)), int(lanes)))

    def set_persistent(self, persistent):
        """Persistent form of the stage launches (resident workgroups walk the tiles)."""
        self._check(self.lib.dg_batch_set_persistent(self._b, int(bool(persistent))))

    def set_xcd_bins(self, xcd_bins):
        """Tile list in per-XCD bins of instances (True) or all tiles heaviest first."""
        self._check(self.lib.dg_batch_set_xcd_bins(self._b, int(bool(xcd_bins))))

    def set_by_degree(self, by_degree):
        """One launch per degree present per stage, each with the single-solve
        kernel of its degree (True), or the all-degree kernel (False)."""
        self._check(self.lib.dg_batch_set_by_degree(self._b, int(bool(by_degree))))

    def by_degree(self):
        v = C.c_int()
        self._check(self.lib.dg_batch_get_by_degree(self._b, C.byref(v)))
        return bool(v.value)

    def set_degree_lanes(self, p, lanes):
        """By-degree form: lanes per workgroup of degree p's launch (0 = common size)."""
        self._check(self.lib.dg_batch_set_degree_lanes(self._b, int(p), int(lanes)))

    def degree_lanes(self, p):
        v = C.c_int()
        self._check(self.lib.dg_batch_get_degree_lanes(self._b, int(p), C.byref(v)))
        return v.value

    def options(self):
        """(use_graph, lanes, persistent, xcd_bins) currently in use."""
        g, l, p, x = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        self._check(self.lib.dg_batch_get_options(self._b, C.byref(g), C.byref(l), C.byref(p), C.byref(x)))
        return bool(g.value), l.value, bool(p.value), bool(x.value)

    def autotune(self, Qs, cfl):
        """Time the workgroup sizes and orders in the launch forms (by degree
        only for a batch with AV / limiter or the SD scheme), round-robin over
        three rounds, then each degree's own size, and keep the fastest; Qs
        are advanced (9 steps per timing).  Returns the common lanes per
        workgroup (see options(), by_degree() and degree_lanes() for the
        schedule)."""
        cfl = [float(cfl)] * self.n if isinstance(cfl, (int, float)) else [float(c) for c in cfl]
        g = C.c_int()
        self._check(self.lib.dg_batch_autotune(self._b, self._ptrs(Qs), (C.c_double * self.n)(*cfl), C.byref(g)))
        return g.value
