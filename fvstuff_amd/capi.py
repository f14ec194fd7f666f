"""ctypes mirror of include/dgsem.h (structs, enums, status codes).

Only type definitions live here; fvstuff_amd.dgsem loads the HIP library.
"""
import ctypes as C

DG_OK = 0
DG_ERR_INVALID_ARG = 1
DG_ERR_HIP = 2
DG_ERR_NONFINITE = 3
DG_ERR_DT_UNDERFLOW = 4
DG_ERR_BLOWUP = 5

STATUS_NAMES = {
    DG_OK: "ok",
    DG_ERR_INVALID_ARG: "invalid argument",
    DG_ERR_HIP: "HIP runtime error",
    DG_ERR_NONFINITE: "non-finite wave speed",
    DG_ERR_DT_UNDERFLOW: "dt underflow",
    DG_ERR_BLOWUP: "blow-up",
}

VOL_CHANDRASHEKAR, VOL_RANOCHA, VOL_COLLOCATION = 0, 1, 2
SURF_HLLC, SURF_RUSANOV = 0, 1
BC_WALL, BC_PERIODIC = 0, 1
FP64, MIXED = 0, 1
IC_CIRCULAR_SOD, IC_VORTEX, IC_KELVIN_HELMHOLTZ = 0, 1, 2
SCHEME_DGSEM, SCHEME_SD = 0, 1

_NAMES = {
    "volume_flux": {"chandrashekar": 0, "ranocha": 1, "collocation": 2},
    "surface_flux": {"hllc": 0, "rusanov": 1},
    "bc": {"wall": 0, "periodic": 1},
    "precision": {"fp64": 0, "mixed": 1},
    "scheme": {"dgsem": 0, "sd": 1},
}


class DGConfig(C.Structure):
    _fields_ = [
        ("K", C.c_int), ("p", C.c_int),
        ("x0", C.c_double), ("y0", C.c_double), ("L", C.c_double),
        ("gamma", C.c_double),
        ("volume_flux", C.c_int), ("surface_flux", C.c_int), ("bc", C.c_int),
        ("av", C.c_int), ("limiter", C.c_int), ("precision", C.c_int),
        ("av_cav", C.c_double), ("av_ksensor", C.c_double),
        ("eps_rho", C.c_double), ("eps_p", C.c_double),
        ("elems_per_block", C.c_int), ("xcd_remap", C.c_int),
        ("scheme", C.c_int), ("mfma", C.c_int),
    ]


class DGICParams(C.Structure):
    _fields_ = [
        ("kind", C.c_int),
        ("rho_hi", C.c_double), ("p_hi", C.c_double),
        ("vortex_eps", C.c_double), ("xc", C.c_double), ("yc", C.c_double),
        ("u0", C.c_double), ("v0", C.c_double),
        ("seed", C.c_ulonglong), ("noise_amp", C.c_double),
    ]


def make_config(K=50, p=3, x0=0.0, y0=0.0, L=1.0, gamma=1.4,
                volume_flux="chandrashekar", surface_flux="hllc", bc="wall",
                av=True, limiter=True, precision="fp64", av_cav=0.25,
                av_ksensor=0.1, eps_rho=1e-6, eps_p=1e-6,
                elems_per_block=0, xcd_remap=-1, scheme="dgsem", mfma=-1):
    """A dg_config; defaults are the reference's compile-time defaults
    (dgsem_lobatto_2d.cu:61-126, Makefile DGSEM_ORDER=3, main :915-922)."""
    def enum(field, v):
        return _NAMES[field][v] if isinstance(v, str) else int(v)
    return DGConfig(int(K), int(p), float(x0), float(y0), float(L), float(gamma),
                    enum("volume_flux", volume_flux), enum("surface_flux", surface_flux),
                    enum("bc", bc), int(bool(av)), int(bool(limiter)),
                    enum("precision", precision), float(av_cav), float(av_ksensor),
                    float(eps_rho), float(eps_p), int(elems_per_block), int(xcd_remap),
                    enum("scheme", scheme), int(mfma))


def sd_config(K=50, p=3, precision="fp64", **kw):
    """The spectral-difference sibling (fr_lobatto_2d.cu) with its driver's
    defaults: circular Sod on [0,1]^2, reflecting walls, HLLC faces, nodal
    positivity repair, no artificial viscosity (main :835-838, CFL 0.1/p)."""
    args = dict(K=K, p=p, x0=0.0, y0=0.0, L=1.0, surface_flux="hllc", bc="wall",
                av=False, limiter=True, precision=precision, scheme="sd")
    args.update(kw)
    return make_config(**args)


def make_ic(kind="vortex", rho_hi=11.0, p_hi=10.0, vortex_eps=5.0, xc=5.0, yc=5.0,
            u0=1.0, v0=1.0, seed=1234, noise_amp=1e-3):
    kinds = {"sod": IC_CIRCULAR_SOD, "vortex": IC_VORTEX, "kh": IC_KELVIN_HELMHOLTZ}
    k = kinds[kind] if isinstance(kind, str) else int(kind)
    return DGICParams(k, float(rho_hi), float(p_hi), float(vortex_eps), float(xc), float(yc),
                      float(u0), float(v0), int(seed), float(noise_amp))


def vortex_config(K=32, p=3, precision="fp64", **kw):
    """Isentropic vortex on periodic [0,10]^2 with the Ranocha EC volume flux
    and Rusanov faces, no AV / limiter (they are inactive on smooth data)."""
    args = dict(K=K, p=p, x0=0.0, y0=0.0, L=10.0, volume_flux="ranocha",
                surface_flux="rusanov", bc="periodic", av=False, limiter=False,
                precision=precision)
    args.update(kw)
    return make_config(**args)


def kh_config(K=64, p=3, precision="mixed", **kw):
    """Kelvin-Helmholtz step shear layer on periodic [-1,1]^2."""
    args = dict(K=K, p=p, x0=-1.0, y0=-1.0, L=2.0, volume_flux="ranocha",
                surface_flux="rusanov", bc="periodic", av=False, limiter=False,
                precision=precision)
    args.update(kw)
    return make_config(**args)


def sod_ensemble(n=256, seed=2024, precision="mixed", k_range=(16, 128), degrees=(3, 5, 7),
                 rho_range=(2.0, 11.0), p_range=(2.0, 10.0)):
    """Seed-drawn instances of the reference's default configuration
    (dgsem_lobatto_2d.cu:61-126, :915-922: circular Sod on the unit square
    with walls, Chandrashekar volume flux, HLLC faces, Ducros AV, Zhang-Shu
    limiter): K ~ U{k_range}, p ~ U{degrees}, inner density ~ U[rho_range]
    and pressure ~ U[p_range] (outer state 1, 1).  Returns (configs, ics,
    cfls).  The reference's CFL 0.6 holds at p = 3 only: in the FP64 oracle,
    K = 16 Sod blows up within 300 steps at CFL 0.6 for p = 5 and at 0.4 for
    p = 7; 0.5 (p <= 3), 0.3 (p = 5) and 0.2 (p = 7) run 500 steps at
    K = 16 and 32 for both ends of the ratio ranges."""
    import numpy as np
    rng = np.random.default_rng(seed)
    cfgs, ics, cfls = [], [], []
    for _ in range(n):
        K = int(rng.integers(k_range[0], k_range[1] + 1))
        p = int(rng.choice(degrees))
        rho_hi = float(rng.uniform(*rho_range))
        p_hi = float(rng.uniform(*p_range))
        cfgs.append(make_config(K=K, p=p, precision=precision))
        ics.append(make_ic("sod", rho_hi=rho_hi, p_hi=p_hi))
        cfls.append(0.5 if p <= 3 else (0.3 if p <= 5 else 0.2))
    return cfgs, ics, cfls


def ensemble(n=256, seed=2024, precision="mixed", k_range=(16, 128), degrees=(3, 5, 7),
             eps_range=(1.0, 5.0), mach_range=(0.1, 0.8), gamma=1.4):
    """Seed-drawn isentropic-vortex instances: K ~ U{k_range}, p ~ U{degrees},
    strength ~ U[eps_range], background Mach ~ U[mach_range] along the diagonal
    (|(u0,v0)| = M*sqrt(gamma) for the unit background sound speed).
    Returns (configs, ics, cfls); the CFL follows the degree's SSP-RK3 limit
    with a margin.  The dt formula's wave speed |u|+|v|+c is least
    conservative at low Mach (the 2-D limit scales with (|u|+c)+(|v|+c)): a
    weak, slow vortex grows an unstable mode within a few hundred steps at
    CFL 0.5 (p=3), 0.3 (p=5) and 0.25 (p=7) in the FP64 oracle; 0.4, 0.25 and
    0.2 stay stable over 1500 steps."""
    import numpy as np
    rng = np.random.default_rng(seed)
    cfgs, ics, cfls = [], [], []
    for _ in range(n):
        K = int(rng.integers(k_range[0], k_range[1] + 1))
        p = int(rng.choice(degrees))
        eps = float(rng.uniform(*eps_range))
        M = float(rng.uniform(*mach_range))
        speed = M * gamma ** 0.5 / 2 ** 0.5
        cfgs.append(vortex_config(K=K, p=p, precision=precision))
        ics.append(make_ic("vortex", vortex_eps=eps, u0=speed, v0=speed))
        cfls.append(0.4 if p <= 3 else (0.25 if p <= 5 else 0.2))
    return cfgs, ics, cfls
