"""Pins the oracle's pointwise physics (dgsem_lobatto_2d.cu:184-355 and the
closed-form Ranocha / Rusanov fluxes) by properties that fail on a dropped
term, a flipped sign or a swapped argument."""
import math

import numpy as np
import pytest

from oracle import oracle as O

G = 1.4
RANOCHA, CHANDRA = 1, 0
HLLC, RUSANOV = 0, 1
rng = np.random.default_rng(7)


def rand_prim(n):
    r = rng.uniform(0.2, 3.0, n); u = rng.uniform(-2, 2, n); v = rng.uniform(-2, 2, n)
    p = rng.uniform(0.2, 4.0, n)
    return np.stack([r, u, v, p], 1)


def cons(W):
    r, u, v, p = W
    return np.array([r, r * u, r * v, p / (G - 1) + 0.5 * r * (u * u + v * v)])


def entropy_vars(W):
    # keep_fv_euler_2d.cu:123-131
    r, u, v, p = W
    s = math.log(p / r ** G)
    return np.array([(G - s) / (G - 1) - r * (u * u + v * v) / (2 * p), r * u / p, r * v / p, -r / p])


def logmean_ld(a, b):
    # (a-b)/ln(a/b) with ln(a/b) = log1p((a-b)/b): a-b is exact, no cancellation
    a = np.longdouble(a); b = np.longdouble(b)
    return float((a - b) / np.log1p((a - b) / b))


@pytest.mark.parametrize("ratio", [1.0 + 1e-9, 1.0 + 1e-4, 1.005, 1.03, 1.5, 7.0, 1e3])
def test_logmean_both_branches(ratio):
    for a in [1e-3, 0.7, 2.0, 50.0]:
        b = a * ratio
        ref = logmean_ld(a, b) if ratio != 1.0 else a
        assert abs(O.logmean(a, b) - ref) <= 4e-15 * ref
        assert abs(O.logmean(b, a) - ref) <= 4e-15 * ref
    assert O.logmean(1.3, 1.3) == 1.3


@pytest.mark.parametrize("kind", [RANOCHA, CHANDRA])
def test_ec_flux_consistency_symmetry(kind):
    for W in rand_prim(20):
        for d, n in [(0, (1, 0)), (1, (0, 1))]:
            F = O.volume_flux(kind, W, W, d)
            np.testing.assert_allclose(F, O.physical_flux(W, *n), rtol=1e-13, atol=1e-13)
    for WL, WR in zip(rand_prim(20), rand_prim(20)):
        for d in (0, 1):
            np.testing.assert_allclose(O.volume_flux(kind, WL, WR, d), O.volume_flux(kind, WR, WL, d),
                                       rtol=1e-14, atol=1e-14)


@pytest.mark.parametrize("kind", [RANOCHA, CHANDRA])
def test_ec_flux_tadmor_identity(kind):
    """(v_R - v_L) . F# = psi_R - psi_L with psi = rho u_n (entropy potential)."""
    for WL, WR in zip(rand_prim(50), rand_prim(50)):
        for d in (0, 1):
            F = O.volume_flux(kind, WL, WR, d)
            lhs = np.dot(entropy_vars(WR) - entropy_vars(WL), F)
            psi = lambda W: W[0] * W[1 + d]
            rhs = psi(WR) - psi(WL)
            assert abs(lhs - rhs) < 1e-11 * (1 + abs(rhs)), (lhs, rhs)


def test_ranocha_kinetic_energy_preserving():
    """f_rhou - u_avg f_rho = p_avg (Jameson's KEP condition)."""
    for WL, WR in zip(rand_prim(30), rand_prim(30)):
        F = O.volume_flux(RANOCHA, WL, WR, 0)
        uav = 0.5 * (WL[1] + WR[1]); pav = 0.5 * (WL[3] + WR[3])
        assert abs(F[1] - uav * F[0] - pav) < 1e-13 * (1 + abs(pav))
        G_ = O.volume_flux(RANOCHA, WL, WR, 1)
        vav = 0.5 * (WL[2] + WR[2])
        assert abs(G_[2] - vav * G_[0] - pav) < 1e-13 * (1 + abs(pav))


@pytest.mark.parametrize("kind", [HLLC, RUSANOV])
def test_riemann_consistency(kind):
    for W in rand_prim(20):
        for n in [(1.0, 0.0), (0.0, 1.0)]:
            np.testing.assert_allclose(O.surface_flux(kind, W, W, *n), O.physical_flux(W, *n),
                                       rtol=1e-13, atol=1e-13)


def test_hllc_supersonic_upwinding():
    # :306-309: SL >= 0 -> left flux, SR <= 0 -> right flux
    WL = np.array([1.0, 5.0, 0.3, 1.0]); WR = np.array([0.5, 4.0, -0.2, 0.8])
    np.testing.assert_array_equal(O.surface_flux(HLLC, WL, WR, 1.0, 0.0), O.physical_flux(WL, 1.0, 0.0))
    WL2 = np.array([1.0, -5.0, 0.3, 1.0]); WR2 = np.array([0.5, -4.0, -0.2, 0.8])
    np.testing.assert_array_equal(O.surface_flux(HLLC, WL2, WR2, 1.0, 0.0), O.physical_flux(WR2, 1.0, 0.0))
    # y-normal: rotation of the x case
    WLy = np.array([1.0, 0.3, 5.0, 1.0]); WRy = np.array([0.5, -0.2, 4.0, 0.8])
    np.testing.assert_array_equal(O.surface_flux(HLLC, WLy, WRy, 0.0, 1.0), O.physical_flux(WLy, 0.0, 1.0))


def test_hllc_rotation_invariance():
    for WL, WR in zip(rand_prim(20), rand_prim(20)):
        Fx = O.surface_flux(HLLC, WL, WR, 1.0, 0.0)
        sw = lambda W: np.array([W[0], W[2], W[1], W[3]])
        Gy = O.surface_flux(HLLC, sw(WL), sw(WR), 0.0, 1.0)
        np.testing.assert_allclose(Gy, [Fx[0], Fx[2], Fx[1], Fx[3]], rtol=1e-13, atol=1e-13)


def test_rusanov_brute_force():
    for WL, WR in zip(rand_prim(20), rand_prim(20)):
        for nx, ny in [(1.0, 0.0), (0.0, 1.0)]:
            fl = O.physical_flux(WL, nx, ny); fr = O.physical_flux(WR, nx, ny)
            cl = math.sqrt(G * WL[3] / WL[0]); cr = math.sqrt(G * WR[3] / WR[0])
            lam = max(abs(WL[1] * nx + WL[2] * ny) + cl, abs(WR[1] * nx + WR[2] * ny) + cr)
            ref = 0.5 * (fl + fr) - 0.5 * lam * (cons(WR) - cons(WL))
            np.testing.assert_allclose(O.surface_flux(RUSANOV, WL, WR, nx, ny), ref, rtol=1e-13, atol=1e-13)


def test_cons_to_prim_sanitize():
    W = rand_prim(1)[0]
    np.testing.assert_allclose(O.cons_to_prim(cons(W)), W, rtol=1e-14)
    # negative pressure floored (:215-221), density floored, velocity clamped (:347-355)
    U = np.array([1.0, 3.0, 0.0, 1.0])      # KE 4.5 > E
    assert O.cons_to_prim(U)[3] == pytest.approx(1e-14, rel=1e-6)
    assert O.cons_to_prim(np.array([-1.0, 0.0, 0.0, 1.0]))[0] == 1e-14
    Wv = O.cons_to_prim(np.array([1e-6, 1.0, 1.0, 1.0]))
    assert Wv[1] == 0.0 and Wv[2] == 0.0     # rho < 1e-5: velocity zeroed
    assert O.cons_to_prim(np.array([1.0, 2e4, -3e4, 1e10]))[1:3].tolist() == [1e4, -1e4]


def test_vortex_closed_form():
    # q1_dg_euler_2d.cu:2166-2188 at the centre and far away
    W = O.vortex_exact(5.0, 5.0)
    T = 1 - (G - 1) * 25 / (8 * G * math.pi ** 2) * math.e
    np.testing.assert_allclose(W, [T ** (1 / (G - 1)), 1.0, 1.0, T ** (G / (G - 1))], rtol=1e-14)
    W = O.vortex_exact(5.0 + 1.0, 5.0)
    f = 5 / (2 * math.pi)
    assert abs(W[2] - (1.0 + f * 1.0)) < 1e-14 and abs(W[1] - 1.0) < 1e-14
    # isentropic: p = rho^gamma everywhere
    for x, y in rng.uniform(0, 10, (10, 2)):
        W = O.vortex_exact(x, y)
        assert abs(W[3] - W[0] ** G) < 1e-14


def test_kh_noise_range_and_determinism():
    vals = np.array([O.kh_noise(1234, n, 1e-3) for n in range(2000)])
    assert vals.min() >= -1e-3 and vals.max() < 1e-3
    assert abs(vals.mean()) < 1e-4 and vals.std() > 5e-4
    assert O.kh_noise(1234, 17, 1e-3) == O.kh_noise(1234, 17, 1e-3)
    assert O.kh_noise(1235, 17, 1e-3) != O.kh_noise(1234, 17, 1e-3)
