"""Closed-form pins of the oracle's shock sensor (compute_sensor,
dgsem_lobatto_2d.cu:662-689) and artificial viscosity (:483-521).

State: rho = 1, p = 1, v = 0 and u = a*x (a uniform compression for a < 0),
which the LGL operators differentiate exactly.  Then, in every element,
div u = a, c^2 = gamma, theta = a^2 / (a^2 + k c^2 / len^2 + 1e-30) with
len = h/(2p+1), the element wave speed is max|u| + sqrt(gamma) and the
element viscosity nu = c_av * theta * len * lam.

The AV term of one line is -nu jac^2 W^-1 D^T W D U (jac = 2/h).  With the
summation-by-parts identity W D + D^T W = B = diag(-1, 0, ..., 0, 1) it is
-nu jac^2 (W^-1 B U_xi - U_xixi), exact for polynomial U: for rho*u = a x
(linear) only the end nodes see it, for E = p/(g-1) + rho u^2/2 (quadratic)
every node does.  A wrong sign, scale, transpose or weight in either AV pass
fails these."""
import math

import numpy as np
import pytest

from fvstuff_amd.capi import make_config
from oracle.oracle import Oracle, lgl
from tests.layout import as5, node_xy, prim_to_cons

GAMMA, KSENS, CAV = 1.4, 0.1, 0.25


def linear_compression(K, p, a):
    x, _ = lgl(p)
    X, _ = node_xy(K, p, 0.0, 0.0, 1.0, x)
    r = np.ones_like(X)
    U = prim_to_cons(r, a * X, 0 * X, r, GAMMA)
    return np.ascontiguousarray(U.reshape(-1))


def theta_closed(K, p, a):
    h = 1.0 / K
    ln = h / (2 * p + 1)
    return a * a / (a * a + KSENS * GAMMA / (ln * ln) + 1e-30)


@pytest.mark.parametrize("p", [1, 2, 3, 5])
@pytest.mark.parametrize("a", [-0.7, 0.4])
def test_sensor_closed_form(p, a):
    """compute_sensor does not clip the divergence: expansion and
    compression of the same rate give the same theta."""
    K = 5
    o = Oracle(make_config(K=K, p=p, av=True, limiter=False))
    th = o.sensor(linear_compression(K, p, a))
    np.testing.assert_allclose(th, theta_closed(K, p, a), rtol=1e-10)


@pytest.mark.parametrize("p", [2, 3, 5])
def test_av_term_closed_form(p):
    K, a = 4, -0.7
    h = 1.0 / K
    jac = 2.0 / h
    ln = h / (2 * p + 1)
    on = Oracle(make_config(K=K, p=p, av=True, limiter=False))
    off = Oracle(make_config(K=K, p=p, av=False, limiter=False))
    Q = linear_compression(K, p, a)
    R_on, lam_on = on.compute_rhs(Q)
    R_off, _ = off.compute_rhs(Q)
    av = as5(R_on - R_off, K, p)                       # (var, ej, ei, jn, in)
    _, w = lgl(p)
    th = theta_closed(K, p, a)
    for ei in range(K):
        xl, xr = ei * h, (ei + 1) * h
        lam = abs(a) * max(abs(xl), abs(xr)) + math.sqrt(GAMMA)
        np.testing.assert_allclose(lam_on.reshape(K, K)[:, ei], lam, rtol=1e-13)
        nu = CAV * th * ln * lam
        # rho*u = a x: U_xi = a h/2, U_xixi = 0
        m = np.zeros(p + 1)
        m[0] = -(a * h / 2) / w[0]
        m[p] = (a * h / 2) / w[p]
        # E = 1/(g-1) + a^2 x^2 / 2: U_xi = a^2 x h/2, U_xixi = a^2 h^2/4
        e = np.full(p + 1, -a * a * h * h / 4)
        e[0] += -(a * a * xl * h / 2) / w[0]
        e[p] += (a * a * xr * h / 2) / w[p]
        scale = nu * jac * jac
        for ej in range(K):
            for jn in range(p + 1):
                np.testing.assert_allclose(av[1, ej, ei, jn], -scale * m, rtol=1e-9, atol=1e-12 * scale)
                np.testing.assert_allclose(av[3, ej, ei, jn], -scale * e, rtol=1e-9, atol=1e-12 * scale)
    np.testing.assert_allclose(av[0], 0.0, atol=1e-12)   # rho is constant
    np.testing.assert_allclose(av[2], 0.0, atol=1e-12)   # rho*v = 0
