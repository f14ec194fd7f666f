"""Pins the oracle's LGL nodes, GLL weights and differentiation matrix.

Reference values: dgsem_lobatto_2d.cu:145-157 (p <= 3 tables) and
fsbp/run_fsbp.py:14-15 (p = 3 literals).  For p > 3, which the reference does
not tabulate, the closed form pins them: the interior nodes are the roots of
P_p' (computed here independently with numpy's Legendre class), the weights are
2/(p(p+1)P_p(x)^2), the rule integrates degree 2p-1 exactly, and D is the exact
derivative on polynomials of degree <= p with the SBP property.
"""
import math

import numpy as np
import pytest
from numpy.polynomial import legendre as npleg

from oracle import oracle as O

P_ALL = list(range(1, 9))


def test_reference_tables_p1_to_p3():
    x1, w1 = O.lgl(1)
    np.testing.assert_array_equal(x1, [-1.0, 1.0]); np.testing.assert_array_equal(w1, [1.0, 1.0])
    x2, w2 = O.lgl(2)
    np.testing.assert_array_equal(x2, [-1.0, 0.0, 1.0])
    np.testing.assert_allclose(w2, [1 / 3, 4 / 3, 1 / 3], rtol=0, atol=1e-16)
    x3, w3 = O.lgl(3)
    # fsbp/run_fsbp.py:14-15
    np.testing.assert_allclose(x3, [-1.0, -1 / math.sqrt(5), 1 / math.sqrt(5), 1.0], rtol=0, atol=1e-16)
    np.testing.assert_allclose(w3, [1 / 6, 5 / 6, 5 / 6, 1 / 6], rtol=0, atol=1e-16)


@pytest.mark.parametrize("p", P_ALL)
def test_lgl_closed_form(p):
    x, w = O.lgl(p)
    interior = np.sort(npleg.Legendre.basis(p).deriv().roots().real)
    np.testing.assert_allclose(x[1:-1], interior, rtol=0, atol=2e-15)
    assert x[0] == -1.0 and x[-1] == 1.0
    Pp = npleg.Legendre.basis(p)(x)
    np.testing.assert_allclose(w, 2.0 / (p * (p + 1) * Pp ** 2), rtol=1e-14)
    assert abs(w.sum() - 2.0) < 1e-14
    for k in range(2 * p):   # exact to degree 2p-1
        exact = (1 - (-1) ** (k + 1)) / (k + 1)
        assert abs(np.dot(w, x ** k) - exact) < 1e-14, k
    # degree 2p is NOT integrated exactly by Lobatto (pins that it is Lobatto, not Gauss)
    exact = (1 - (-1) ** (2 * p + 1)) / (2 * p + 1)
    assert abs(np.dot(w, x ** (2 * p)) - exact) > 1e-6


@pytest.mark.parametrize("p", P_ALL)
def test_diff_matrix(p):
    x, w = O.lgl(p)
    D = O.diff_matrix(p)
    for k in range(p + 1):
        dk = k * x ** (k - 1) if k > 0 else np.zeros_like(x)
        np.testing.assert_allclose(D @ x ** k, dk, rtol=0, atol=1e-12 * (p + 1) ** 2)
    np.testing.assert_allclose(D.sum(axis=1), 0.0, atol=1e-13)
    # SBP: W D + D^T W = diag(-1, 0, ..., 0, 1)
    W = np.diag(w)
    B = np.zeros((p + 1, p + 1)); B[0, 0] = -1.0; B[-1, -1] = 1.0
    np.testing.assert_allclose(W @ D + D.T @ W, B, atol=1e-13 * (p + 1))
    assert abs(D[0, 0] + p * (p + 1) / 4.0) < 1e-12
    assert abs(D[-1, -1] - p * (p + 1) / 4.0) < 1e-12
