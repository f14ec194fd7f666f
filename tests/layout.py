"""Helpers for the var-major state layout Q[var][el][jn][in] (el = ej*K + ei),
dgsem_lobatto_2d.cu:126, shared by the CPU and GPU tests."""
import numpy as np


def as5(Q, K, p):
    """View a flat state as (var, ej, ei, jn, in)."""
    n = p + 1
    return np.asarray(Q).reshape(4, K, K, n, n)


def node_xy(K, p, x0, y0, L, xi):
    h = L / K
    n = p + 1
    ei = np.arange(K)[None, :, None, None]
    ej = np.arange(K)[:, None, None, None]
    x = x0 + h * (ei + 0.5 * (xi[None, None, None, :] + 1.0)) + 0 * ej
    y = y0 + h * (ej + 0.5 * (xi[None, None, :, None] + 1.0)) + 0 * ei
    return np.broadcast_to(x, (K, K, n, n)), np.broadcast_to(y, (K, K, n, n))


def prim_to_cons(r, u, v, p, gamma=1.4):
    return np.stack([r, r * u, r * v, p / (gamma - 1) + 0.5 * r * (u * u + v * v)])


def cons_to_prim(U, gamma=1.4):
    r = U[0]; u = U[1] / r; v = U[2] / r
    p = (gamma - 1) * (U[3] - 0.5 * (U[1] ** 2 + U[2] ** 2) / r)
    return r, u, v, p


def gll_weights_2d(w, K, h):
    """Quadrature weights (h/2)^2 w_i w_j broadcast to (ej, ei, jn, in)."""
    ww = 0.25 * h * h * np.outer(w, w)
    return np.broadcast_to(ww, (K, K) + ww.shape)


def rel_l2(A, B, w2d):
    """GLL-weighted relative L2 over all 4 variables."""
    A = np.asarray(A).reshape((4,) + w2d.shape); B = np.asarray(B).reshape((4,) + w2d.shape)
    num = np.sqrt(np.sum(w2d[None] * (A - B) ** 2))
    den = np.sqrt(np.sum(w2d[None] * B ** 2))
    return num / den


def entropy_vars(U, gamma=1.4):
    r, u, v, p = cons_to_prim(U, gamma)
    s = np.log(p) - gamma * np.log(r)
    return np.stack([(gamma - s) / (gamma - 1) - r * (u * u + v * v) / (2 * p), r * u / p, r * v / p, -r / p])
