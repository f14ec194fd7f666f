"""Pins of the oracle's spectral-difference sibling (fr_lobatto_2d.cu): the
flux-node operators (:76-166), the SD right-hand side (:332-525), the nodal
positivity repair (:589-615), its IC (:633-657) and step loop (:937-990),
checked against closed forms and properties the discretisation must have.

The reference's own construction of the derivative matrix (:120-166) is not
finite in its end rows (a solution node coincides with a flux node there:
0 * (1/0)); the oracle builds the matrix the reference defines, the
derivative of the flux nodes' Lagrange basis at the solution nodes, which
these tests pin by polynomial exactness.  test_reference_end_rows_singular
restates the construction as written to show where it breaks."""
import math

import numpy as np
import pytest

from fvstuff_amd.capi import make_ic, sd_config
from oracle import oracle as O
from oracle.oracle import Oracle, lgl
from tests.layout import as5, node_xy, prim_to_cons, gll_weights_2d, rel_l2

PS = [1, 2, 3, 4, 5, 6, 7]
GL_LITERALS = {1: [0.0], 2: [-0.5773502691896258, 0.5773502691896258],
               3: [-0.7745966692414834, 0.0, 0.7745966692414834]}   # fr_lobatto_2d.cu:84-89


def ops(p):
    return Oracle(sd_config(K=2, p=p)).sd_operators()


def uniform_state(o, r=1.3, u=0.4, v=-0.2, p=0.9):
    n = o.NN
    U = prim_to_cons(*(np.full((o.K, o.K, n, n), a) for a in (r, u, v, p)))
    return np.ascontiguousarray(U.reshape(-1))


@pytest.mark.parametrize("p", PS)
def test_flux_nodes(p):
    xf, _, _ = ops(p)
    assert xf[0] == -1.0 and xf[-1] == 1.0 and len(xf) == p + 2
    np.testing.assert_allclose(xf[1:-1], np.polynomial.legendre.leggauss(p)[0], rtol=0, atol=2e-16)
    if p in GL_LITERALS:
        assert list(xf[1:-1]) == GL_LITERALS[p]
    np.testing.assert_array_equal(xf, -xf[::-1])


@pytest.mark.parametrize("p", PS)
def test_interpolation_exact_to_degree_p(p):
    xs, _ = lgl(p)
    xf, Isf, _ = ops(p)
    np.testing.assert_allclose(Isf.sum(1), 1.0, rtol=0, atol=1e-14)
    for d in range(p + 1):
        np.testing.assert_allclose(Isf @ xs ** d, xf ** d, rtol=0, atol=5e-14)
    # the end rows pick the end solution nodes (they coincide with +-1)
    np.testing.assert_allclose(Isf[0], np.eye(p + 1)[0], atol=1e-15)
    np.testing.assert_allclose(Isf[-1], np.eye(p + 1)[-1], atol=1e-15)


@pytest.mark.parametrize("p", PS)
def test_derivative_exact_to_degree_p_plus_1(p):
    """Dfs differentiates every polynomial of degree <= p+1 (the flux
    polynomial through p+2 flux nodes) exactly at the solution nodes."""
    xs, _ = lgl(p)
    xf, _, Dfs = ops(p)
    assert np.all(np.isfinite(Dfs))
    for d in range(p + 2):
        ref = d * xs ** (d - 1) if d > 0 else np.zeros_like(xs)
        np.testing.assert_allclose(Dfs @ xf ** d, ref, rtol=0, atol=2e-12 * max(1, d * d))


@pytest.mark.parametrize("p", PS)
def test_derivative_telescopes(p):
    """sum_i w_i Dfs[i, k] = [k = last] - [k = first]: the LGL quadrature of
    the derivative of a degree-(p+1) polynomial is exact, which makes the
    scheme conservative (the face fluxes are all that survive)."""
    _, w = lgl(p)
    _, _, Dfs = ops(p)
    e = np.zeros(p + 2); e[0] = -1.0; e[-1] = 1.0
    np.testing.assert_allclose(w @ Dfs, e, rtol=0, atol=5e-14)


@pytest.mark.parametrize("p", [1, 2, 3])
def test_reference_end_rows_singular(p):
    """The construction of fr_lobatto_2d.cu:120-166 as written: finite and
    equal to the oracle in the interior rows, NaN (0 * inf) in the end rows,
    and the coincident entry with the sign flipped."""
    xs, _ = lgl(p)
    xf, _, Dfs = ops(p)
    nf = p + 2
    wf = np.array([1.0 / np.prod([xf[k] - xf[m] for m in range(nf) if m != k]) for k in range(nf)])

    def lag(k, x):
        return np.prod([(x - xf[m]) / (xf[k] - xf[m]) for m in range(nf) if m != k])

    A = np.zeros((p + 1, nf))
    with np.errstate(divide="ignore", invalid="ignore"):
        for i in range(p + 1):
            for k in range(nf):
                if abs(xs[i] - xf[k]) < 1e-14:
                    A[i, k] = sum((wf[m] / wf[k]) / (xf[k] - xf[m]) for m in range(nf) if m != k)
                else:
                    A[i, k] = lag(k, xs[i]) * sum(np.float64(1.0) / np.float64(xs[i] - xf[m])
                                                  for m in range(nf) if m != k)
    np.testing.assert_allclose(A[1:-1], Dfs[1:-1], rtol=1e-13, atol=1e-14)
    assert np.isnan(A[0, 1:]).all() and np.isnan(A[-1, :-1]).all()
    assert A[0, 0] == pytest.approx(-Dfs[0, 0], rel=1e-14)
    assert A[-1, -1] == pytest.approx(-Dfs[-1, -1], rel=1e-14)


@pytest.mark.parametrize("surf", ["hllc", "rusanov"])
@pytest.mark.parametrize("p", [1, 3, 4, 7])
def test_free_stream_periodic(surf, p):
    o = Oracle(sd_config(K=4, p=p, bc="periodic", surface_flux=surf))
    R, lam = o.compute_rhs(uniform_state(o))
    assert np.abs(R).max() < 1e-12
    assert np.allclose(lam, 0.6 + math.sqrt(1.4 * 0.9 / 1.3), rtol=1e-14)


@pytest.mark.parametrize("surf", ["hllc", "rusanov"])
def test_free_stream_wall_at_rest(surf):
    o = Oracle(sd_config(K=5, p=3, surface_flux=surf))
    R, _ = o.compute_rhs(uniform_state(o, u=0.0, v=0.0))
    assert np.abs(R).max() < 1e-12


@pytest.mark.parametrize("surf", ["hllc", "rusanov"])
@pytest.mark.parametrize("p", [2, 3, 5])
def test_discrete_conservation_periodic(surf, p):
    cfg = sd_config(K=6, p=p, x0=0.0, y0=0.0, L=10.0, bc="periodic", surface_flux=surf)
    o = Oracle(cfg)
    Q = o.init_state(make_ic("vortex"))
    R, _ = o.compute_rhs(Q)
    _, w = lgl(p)
    w2 = gll_weights_2d(w, o.K, o.h)
    R5 = as5(R, o.K, p)
    for q in range(4):
        tot = np.sum(w2 * R5[q])
        assert abs(tot) < 1e-12 * np.sum(w2 * np.abs(R5[q])) + 1e-14, (q, tot)


def test_wall_box_conserves_mass_and_energy():
    """Reflecting walls carry no mass or energy flux (the mirrored HLLC
    state has a zero contact speed), so a closed box keeps both."""
    o = Oracle(sd_config(K=8, p=3))
    Q = o.init_state(make_ic("sod"))
    R, _ = o.compute_rhs(Q)
    _, w = lgl(3)
    w2 = gll_weights_2d(w, o.K, o.h)
    R5 = as5(R, o.K, 3)
    for q in (0, 3):
        assert abs(np.sum(w2 * R5[q])) < 1e-12 * np.sum(w2 * np.abs(R5[q])), q


def test_sod_ic_width():
    """Circular Sod with the tanh front of width 2h (fr_lobatto_2d.cu:633-657)."""
    K, p = 6, 3
    o = Oracle(sd_config(K=K, p=p))
    Q5 = as5(o.init_state(make_ic("sod")), K, p)
    x, _ = lgl(p)
    X, Y = node_xy(K, p, 0.0, 0.0, 1.0, x)
    phi = 0.5 * (1 + np.tanh((0.25 - np.hypot(X - 0.5, Y - 0.5)) / (2.0 / K)))
    np.testing.assert_allclose(Q5[0], 0.125 + 10.875 * phi, rtol=1e-14)
    np.testing.assert_allclose(Q5[3], (0.1 + 9.9 * phi) / 0.4, rtol=1e-14)
    assert np.all(Q5[1] == 0) and np.all(Q5[2] == 0)


def test_nodal_repair():
    o = Oracle(sd_config(K=4, p=2))
    Q = o.init_state(make_ic("sod"))
    np.testing.assert_array_equal(o.limit(Q), Q)          # identity on admissible states
    Q5 = as5(Q, 4, 2).copy()
    Q5[0, 1, 2, 1, 1] = -0.05; Q5[1, 1, 2, 1, 1] = 0.3     # negative density, moving
    Q5[3, 2, 0, 0, 2] = 1e-9                                # negative pressure
    Ql = as5(o.limit(np.ascontiguousarray(Q5.reshape(-1))), 4, 2)
    n1 = (slice(None), 1, 2, 1, 1)
    # density floored, momentum zeroed; the pressure of the kept energy is
    # then far above eps_p, so E stays
    np.testing.assert_allclose(Ql[n1], [1e-6, 0.0, 0.0, Q5[n1][3]], rtol=1e-15)
    r, mx, my, E = Ql[:, 2, 0, 0, 2]
    assert 0.4 * (E - 0.5 * (mx * mx + my * my) / r) == pytest.approx(1e-6, rel=1e-9)
    changed = np.any(Ql != Q5, axis=0)
    assert changed.sum() == 2                               # nodal: nothing else moves


def test_dt_and_step():
    o = Oracle(sd_config(K=8, p=3))
    Q = o.init_state(make_ic("sod"))
    _, lam = o.compute_rhs(Q)
    U = as5(Q, 8, 3)
    r = U[0]; u = U[1] / r; v = U[2] / r; p = 0.4 * (U[3] - 0.5 * r * (u * u + v * v))
    np.testing.assert_allclose(lam, (np.abs(u) + np.abs(v) + np.sqrt(1.4 * p / r)).max((2, 3)).reshape(-1),
                               rtol=1e-14)
    cfl = 0.1 / 3                                          # the driver's CFL, main :838
    rc, t, dt = o.advance(Q, 1, cfl)
    assert rc == 0 and dt == pytest.approx(cfl * o.h / (lam.max() * 4), rel=1e-15) and t == dt


def test_sod_run_conserves_and_stays_symmetric():
    """A few hundred steps of the reference's default problem: mass and
    energy stay at their initial totals, and the solution keeps the
    problem's mirror symmetry x <-> 1-x."""
    K, p = 12, 3
    o = Oracle(sd_config(K=K, p=p))
    Q = o.init_state(make_ic("sod"))
    _, w = lgl(p)
    w2 = gll_weights_2d(w, K, o.h)
    tot0 = [np.sum(w2 * as5(Q, K, p)[q]) for q in range(4)]
    rc, t, _ = o.advance(Q, 200, 0.1 / p)
    assert rc == 0 and t > 0.01
    Q5 = as5(Q, K, p)
    for q in (0, 3):
        assert abs(np.sum(w2 * Q5[q]) - tot0[q]) < 1e-12 * tot0[q], q
    mirror = Q5[0][:, ::-1, :, ::-1]
    np.testing.assert_allclose(Q5[0], mirror, rtol=1e-10)


@pytest.mark.parametrize("p", [2, 3])
def test_vortex_convergence_order(p):
    """SD on the periodic isentropic vortex: the error against the advected
    exact vortex drops at about h^(p+1)."""
    errs = []
    for K in ((20, 40) if p == 2 else (10, 20)):    # p = 2 is pre-asymptotic at K = 10 (rate 2.49)
        cfg = sd_config(K=K, p=p, x0=0.0, y0=0.0, L=10.0, bc="periodic", surface_flux="rusanov",
                        limiter=False)
        o = Oracle(cfg)
        Q = o.init_state(make_ic("vortex"))
        rc, t, _ = o.advance(Q, 100000, 0.2, t_end=0.5)
        assert rc == 0
        x, w = lgl(p)
        X, Y = node_xy(K, p, 0.0, 0.0, 10.0, x)
        ex = np.zeros((4,) + X.shape)
        for idx in np.ndindex(X.shape):
            W = O.vortex_exact((X[idx] - t) % 10.0, (Y[idx] - t) % 10.0)
            ex[(slice(None),) + idx] = prim_to_cons(*W)
        errs.append(rel_l2(Q, ex, gll_weights_2d(w, K, o.h)))
    rate = math.log2(errs[0] / errs[1])
    assert rate > p + 0.5, (errs, rate)
    assert errs[1] < 3e-3
