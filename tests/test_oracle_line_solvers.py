"""Pin the line solvers of the oracle.

The reference's own line-solver tests solve u''(t) = -2 on [0, 1] in closed form
(unit_tests/sparse_solvers/lss_thomas_lu_solver_t.hpp:12-243, lss_double_sweep_solver_t.hpp:49-236);
central differences are exact for quadratics, so those cases pin the Dirichlet/Neumann/Robin folding
and its sign conventions (Neumann value g means u' = -g, Robin (l, g) means u' + l u + g = 0).
LAPACK's banded solver pins the plain algorithm on the batched-system distribution.
"""
import numpy as np
import pytest
import scipy.linalg

D, N, R = 0, 1, 2


def _kat(n, fp):
    dt = np.float64 if fp == "f64" else np.float32
    h = dt(1.0) / dt(n - 1)
    t = np.arange(n, dtype=np.float64) / (n - 1)
    ones = np.ones(n, dt)
    return ones, -2 * ones, ones, np.full(n, dt(-2.0) * h * h, dt), t


@pytest.mark.parametrize("solver", ["thomas_lu", "double_sweep"])
@pytest.mark.parametrize("fp,tol", [("f64", 1e-10), ("f32", 2e-3)])
def test_u2_dirichlet(oracle, solver, fp, tol):
    a, b, c, f, t = _kat(100, fp)
    x = getattr(oracle, solver)(a, b, c, f, (D, 0, 0.0), (D, 0, 0.0), fp)
    np.testing.assert_allclose(x, t * (1 - t), atol=tol)


@pytest.mark.parametrize("solver", ["thomas_lu", "double_sweep"])
@pytest.mark.parametrize("fp,tol", [("f64", 1e-10), ("f32", 2e-3)])
def test_u2_dirichlet_robin(oracle, solver, fp, tol):
    # u(0) = 1, u'(1) + u(1) = 0  ->  u = -t^2 + t + 1
    a, b, c, f, t = _kat(100, fp)
    x = getattr(oracle, solver)(a, b, c, f, (D, 0, 1.0), (R, 1.0, 0.0), fp)
    np.testing.assert_allclose(x, -t * t + t + 1, atol=tol)


@pytest.mark.parametrize("solver", ["thomas_lu", "double_sweep"])
@pytest.mark.parametrize("fp,tol", [("f64", 1e-8), ("f32", 5e-2)])
def test_u2_neumann_robin(oracle, solver, fp, tol):
    # u'(0) - 1 = 0 (Neumann value -1), u'(1) + 2 u(1) = 0  ->  u = -t^2 + t + 1/2, N = 1000
    a, b, c, f, t = _kat(1000, fp)
    x = getattr(oracle, solver)(a, b, c, f, (N, 0, -1.0), (R, 2.0, 0.0), fp)
    np.testing.assert_allclose(x, -t * t + t + 0.5, atol=tol)


def _random_systems(rng, B, n, dt=np.float64):
    a = rng.uniform(-1, 0, (B, n))
    c = rng.uniform(-1, 0, (B, n))
    b = 2.5 + np.abs(a) + np.abs(c)
    d = rng.standard_normal((B, n))
    return (v.astype(dt) for v in (a, b, c, d))


@pytest.mark.parametrize("n", [4, 17, 256])
def test_batched_thomas_matches_lapack(oracle, n):
    rng = np.random.default_rng(1)
    B = 64
    a, b, c, d = _random_systems(rng, B, n)
    x = oracle.thomas_batched(a, b, c, d, B, n, strided=False).reshape(B, n)
    for k in range(B):
        ab = np.zeros((3, n))
        ab[0, 1:] = c[k, :-1]
        ab[1] = b[k]
        ab[2, :-1] = a[k, 1:]
        ref = scipy.linalg.solve_banded((1, 1), ab, d[k])
        np.testing.assert_allclose(x[k], ref, rtol=1e-13, atol=1e-14)


def test_batched_layouts_agree(oracle):
    rng = np.random.default_rng(2)
    B, n = 50, 33
    a, b, c, d = _random_systems(rng, B, n)
    xc = oracle.thomas_batched(a, b, c, d, B, n, strided=False).reshape(B, n)
    xs = oracle.thomas_batched(a.T.copy(), b.T.copy(), c.T.copy(), d.T.copy(), B, n, strided=True).reshape(n, B)
    assert np.array_equal(xc, xs.T)


def test_thomas_and_double_sweep_agree(oracle):
    rng = np.random.default_rng(3)
    for lo, hi in [((D, 0, 0.3), (D, 0, -0.7)), ((N, 0, 0.2), (R, 1.5, 0.1)), ((R, 0.5, -0.2), (N, 0, 0.4))]:
        a, b, c, d = (v[0] for v in _random_systems(rng, 1, 40))
        x1 = oracle.thomas_lu(a, b, c, d, lo, hi)
        x2 = oracle.double_sweep(a, b, c, d, lo, hi)
        np.testing.assert_allclose(x1, x2, rtol=1e-13, atol=1e-14)
