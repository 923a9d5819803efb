"""The oracle's heat-source term (rhs_intermed_1_source, lss_heat_douglas_rachford_method.hpp:70-104 and
the same function of the CS/MCS/HV files).  The reference defines it but never steps it (its loop
returns early when a source is set, implicit_kernel.hpp:1032-1044), so no reference output covers it;
these tests pin it by a closed form instead.

With every coefficient zero the splitting operators vanish: each line system is the identity, every
later stage of CS/MCS/HV reproduces Y1, and the v = 0 row (explicit boundary in D, E, F) stays put.  A
step then only adds the source term, so after the loop every interior point holds
    u0 + sum over steps of k (theta F(t_next) + (1 - theta) F(t))
with t the step's loop time and t_next = t + k (Forward) or t - k (Backward).  A term taken at the
wrong time, a wrong theta weight, a missing k, the source added in a second stage or at the physical
point of another grid index all change that sum."""
import numpy as np
import pytest

pytestmark = pytest.mark.filterwarnings("ignore")


def _zero(t, x, y):
    return 0.0


def _src(t, x, y):
    return (1.0 + 3.0 * t) * x * y + 0.25 * y * y - 0.1 * t


def _pde(O):
    return O.custom_pde(_zero, _zero, _zero, _zero, _zero, _zero, lambda x, y: x + 2.0 * y,
                        lambda t, y: 0.0, lambda t, y: 0.0, lambda t, x: 0.0, src=_src)


@pytest.mark.parametrize("splitting", [0, 1, 2, 3])
@pytest.mark.parametrize("direction", [0, 1])
@pytest.mark.parametrize("scheme,theta", [(0, 1.0), (1, 0.3), (2, 0.5)])
def test_zero_operator_leaves_the_source_quadrature(oracle, splitting, direction, scheme, theta):
    O = oracle
    cfg = O.Cfg(x_lo=0.0, x_hi=2.0, y_lo=0.0, y_hi=1.5, nx=14, ny=11, nt=9, t_lo=0.1, t_hi=0.9,
                accumulation_point=1.0, alpha_scale=3.0, beta_scale=15.0, nonuniform=True, splitting=splitting,
                weighting=0.5, implicit_scheme=scheme, tridiagonal=O.THOMAS_LU, direction=direction)
    _, _, xs, ys, times = O.grid(cfg, "f64")
    k = (cfg.t_hi - cfg.t_lo) / (cfg.nt - 1)
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    want = X + 2.0 * Y
    for t in times:
        tn = t + k if direction == O.FORWARD else t - k
        want = want + k * (theta * _src(tn, X, Y) + (1.0 - theta) * _src(t, X, Y))
    got = O.solve(_pde(O), cfg, "f64")
    inner = (slice(1, -1), slice(1, -1))
    np.testing.assert_allclose(got[inner], want[inner], rtol=1e-13, atol=1e-13)
    # the v = 0 row keeps its initial values (its explicit stage has no source)
    np.testing.assert_array_equal(got[1:-1, 0], (X + 2.0 * Y)[1:-1, 0])


def test_float_source_follows_the_same_quadrature(oracle):
    O = oracle
    cfg = O.Cfg(x_lo=0.0, x_hi=2.0, y_lo=0.0, y_hi=1.5, nx=12, ny=10, nt=7, t_lo=0.0, t_hi=0.6,
                splitting=O.MOD_CRAIG_SNEYD, implicit_scheme=O.CRANK_NICOLSON, direction=O.BACKWARD)
    _, _, xs, ys, times = O.grid(cfg, "f32")
    k = (cfg.t_hi - cfg.t_lo) / (cfg.nt - 1)
    X, Y = np.meshgrid(xs.astype(np.float64), ys.astype(np.float64), indexing="ij")
    want = X + 2.0 * Y
    for t in times.astype(np.float64):
        want = want + k * (0.5 * _src(t - k, X, Y) + 0.5 * _src(t, X, Y))
    got = O.solve(_pde(O), cfg, "f32").astype(np.float64)
    np.testing.assert_allclose(got[1:-1, 1:-1], want[1:-1, 1:-1], rtol=2e-6, atol=2e-6)


def test_source_is_linear_on_top_of_the_homogeneous_solve(oracle):
    """With a real operator the source enters linearly: solve(L, u0, F) - solve(L, u0, 0) equals
    solve(L, 0, F) with zero boundary data (the loop is affine in (u0, boundaries, F))."""
    O = oracle
    a = lambda t, x, y: 0.5 * y * x * x
    b = lambda t, x, y: 0.02 * y
    c = lambda t, x, y: 0.01 * x * y
    d = lambda t, x, y: 0.03 * x
    e = lambda t, x, y: 2.0 * (0.2 - y)
    f = lambda t, x, y: -0.03
    init = lambda x, y: max(0.0, x - 1.0)
    cfg = O.Cfg(x_lo=0.0, x_hi=2.0, y_lo=0.0, y_hi=1.2, nx=20, ny=14, nt=11, t_hi=0.5, accumulation_point=1.0,
                alpha_scale=3.0, beta_scale=15.0, nonuniform=True, splitting=O.HUNDSDORFER_VERWER,
                direction=O.BACKWARD)
    full = O.solve(O.custom_pde(a, b, c, d, e, f, init, lambda t, y: 0.0, lambda t, y: -1.0, lambda t, x: x,
                                src=_src), cfg, "f64")
    hom = O.solve(O.custom_pde(a, b, c, d, e, f, init, lambda t, y: 0.0, lambda t, y: -1.0, lambda t, x: x), cfg,
                  "f64")
    only = O.solve(O.custom_pde(a, b, c, d, e, f, lambda x, y: 0.0, lambda t, y: 0.0, lambda t, y: 0.0,
                                lambda t, x: 0.0, src=_src), cfg, "f64")
    np.testing.assert_allclose(full - hom, only, rtol=0, atol=1e-12 * np.abs(full).max())
    assert np.abs(only).max() > 1e-2  # the source moves the solution
