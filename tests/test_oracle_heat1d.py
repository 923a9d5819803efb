"""Pins of the CPU restatement of the 1-D implicit heat solver (oracle/lss_oracle_heat1d.inc).

* The surfaces the reference's own 1-D print tests committed (tests/golden/reference_outputs/,
  6 significant digits): Black-Scholes (Euler, Crank-Nicolson; final level and every level), the pure
  heat equation and an advection-diffusion equation.  Double values agree to print precision.
* The closed forms those tests print beside them: black_scholes_exact (common/lss_utility.hpp:70-110)
  and the heat-equation series; the scheme converges to them (error falling with the grid).
* What the mathematics fixes: one step on a tiny grid against the dense-matrix theta scheme (brute
  force), which catches a dropped term, a wrong sign or a transposed operand in the coefficients.
"""
import numpy as np
import pytest

from tests.golden import reference_outputs as R


def _pde(O, case):
    if case["model"] == "bs":
        return O.bs_heat1d_pde(R.BS["strike"], R.BS["rate"], R.BS["sig"], R.BS["maturity"], case["cfg"]["x_hi"])
    zero = lambda t, x: 0.0  # noqa: E731
    dz = (O.BC_DIRICHLET, lambda t: 0.0)
    if case["model"] == "heat":
        return O.custom_heat1d_pde(lambda t, x: 1.0, zero, zero, lambda x: x, dz, dz)
    return O.custom_heat1d_pde(lambda t, x: 1.0, lambda t, x: -1.0, zero, lambda x: 1.0, dz, dz)


@pytest.mark.parametrize("case", R.CASES1D, ids=[c["stem"] for c in R.CASES1D])
def test_oracle_matches_reference_1d_outputs(oracle, case):
    xs, ts, vals = R.load1d(case["stem"], "approx", "f64")
    cfg = oracle.Heat1dCfg(**case["cfg"])
    gx, _ = oracle.heat1d_grid(cfg)
    np.testing.assert_allclose(gx, xs, rtol=5.01e-6, atol=1e-12)
    out, surf = oracle.heat1d_solve(_pde(oracle, case), cfg, "f64", surfaces=True)
    got = surf if case["surf"] else out
    assert got.shape == vals.shape
    err = np.abs(got - vals)
    # 6 printed digits: half a unit in the 6th digit, plus the print of values that are zero to rounding
    assert np.all(err <= 5.01e-6 * np.abs(vals) + 1e-13 * np.abs(vals).max()), err.max()
    if case["surf"]:  # the last level of a surface solve is the final level of the plain solve
        assert np.array_equal(surf[0 if cfg.direction == oracle.BACKWARD else -1], out)


@pytest.mark.parametrize("case", R.CASES1D, ids=[c["stem"] for c in R.CASES1D])
def test_oracle_float_matches_reference_1d_outputs(oracle, case):
    _, _, vals = R.load1d(case["stem"], "approx", "f32")
    out, surf = oracle.heat1d_solve(_pde(oracle, case), oracle.Heat1dCfg(**case["cfg"]), "f32", surfaces=True)
    got = (surf if case["surf"] else out).astype(np.float64)
    assert np.abs(got - vals).max() <= 2e-5 * np.abs(vals).max()


def _bs_exact(x, tau, K=10.0, r=0.2, sig=0.25):
    """black_scholes_exact::call (common/lss_utility.hpp:89-96) with the exact normal CDF."""
    from scipy.stats import norm
    x = np.asarray(x, dtype=np.float64)
    with np.errstate(divide="ignore"):
        d1 = (np.log(x / K) + (r + 0.5 * sig * sig) * tau) / (sig * np.sqrt(tau))
    d2 = d1 - sig * np.sqrt(tau)
    return norm.cdf(d1) * x - norm.cdf(d2) * K * np.exp(-r * tau)


def test_reference_bench_files_are_black_scholes_exact():
    """The committed _bench_ files are black_scholes_exact with the reference's norm_cdf, a rational
    approximation of about 1e-4 accuracy (common/lss_utility.hpp:49-65); they agree with the exact
    closed form to that accuracy, which is why the scheme is pinned against the exact one below."""
    xs, _, bench = R.load1d("ImplBlackScholesEquationDirichletBCThomasLUSolverCrankNicolson", "bench")
    assert np.abs(bench - _bs_exact(xs, 1.0)).max() <= 2e-3


@pytest.mark.parametrize("scheme", ["euler", "cn"])
def test_oracle_converges_to_black_scholes_exact(oracle, scheme):
    """Second order in space; first (Euler) or second (Crank-Nicolson) order in time: refining both
    grids cuts the error against the closed form, near the money, by the expected factor."""
    errs = []
    for n in (100, 200, 400):
        cfg = oracle.Heat1dCfg(x_lo=0.0, x_hi=20.0, n=n, nt=n, t_hi=1.0, accumulation_point=10.0, alpha_scale=3.0,
                               nonuniform=True, implicit_scheme=oracle.EULER if scheme == "euler" else
                               oracle.CRANK_NICOLSON, direction=oracle.BACKWARD)
        u = oracle.heat1d_solve(oracle.bs_heat1d_pde(10.0, 0.2, 0.25, 1.0, 20.0), cfg)
        xs, _ = oracle.heat1d_grid(cfg)
        m = (xs > 5.0) & (xs < 15.0)
        errs.append(np.abs(u[m] - _bs_exact(xs[m], 1.0)).max())
    print(scheme, errs)
    assert errs[0] < 2e-2 and errs[2] < errs[1] < errs[0]
    ratio = errs[1] / errs[2]
    assert ratio > (1.7 if scheme == "euler" else 3.0)


def test_heat_series_closed_form(oracle):
    """u_t = u_xx, u(x, 0) = x, u(0) = u(1) = 0 has the series 2/pi sum (-1)^(i+1) e^(-(i pi)^2 t)
    sin(i pi x) / i; Crank-Nicolson on a fine grid matches it (the jump of the initial data at x = 1
    leaves a slowly damped Crank-Nicolson error of a few 1e-4)."""
    n = 401
    cfg = oracle.Heat1dCfg(x_lo=0.0, x_hi=1.0, n=n, nt=401, t_hi=0.1, implicit_scheme=oracle.CRANK_NICOLSON,
                           direction=oracle.FORWARD)
    dz = (oracle.BC_DIRICHLET, lambda t: 0.0)
    zero = lambda t, x: 0.0  # noqa: E731
    u = oracle.heat1d_solve(oracle.custom_heat1d_pde(lambda t, x: 1.0, zero, zero, lambda x: x, dz, dz), cfg)
    xs, _ = oracle.heat1d_grid(cfg)
    i = np.arange(1, 200)[:, None]
    exact = 2 / np.pi * np.sum((-1.0) ** (i + 1) * np.exp(-(i * np.pi) ** 2 * 0.1) * np.sin(i * np.pi * xs) / i,
                               axis=0)
    assert np.abs(u - exact).max() <= 5e-4


def _dense_step(A, B, D, theta, prev, lo, hi):
    """One theta step written as dense matrices (Dirichlet ends): (I - theta L) u1 = (I + (1-theta) L) u0
    with L u_t = A u_{t-1} - 2 B u_t + D u_{t+1}; brute force for tiny grids."""
    n = len(prev)
    Lm = np.zeros((n, n))
    for t in range(1, n - 1):
        Lm[t, t - 1], Lm[t, t], Lm[t, t + 1] = A[t], -2 * B[t], D[t]
    rhs = prev + (1 - theta) * (Lm @ prev)
    M = np.eye(n) - theta * Lm
    M[0, :], M[-1, :] = 0, 0
    M[0, 0] = M[-1, -1] = 1
    rhs[0], rhs[-1] = lo, hi
    return np.linalg.solve(M, rhs)


@pytest.mark.parametrize("theta_scheme", ["euler", "cn"])
def test_one_step_equals_dense_theta_scheme(oracle, theta_scheme):
    """A single step on a tiny uniform grid against the dense-matrix form of the theta scheme with the
    reference's A = lambda a - gamma b, B = lambda a - delta c, D = lambda a + gamma b (brute force)."""
    n, T = 9, 0.3
    cfg = oracle.Heat1dCfg(x_lo=0.0, x_hi=1.0, n=n, nt=2, t_hi=T, implicit_scheme=oracle.EULER if
                           theta_scheme == "euler" else oracle.CRANK_NICOLSON, direction=oracle.FORWARD)
    # a uniform physical grid (the sinh map of grid_enum::Uniform is not equispaced) needs the
    # transform's own derivatives, so take the oracle's grid and feed coefficients in zeta units
    a = lambda t, x: 0.3 + 0.1 * x  # noqa: E731
    b = lambda t, x: 0.2 - x  # noqa: E731
    c = lambda t, x: -0.4  # noqa: E731
    init = lambda x: np.sin(3 * x) + 0.5  # noqa: E731
    lo, hi = 0.25, -0.5
    pde = oracle.custom_heat1d_pde(a, b, c, init, (oracle.BC_DIRICHLET, lambda t: lo),
                                   (oracle.BC_DIRICHLET, lambda t: hi))
    u = oracle.heat1d_solve(pde, cfg)
    xs, _ = oracle.heat1d_grid(cfg)
    # the transform: X(zeta) = 0 + 1 * sinh(c0 zeta + c1 (1 - zeta)), c0 = asinh(1), c1 = 0
    c0 = np.arcsinh(1.0)
    z = np.linspace(0, 1, n)
    X, dX, ddX = np.sinh(c0 * z), c0 * np.cosh(c0 * z), c0 * c0 * np.sinh(c0 * z)
    np.testing.assert_allclose(xs, X, rtol=1e-14, atol=1e-15)
    at = np.array([a(0, x) for x in X]) / dX ** 2
    bt = np.array([b(0, x) for x in X]) / dX - np.array([a(0, x) for x in X]) * ddX / dX ** 3
    ct = np.array([c(0, x) for x in X])
    h, k = 1.0 / (n - 1), T
    lam, gam, dl = k / h / h, k / (2 * h), 0.5 * k
    A, B, D = lam * at - gam * bt, lam * at - dl * ct, lam * at + gam * bt
    theta = 1.0 if theta_scheme == "euler" else 0.5
    ref = _dense_step(A, B, D, theta, np.array([init(x) for x in X]), lo, hi)
    np.testing.assert_allclose(u, ref, rtol=1e-12, atol=1e-13)
