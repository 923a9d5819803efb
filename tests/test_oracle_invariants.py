"""Properties the mathematics fixes, used to pin oracle paths the reference surfaces leave open."""
import numpy as np
import pytest


def _const_pde(O, c0):
    z = lambda t, x, y: 0.0
    return O.custom_pde(a=lambda t, s, v: 0.5 * v * s * s, b=lambda t, s, v: 0.5 * 0.09 * v,
                        c=lambda t, s, v: 0.2 * 0.3 * v * s, d=z, e=lambda t, s, v: 2.0 * (0.2 - v), f=z,
                        init=lambda s, v: c0, h_lower=lambda t, v: c0, h_upper=lambda t, v: 0.0,
                        v_upper=lambda t, s: c0)


@pytest.mark.parametrize("splitting,weighting", [(0, 0.3), (1, 0.3), (2, "theta"), (3, "theta")])
@pytest.mark.parametrize("theta_scheme,theta", [(0, 1.0), (2, 0.5), (3, 0.8)])
def test_constants_are_preserved_when_f_is_zero(oracle, splitting, weighting, theta_scheme, theta):
    # With f = d = 0 every stage operator has unit row sums (the v=0 row and the S_max
    # extrapolation with a zero Neumann value too), so a constant surface stays constant; a wrong
    # sign or a dropped stencil term breaks this.  MCS and HV read the S_min/S_max rows of the
    # intermediate Y2, which the reference leaves at zero (lss_heat_modified_craig_sneyd_method.hpp:304,
    # lss_heat_hundsdorfer_verwer_method.hpp:331); those contributions cancel exactly only when
    # zeta == theta, so the invariant is checked there and its failure at zeta != theta below.
    # k must be small enough for the explicit v = 0 row (explicit_boundary.hpp:75-83) not to
    # amplify rounding.
    w = theta if weighting == "theta" else weighting
    cfg = oracle.Cfg(10.0, 200.0, 0.0, 1.2, 17, 13, 60, accumulation_point=100.0, nonuniform=True,
                     splitting=splitting, weighting=w, implicit_scheme=theta_scheme)
    u = oracle.solve(_const_pde(oracle, 3.25), cfg, "f64")
    np.testing.assert_allclose(u, 3.25, rtol=1e-13, atol=0)


@pytest.mark.parametrize("splitting", [2, 3])
def test_zero_y2_border_breaks_constants_when_weighting_differs_from_theta(oracle, splitting):
    cfg = oracle.Cfg(10.0, 200.0, 0.0, 1.2, 17, 13, 60, accumulation_point=100.0, nonuniform=True,
                     splitting=splitting, weighting=0.3)
    u = oracle.solve(_const_pde(oracle, 3.25), cfg, "f64")
    assert np.abs(u - 3.25).max() > 1e-3


def test_cs_equals_mcs_when_weighting_equals_theta(oracle):
    H = oracle.heston_pde(oracle.Heston())
    kw = dict(x_lo=50.0, x_hi=200.0, y_lo=0.0, y_hi=1.2, nx=30, ny=20, nt=25, accumulation_point=100.0,
              nonuniform=True, weighting=0.5, implicit_scheme=2)
    cs = oracle.solve(H, oracle.Cfg(**kw, splitting=1), "f64")
    mcs = oracle.solve(H, oracle.Cfg(**kw, splitting=2), "f64")
    assert np.array_equal(cs, mcs)
    mcs2 = oracle.solve(H, oracle.Cfg(**dict(kw, weighting=0.3), splitting=2), "f64")
    assert not np.array_equal(cs, mcs2)


def test_solution_is_linear_in_payoff_and_boundary_data(oracle):
    def pde(k):
        return oracle.custom_pde(a=lambda t, s, v: 0.5 * v * s * s, b=lambda t, s, v: 0.5 * 0.09 * v,
                                 c=lambda t, s, v: 0.06 * v * s, d=lambda t, s, v: 0.03 * s,
                                 e=lambda t, s, v: 2.0 * (0.2 - v), f=lambda t, s, v: -0.03,
                                 init=lambda s, v: k * max(s - 100.0, 0.0) + (1 - k) * np.sin(s / 30.0),
                                 h_lower=lambda t, v: 2.0 * k, h_upper=lambda t, v: -1.0 * k,
                                 v_upper=lambda t, s: k * s + (1 - k))
    cfg = oracle.Cfg(20.0, 200.0, 0.0, 1.0, 24, 16, 20, accumulation_point=100.0, nonuniform=True, splitting=3)
    u0, u1, uh = (oracle.solve(pde(k), cfg, "f64") for k in (0.0, 1.0, 0.5))
    np.testing.assert_allclose(uh, 0.5 * (u0 + u1), rtol=0, atol=1e-11 * np.abs(u1).max())


def test_forward_and_backward_agree_for_time_independent_data(oracle):
    H = oracle.heston_pde(oracle.Heston())
    kw = dict(x_lo=50.0, x_hi=200.0, y_lo=0.0, y_hi=1.2, nx=25, ny=14, nt=11, accumulation_point=100.0,
              nonuniform=True, splitting=0)
    a = oracle.solve(H, oracle.Cfg(**kw, direction=0), "f64")
    b = oracle.solve(H, oracle.Cfg(**kw, direction=1), "f64")
    assert np.array_equal(a, b)


@pytest.mark.parametrize("n,fp,bad", [(50, "f64", 2), (100, "f64", 4), (150, "f64", 6), (32, "f64", 0),
                                      (64, "f64", 0), (256, "f64", 0), (512, "f64", 0), (1024, "f64", 0)])
def test_index_maps_truncate_like_ieee(oracle, n, fp, bad):
    # index_of(h*i) = size_t((h*i)/h) (lss_grid_config.hpp:56-59) can land on i-1; the maps the
    # oracle uses must be exactly those of IEEE arithmetic in the solve's precision
    cfg = oracle.Cfg(0.0, 1.0, 0.0, 1.0, n, 8, 3)
    idx1, _, _, _, _ = oracle.grid(cfg, fp)
    dt = np.float64 if fp == "f64" else np.float32
    h = dt(1.0) / dt(n - 1)
    expect = np.floor((h * np.arange(n, dtype=dt)) / h).astype(np.uint32)
    assert np.array_equal(idx1, expect)
    assert np.count_nonzero(idx1 != np.arange(n)) == bad


def test_float_index_maps_hit_last_point(oracle):
    for n in (64, 256):
        idx1, _, _, _, _ = oracle.grid(oracle.Cfg(0.0, 1.0, 0.0, 1.0, n, 8, 3), "f32")
        assert idx1[n - 1] == n - 2


def test_loop_times_accumulate_like_the_reference(oracle):
    # Backward: time = T - k, then time -= k (implicit_kernel.hpp:191-205)
    cfg = oracle.Cfg(0.0, 1.0, 0.0, 1.0, 8, 8, 101, t_lo=0.0, t_hi=1.0, direction=1)
    *_, times = oracle.grid(cfg, "f64")
    k = (1.0 - 0.0) / 100.0
    t, expect = 1.0 - k, []
    for _ in range(100):
        expect.append(t)
        t -= k
    assert np.array_equal(times, np.array(expect))
    assert times[-1] != 0.0 and abs(times[-1]) < 1e-14


@pytest.mark.parametrize("splitting", [0, 1, 2, 3])
def test_threaded_oracle_equals_the_serial_one(oracle, splitting):
    """The oracle's OpenMP line loops (used for large grids) give the serial result bit for bit."""
    import dataclasses
    from tests.helpers import config0, oracle_solve
    m = dataclasses.replace(config0(nx=130, ny=70, nt=9), splitting=splitting, weighting=0.6)
    try:
        oracle.set_parallel_min(1 << 40)
        serial = oracle_solve(oracle, m)
        oracle.set_parallel_min(0)
        threaded = oracle_solve(oracle, m)
    finally:
        oracle.set_parallel_min(262144)
    assert np.array_equal(serial, threaded)
