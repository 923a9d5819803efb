"""Pin the CPU restatement (oracle/) against the surfaces the reference's own tests committed.

The reference prints with 6 significant digits (common/lss_print.hpp:119-160), so a correct
restatement agrees elementwise to within half a unit in the 6th digit (relative 5e-6) in double.
Float surfaces were produced with another libm; they agree to a few 1e-6 of the surface maximum
after 200 steps, about 2.6e-5 after the explicit cases' 8000 (including the float-only index-map
artefacts near v_max, 1e-3 of the maximum there, which a wrong index map would miss by O(1)).
"""
import numpy as np
import pytest

from tests.golden import reference_outputs as R


def _pde(O, model):
    kind, kw = model
    if kind == "heston":
        return O.heston_pde(O.Heston(**kw))
    return O.sabr_pde(O.Sabr(**kw))


def _cfg(O, case):
    kw = dict(case["cfg"])
    return O.Cfg(**kw, cached=True, coeff_time_independent=(case["model"][0] == "heston"))


@pytest.mark.parametrize("case", R.CASES, ids=[c["stem"] for c in R.CASES])
def test_double_surface_matches_reference(oracle, case):
    xs, ys, vals = R.load(case["stem"], "f64")
    cfg = _cfg(oracle, case)
    u = oracle.solve(_pde(oracle, case["model"]), cfg, "f64")
    assert u.shape == vals.shape
    _, _, gx, gy, _ = oracle.grid(cfg, "f64")
    np.testing.assert_allclose(gx, xs, rtol=5.01e-6, atol=1e-9)
    np.testing.assert_allclose(gy, ys, rtol=5.01e-6, atol=1e-9)
    err = np.abs(u - vals)
    assert np.all(err <= 5.01e-6 * np.abs(vals) + 1e-12 * np.abs(vals).max()), err.max()


@pytest.mark.parametrize("case", [c for c in R.CASES if c["floats"]], ids=[c["stem"] for c in R.CASES if c["floats"]])
def test_float_surface_matches_reference(oracle, case):
    _, _, vals = R.load(case["stem"], "f32")
    u = oracle.solve(_pde(oracle, case["model"]), _cfg(oracle, case), "f32").astype(np.float64)
    scale = np.abs(vals).max()
    # float rounding accumulates over the steps like a random walk: the 200-step surfaces agree to
    # 1.5e-5 of the maximum, the 8000-step explicit ones to sqrt(40) times that
    steps = case["cfg"]["nt"] - 1
    assert np.abs(u - vals).max() <= 1.5e-5 * max(1.0, (steps / 200) ** 0.5) * scale


def test_cached_mode_is_bitwise_faithful(oracle):
    case = R.CASES[1]
    kw = dict(case["cfg"], nx=40, ny=20, nt=30)
    pde = _pde(oracle, case["model"])
    a = oracle.solve(pde, oracle.Cfg(**kw, cached=False), "f64")
    b = oracle.solve(pde, oracle.Cfg(**kw, cached=True, coeff_time_independent=True), "f64")
    assert np.array_equal(a, b)
    s = R.CASES[-1]
    kw = dict(s["cfg"], nx=30, ny=16, nt=20)
    pde = _pde(oracle, s["model"])
    a = oracle.solve(pde, oracle.Cfg(**kw, cached=False), "f32")
    b = oracle.solve(pde, oracle.Cfg(**kw, cached=True), "f32")
    assert np.array_equal(a, b)
