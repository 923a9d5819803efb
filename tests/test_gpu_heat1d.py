"""GPU 1-D implicit heat / Black-Scholes solver (include/lss_heat1d.h) against the oracle.

* The exact kernel (one lane per problem, the reference's Thomas LU / double-sweep order, no
  contraction) equals the oracle's restatement of the reference's host solver bit for bit, final level
  and every level, fp64 and fp32, Euler and Crank-Nicolson, Forward and Backward, both grids.
* The partitioned kernel (one wavefront per problem, the whole loop in registers, fused multiply-adds)
  stays within 1e-12 normwise of the oracle in fp64 and 1e-5 in fp32, for line lengths that take
  every chunk length 1..16.
* Books of mixed problems in one launch give each problem its single-problem result.
* On fine grids the GPU solution converges to black_scholes_exact (common/lss_utility.hpp:70-110).
"""
import numpy as np
import pytest

import linear_system_solvers_amd as L
from linear_system_solvers_amd import heat1d as H

pytestmark = pytest.mark.gpu


def _oracle(O, m: H.BsModel, fp="f64", surfaces=False):
    cfg = O.Heat1dCfg(x_lo=m.x_lo, x_hi=m.x_hi, n=m.n, nt=m.nt, t_lo=m.t_lo, t_hi=m.t_hi,
                      accumulation_point=m.accumulation_point, alpha_scale=m.alpha_scale,
                      nonuniform=m.grid == L.NONUNIFORM, implicit_scheme=m.implicit_scheme,
                      tridiagonal=O.DOUBLE_SWEEP if m.tridiagonal == L.DOUBLE_SWEEP else O.THOMAS_LU,
                      direction=m.direction)
    return O.heat1d_solve(O.bs_heat1d_pde(m.strike, m.rate, m.sig, m.maturity, m.x_hi), cfg, fp, surfaces=surfaces)


def _normwise(g, c):
    return float(np.abs(np.asarray(g, np.float64) - c).max() / np.abs(c).max())


CASES = [
    dict(),                                                                   # print test :209 (CN, default hints)
    dict(implicit_scheme=L.EULER),                                            # :105
    dict(implicit_scheme=L.EULER, accumulation_point=10.0),                   # :327
    dict(accumulation_point=10.0, grid=L.NONUNIFORM),                         # :439
    dict(direction=L.FORWARD, tridiagonal=L.DOUBLE_SWEEP),
    dict(n=37, nt=11, grid=L.NONUNIFORM, accumulation_point=10.0, tridiagonal=L.DOUBLE_SWEEP),
]


@pytest.mark.parametrize("fp", [L.FP64, L.FP32])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_exact_kernel_is_bit_identical_to_the_oracle(oracle, fp, case):
    m = H.BsModel(**CASES[case])
    plan = H.Heat1dPlan.from_bs_models([m], fp)
    outs, surf = plan.run_host(surfaces=True)
    o, osurf = _oracle(oracle, m, "f64" if fp == L.FP64 else "f32", surfaces=True)
    print(f"case {case} fp {fp}: max |gpu - oracle| {np.abs(outs[0] - o).max():.3e}", flush=True)
    assert np.array_equal(outs[0], o)
    assert np.array_equal(surf[0], osurf)
    np.testing.assert_array_equal(plan.grid(0), oracle.heat1d_grid(
        oracle.Heat1dCfg(x_lo=m.x_lo, x_hi=m.x_hi, n=m.n, nt=m.nt, accumulation_point=m.accumulation_point,
                         alpha_scale=m.alpha_scale, nonuniform=m.grid == L.NONUNIFORM),
        "f64" if fp == L.FP64 else "f32")[0])
    plan.close()


@pytest.mark.parametrize("fp,tol", [(L.FP64, 1e-12), (L.FP32, 1e-5)])
@pytest.mark.parametrize("n", [10, 66, 100, 130, 152, 300, 514, 702, 1026])
def test_partitioned_kernel_within_tolerance(oracle, fp, tol, n):
    """f64: within 1e-12 of the oracle.  f32: the partitioned elimination rounds differently from the
    sequential Thomas sweep, and Crank-Nicolson carries that difference through the loop undamped; the
    bar is 1e-5 or the f32 oracle's own distance from the f64 oracle, whichever is larger, and the GPU's
    distance from the f64 oracle within 1.5 times the f32 oracle's (as accurate as the reference's f32)."""
    # unknowns n - 2 take chunk lengths 1, 1, 2, 2, 3, 6, 8, 16, 16
    for extra in (dict(), dict(implicit_scheme=L.EULER, grid=L.NONUNIFORM, accumulation_point=10.0,
                               direction=L.FORWARD)):
        m = H.BsModel(n=n, nt=60, line_solver=L.LINE_PARTITIONED, **extra)
        plan = H.Heat1dPlan.from_bs_models([m], fp)
        outs, surf = plan.run_host(surfaces=True)
        o, osurf = _oracle(oracle, m, "f64" if fp == L.FP64 else "f32", surfaces=True)
        e, es = _normwise(outs[0], o), _normwise(surf[0], osurf)
        bar, extra_msg = tol, ""
        if fp == L.FP32:
            o64, osurf64 = _oracle(oracle, m, "f64", surfaces=True)
            own = max(_normwise(o, o64), _normwise(osurf, osurf64))
            g64 = max(_normwise(outs[0], o64), _normwise(surf[0], osurf64))
            bar = max(tol, own)
            extra_msg = f"; f32 oracle vs f64 oracle {own:.3e}, partitioned vs f64 oracle {g64:.3e}"
            assert g64 <= 1.5 * own + tol
        print(f"n {n} {extra}: partitioned vs oracle {e:.3e} (surface {es:.3e}){extra_msg}", flush=True)
        assert e <= bar and es <= bar
        plan.close()


def test_book_of_mixed_problems_gives_each_its_single_result(oracle):
    rng = np.random.default_rng(11)
    models = []
    for q in range(300):
        models.append(H.BsModel(strike=float(rng.uniform(5, 15)), rate=float(rng.uniform(0, 0.1)),
                                sig=float(rng.uniform(0.1, 0.6)), maturity=1.0, n=int(rng.integers(8, 600)),
                                nt=int(rng.integers(2, 120)), accumulation_point=float(rng.uniform(5, 15)),
                                grid=L.NONUNIFORM if q % 2 else L.UNIFORM,
                                implicit_scheme=L.EULER if q % 3 == 0 else L.CRANK_NICOLSON,
                                tridiagonal=L.DOUBLE_SWEEP if q % 5 == 0 else L.THOMAS_LU,
                                direction=L.FORWARD if q % 7 == 0 else L.BACKWARD))
    for solver in (L.LINE_EXACT, L.LINE_PARTITIONED):
        for m in models:
            m.line_solver = solver
        plan = H.Heat1dPlan.from_bs_models(models, L.FP64)
        outs = plan.run_host()
        worst = 0.0
        for q in range(0, len(models), 7):
            o = _oracle(oracle, models[q])
            if solver == L.LINE_EXACT:
                assert np.array_equal(outs[q], o), q
            else:
                worst = max(worst, _normwise(outs[q], o))
        print(f"book of {len(models)}, line solver {solver}: worst normwise {worst:.3e}", flush=True)
        assert worst <= 1e-12
        plan.close()


def test_device_run_with_surfaces_equals_host_run():
    torch = pytest.importorskip("torch")
    models = [H.BsModel(n=200, nt=50, line_solver=L.LINE_PARTITIONED, strike=8.0 + q) for q in range(5)]
    plan = H.Heat1dPlan.from_bs_models(models, L.FP64)
    host_out, host_surf = plan.run_host(surfaces=True)
    ins = [torch.from_numpy(plan.initial(p)).cuda() for p in range(plan.count)]
    outs = [torch.empty_like(u) for u in ins]
    surfs = [torch.empty((50, 200), dtype=torch.float64, device="cuda") for _ in ins]
    stream = torch.cuda.current_stream().cuda_stream
    for _ in range(2):  # the second run reuses the device pointer table
        plan.run_dev([u.data_ptr() for u in ins], [u.data_ptr() for u in outs], [s.data_ptr() for s in surfs], stream)
    torch.cuda.synchronize()
    for p in range(plan.count):
        assert np.array_equal(outs[p].cpu().numpy(), host_out[p])
        assert np.array_equal(surfs[p].cpu().numpy(), host_surf[p])
    plan.close()


def test_gpu_converges_to_black_scholes_exact():
    from scipy.stats import norm
    errs = []
    for n in (200, 400, 800):
        m = H.BsModel(n=n, nt=n, grid=L.NONUNIFORM, accumulation_point=10.0, line_solver=L.LINE_PARTITIONED)
        plan = H.Heat1dPlan.from_bs_models([m], L.FP64)
        u = plan.run_host()[0]
        x = plan.grid(0)
        sel = (x > 5) & (x < 15)
        xs = x[sel]
        d1 = (np.log(xs / 10.0) + (0.2 + 0.5 * 0.25 ** 2)) / 0.25
        exact = norm.cdf(d1) * xs - norm.cdf(d1 - 0.25) * 10.0 * np.exp(-0.2)
        errs.append(np.abs(u[sel] - exact).max())
        plan.close()
    print("errors vs black_scholes_exact", errs, flush=True)
    assert errs[2] < errs[1] < errs[0] < 5e-3
    assert errs[1] / errs[2] > 3.0


def test_invalid_descriptions_are_rejected():
    with pytest.raises(L.LssError):
        H.Heat1dPlan.from_bs_models([H.BsModel(n=3)])
    with pytest.raises(L.LssError):
        H.Heat1dPlan.from_bs_models([H.BsModel(implicit_scheme=L.THETA_30)])
    with pytest.raises(L.LssError):
        H.Heat1dPlan.from_bs_models([H.BsModel(n=1100, line_solver=L.LINE_PARTITIONED)])
    plan = H.Heat1dPlan.from_bs_models([H.BsModel()])
    with pytest.raises(L.LssError):
        plan.run_host([np.zeros(99)])
    plan.close()
