"""The reference's committed ADI surfaces and the exact configuration of the test that wrote each.

Each entry gives the file stem under tests/golden/reference_outputs/, the writing test
(file:line in the reference), the PDE data and the discretization/scheme, in the reference-API
meaning used by oracle.Cfg and by the drop-in (linear_system_solvers_amd.heston.Problem).
"""
import os

import numpy as np

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_outputs")

DR, CS, MCS, HV = 0, 1, 2, 3
EXPLICIT_EULER = 4  # explicit_solvers::general_heston_equation with the Euler scheme
EULER, THETA_30, CN, THETA_80 = 0, 1, 2, 3
DOUBLE_SWEEP, THOMAS_LU = 1, 3

# model: ("heston", dict(strike, rate, sig_sig, sig_kappa, sig_theta, rho)) or ("sabr", dict(...))
CASES = [
    dict(stem="ImplHestonEquationThomasLUSolverDouglasRachfordCrankNicolson",
         where="unit_tests/common/lss_print_t.hpp:2104-2210",
         model=("heston", dict(strike=100.0, rate=0.03, sig_sig=0.3, sig_kappa=2.0, sig_theta=0.2, rho=0.2)),
         cfg=dict(x_lo=50.0, x_hi=150.0, y_lo=0.0, y_hi=1.2, nx=150, ny=50, nt=200, nonuniform=False,
                  accumulation_point=0.0, alpha_scale=3.0, beta_scale=50.0, splitting=DR, weighting=0.5,
                  implicit_scheme=CN, tridiagonal=THOMAS_LU),
         floats=True),
    dict(stem="ImplHestonEquationThomasLUSolverCrankNicolson",
         where="unit_tests/common/lss_print_t.hpp:1850-1962",
         model=("heston", dict(strike=100.0, rate=0.03, sig_sig=0.3, sig_kappa=2.0, sig_theta=0.2, rho=0.2)),
         cfg=dict(x_lo=30.0, x_hi=200.0, y_lo=0.0, y_hi=1.2, nx=150, ny=50, nt=200, nonuniform=True,
                  accumulation_point=100.0, alpha_scale=3.0, beta_scale=50.0, splitting=DR, weighting=0.5,
                  implicit_scheme=CN, tridiagonal=THOMAS_LU),
         floats=True),
    dict(stem="ImplHestonEquationThomasLUSolverCraigSneydCrankNicolson",
         where="unit_tests/common/lss_print_t.hpp:2224-2333",
         model=("heston", dict(strike=100.0, rate=0.03, sig_sig=0.3, sig_kappa=2.0, sig_theta=0.2, rho=0.2)),
         cfg=dict(x_lo=50.0, x_hi=200.0, y_lo=0.0, y_hi=1.2, nx=150, ny=50, nt=200, nonuniform=True,
                  accumulation_point=100.0, alpha_scale=3.0, beta_scale=50.0, splitting=CS, weighting=0.8,
                  implicit_scheme=THETA_80, tridiagonal=THOMAS_LU),
         floats=True),
    dict(stem="ImplHestonEquationThomasLUSolverModCraigSneydCrankNicolson",
         where="unit_tests/common/lss_print_t.hpp:2347-2456",
         model=("heston", dict(strike=100.0, rate=0.03, sig_sig=0.3, sig_kappa=2.0, sig_theta=0.2, rho=0.2)),
         cfg=dict(x_lo=50.0, x_hi=180.0, y_lo=0.0, y_hi=1.2, nx=150, ny=50, nt=200, nonuniform=True,
                  accumulation_point=100.0, alpha_scale=3.0, beta_scale=50.0, splitting=MCS, weighting=0.5,
                  implicit_scheme=CN, tridiagonal=THOMAS_LU),
         floats=True),
    dict(stem="ImplHestonEquationThomasLUSolverHundsdorferVerwerCrankNicolson",
         where="unit_tests/common/lss_print_t.hpp:2470-2579",
         model=("heston", dict(strike=100.0, rate=0.03, sig_sig=0.3, sig_kappa=2.0, sig_theta=0.2, rho=0.2)),
         cfg=dict(x_lo=50.0, x_hi=200.0, y_lo=0.0, y_hi=1.2, nx=150, ny=50, nt=200, nonuniform=True,
                  accumulation_point=100.0, alpha_scale=3.0, beta_scale=50.0, splitting=HV, weighting=0.5,
                  implicit_scheme=CN, tridiagonal=THOMAS_LU),
         floats=False),
    dict(stem="ImplHestonEquationCUDAQRSolverCrankNicolson",
         where="unit_tests/common/lss_print_t.hpp:1724-1836 (Device cuSOLVER QR; LU-equivalent at 6 digits)",
         model=("heston", dict(strike=10.0, rate=0.03, sig_sig=0.3, sig_kappa=2.0, sig_theta=0.2, rho=0.2)),
         cfg=dict(x_lo=0.0, x_hi=20.0, y_lo=0.0, y_hi=1.0, nx=100, ny=50, nt=150, nonuniform=True,
                  accumulation_point=10.0, alpha_scale=10.0 / 3.0, beta_scale=1.0, splitting=DR, weighting=0.5,
                  implicit_scheme=CN, tridiagonal=THOMAS_LU),
         floats=True),
    dict(stem="ImplHestonEquationDSSolverCrankNicolsonFromBuilder",
         where="unit_tests/builders/lss_2d_general_heston_equation_builder_t.hpp:19-163",
         model=("heston", dict(strike=10.0, rate=0.03, sig_sig=0.3, sig_kappa=2.0, sig_theta=0.2, rho=0.2)),
         cfg=dict(x_lo=0.0, x_hi=20.0, y_lo=0.0, y_hi=1.5, nx=100, ny=50, nt=150, nonuniform=True,
                  accumulation_point=10.0, alpha_scale=3.0, beta_scale=60.0, splitting=DR, weighting=0.5,
                  implicit_scheme=CN, tridiagonal=DOUBLE_SWEEP),
         floats=True),
    dict(stem="ImplSABREquationDoubleSweepSolverCrankNicolson",
         where="unit_tests/common/lss_print_t.hpp:1976-2090",
         model=("sabr", dict(strike=100.0, maturity=1.0, rate=0.03, sig_sig=0.081, rho=0.6, beta=0.7)),
         cfg=dict(x_lo=50.0, x_hi=200.0, y_lo=0.0, y_hi=1.2, nx=100, ny=50, nt=200, nonuniform=True,
                  accumulation_point=100.0, alpha_scale=3.0, beta_scale=50.0, splitting=DR, weighting=0.5,
                  implicit_scheme=CN, tridiagonal=DOUBLE_SWEEP),
         floats=True),
    # explicit Euler (host_expl_bwd_euler_solver_config); the reference's Device runs of the first and
    # third case (ExplHestonEquationSolverEulerDevice, ExplSABREquationEulerOnDevice) printed the
    # same digits as these Host runs
    dict(stem="ExplHestonEquationSolverEulerHost",
         where="unit_tests/common/lss_print_t.hpp:2599-2703",
         model=("heston", dict(strike=100.0, rate=0.03, sig_sig=0.3, sig_kappa=2.0, sig_theta=0.2, rho=0.8)),
         cfg=dict(x_lo=50.0, x_hi=200.0, y_lo=0.0, y_hi=1.0, nx=50, ny=30, nt=8000, nonuniform=True,
                  accumulation_point=100.0, alpha_scale=3.0, beta_scale=50.0, splitting=EXPLICIT_EULER,
                  weighting=0.5, implicit_scheme=CN, tridiagonal=THOMAS_LU),
         floats=True),
    dict(stem="ExplHestonEquationEulerHostFromBuilder",
         where="unit_tests/builders/lss_2d_general_heston_equation_builder_t.hpp:171-310",
         model=("heston", dict(strike=10.0, rate=0.03, sig_sig=0.3, sig_kappa=2.0, sig_theta=0.2, rho=0.8)),
         cfg=dict(x_lo=0.0, x_hi=20.0, y_lo=0.0, y_hi=1.5, nx=50, ny=30, nt=8000, nonuniform=True,
                  accumulation_point=10.0, alpha_scale=3.0, beta_scale=60.0, splitting=EXPLICIT_EULER,
                  weighting=0.5, implicit_scheme=CN, tridiagonal=THOMAS_LU),
         floats=True),
    dict(stem="ExplSABREquationEulerOnHost",
         where="unit_tests/common/lss_print_t.hpp:2835-2944",
         model=("sabr", dict(strike=100.0, maturity=1.0, rate=0.03, sig_sig=0.081, rho=0.6, beta=0.7)),
         cfg=dict(x_lo=50.0, x_hi=200.0, y_lo=0.0, y_hi=1.2, nx=50, ny=40, nt=8000, nonuniform=True,
                  accumulation_point=100.0, alpha_scale=3.0, beta_scale=50.0, splitting=EXPLICIT_EULER,
                  weighting=0.5, implicit_scheme=CN, tridiagonal=THOMAS_LU),
         floats=True),
]


def load(stem: str, fp: str = "f64"):
    """(xs, ys, values) parsed from the lss_print text format."""
    path = os.path.join(HERE, f"{stem}_{'double' if fp == 'f64' else 'float'}.txt")
    with open(path) as fh:
        lines = fh.read().split("\n")
    assert lines[0] == "SPACE_POINTS_X" and lines[2] == "SPACE_POINTS_Y" and lines[4] == "VALUES"
    xs = np.array([float(v) for v in lines[1].split(",")])
    ys = np.array([float(v) for v in lines[3].split(",")])
    vals = np.array([[float(v) for v in row.split(",")] for row in lines[5:] if row.strip()])
    return xs, ys, vals


# ---------------------------------------------------------------------------------------------
# One-dimensional implicit heat equation (general_heat_equation): the reference's 1-D print tests.
# Each writes <stem>_approx_<type>.txt (the solver's final level, or every level for "Surf") and
# <stem>_bench_<type>.txt (the closed form: black_scholes_exact or the heat series).
# model "bs": a = 0.5 sig^2 x^2, b = r x, c = -r, terminal max(0, x - K), Dirichlet 0 and
#   x_hi - K exp(-r (T - t)); "heat": a = 1, b = 0, c = 0, initial x, Dirichlet 0, 0;
#   "advdiff": a = 1, b = -1, c = 0, initial 1, Dirichlet 0, 0.
# ---------------------------------------------------------------------------------------------
FORWARD, BACKWARD = 0, 1
BS = dict(strike=10.0, rate=0.2, sig=0.25, maturity=1.0)
CASES1D = [
    dict(stem="ImplBlackScholesEquationDirichletBCThomasLUSolverEuler", where="unit_tests/common/lss_print_t.hpp:105-207",
         model="bs", surf=False,
         cfg=dict(x_lo=0.0, x_hi=20.0, n=100, nt=100, t_hi=1.0, accumulation_point=0.0, alpha_scale=3.0,
                  nonuniform=False, implicit_scheme=EULER, tridiagonal=THOMAS_LU, direction=BACKWARD)),
    dict(stem="ImplBlackScholesEquationDirichletBCThomasLUSolverCrankNicolson",
         where="unit_tests/common/lss_print_t.hpp:209-311", model="bs", surf=False,
         cfg=dict(x_lo=0.0, x_hi=20.0, n=100, nt=100, t_hi=1.0, accumulation_point=0.0, alpha_scale=3.0,
                  nonuniform=False, implicit_scheme=CN, tridiagonal=THOMAS_LU, direction=BACKWARD)),
    dict(stem="ImplBlackScholesEquationDirichletBCThomasLUSolverEulerSurf",
         where="unit_tests/common/lss_print_t.hpp:327-437", model="bs", surf=True,
         cfg=dict(x_lo=0.0, x_hi=20.0, n=100, nt=100, t_hi=1.0, accumulation_point=10.0, alpha_scale=3.0,
                  nonuniform=False, implicit_scheme=EULER, tridiagonal=THOMAS_LU, direction=BACKWARD)),
    dict(stem="ImplBlackScholesEquationDirichletBCThomasLUSolverSurf",
         where="unit_tests/common/lss_print_t.hpp:439-552", model="bs", surf=True,
         cfg=dict(x_lo=0.0, x_hi=20.0, n=100, nt=100, t_hi=1.0, accumulation_point=10.0, alpha_scale=3.0,
                  nonuniform=True, implicit_scheme=CN, tridiagonal=THOMAS_LU, direction=BACKWARD)),
    # the reference ran these two through cuSOLVER QR on the device; the oracle runs Thomas LU
    dict(stem="ImplPureHeatEquationDirichletBCCUDASolverDeviceQREulerSurf",
         where="unit_tests/common/lss_print_t.hpp:568-683", model="heat", surf=True,
         cfg=dict(x_lo=0.0, x_hi=1.0, n=100, nt=100, t_hi=0.1, accumulation_point=0.0, alpha_scale=3.0,
                  nonuniform=False, implicit_scheme=EULER, tridiagonal=THOMAS_LU, direction=FORWARD)),
    dict(stem="ImplPureHeatEquationDirichletBCCUDASolverDeviceQRCrankNicolsonSurf",
         where="unit_tests/common/lss_print_t.hpp:685-802", model="heat", surf=True,
         cfg=dict(x_lo=0.0, x_hi=1.0, n=100, nt=100, t_hi=0.1, accumulation_point=0.5, alpha_scale=3.0,
                  nonuniform=True, implicit_scheme=CN, tridiagonal=THOMAS_LU, direction=FORWARD)),
    dict(stem="ImplAdvDiffEquationDirichletBCThomasLUSolverCN", where="unit_tests/common/lss_print_t.hpp:949-1075",
         model="advdiff", surf=True,
         cfg=dict(x_lo=0.0, x_hi=1.0, n=100, nt=100, t_hi=0.1, accumulation_point=0.5, alpha_scale=3.0,
                  nonuniform=True, implicit_scheme=CN, tridiagonal=THOMAS_LU, direction=FORWARD)),
]


def load1d(stem: str, kind: str = "approx", fp: str = "f64"):
    """(xs, times or None, values) of a 1-D print file: values (n,) or, with TIME_POINTS, (nt, n)."""
    path = os.path.join(HERE, f"{stem}_{kind}_{'double' if fp == 'f64' else 'float'}.txt")
    with open(path) as fh:
        lines = [ln for ln in fh.read().split("\n")]
    assert lines[0] == "SPACE_POINTS"
    xs = np.array([float(v) for v in lines[1].split(",")])
    if lines[2] == "TIME_POINTS":
        ts = np.array([float(v) for v in lines[3].split(",")])
        assert lines[4] == "VALUES"
        vals = np.array([[float(v) for v in row.split(",")] for row in lines[5:] if row.strip()])
        return xs, ts, vals
    assert lines[2] == "VALUES"
    return xs, None, np.array([float(v) for v in lines[3].split(",")])
