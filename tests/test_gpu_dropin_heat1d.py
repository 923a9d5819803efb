"""The C++ 1-D drop-in (include/lss/heat_equation_1d.hpp) running reference-style user code on the GPU.

tests/cpp/dropin_heat1d.cpp writes each case as the reference's print tests write it and prints in the
reference's 1-D formats.  Printed with 6 digits it must match the reference's committed outputs;
printed with round-trip precision it must equal the oracle's result bit for bit (the drop-in samples the
closures in the reference's order and runs the exact kernel in the selected host solver's order).
"""
import os
import subprocess

import numpy as np
import pytest

from tests.golden import reference_outputs as R

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "dropin_heat1d")

# case -> reference output written by the same test code
STEMS = {0: "ImplBlackScholesEquationDirichletBCThomasLUSolverEuler",
         1: "ImplBlackScholesEquationDirichletBCThomasLUSolverCrankNicolson",
         2: "ImplBlackScholesEquationDirichletBCThomasLUSolverEulerSurf",
         3: "ImplBlackScholesEquationDirichletBCThomasLUSolverSurf",
         4: "ImplPureHeatEquationDirichletBCCUDASolverDeviceQREulerSurf",
         5: "ImplPureHeatEquationDirichletBCCUDASolverDeviceQRCrankNicolsonSurf",
         6: "ImplAdvDiffEquationDirichletBCThomasLUSolverCN"}


def _run(case, fp, full=False):
    if not os.path.exists(BIN):
        pytest.fail("tests/cpp/dropin_heat1d not built (run __graft_entry__.build())")
    out = subprocess.run([BIN, str(case), fp] + (["full"] if full else []), capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr
    lines = [ln for ln in out.stdout.split("\n") if ln.strip()]
    assert lines[0] == "SPACE_POINTS"
    vi = lines.index("VALUES")
    vals = np.array([[float(v) for v in row.split(",")] for row in lines[vi + 1:]])
    return vals[0] if vals.shape[0] == 1 else vals


def _oracle_case(O, case, fp):
    if case in STEMS:
        c = next(c for c in R.CASES1D if c["stem"] == STEMS[case])
        from tests.test_oracle_heat1d import _pde
        out, surf = O.heat1d_solve(_pde(O, c), O.Heat1dCfg(**c["cfg"]), fp, surfaces=True)
        return surf if c["surf"] else out
    # cases 7 and 8 of dropin_heat1d.cpp
    a = lambda t, x: 1.0 + 0.5 * x  # noqa: E731
    b = lambda t, x: 0.3  # noqa: E731
    c = lambda t, x: -0.2  # noqa: E731
    init = lambda x: x * x  # noqa: E731
    if case == 7:
        pde = O.custom_heat1d_pde(a, b, c, init, (O.BC_ROBIN, lambda t: 0.5, lambda t: -0.1 * t),
                                  (O.BC_NEUMANN, lambda t: 0.3), src=lambda t, x: t * x * (1.0 - x))
        cfg = O.Heat1dCfg(x_lo=0.0, x_hi=1.0, n=60, nt=80, t_hi=0.5, accumulation_point=0.5, alpha_scale=3.0,
                          nonuniform=True, implicit_scheme=O.CRANK_NICOLSON, tridiagonal=O.DOUBLE_SWEEP,
                          direction=O.FORWARD)
        return O.heat1d_solve(pde, cfg, fp, surfaces=True)[1]
    pde = O.custom_heat1d_pde(a, b, c, init, (O.BC_NEUMANN, lambda t: -0.4),
                              (O.BC_ROBIN, lambda t: 0.8, lambda t: 0.25 * t))
    cfg = O.Heat1dCfg(x_lo=0.0, x_hi=1.0, n=60, nt=80, t_hi=0.5, accumulation_point=0.5, alpha_scale=3.0,
                      nonuniform=True, implicit_scheme=O.EULER, tridiagonal=O.THOMAS_LU, direction=O.BACKWARD)
    return O.heat1d_solve(pde, cfg, fp)


@pytest.mark.parametrize("case", sorted(STEMS))
def test_dropin_matches_the_reference_outputs(case):
    for fp, rtol in (("double", 5.01e-6), ("float", None)):
        got = _run(case, fp)
        _, _, ref = R.load1d(STEMS[case], "approx", "f64" if fp == "double" else "f32")
        assert got.shape == ref.shape
        if rtol is not None:
            assert np.all(np.abs(got - ref) <= rtol * np.abs(ref) + 1e-13 * np.abs(ref).max())
        else:
            assert np.abs(got - ref).max() <= 2e-5 * np.abs(ref).max()


@pytest.mark.parametrize("fp", ["double", "float"])
@pytest.mark.parametrize("case", range(9))
def test_dropin_full_precision_equals_oracle(oracle, case, fp):
    got = _run(case, fp, full=True)
    o = _oracle_case(oracle, case, "f64" if fp == "double" else "f32")
    got = got.astype(np.float64 if fp == "double" else np.float32)
    print(f"case {case} {fp}: max |drop-in - oracle| {np.abs(got.astype(np.float64) - o).max():.3e}", flush=True)
    assert got.shape == o.shape
    assert np.array_equal(got, o.astype(got.dtype))
