"""The C++ drop-in (include/lss/lss.hpp) running reference-style user code on the GPU.

tests/cpp/dropin_heston.cpp writes each case the way the reference's tests do and prints the surface in
the reference's format (6 significant digits).  The drop-in uses FIELDS coefficients and the exact
line order, which reproduce the reference bit for bit, so its printed surface must equal the oracle's
values printed the same way, and agree with the reference's committed surface to print precision.
"""
import os
import subprocess

import numpy as np
import pytest

from tests.golden import reference_outputs as R

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "dropin_heston")

# case -> reference surface written by the same test code
CASES = {0: "ImplHestonEquationDSSolverCrankNicolsonFromBuilder",
         1: "ImplHestonEquationThomasLUSolverDouglasRachfordCrankNicolson",
         2: "ImplSABREquationDoubleSweepSolverCrankNicolson",
         3: "ImplHestonEquationCUDAQRSolverCrankNicolson",
         4: "ImplHestonEquationThomasLUSolverHundsdorferVerwerCrankNicolson",
         5: "ExplHestonEquationSolverEulerHost"}


def _run(case, fp):
    out = subprocess.run([BIN, str(case), fp], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    lines = out.stdout.split("\n")
    assert lines[0] == "SPACE_POINTS_X" and lines[2] == "SPACE_POINTS_Y" and lines[4] == "VALUES"
    return np.array([[float(v) for v in r.split(",")] for r in lines[5:] if r.strip()])


@pytest.mark.parametrize("case", sorted(CASES))
def test_dropin_double_matches_reference_surface_and_oracle(oracle, case):
    if not os.path.exists(BIN):
        pytest.fail("tests/cpp/dropin_heston not built (run __graft_entry__.build())")
    got = _run(case, "double")
    stem = CASES[case]
    _, _, ref = R.load(stem, "f64")
    assert got.shape == ref.shape
    assert np.all(np.abs(got - ref) <= 5.01e-6 * np.abs(ref) + 1e-12 * np.abs(ref).max())
    c = next(c for c in R.CASES if c["stem"] == stem)
    from tests.helpers import model_from_case, oracle_solve
    o = oracle_solve(oracle, model_from_case(c))
    printed = np.vectorize(lambda v: float(format(v, ".6g")))(o)
    assert np.array_equal(got, printed)


@pytest.mark.parametrize("case", [0, 1, 2, 3, 5])
def test_dropin_float_matches_reference_surface(case):
    if not os.path.exists(BIN):
        pytest.fail("tests/cpp/dropin_heston not built (run __graft_entry__.build())")
    got = _run(case, "float")
    _, _, ref = R.load(CASES[case], "f32")
    steps = next(c for c in R.CASES if c["stem"] == CASES[case])["cfg"]["nt"] - 1
    # float rounding accumulates over the steps (tests/test_oracle_reference_outputs.py)
    assert np.abs(got - ref).max() <= 1.5e-5 * max(1.0, (steps / 200) ** 0.5) * np.abs(ref).max()


def _run_full(case, fp):
    out = subprocess.run([BIN, str(case), fp, "full"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    lines = out.stdout.split("\n")
    assert lines[4] == "VALUES"
    return np.array([[float(v) for v in r.split(",")] for r in lines[5:] if r.strip()])


@pytest.mark.parametrize("fp", ["double", "float"])
@pytest.mark.parametrize("case", sorted(CASES))
def test_dropin_full_precision_equals_oracle(oracle, case, fp):
    """The surface printed with round-trip precision (17 significant digits for double, 9 for float)
    equals the oracle's surface bit for bit: the drop-in's C++ sampling of the user's closures
    (FIELDS coefficients, boundary sets, index maps, loop times) and the exact line order reproduce
    the reference's host arithmetic."""
    if not os.path.exists(BIN):
        pytest.fail("tests/cpp/dropin_heston not built (run __graft_entry__.build())")
    got = _run_full(case, fp)
    c = next(c for c in R.CASES if c["stem"] == CASES[case])
    from tests.helpers import model_from_case, oracle_solve
    o = oracle_solve(oracle, model_from_case(c), "f64" if fp == "double" else "f32")
    got = got.astype(np.float64 if fp == "double" else np.float32)
    diff = np.abs(got.astype(np.float64) - o.astype(np.float64)).max()
    print(f"case {case} {fp}: max |drop-in - oracle| = {diff:.3e}", flush=True)
    assert got.shape == o.shape
    assert np.array_equal(got, o.astype(got.dtype))
