"""CPU tests of bench.py's book check (check_all): how it classifies a surface against the oracle.

The fast path's surfaces are replaced here by the oracle's own surfaces, changed on purpose, so
that each branch of the classification is exercised without a GPU.
"""
import dataclasses

import numpy as np

import bench
import linear_system_solvers_amd as L
from oracle import oracle as O
from tests.helpers import config0, oracle_solve


def _models():
    stable = config0(nx=24, ny=12, nt=11)
    # a step far past the explicit v = 0 row's limit: the solution grows without bound
    unstable = config0(nx=24, ny=40, nt=11, t_hi=3.0, splitting=L.CRAIG_SNEYD)
    return [stable, unstable]


def test_check_all_classifies_each_case():
    models = _models()
    ref = [oracle_solve(O, m) for m in models]
    assert np.abs(ref[1]).max() > 10 * models[1].x_hi, "the second problem must be unstable"
    exact = [r.copy() for r in ref]
    outs = [r * (1 + 1e-12) for r in ref]
    res = bench.check_all(models, outs, lambda s: None, exact_outs=exact, threads=2)
    assert res["exact_path_bit_identical"] == 2 and res["failed"] == 0
    assert res["stable"]["problems"] == 1 and res["stable"]["within_tolerance"] == 1
    assert res["unstable"]["problems"] == 1


def test_check_all_flags_an_exact_path_difference_and_a_wrong_stable_surface():
    models = _models()
    ref = [oracle_solve(O, m) for m in models]
    exact = [r.copy() for r in ref]
    exact[0][3, 3] = np.nextafter(exact[0][3, 3], np.inf)  # one ulp off
    outs = [r.copy() for r in ref]
    outs[0] = outs[0] * (1 + 1e-8)  # a stable problem off by 1e-8: far above its sensitivity
    res = bench.check_all(models, outs, lambda s: None, exact_outs=exact, threads=2)
    assert res["exact_path_bit_identical"] == 1 and res["exact_path_mismatch_sample"] == [0]
    assert res["failed"] == 1 and res["failed_sample"][0][0] == 0


def test_check_all_accepts_a_difference_within_the_reference_solvers_own_spread():
    models = _models()
    ref = [oracle_solve(O, m) for m in models]
    # the unstable problem solved with the reference's other tridiagonal solver: a legitimate
    # rounding difference, amplified by the problem itself
    other = oracle_solve(O, dataclasses.replace(models[1], tridiagonal=L.DOUBLE_SWEEP))
    res = bench.check_all(models, [ref[0], other], lambda s: None, threads=2)
    assert res["failed"] == 0
