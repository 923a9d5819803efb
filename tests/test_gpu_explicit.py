"""GPU tests of the explicit Euler solver (LSS_EXPLICIT_EULER) through the C-ABI.

The reference's explicit solver (explicit_solvers::general_heston_equation with the Euler scheme,
lss_heston_euler_solver_scheme.hpp:46-92) steps the interior with explicit_heston_scheme::rhs and
then runs the same explicit boundary solves A and B as the ADI schemes.  With FIELDS coefficients
the GPU path keeps the reference's operation order, so it must equal the oracle bit for bit; the
oracle itself is pinned to the reference's committed explicit surfaces
(tests/test_oracle_reference_outputs.py), and so is the GPU result here.
"""
import numpy as np
import pytest

import linear_system_solvers_amd as L
from tests.golden import reference_outputs as R
from tests.helpers import model_from_case, oracle_solve

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

EXPLICIT = [c for c in R.CASES if c["cfg"]["splitting"] == R.EXPLICIT_EULER]


@pytest.mark.parametrize("fp", [L.FP64, L.FP32])
@pytest.mark.parametrize("case", EXPLICIT, ids=[c["stem"] for c in EXPLICIT])
def test_fields_is_bit_identical_to_the_oracle_and_matches_the_reference(oracle, case, fp):
    m = model_from_case(case)
    plan = L.Plan.from_models([m], fp)
    g = plan.run_host()[0]
    plan.close()
    c = oracle_solve(oracle, m, "f64" if fp == L.FP64 else "f32")
    print(f"{case['stem']} fp {fp}: bit-identical {np.array_equal(g, c)}, "
          f"max diff {np.abs(g.astype(np.float64) - c).max():.3e}", flush=True)
    assert np.array_equal(g, c)
    _, _, vals = R.load(case["stem"], "f64" if fp == L.FP64 else "f32")
    if fp == L.FP64:
        assert np.all(np.abs(g - vals) <= 5.01e-6 * np.abs(vals) + 1e-12 * np.abs(vals).max())


def test_batched_explicit_problems_each_equal_their_oracle_solve(oracle):
    """Problems of different sizes, step counts, grids and models (Heston, and SABR with time-dependent
    coefficients) stepped together in one plan, FIELDS and SEPARABLE mixed."""
    rng = np.random.default_rng(21)
    base = [model_from_case(c) for c in EXPLICIT]
    models = []
    for q in range(9):
        m = model_from_case(EXPLICIT[q % len(EXPLICIT)])
        m.nx = int(rng.integers(8, 70))
        m.ny = int(rng.integers(6, 45))
        m.nt = int(rng.integers(20, 400))
        m.t_hi = m.t_hi * (m.nt - 1) / (base[q % len(base)].nt - 1)  # keep the reference cases' time step
        m.grid = L.NONUNIFORM if q % 3 else L.UNIFORM
        m.coeff_mode = L.COEFF_SEPARABLE if q % 2 else L.COEFF_FIELDS
        models.append(m)
    outs = L.Plan.from_models(models).run_host()
    for q, (m, g) in enumerate(zip(models, outs)):
        c = oracle_solve(oracle, m)
        norm = float(np.abs(g - c).max() / np.abs(c).max())
        print(f"explicit problem {q}: {m.nx}x{m.ny}x{m.steps} kind {m.kind} mode {m.coeff_mode}: normwise {norm:.2e}, "
              f"bit-identical {np.array_equal(g, c)}", flush=True)
        if m.coeff_mode == L.COEFF_FIELDS:
            assert np.array_equal(g, c)
        else:
            assert norm <= 1e-12


def test_graph_replay_equals_direct_launches():
    m = model_from_case(EXPLICIT[0], nt=301, coeff_mode=L.COEFF_SEPARABLE)
    m.t_hi = m.t_hi * 300 / 7999
    plan = L.Plan.from_models([m])
    u0 = torch.from_numpy(plan.initial(0)).cuda()
    outs = []
    for graph in (False, True, True):
        o = torch.empty_like(u0)
        plan.run_dev([u0.data_ptr()], [o.data_ptr()], graph=graph)
        torch.cuda.synchronize()
        outs.append(o.cpu().numpy())
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])


def test_explicit_and_adi_problems_do_not_share_a_plan_and_explicit_plans_have_no_adjoint():
    e = model_from_case(EXPLICIT[0], nt=11)
    a = model_from_case(EXPLICIT[0], nt=11, splitting=L.DOUGLAS_RACHFORD)
    with pytest.raises(L.LssError):
        L.Plan.from_models([e, a])
    plan = L.Plan.from_models([e])
    u0 = torch.from_numpy(plan.initial(0)).cuda()
    o = torch.empty_like(u0)
    with pytest.raises(L.LssError):
        plan.forward([u0.data_ptr()], [o.data_ptr()], ckpt_every=4)


@pytest.mark.parametrize("fp", [L.FP64, L.FP32])
@pytest.mark.parametrize("mode", [L.COEFF_FIELDS, L.COEFF_SEPARABLE])
def test_resident_loop_equals_the_step_by_step_kernel(oracle, fp, mode):
    """A plan of small explicit problems with time-independent data runs the whole loop in one launch,
    the time levels resident in LDS; lss_adi_run_levels steps the same plan launch by launch.  Both
    give the same bits, at the residency limits (1536 points, ny = 256, nx = 4) too, and equal the
    oracle (bit for bit with FIELDS)."""
    base = model_from_case(EXPLICIT[0])
    dt = base.t_hi / (base.nt - 1)
    for nx, ny, nt, dirn in [(50, 30, 401, L.BACKWARD), (4, 30, 101, L.BACKWARD), (6, 256, 21, L.BACKWARD),
                             (48, 32, 301, L.BACKWARD), (13, 7, 1000, L.BACKWARD), (50, 30, 201, L.FORWARD)]:
        m = model_from_case(EXPLICIT[0], nx=nx, ny=ny, nt=nt, coeff_mode=mode, direction=dirn)
        m.t_hi = m.t_lo + dt * (nt - 1)
        plan = L.Plan.from_models([m], fp)
        assert plan.kernel_config()["explicit_resident"] == 1
        g = plan.run_host()[0]
        lv = plan.run_levels_host()[0]
        plan.close()
        same = np.array_equal(g, lv[-1], equal_nan=True)
        print(f"explicit {nx}x{ny}x{nt - 1} dir {dirn} fp {fp} mode {mode}: resident equals step-by-step {same}",
              flush=True)
        assert same
        if (nx, ny) in ((50, 30), (13, 7)):  # (both directions of the reference's case)
            c = oracle_solve(oracle, m, "f64" if fp == L.FP64 else "f32")
            assert np.isfinite(c).all()
            if mode == L.COEFF_FIELDS:
                assert np.array_equal(g, c)
            else:
                assert np.abs(g - c).max() <= (1e-12 if fp == L.FP64 else 1e-5) * np.abs(c).max()


def test_plans_outside_the_residency_limits_step_launch_by_launch(oracle):
    """More than 1536 points, or coefficients sampled per step (SABR's D(t)), keep the per-step kernel."""
    big = model_from_case(EXPLICIT[0], nx=40, ny=40, nt=51)
    big.t_hi = big.t_hi * 50 / 7999
    sabr = [model_from_case(c, nt=51) for c in EXPLICIT if c["model"][0] == "sabr"][0]
    sabr.t_hi = sabr.t_hi * 50 / 7999
    for m in (big, sabr):
        plan = L.Plan.from_models([m])
        assert plan.kernel_config()["explicit_resident"] == 0
        g = plan.run_host()[0]
        plan.close()
        assert np.array_equal(g, oracle_solve(oracle, m))
