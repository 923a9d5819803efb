"""The drop-in's additions beside solve(container_2d&), run from reference-style C++ (tests/cpp/dropin_extras.cpp):
every time level into a container_3d, the batched_heston_book class and the heston_adjoint pair."""
import os
import subprocess

import numpy as np
import pytest

import linear_system_solvers_amd as L

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "dropin_extras")


def _run(*args):
    if not os.path.exists(BIN):
        pytest.fail("tests/cpp/dropin_extras not built (run __graft_entry__.build())")
    out = subprocess.run([BIN, *args], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    return out.stdout.split("\n")


def _blocks(lines):
    blocks, cur, name = {}, None, None
    for ln in lines:
        if ln.startswith("LAYER") or ln == "FINAL":
            name = ln
            cur = blocks.setdefault(name, [])
        elif ln.strip():
            cur.append([float(v) for v in ln.split(",")])
    return {k: np.array(v) for k, v in blocks.items()}


@pytest.mark.parametrize("fp", ["double", "float"])
def test_every_time_level_matches_oracle_steps(oracle, fp):
    """solve(container_3d&) on a Backward Heston DR problem: the last layer is the payoff, layer
    nt - 1 - s the surface after s loop steps (the oracle's loop body applied s times), layer 0 the
    final surface of solve(container_2d&)."""
    b = _blocks(_run("levels", fp))
    nt = 31
    layers = [b[f"LAYER {l}"] for l in range(nt)]
    assert np.array_equal(layers[0], b["FINAL"])
    dt = np.float64 if fp == "double" else np.float32
    ofp = "f64" if fp == "double" else "f32"
    pde = oracle.heston_pde(oracle.Heston(strike=100.0, rate=0.03, sig_sig=0.3, sig_kappa=2.0, sig_theta=0.2,
                                          rho=0.2))
    cfg = oracle.Cfg(x_lo=0.0, x_hi=200.0, y_lo=0.0, y_hi=1.2, nx=48, ny=24, nt=nt, t_hi=0.25,
                     accumulation_point=100.0, alpha_scale=3.0, beta_scale=15.0, nonuniform=True,
                     splitting=oracle.DOUGLAS_RACHFORD, weighting=0.5, implicit_scheme=oracle.CRANK_NICOLSON,
                     tridiagonal=oracle.THOMAS_LU, direction=oracle.BACKWARD)
    _, _, xs, ys, times = oracle.grid(cfg, ofp)
    payoff = np.maximum(0.0, xs.astype(np.float64) - 100.0)[:, None] * np.ones((1, len(ys)))
    assert np.array_equal(layers[nt - 1].astype(dt), payoff.astype(dt))
    prev = layers[nt - 1].astype(dt)
    for s, t in enumerate(times):
        prev = oracle.step(pde, cfg, prev, float(t), ofp)
        got = layers[nt - 2 - s].astype(dt)
        assert np.array_equal(got, prev), (s, np.abs(got.astype(np.float64) - prev).max())
    assert np.array_equal(prev, oracle.solve(pde, cfg, ofp))


def test_batched_book_class_equals_single_solves():
    out = [ln.split() for ln in _run("book") if ln.startswith("problem")]
    assert len(out) == 3
    for row in out:
        print(" ".join(row), flush=True)
        assert float(row[3]) == 0.0      # exact lines: the single solves' bits
        assert float(row[5]) <= 1e-12    # partitioned lines: within 1e-12 normwise


def test_adjoint_class_equals_the_library_adjoint():
    """heston_adjoint (C++, host pointers) against lss_adi_forward / lss_adi_backward through the Python
    binding on the same problem (device pointers): the same kernels, so the same numbers."""
    torch = pytest.importorskip("torch")
    lines = _run("adjoint")
    price = float(next(ln for ln in lines if ln.startswith("price")).split()[1])
    grad = [float(v) for v in next(ln for ln in lines if ln.startswith("grad")).split()[1:]]
    k = min(0.01, 1.5 / 23)
    m = L.Model(kind=L.MODEL_HESTON, strike=100.0, rate=0.02, sig_sig=0.8, sig_kappa=3.0, sig_theta=0.2, rho=0.2,
                maturity=k * 12, x_lo=0.0, x_hi=200.0, y_lo=0.0, y_hi=1.2, t_lo=0.0, t_hi=k * 12, nx=40, ny=24, nt=13,
                grid=L.NONUNIFORM, accumulation_point=100.0, alpha_scale=3.0, beta_scale=50.0,
                splitting=L.MODIFIED_CRAIG_SNEYD, weighting=0.5, implicit_scheme=L.CRANK_NICOLSON,
                tridiagonal=L.THOMAS_LU, direction=L.BACKWARD, coeff_mode=L.COEFF_SEPARABLE,
                line_solver=L.LINE_PARTITIONED)
    plan = L.Plan.from_models([m])
    xs, ys = plan.grid(0)
    w = L.bilinear_weights(xs, ys, 100.0, 0.04)
    u0 = torch.from_numpy(plan.initial(0)).cuda()
    out = torch.empty_like(u0)
    plan.forward([u0.data_ptr()], [out.data_ptr()], 4)
    lam = torch.from_numpy(w).cuda()
    dp = plan.backward([lam.data_ptr()])
    torch.cuda.synchronize()
    ref_price = float(np.sum(w * out.cpu().numpy()))
    print(f"C++ price {price:.17g} python {ref_price:.17g}; C++ grad {grad}; python {dp[0, :5]}", flush=True)
    assert price == pytest.approx(ref_price, rel=1e-14)
    np.testing.assert_allclose(grad, dp[0, :5], rtol=1e-13, atol=1e-15)
    plan.close()


def _source_pde(O):
    """The data of dropin_extras.cpp's heston(..., with_source = true), in the same operation order."""
    rate, sig_sig, sig_kappa, sig_theta, rho, strike = 0.03, 0.3, 2.0, 0.2, -0.4, 100.0
    return O.custom_pde(lambda t, s, v: (0.5 * v * s * s), lambda t, s, v: (0.5 * sig_sig * sig_sig * v),
                        lambda t, s, v: (rho * sig_sig * v * s), lambda t, s, v: (rate * s),
                        lambda t, s, v: (sig_kappa * (sig_theta - v)), lambda t, s, v: (-rate),
                        lambda s, v: max(0.0, s - strike), lambda t, v: 0.0, lambda t, s: -1.0, lambda t, s: s,
                        src=lambda t, s, v: (0.05 * v * s * (1.0 + t)))


@pytest.mark.parametrize("fp", ["double", "float"])
@pytest.mark.parametrize("direction", ["fwd", "bwd"])
@pytest.mark.parametrize("scheme", [0, 1, 2, 3])
def test_heat_source_matches_oracle_bit_for_bit(oracle, scheme, direction, fp):
    """general_heston_equation::solve with a heat source: the first stage of each scheme adds
    theta rho F(t_next) + (1 - theta) rho F(t) (rhs_intermed_1_source); the exact lines give the
    oracle's bits."""
    tri = "ds" if scheme % 2 else "tlu"
    got = np.array([[float(v) for v in ln.split(",")] for ln in _run("source", str(scheme), direction, fp, tri)
                    if ln.strip()])
    O = oracle
    cfg = O.Cfg(x_lo=0.0, x_hi=200.0, y_lo=0.0, y_hi=1.2, nx=40, ny=20, nt=17, t_hi=0.25, accumulation_point=100.0,
                alpha_scale=3.0, beta_scale=15.0, nonuniform=True, splitting=scheme, weighting=0.5,
                implicit_scheme=O.CRANK_NICOLSON, tridiagonal=O.DOUBLE_SWEEP if tri == "ds" else O.THOMAS_LU,
                direction=O.FORWARD if direction == "fwd" else O.BACKWARD)
    ofp = "f64" if fp == "double" else "f32"
    want = O.solve(_source_pde(O), cfg, ofp)
    dt = np.float64 if fp == "double" else np.float32
    assert np.array_equal(got.astype(dt), want), np.abs(got - want.astype(np.float64)).max()
    # the source matters: without it the surface differs
    hom = O.solve(O.heston_pde(O.Heston(strike=100.0, rate=0.03, sig_sig=0.3, sig_kappa=2.0, sig_theta=0.2,
                                        rho=-0.4)), cfg, ofp)
    assert np.abs(want.astype(np.float64) - hom).max() > 1e-3


def test_heat_source_partitioned_book_matches_single_solves():
    out = [ln.split() for ln in _run("source_book") if ln.startswith("problem")]
    assert len(out) == 4
    for row in out:
        print(" ".join(row), flush=True)
        assert float(row[3]) <= 1e-12
