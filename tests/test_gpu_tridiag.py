"""GPU batched tridiagonal kernel (the S-line / v-line solver in isolation) against the oracle's Thomas LU.

Systems: a, c ~ U[-1, 0], b = 2.5 + |a| + |c|, d ~ N(0, 1), seed 1; a[0] and c[n-1] ignored
(Neumann(0) folding reduces the reference's Thomas LU to the plain algorithm).  Tolerance 1e-13
relative (FP64), 1e-5 (FP32).
"""
import numpy as np
import pytest

import linear_system_solvers_amd as L

pytestmark = pytest.mark.gpu


def _systems(B, n, dt, seed=1):
    rng = np.random.default_rng(seed)
    a = rng.uniform(-1, 0, (B, n))
    c = rng.uniform(-1, 0, (B, n))
    b = 2.5 + np.abs(a) + np.abs(c)
    d = rng.standard_normal((B, n))
    return [v.astype(dt) for v in (a, b, c, d)]


@pytest.mark.parametrize("fp", [L.FP64, L.FP32])
@pytest.mark.parametrize("layout", [L.LAYOUT_CONTIGUOUS, L.LAYOUT_STRIDED])
@pytest.mark.parametrize("B,n", [(1, 4), (37, 5), (1000, 256), (65536, 256)])
def test_tridiag_matches_oracle(oracle, fp, layout, B, n):
    torch = pytest.importorskip("torch")
    dt = L.np_dtype(fp)
    a, b, c, d = _systems(B, n, dt)
    if layout == L.LAYOUT_STRIDED:
        a, b, c, d = (np.ascontiguousarray(v.T) for v in (a, b, c, d))
    ta, tb, tc, td = (torch.from_numpy(v).cuda() for v in (a, b, c, d))
    tx = torch.empty_like(td)
    L.tridiag_batched(fp, layout, B, n, ta.data_ptr(), tb.data_ptr(), tc.data_ptr(), td.data_ptr(), tx.data_ptr(),
                      torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    x = tx.cpu().numpy()
    ref = oracle.thomas_batched(a, b, c, d, B, n, strided=(layout == L.LAYOUT_STRIDED),
                                fp="f64" if fp == L.FP64 else "f32").reshape(x.shape)
    err = np.abs(x.astype(np.float64) - ref).max() / np.abs(ref).max()
    print(f"B={B} n={n} layout={layout} fp={fp}: err {err:.2e} bit-identical {np.array_equal(x, ref)}", flush=True)
    assert err <= (1e-13 if fp == L.FP64 else 1e-5)


# ---------------------------------------------------------------------------------------------
# End conditions (lss_tridiag_batched_bc): the reference's known answers for u'' = -2 on [0, 1]
# (unit_tests/sparse_solvers/lss_thomas_lu_solver_t.hpp:12-243; central differences are exact for
# quadratics) and the oracle's thomas_lu with the same folding on random systems.
# ---------------------------------------------------------------------------------------------
def _solve_bc(fp, layout, a, b, c, d, lower, upper):
    torch = pytest.importorskip("torch")
    B, n = a.shape
    if layout == L.LAYOUT_STRIDED:
        a, b, c, d = (np.ascontiguousarray(v.T) for v in (a, b, c, d))
    ta, tb, tc, td = (torch.from_numpy(np.ascontiguousarray(v)).cuda() for v in (a, b, c, d))
    tx = torch.full_like(td, float("nan"))
    L.tridiag_batched_bc(fp, layout, B, n, ta.data_ptr(), tb.data_ptr(), tc.data_ptr(), td.data_ptr(), tx.data_ptr(),
                         lower, upper, torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    x = tx.cpu().numpy().astype(np.float64)
    return x.T if layout == L.LAYOUT_STRIDED else x


KAT = [  # (n, lower, upper, exact(t), f64 tolerance)
    (100, (L.BC_DIRICHLET, 0.0), (L.BC_DIRICHLET, 0.0), lambda t: t * (1 - t), 1e-10),
    (100, (L.BC_DIRICHLET, 1.0), (L.BC_ROBIN, 1.0, 0.0), lambda t: -t * t + t + 1, 1e-10),
    (1000, (L.BC_NEUMANN, -1.0), (L.BC_ROBIN, 2.0, 0.0), lambda t: -t * t + t + 0.5, 1e-8),
    (2000, (L.BC_NEUMANN, -1.0), (L.BC_ROBIN, 2.0, 0.0), lambda t: -t * t + t + 0.5, 1e-7),
]


@pytest.mark.parametrize("fp", [L.FP64, L.FP32])
@pytest.mark.parametrize("layout", [L.LAYOUT_CONTIGUOUS, L.LAYOUT_STRIDED])
@pytest.mark.parametrize("case", range(len(KAT)))
def test_end_conditions_reproduce_the_known_answers(fp, layout, case):
    n, lower, upper, exact, tol = KAT[case]
    dt = L.np_dtype(fp)
    B = 3
    h = dt(1.0) / dt(n - 1)
    a = np.ones((B, n), dt)
    b = np.full((B, n), -2.0, dt)
    c = np.ones((B, n), dt)
    d = np.full((B, n), dt(-2.0) * h * h, dt)
    x = _solve_bc(fp, layout, a, b, c, d, lower, upper)
    t = np.arange(n) / (n - 1)
    err = np.abs(x - exact(t)[None, :]).max()
    print(f"n={n} {lower} {upper} layout={layout} fp={fp}: max err {err:.2e}", flush=True)
    assert err <= (tol if fp == L.FP64 else 2e-1 if n >= 1000 else 2e-3)


@pytest.mark.parametrize("fp", [L.FP64, L.FP32])
@pytest.mark.parametrize("layout", [L.LAYOUT_CONTIGUOUS, L.LAYOUT_STRIDED])
@pytest.mark.parametrize("n", [5, 17, 256, 1500])
def test_end_conditions_match_the_oracle(oracle, fp, layout, n):
    rng = np.random.default_rng(n)
    dt = L.np_dtype(fp)
    ends = [((L.BC_DIRICHLET, 0.7), (L.BC_DIRICHLET, -0.3)), ((L.BC_NEUMANN, 0.4), (L.BC_ROBIN, 1.5, -0.2)),
            ((L.BC_ROBIN, 0.8, 0.1), (L.BC_NEUMANN, -0.6)), ((L.BC_DIRICHLET, 1.1), (L.BC_ROBIN, 0.5, 0.3))]
    ofp = "f64" if fp == L.FP64 else "f32"
    for lower, upper in ends:
        a, b, c, d = _systems(40, n, dt, seed=n + 1)
        x = _solve_bc(fp, layout, a, b, c, d, lower, upper)
        lo = (lower[0], lower[1] if len(lower) == 3 else 0.0, lower[-1])
        hi = (upper[0], upper[1] if len(upper) == 3 else 0.0, upper[-1])
        ref = np.stack([oracle.thomas_lu(a[k], b[k], c[k], d[k], lo, hi, ofp) for k in range(a.shape[0])])
        err = np.abs(x - ref).max() / np.abs(ref).max()
        print(f"n={n} {lower} {upper} layout={layout} fp={fp}: err {err:.2e}", flush=True)
        assert np.isfinite(x).all()
        assert err <= (1e-13 if fp == L.FP64 else 1e-5)


def test_end_condition_type_is_checked():
    torch = pytest.importorskip("torch")
    t = torch.zeros((2, 8), dtype=torch.float64, device="cuda")
    with pytest.raises(L.LssError):
        L.tridiag_batched_bc(L.FP64, L.LAYOUT_CONTIGUOUS, 2, 8, t.data_ptr(), t.data_ptr(), t.data_ptr(),
                             t.data_ptr(), t.data_ptr(), (7, 0.0), (L.BC_DIRICHLET, 0.0))
