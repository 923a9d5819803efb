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
