"""lss_tridiag_batched_noalloc (static per-device scratch, no device allocation) against the allocating
entry points lss_tridiag_batched / lss_tridiag_batched_bc, through the C-ABI.

The same kernels run in both, so the solutions must be bit-identical.  Two calls in a row use the
shared scratch array, the second with a smaller input; both run the one-lane-per-system kernel (the
only one that uses scratch).  Sizes beyond the static maximum are rejected on the host.
"""
import numpy as np
import pytest

import linear_system_solvers_amd as L

pytestmark = pytest.mark.gpu


def _systems(B, n, dt, seed):
    rng = np.random.default_rng(seed)
    a = rng.uniform(-1, 0, (B, n))
    c = rng.uniform(-1, 0, (B, n))
    b = 2.5 + np.abs(a) + np.abs(c)
    d = rng.standard_normal((B, n))
    return [v.astype(dt) for v in (a, b, c, d)]


def _run(fp, layout, B, n, seed, lower=None, upper=None):
    torch = pytest.importorskip("torch")
    dt = L.np_dtype(fp)
    a, b, c, d = _systems(B, n, dt, seed)
    if layout == L.LAYOUT_STRIDED:
        a, b, c, d = (np.ascontiguousarray(v.T) for v in (a, b, c, d))
    ta, tb, tc, td = (torch.from_numpy(v).cuda() for v in (a, b, c, d))
    ptrs = [t.data_ptr() for t in (ta, tb, tc, td)]
    stream = torch.cuda.current_stream().cuda_stream
    x_ref = torch.full_like(td, float("nan"))
    x_new = torch.full_like(td, float("nan"))
    if lower is None:
        L.tridiag_batched(fp, layout, B, n, *ptrs, x_ref.data_ptr(), stream)
    else:
        L.tridiag_batched_bc(fp, layout, B, n, *ptrs, x_ref.data_ptr(), lower, upper, stream)
    L.tridiag_batched_noalloc(fp, layout, B, n, *ptrs, x_new.data_ptr(), lower, upper, stream)
    torch.cuda.synchronize()
    return x_ref.cpu().numpy(), x_new.cpu().numpy()


def test_noalloc_matches_the_allocating_entry_points_on_two_calls():
    # call 1: 2048 contiguous fp64 systems of 3000 unknowns (47 MiB of scratch), plain ends
    print("noalloc call 1: fp64 contiguous B=2048 n=3000", flush=True)
    ref, new = _run(L.FP64, L.LAYOUT_CONTIGUOUS, 2048, 3000, seed=11)
    assert np.isfinite(ref).all()
    assert np.array_equal(ref, new)
    # call 2, smaller: 600 strided fp32 systems of 1100 unknowns with Dirichlet / Robin ends
    print("noalloc call 2: fp32 strided B=600 n=1100 Dirichlet/Robin", flush=True)
    ref, new = _run(L.FP32, L.LAYOUT_STRIDED, 600, 1100, seed=12, lower=(L.BC_DIRICHLET, 0.5),
                    upper=(L.BC_ROBIN, 0.3, -0.2))
    assert np.isfinite(ref).all()
    assert np.array_equal(ref, new)


def test_noalloc_partitioned_sizes_match():
    # systems that fit the partitioned kernels use no scratch at all; results equal as well
    print("noalloc partitioned: fp64 strided B=4096 n=256", flush=True)
    ref, new = _run(L.FP64, L.LAYOUT_STRIDED, 4096, 256, seed=13, lower=(L.BC_NEUMANN, 0.1),
                    upper=(L.BC_DIRICHLET, 1.0))
    assert np.array_equal(ref, new)


def test_noalloc_rejects_sizes_beyond_the_static_scratch():
    # checked on the host before anything is launched
    B, n = 4096, 4096  # 128 MiB of fp64 scratch
    assert B * n * 8 > L.TRIDIAG_SCRATCH_BYTES
    with pytest.raises(L.LssError, match="exceeds its static maximum"):
        L.tridiag_batched_noalloc(L.FP64, L.LAYOUT_CONTIGUOUS, B, n, 1, 1, 1, 1, 1)
