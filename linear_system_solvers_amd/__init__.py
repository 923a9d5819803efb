"""linear_system_solvers_amd -- MI355X (gfx950) Heston/SABR ADI time-stepper.

Python front end of liblss_amd.so through its C-ABI (include/lss_amd.h).  Every compute call goes
to the HIP library; there is no CPU path here.  If the library has not been built, importing works
but every entry point raises LssUnavailable.

Reference-API meaning of the model fields: see ``Model`` (pde_discretization_config_2d,
grid_config_hints_2d, splitting_method_config and heat_implicit_solver_config of
MichalSara99/linear_system_solvers).
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import os
from typing import Optional, Sequence

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "liblss_amd.so")

FP64, FP32 = 0, 1
DOUGLAS_RACHFORD, CRAIG_SNEYD, MODIFIED_CRAIG_SNEYD, HUNDSDORFER_VERWER = 0, 1, 2, 3
EXPLICIT_EULER = 4  # the explicit solver (Euler scheme) instead of an ADI splitting
EULER, THETA_30, CRANK_NICOLSON, THETA_80 = 0, 1, 2, 3
CUDA_SOLVER, DOUBLE_SWEEP, SOR_SOLVER, THOMAS_LU = 0, 1, 2, 3
FORWARD, BACKWARD = 0, 1
UNIFORM, NONUNIFORM = 0, 1
COEFF_FIELDS, COEFF_SEPARABLE = 0, 1
LINE_EXACT, LINE_PARTITIONED = 0, 1
LAYOUT_CONTIGUOUS, LAYOUT_STRIDED = 0, 1
BC_DIRICHLET, BC_NEUMANN, BC_ROBIN = 0, 1, 2
MODEL_HESTON, MODEL_SABR = 0, 1

# every symbol include/lss_amd.h declares
EXPORTS = ("lss_last_error", "lss_version", "lss_adi_plan_create", "lss_adi_plan_destroy", "lss_adi_run_dev",
           "lss_adi_run_host", "lss_model_plan_create", "lss_model_initial", "lss_model_grid", "lss_plan_problem_shape",
           "lss_tridiag_batched", "lss_adi_run_profiled", "lss_adi_set_param_terms", "lss_plan_param_count",
           "lss_adi_forward", "lss_adi_backward", "lss_adi_autotune",
           "lss_adi_set_mapping", "lss_plan_kernel_config", "lss_tridiag_batched_bc", "lss_adi_set_adjoint_overlap",
           "lss_tridiag_batched_noalloc", "lss_adi_run_levels", "lss_adi_forward_host", "lss_adi_backward_host")
TRIDIAG_SCRATCH_BYTES = 64 << 20  # static scratch of lss_tridiag_batched_noalloc (LSS_TRIDIAG_SCRATCH_BYTES)
MAX_PARAMS = 6
MAX_TERMS = 12
HESTON_PARAMS = ("kappa", "theta_v", "sigma", "rho", "r")
# SABR: nu, rho, beta and "alpha_axis", the scale lam of the alpha axis about y_lo (grid and price point
# stretched together): dPrice/dalpha0 = d/dlam / alpha0 when y_lo = 0 and y_hi is a multiple of alpha0
SABR_PARAMS = ("nu", "rho", "beta", "alpha_axis")
KERNEL_CONFIG_LEN = 13
KERNEL_KINDS = ("bnd_a", "xsweep1", "vsweep1", "hv_y3", "xsweep2", "vsweep2", "bnd_b", "euler")


class LssUnavailable(RuntimeError):
    pass


class LssError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"lss error {code}: {msg}")
        self.code = code


class _Model(C.Structure):
    _fields_ = [("kind", C.c_int32), ("pad", C.c_int32),
                *[(n, C.c_double) for n in ("strike", "rate", "sig_sig", "sig_kappa", "sig_theta", "rho", "beta",
                                            "maturity", "bnd_low", "bnd_high", "x_lo", "x_hi", "y_lo", "y_hi",
                                            "t_lo", "t_hi")],
                ("nx", C.c_int64), ("ny", C.c_int64), ("nt", C.c_int64),
                ("accumulation_point", C.c_double), ("alpha_scale", C.c_double), ("beta_scale", C.c_double),
                ("grid", C.c_int32), ("splitting", C.c_int32), ("weighting", C.c_double),
                ("implicit_scheme", C.c_int32), ("tridiagonal", C.c_int32), ("direction", C.c_int32),
                ("coeff_mode", C.c_int32), ("line_solver", C.c_int32)]


class _Desc(C.Structure):
    _fields_ = [("fp", C.c_int32), ("splitting", C.c_int32), ("tridiagonal", C.c_int32), ("coeff_mode", C.c_int32),
                ("line_solver", C.c_int32), ("pad", C.c_int32), ("nx", C.c_int64), ("ny", C.c_int64), ("steps", C.c_int64),
                *[(n, C.c_double) for n in ("theta", "weighting", "alpha", "beta", "gamma", "delta", "ni", "rho",
                                            "h1")],
                ("idx1", C.c_void_p), ("idx2", C.c_void_p), ("coeff_sets", C.c_int64), ("coeff", C.c_void_p),
                ("bnd_sets", C.c_int64), ("bnd", C.c_void_p), ("src_sets", C.c_int64), ("src", C.c_void_p)]


@dataclasses.dataclass
class Model:
    """One Heston or SABR problem with the reference tests' data (see include/lss_amd.h, lss_model)."""
    kind: int = MODEL_HESTON
    strike: float = 100.0
    rate: float = 0.03
    sig_sig: float = 0.3
    sig_kappa: float = 2.0
    sig_theta: float = 0.2
    rho: float = 0.2
    beta: float = 0.7
    maturity: float = 1.0
    bnd_low: float = 0.0
    bnd_high: float = -1.0
    x_lo: float = 50.0
    x_hi: float = 200.0
    y_lo: float = 0.0
    y_hi: float = 1.2
    t_lo: float = 0.0
    t_hi: float = 1.0
    nx: int = 100
    ny: int = 50
    nt: int = 101
    accumulation_point: float = 0.0
    alpha_scale: float = 3.0
    beta_scale: float = 50.0
    grid: int = UNIFORM
    splitting: int = DOUGLAS_RACHFORD
    weighting: float = 0.5
    implicit_scheme: int = CRANK_NICOLSON
    tridiagonal: int = THOMAS_LU
    direction: int = BACKWARD
    coeff_mode: int = COEFF_FIELDS
    line_solver: int = LINE_EXACT

    def c(self) -> _Model:
        m = _Model()
        for f in dataclasses.fields(self):
            setattr(m, f.name, getattr(self, f.name))
        return m

    @property
    def steps(self) -> int:
        return self.nt - 1


_lib = None


def available() -> bool:
    return os.path.exists(LIB_PATH)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LssUnavailable(f"{LIB_PATH} is not built (run __graft_entry__.build())")
        # The library and PyTorch both need libamdhip64.so.7; whichever is loaded first serves the
        # whole process.  Load torch's first when torch is present so both share one HIP runtime.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        L.lss_last_error.restype = C.c_char_p
        L.lss_adi_plan_create.argtypes = [C.POINTER(_Desc), C.c_int64, C.POINTER(C.c_void_p)]
        L.lss_adi_plan_destroy.argtypes = [C.c_void_p]
        L.lss_adi_run_dev.argtypes = [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.c_void_p,
                                      C.c_int32]
        L.lss_adi_run_host.argtypes = [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)]
        L.lss_model_plan_create.argtypes = [C.POINTER(_Model), C.c_int64, C.c_int32, C.POINTER(C.c_void_p)]
        L.lss_model_initial.argtypes = [C.c_void_p, C.c_int64, C.c_void_p]
        L.lss_model_grid.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]
        L.lss_plan_problem_shape.argtypes = [C.c_void_p, C.c_int64, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                             C.POINTER(C.c_int64)]
        L.lss_adi_run_profiled.argtypes = [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.c_void_p,
                                           C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_int64)]
        L.lss_tridiag_batched.argtypes = [C.c_int32, C.c_int32, C.c_int64, C.c_int64, C.c_void_p, C.c_void_p,
                                          C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.lss_tridiag_batched_bc.argtypes = [C.c_int32, C.c_int32, C.c_int64, C.c_int64, C.c_void_p, C.c_void_p,
                                             C.c_void_p, C.c_void_p, C.POINTER(_LineBC), C.POINTER(_LineBC),
                                             C.c_void_p, C.c_void_p]
        L.lss_tridiag_batched_noalloc.argtypes = [C.c_int32, C.c_int32, C.c_int64, C.c_int64, C.c_void_p,
                                                  C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(_LineBC),
                                                  C.POINTER(_LineBC), C.c_void_p, C.c_void_p]
        L.lss_adi_set_param_terms.argtypes = [C.c_void_p, C.c_int64, C.c_int32, C.c_int32, C.POINTER(C.c_int32),
                                              C.POINTER(C.c_int32), C.c_int64, C.c_void_p]
        L.lss_plan_param_count.argtypes = [C.c_void_p, C.c_int64, C.POINTER(C.c_int32)]
        L.lss_adi_set_mapping.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32]
        L.lss_adi_set_adjoint_overlap.argtypes = [C.c_void_p, C.c_int32]
        L.lss_plan_kernel_config.argtypes = [C.c_void_p, C.POINTER(C.c_int32)]
        L.lss_adi_autotune.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.POINTER(C.c_int32),
                                       C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
        L.lss_adi_forward.argtypes = [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.c_int32,
                                      C.c_void_p]
        L.lss_adi_backward.argtypes = [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                       C.POINTER(C.c_double), C.c_void_p]
        L.lss_adi_run_levels.argtypes = [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.c_void_p,
                                         C.c_int32]
        L.lss_adi_forward_host.argtypes = [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.c_int32]
        L.lss_adi_backward_host.argtypes = [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                            C.POINTER(C.c_double)]
        _lib = L
    return _lib


def _check(rc: int) -> None:
    if rc != 0:
        raise LssError(rc, lib().lss_last_error().decode())


def np_dtype(fp: int):
    return np.float64 if fp == FP64 else np.float32


class Plan:
    """A set of ADI problems stepped together on the GPU (one problem: the single-problem stepper)."""

    def __init__(self, handle: int, fp: int, count: int):
        self._h = C.c_void_p(handle)
        self.fp = fp
        self.count = count

    @classmethod
    def from_models(cls, models: Sequence[Model], fp: int = FP64) -> "Plan":
        arr = (_Model * len(models))(*[m.c() for m in models])
        h = C.c_void_p()
        _check(lib().lss_model_plan_create(arr, len(models), fp, C.byref(h)))
        return cls(h.value, fp, len(models))

    def close(self) -> None:
        if self._h is not None and self._h.value:
            lib().lss_adi_plan_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def shape(self, p: int):
        nx, ny, st = C.c_int64(), C.c_int64(), C.c_int64()
        _check(lib().lss_plan_problem_shape(self._h, p, C.byref(nx), C.byref(ny), C.byref(st)))
        return nx.value, ny.value, st.value

    def initial(self, p: int) -> np.ndarray:
        nx, ny, _ = self.shape(p)
        out = np.zeros((nx, ny), dtype=np_dtype(self.fp))
        _check(lib().lss_model_initial(self._h, p, out.ctypes.data))
        return out

    def grid(self, p: int):
        """Physical coordinates (xs, ys) of problem p of a model plan."""
        nx, ny, _ = self.shape(p)
        xs = np.zeros(nx, dtype=np_dtype(self.fp))
        ys = np.zeros(ny, dtype=np_dtype(self.fp))
        _check(lib().lss_model_grid(self._h, p, xs.ctypes.data, ys.ctypes.data))
        return xs, ys

    def run_host(self, inits: Optional[Sequence[np.ndarray]] = None) -> list:
        """Synchronous solve from host arrays (default: the model plan's sampled initial surfaces)."""
        dt = np_dtype(self.fp)
        if inits is None:
            inits = [self.initial(p) for p in range(self.count)]
        if len(inits) != self.count:
            raise LssError(-1, f"run_host needs {self.count} surfaces, got {len(inits)}")
        for p, a in enumerate(inits):
            nx, ny, _ = self.shape(p)
            if np.shape(a) != (nx, ny):
                raise LssError(-1, f"surface {p} has shape {np.shape(a)}, the plan's problem is {(nx, ny)}")
        ins = [np.ascontiguousarray(a, dtype=dt) for a in inits]
        outs = [np.zeros_like(a) for a in ins]
        pin = (C.c_void_p * self.count)(*[a.ctypes.data for a in ins])
        pout = (C.c_void_p * self.count)(*[a.ctypes.data for a in outs])
        _check(lib().lss_adi_run_host(self._h, pin, pout))
        return outs

    def run_levels_host(self, inits: Optional[Sequence[np.ndarray]] = None) -> list:
        """Every time level from host arrays: per problem an array (steps + 1, nx, ny), row s = the surface
        after s steps (row 0 = the initial surface)."""
        dt = np_dtype(self.fp)
        if inits is None:
            inits = [self.initial(p) for p in range(self.count)]
        shapes = [self.shape(p) for p in range(self.count)]
        if len(inits) != self.count or any(np.shape(a) != sh[:2] for a, sh in zip(inits, shapes)):
            raise LssError(-1, "run_levels_host needs one surface of the problem's shape per problem")
        ins = [np.ascontiguousarray(a, dtype=dt) for a in inits]
        lv = [np.zeros((st + 1, nx, ny), dtype=dt) for nx, ny, st in shapes]
        pin = (C.c_void_p * self.count)(*[a.ctypes.data for a in ins])
        plv = (C.c_void_p * self.count)(*[a.ctypes.data for a in lv])
        _check(lib().lss_adi_run_levels(self._h, pin, plv, None, 1))
        return lv

    def run_levels_dev(self, in_ptrs: Sequence[int], level_ptrs: Sequence[int], stream: int = 0) -> None:
        """Enqueue a solve that writes every level ((steps + 1) * nx * ny values per problem, device)."""
        pin = (C.c_void_p * self.count)(*in_ptrs)
        plv = (C.c_void_p * self.count)(*level_ptrs)
        _check(lib().lss_adi_run_levels(self._h, pin, plv, C.c_void_p(stream), 0))

    def run_dev(self, in_ptrs: Sequence[int], out_ptrs: Sequence[int], stream: int = 0, graph: bool = True) -> None:
        """Enqueue a solve on `stream` from/to device pointers (e.g. torch tensors' data_ptr())."""
        pin = (C.c_void_p * self.count)(*in_ptrs)
        pout = (C.c_void_p * self.count)(*out_ptrs)
        _check(lib().lss_adi_run_dev(self._h, pin, pout, C.c_void_p(stream), int(graph)))


    def autotune(self, stream: int = 0, again: bool = False):
        """Tile/lane mapping chosen by the plan's autotuner:
        (S-lines per x-sweep group, v-lines per v-sweep group, XCD-contiguous tile order)."""
        x, v, z = C.c_int32(), C.c_int32(), C.c_int32()
        _check(lib().lss_adi_autotune(self._h, C.c_void_p(stream), int(again), C.byref(x), C.byref(v), C.byref(z)))
        return x.value, v.value, z.value

    def kernel_config(self) -> dict:
        c = (C.c_int32 * KERNEL_CONFIG_LEN)()
        _check(lib().lss_plan_kernel_config(self._h, c))
        keys = ("line_solver", "s_lines_per_group", "v_lines_per_group", "xcd_order", "factor_pattern",
                "s_chunk_bound", "v_chunk_bound", "tuned", "adjoint_pattern", "adjoint_s_lines_per_group", "adjoint_overlap", "graph_groups",
                "explicit_resident")
        return dict(zip(keys, list(c)))

    def set_mapping(self, s_lines_per_group: int, v_lines_per_group: int, xcd_order: int = 0) -> None:
        _check(lib().lss_adi_set_mapping(self._h, s_lines_per_group, v_lines_per_group, xcd_order))

    def set_adjoint_overlap(self, mode: int) -> None:
        """-1 automatic (default), 0 never, 1 always: recomputation of the next segment on a second
        stream during the backward pass; takes effect at the next forward()."""
        _check(lib().lss_adi_set_adjoint_overlap(self._h, mode))

    # ---- reverse mode -------------------------------------------------------------------------
    def param_count(self, p: int) -> int:
        n = C.c_int32()
        _check(lib().lss_plan_param_count(self._h, p, C.byref(n)))
        return n.value

    def set_param_terms(self, p: int, nparam: int, term_param: Sequence[int], term_coeff: Sequence[int],
                        factors: np.ndarray) -> None:
        """Derivative terms of problem p: d q_{term_coeff[t]}(i, j) / d param_{term_param[t]} = X_t[i] Y_t[j];
        factors has shape (sets, nterm, nx + ny)."""
        nt = len(term_param)
        tp = (C.c_int32 * max(nt, 1))(*term_param)
        tq = (C.c_int32 * max(nt, 1))(*term_coeff)
        f = np.ascontiguousarray(factors, dtype=np_dtype(self.fp))
        sets = f.shape[0] if nt else 1
        _check(lib().lss_adi_set_param_terms(self._h, p, nparam, nt, tp, tq, sets, f.ctypes.data))

    def forward(self, in_ptrs: Sequence[int], out_ptrs: Sequence[int], ckpt_every: int = 32,
                stream: int = 0) -> None:
        """Time loop from/to device pointers that keeps every ckpt_every-th level for backward()."""
        pin = (C.c_void_p * self.count)(*in_ptrs)
        pout = (C.c_void_p * self.count)(*out_ptrs)
        _check(lib().lss_adi_forward(self._h, pin, pout, int(ckpt_every), C.c_void_p(stream)))

    def backward(self, lam_out_ptrs: Sequence[int], lam_in_ptrs: Optional[Sequence[int]] = None,
                 stream: int = 0) -> np.ndarray:
        """Adjoint of the last forward(): lam_out = dJ/du_out (device), writes dJ/du_in into lam_in_ptrs
        (device, entries may be 0 to skip) and returns dJ/dparams, shape (count, MAX_PARAMS)."""
        pl = (C.c_void_p * self.count)(*lam_out_ptrs)
        pi = (C.c_void_p * self.count)(*lam_in_ptrs) if lam_in_ptrs is not None else None
        dp = (C.c_double * (self.count * MAX_PARAMS))()
        _check(lib().lss_adi_backward(self._h, pl, pi, dp, C.c_void_p(stream)))
        return np.frombuffer(dp, dtype=np.float64).reshape(self.count, MAX_PARAMS).copy()


def bilinear_weights(xs: np.ndarray, ys: np.ndarray, x0: float, y0: float) -> np.ndarray:
    """Weights w with sum(w * U) = the bilinear interpolant of the surface U (rows xs, columns ys) at
    (x0, y0); also the adjoint seed dJ/dU of that price."""
    i = int(np.clip(np.searchsorted(xs, x0, side="right") - 1, 0, len(xs) - 2))
    j = int(np.clip(np.searchsorted(ys, y0, side="right") - 1, 0, len(ys) - 2))
    wx = (x0 - xs[i]) / (xs[i + 1] - xs[i])
    wy = (y0 - ys[j]) / (ys[j + 1] - ys[j])
    w = np.zeros((len(xs), len(ys)), dtype=np.float64)
    w[i, j] = (1 - wx) * (1 - wy)
    w[i + 1, j] = wx * (1 - wy)
    w[i, j + 1] = (1 - wx) * wy
    w[i + 1, j + 1] = wx * wy
    return w


def run_profiled(plan: "Plan", in_ptrs: Sequence[int], out_ptrs: Sequence[int], stream: int = 0) -> dict:
    """Per-kernel-kind event timing of one solve: {kind: (ms, algorithmic_bytes, launches)}."""
    n = len(KERNEL_KINDS)
    ms, by, nl = (C.c_double * n)(), (C.c_double * n)(), (C.c_int64 * n)()
    pin = (C.c_void_p * plan.count)(*in_ptrs)
    pout = (C.c_void_p * plan.count)(*out_ptrs)
    _check(lib().lss_adi_run_profiled(plan._h, pin, pout, C.c_void_p(stream), ms, by, nl))
    return {k: (ms[i], by[i], nl[i]) for i, k in enumerate(KERNEL_KINDS)}


def tridiag_batched(fp: int, layout: int, B: int, n: int, a: int, b: int, c: int, d: int, x: int,
                    stream: int = 0) -> None:
    """Solve B plain tridiagonal systems of length n (device pointers; a[0], c[n-1] ignored)."""
    _check(lib().lss_tridiag_batched(fp, layout, B, n, C.c_void_p(a), C.c_void_p(b), C.c_void_p(c),
                                     C.c_void_p(d), C.c_void_p(x), C.c_void_p(stream)))


class _LineBC(C.Structure):
    _fields_ = [("type", C.c_int32), ("lin", C.c_double), ("val", C.c_double)]


def tridiag_batched_bc(fp: int, layout: int, B: int, n: int, a: int, b: int, c: int, d: int, x: int,
                       lower: tuple, upper: tuple, stream: int = 0) -> None:
    """Solve B tridiagonal systems of length n (device pointers) with the end conditions folded in as
    the reference's thomas_lu_solver does; lower/upper = (BC_DIRICHLET, val), (BC_NEUMANN, val) or
    (BC_ROBIN, lin, val)."""
    def mk(e):
        return _LineBC(int(e[0]), float(e[1]) if len(e) == 3 else 0.0, float(e[-1]))
    lo, hi = mk(lower), mk(upper)
    _check(lib().lss_tridiag_batched_bc(fp, layout, B, n, C.c_void_p(a), C.c_void_p(b), C.c_void_p(c),
                                        C.c_void_p(d), C.byref(lo), C.byref(hi), C.c_void_p(x),
                                        C.c_void_p(stream)))


def tridiag_batched_noalloc(fp: int, layout: int, B: int, n: int, a: int, b: int, c: int, d: int, x: int,
                            lower: Optional[tuple] = None, upper: Optional[tuple] = None, stream: int = 0) -> None:
    """lss_tridiag_batched (lower = upper = None) or lss_tridiag_batched_bc without any device allocation:
    the one-lane-per-system kernel takes its scratch from the library's static per-device array
    (TRIDIAG_SCRATCH_BYTES).  Calls that overlap on one device race on that array."""
    def mk(e):
        return _LineBC(int(e[0]), float(e[1]) if len(e) == 3 else 0.0, float(e[-1]))
    lo = C.byref(mk(lower)) if lower is not None else None
    hi = C.byref(mk(upper)) if upper is not None else None
    _check(lib().lss_tridiag_batched_noalloc(fp, layout, B, n, C.c_void_p(a), C.c_void_p(b), C.c_void_p(c),
                                             C.c_void_p(d), lo, hi, C.c_void_p(x), C.c_void_p(stream)))
