"""1-D implicit heat / Black-Scholes solver of liblss_amd.so (include/lss_heat1d.h) through its C-ABI.

It replaces the device branch of the reference's one_dimensional::implicit_solvers::general_heat_equation
(pde_solvers/one_dimensional/heat_type/lss_1d_general_heat_equation.hpp:64-137).  Every compute call
goes to the HIP library.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
from typing import Optional, Sequence

import numpy as np

from . import (BACKWARD, CRANK_NICOLSON, FP64, LINE_EXACT, THOMAS_LU, UNIFORM, LssError, _check, lib,
               np_dtype)

EXPORTS = ("lss_heat1d_plan_create", "lss_heat1d_plan_destroy", "lss_heat1d_run_dev", "lss_heat1d_run_host",
           "lss_heat1d_problem_shape", "lss_bs_plan_create", "lss_heat1d_initial", "lss_heat1d_grid")


class _BsModel(C.Structure):
    _fields_ = [*[(n, C.c_double) for n in ("strike", "rate", "sig", "maturity", "x_lo", "x_hi", "t_lo", "t_hi")],
                ("n", C.c_int64), ("nt", C.c_int64), ("accumulation_point", C.c_double),
                ("alpha_scale", C.c_double), ("grid", C.c_int32), ("implicit_scheme", C.c_int32),
                ("tridiagonal", C.c_int32), ("direction", C.c_int32), ("line_solver", C.c_int32),
                ("pad", C.c_int32)]


@dataclasses.dataclass
class BsModel:
    """Black-Scholes call of the reference's 1-D tests: a = sig^2 x^2 / 2, b = r x, c = -r, terminal
    max(0, x - K), Dirichlet 0 at x_lo and x_hi - K exp(-r (T - t)) at x_hi; discretization, grid hints
    and solver config in the reference-API meaning (pde_discretization_config_1d, grid_config_hints_1d,
    heat_implicit_solver_config)."""
    strike: float = 10.0
    rate: float = 0.2
    sig: float = 0.25
    maturity: float = 1.0
    x_lo: float = 0.0
    x_hi: float = 20.0
    t_lo: float = 0.0
    t_hi: float = 1.0
    n: int = 100
    nt: int = 100
    accumulation_point: float = 0.0
    alpha_scale: float = 3.0
    grid: int = UNIFORM
    implicit_scheme: int = CRANK_NICOLSON
    tridiagonal: int = THOMAS_LU
    direction: int = BACKWARD
    line_solver: int = LINE_EXACT

    def c(self) -> _BsModel:
        return _BsModel(self.strike, self.rate, self.sig, self.maturity, self.x_lo, self.x_hi, self.t_lo, self.t_hi,
                        self.n, self.nt, self.accumulation_point, self.alpha_scale, self.grid, self.implicit_scheme,
                        self.tridiagonal, self.direction, self.line_solver, 0)

    @property
    def steps(self) -> int:
        return self.nt - 1


def _lib():
    L = lib()
    if not getattr(L, "_h1_ready", False):
        P = C.POINTER(C.c_void_p)
        L.lss_heat1d_plan_destroy.argtypes = [C.c_void_p]
        L.lss_heat1d_run_dev.argtypes = [C.c_void_p, P, P, P, C.c_void_p]
        L.lss_heat1d_run_host.argtypes = [C.c_void_p, P, P, P]
        L.lss_heat1d_problem_shape.argtypes = [C.c_void_p, C.c_int64, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.lss_bs_plan_create.argtypes = [C.POINTER(_BsModel), C.c_int64, C.c_int32, C.POINTER(C.c_void_p)]
        L.lss_heat1d_initial.argtypes = [C.c_void_p, C.c_int64, C.c_void_p]
        L.lss_heat1d_grid.argtypes = [C.c_void_p, C.c_int64, C.c_void_p]
        L._h1_ready = True
    return L


class Heat1dPlan:
    """Independent 1-D problems stepped together on the GPU (one launch for the whole time loop)."""

    def __init__(self, handle: int, fp: int, count: int):
        self._h = C.c_void_p(handle)
        self.fp = fp
        self.count = count

    @classmethod
    def from_bs_models(cls, models: Sequence[BsModel], fp: int = FP64) -> "Heat1dPlan":
        arr = (_BsModel * len(models))(*[m.c() for m in models])
        h = C.c_void_p()
        _check(_lib().lss_bs_plan_create(arr, len(models), fp, C.byref(h)))
        return cls(h.value, fp, len(models))

    def close(self) -> None:
        if self._h is not None and self._h.value:
            _lib().lss_heat1d_plan_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def shape(self, p: int):
        n, st = C.c_int64(), C.c_int64()
        _check(_lib().lss_heat1d_problem_shape(self._h, p, C.byref(n), C.byref(st)))
        return n.value, st.value

    def initial(self, p: int) -> np.ndarray:
        n, _ = self.shape(p)
        out = np.zeros(n, dtype=np_dtype(self.fp))
        _check(_lib().lss_heat1d_initial(self._h, p, out.ctypes.data))
        return out

    def grid(self, p: int) -> np.ndarray:
        n, _ = self.shape(p)
        out = np.zeros(n, dtype=np_dtype(self.fp))
        _check(_lib().lss_heat1d_grid(self._h, p, out.ctypes.data))
        return out

    def run_host(self, inits: Optional[Sequence[np.ndarray]] = None, surfaces: bool = False):
        """Synchronous solve from host arrays; returns the final levels, or (final levels, surfaces of shape
        (steps + 1, n)) with surfaces=True."""
        dt = np_dtype(self.fp)
        if inits is None:
            inits = [self.initial(p) for p in range(self.count)]
        if len(inits) != self.count:
            raise LssError(-1, f"run_host needs {self.count} levels, got {len(inits)}")
        shapes = [self.shape(p) for p in range(self.count)]
        for p, a in enumerate(inits):
            if np.shape(a) != (shapes[p][0],):
                raise LssError(-1, f"level {p} has shape {np.shape(a)}, the plan's problem has {shapes[p][0]} points")
        ins = [np.ascontiguousarray(a, dtype=dt) for a in inits]
        outs = [np.zeros_like(a) for a in ins]
        surf = [np.zeros((st + 1, n), dtype=dt) for n, st in shapes] if surfaces else None
        pin = (C.c_void_p * self.count)(*[a.ctypes.data for a in ins])
        pout = (C.c_void_p * self.count)(*[a.ctypes.data for a in outs])
        psf = (C.c_void_p * self.count)(*[a.ctypes.data for a in surf]) if surfaces else None
        _check(_lib().lss_heat1d_run_host(self._h, pin, pout, psf))
        return (outs, surf) if surfaces else outs

    def run_dev(self, in_ptrs: Sequence[int], out_ptrs: Sequence[int], surf_ptrs: Optional[Sequence[int]] = None,
                stream: int = 0) -> None:
        """Enqueue a solve on `stream` from/to device pointers (surf_ptrs: (steps + 1) * n values each)."""
        pin = (C.c_void_p * self.count)(*in_ptrs)
        pout = (C.c_void_p * self.count)(*out_ptrs)
        psf = (C.c_void_p * self.count)(*surf_ptrs) if surf_ptrs is not None else None
        _check(_lib().lss_heat1d_run_dev(self._h, pin, pout, psf, C.c_void_p(stream)))
