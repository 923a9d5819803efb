"""ctypes front end of the CPU restatement in this directory (TEST INFRASTRUCTURE ONLY).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg import this module; the
product package never does.  It loads ``liblss_oracle.so`` (built by ``make -C oracle``) and exposes
the reference's Host ADI solve, single steps, grid data and line solvers on numpy arrays.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import os
import subprocess
from typing import Callable, Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblss_oracle.so")

# enum values of common/lss_enumerations.hpp
DOUGLAS_RACHFORD, CRAIG_SNEYD, MOD_CRAIG_SNEYD, HUNDSDORFER_VERWER = 0, 1, 2, 3
EXPLICIT_EULER = 4  # the explicit solver (Euler scheme), not a splitting method
EULER, THETA_30, CRANK_NICOLSON, THETA_80 = 0, 1, 2, 3
CUDA_SOLVER, DOUBLE_SWEEP, SOR_SOLVER, THOMAS_LU = 0, 1, 2, 3
FORWARD, BACKWARD = 0, 1
BC_DIRICHLET, BC_NEUMANN, BC_ROBIN = 0, 1, 2

COEF_FN = C.CFUNCTYPE(C.c_double, C.c_void_p, C.c_double, C.c_double, C.c_double)
INIT_FN = C.CFUNCTYPE(C.c_double, C.c_void_p, C.c_double, C.c_double)
BND_FN = C.CFUNCTYPE(C.c_double, C.c_void_p, C.c_double, C.c_double)


class _Pde(C.Structure):
    _fields_ = [("a", COEF_FN), ("b", COEF_FN), ("c", COEF_FN), ("d", COEF_FN), ("e", COEF_FN),
                ("f", COEF_FN), ("init", INIT_FN), ("h_lower", BND_FN), ("h_upper", BND_FN),
                ("v_upper", BND_FN), ("ctx", C.c_void_p), ("src", COEF_FN)]


class _Cfg(C.Structure):
    _fields_ = [("x_lo", C.c_double), ("x_hi", C.c_double), ("y_lo", C.c_double), ("y_hi", C.c_double),
                ("t_lo", C.c_double), ("t_hi", C.c_double), ("nx", C.c_int64), ("ny", C.c_int64),
                ("nt", C.c_int64), ("accumulation_point", C.c_double), ("alpha_scale", C.c_double),
                ("beta_scale", C.c_double), ("nonuniform", C.c_int32), ("splitting", C.c_int32),
                ("weighting", C.c_double), ("implicit_scheme", C.c_int32), ("tridiagonal", C.c_int32),
                ("direction", C.c_int32), ("cached", C.c_int32), ("coeff_time_independent", C.c_int32)]


class _Heston(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("strike", "rate", "sig_sig", "sig_kappa", "sig_theta", "rho", "bnd_low", "bnd_high")]


class _Sabr(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("strike", "maturity", "rate", "sig_sig", "rho", "beta", "bnd_low", "bnd_high")]


@dataclasses.dataclass
class Cfg:
    """Discretization, grid hints and scheme of one solve (reference-API meaning, see lss_oracle.h)."""
    x_lo: float
    x_hi: float
    y_lo: float
    y_hi: float
    nx: int
    ny: int
    nt: int
    t_lo: float = 0.0
    t_hi: float = 1.0
    accumulation_point: float = 0.0
    alpha_scale: float = 3.0
    beta_scale: float = 50.0
    nonuniform: bool = False
    splitting: int = DOUGLAS_RACHFORD
    weighting: float = 0.5
    implicit_scheme: int = CRANK_NICOLSON
    tridiagonal: int = THOMAS_LU
    direction: int = BACKWARD
    cached: bool = True
    coeff_time_independent: bool = False

    def c(self) -> _Cfg:
        return _Cfg(self.x_lo, self.x_hi, self.y_lo, self.y_hi, self.t_lo, self.t_hi, self.nx, self.ny, self.nt,
                    self.accumulation_point, self.alpha_scale, self.beta_scale, int(self.nonuniform),
                    self.splitting, self.weighting, self.implicit_scheme, self.tridiagonal, self.direction,
                    int(self.cached), int(self.coeff_time_independent))


@dataclasses.dataclass
class Heston:
    strike: float = 100.0
    rate: float = 0.03
    sig_sig: float = 0.3
    sig_kappa: float = 2.0
    sig_theta: float = 0.2
    rho: float = 0.2
    bnd_low: float = 0.0
    bnd_high: float = -1.0


@dataclasses.dataclass
class Sabr:
    strike: float = 100.0
    maturity: float = 1.0
    rate: float = 0.03
    sig_sig: float = 0.081
    rho: float = 0.6
    beta: float = 0.7
    bnd_low: float = 0.0
    bnd_high: float = -1.0


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        for sfx, T in (("f64", C.c_double), ("f32", C.c_float)):
            P = C.POINTER(T)
            getattr(L, f"lss_or_solve_{sfx}").argtypes = [C.POINTER(_Pde), C.POINTER(_Cfg), P]
            getattr(L, f"lss_or_step_{sfx}").argtypes = [C.POINTER(_Pde), C.POINTER(_Cfg), P, T, P]
            getattr(L, f"lss_or_grid_{sfx}").argtypes = [C.POINTER(_Cfg), C.POINTER(C.c_uint32),
                                                         C.POINTER(C.c_uint32), P, P, P]
            for nm in ("thomas_lu", "double_sweep"):
                getattr(L, f"lss_or_{nm}_{sfx}").argtypes = [C.c_int64, P, P, P, P, C.c_int32, T, T, C.c_int32,
                                                             T, T, P]
            getattr(L, f"lss_or_thomas_batched_{sfx}").argtypes = [C.c_int64, C.c_int64, C.c_int32, P, P, P, P, P]
        L.lss_or_heston_pde.argtypes = [C.POINTER(_Heston), C.POINTER(_Pde)]
        L.lss_or_set_parallel_min.argtypes = [C.c_int64]
        L.lss_or_sabr_pde.argtypes = [C.POINTER(_Sabr), C.POINTER(_Pde)]
        _lib = L
    return _lib


def set_parallel_min(points: int) -> None:
    """Grids of at least `points` points run the lines of each sweep on OpenMP threads."""
    lib().lss_or_set_parallel_min(int(points))


def _np_dtype(fp: str):
    return np.float64 if fp == "f64" else np.float32


def _ptr(a: np.ndarray, fp: str):
    return a.ctypes.data_as(C.POINTER(C.c_double if fp == "f64" else C.c_float))


class Pde:
    """Holds a lss_or_pde plus whatever keeps its callbacks and context alive."""

    def __init__(self, pde: _Pde, keep):
        self.pde = pde
        self.keep = keep


def heston_pde(p: Heston) -> Pde:
    h = _Heston(*dataclasses.astuple(p))
    pde = _Pde()
    lib().lss_or_heston_pde(C.byref(h), C.byref(pde))
    return Pde(pde, h)


def sabr_pde(p: Sabr) -> Pde:
    s = _Sabr(*dataclasses.astuple(p))
    pde = _Pde()
    lib().lss_or_sabr_pde(C.byref(s), C.byref(pde))
    return Pde(pde, s)


def custom_pde(a: Callable, b: Callable, c: Callable, d: Callable, e: Callable, f: Callable,
               init: Callable, h_lower: Callable, h_upper: Callable, v_upper: Callable,
               src: Optional[Callable] = None) -> Pde:
    """PDE data from Python callables (coefficients and the optional heat source take (t, x, y), init
    (x, y), boundaries (t, s))."""
    fns = [COEF_FN(lambda _c, t, x, y, g=g: float(g(t, x, y))) for g in (a, b, c, d, e, f)]
    fi = INIT_FN(lambda _c, x, y: float(init(x, y)))
    fb = [BND_FN(lambda _c, t, s, g=g: float(g(t, s))) for g in (h_lower, h_upper, v_upper)]
    fs = COEF_FN(lambda _c, t, x, y: float(src(t, x, y))) if src is not None else COEF_FN()
    pde = _Pde(*fns, fi, *fb, None, fs)
    return Pde(pde, (fns, fi, fb, fs))


def solve(pde: Pde, cfg: Cfg, fp: str = "f64") -> np.ndarray:
    out = np.zeros((cfg.nx, cfg.ny), dtype=_np_dtype(fp))
    rc = getattr(lib(), f"lss_or_solve_{fp}")(C.byref(pde.pde), C.byref(cfg.c()), _ptr(out, fp))
    if rc != 0:
        raise ValueError(f"oracle solve failed rc={rc}")
    return out


def step(pde: Pde, cfg: Cfg, prev: np.ndarray, time: float, fp: str = "f64") -> np.ndarray:
    prev = np.ascontiguousarray(prev, dtype=_np_dtype(fp))
    out = np.zeros_like(prev)
    rc = getattr(lib(), f"lss_or_step_{fp}")(C.byref(pde.pde), C.byref(cfg.c()), _ptr(prev, fp), time,
                                            _ptr(out, fp))
    if rc != 0:
        raise ValueError(f"oracle step failed rc={rc}")
    return out


def grid(cfg: Cfg, fp: str = "f64"):
    """(idx1, idx2, xs, ys, times) as the reference computes them in precision fp."""
    dt = _np_dtype(fp)
    idx1 = np.zeros(cfg.nx, np.uint32)
    idx2 = np.zeros(cfg.ny, np.uint32)
    xs = np.zeros(cfg.nx, dt)
    ys = np.zeros(cfg.ny, dt)
    times = np.zeros(cfg.nt - 1, dt)
    rc = getattr(lib(), f"lss_or_grid_{fp}")(C.byref(cfg.c()), idx1.ctypes.data_as(C.POINTER(C.c_uint32)),
                                            idx2.ctypes.data_as(C.POINTER(C.c_uint32)), _ptr(xs, fp),
                                            _ptr(ys, fp), _ptr(times, fp))
    if rc != 0:
        raise ValueError(f"oracle grid failed rc={rc}")
    return idx1, idx2, xs, ys, times


def _line(name: str, a, b, c, f, lo, hi, fp: str):
    dt = _np_dtype(fp)
    a, b, c, f = (np.ascontiguousarray(v, dtype=dt) for v in (a, b, c, f))
    x = np.zeros_like(f)
    lo_t, lo_lin, lo_val = lo
    hi_t, hi_lin, hi_val = hi
    rc = getattr(lib(), f"lss_or_{name}_{fp}")(len(f), _ptr(a, fp), _ptr(b, fp), _ptr(c, fp), _ptr(f, fp), lo_t,
                                              lo_lin, lo_val, hi_t, hi_lin, hi_val, _ptr(x, fp))
    if rc != 0:
        raise ValueError(f"oracle {name} failed rc={rc}")
    return x


def thomas_lu(a, b, c, f, lo, hi, fp: str = "f64") -> np.ndarray:
    """lo/hi = (bc_type, linear_value, value)."""
    return _line("thomas_lu", a, b, c, f, lo, hi, fp)


def double_sweep(a, b, c, f, lo, hi, fp: str = "f64") -> np.ndarray:
    return _line("double_sweep", a, b, c, f, lo, hi, fp)


def thomas_batched(a, b, c, d, B: int, n: int, strided: bool, fp: str = "f64") -> np.ndarray:
    dt = _np_dtype(fp)
    a, b, c, d = (np.ascontiguousarray(v, dtype=dt).reshape(-1) for v in (a, b, c, d))
    x = np.zeros_like(d)
    rc = getattr(lib(), f"lss_or_thomas_batched_{fp}")(B, n, int(strided), _ptr(a, fp), _ptr(b, fp), _ptr(c, fp),
                                                      _ptr(d, fp), _ptr(x, fp))
    if rc != 0:
        raise ValueError(f"oracle batched failed rc={rc}")
    return x


# ---------------------------------------------------------------------------------------------
# One-dimensional implicit heat equation (general_heat_equation, Host ThomasLU / DoubleSweep)
# ---------------------------------------------------------------------------------------------
H1_COEF_FN = C.CFUNCTYPE(C.c_double, C.c_void_p, C.c_double, C.c_double)
H1_INIT_FN = C.CFUNCTYPE(C.c_double, C.c_void_p, C.c_double)
H1_BND_FN = C.CFUNCTYPE(C.c_double, C.c_void_p, C.c_double)


class _Heat1dPde(C.Structure):
    _fields_ = [("a", H1_COEF_FN), ("b", H1_COEF_FN), ("c", H1_COEF_FN), ("src", H1_COEF_FN),
                ("init", H1_INIT_FN), ("lo_type", C.c_int32), ("hi_type", C.c_int32),
                ("lo_val", H1_BND_FN), ("lo_lin", H1_BND_FN), ("hi_val", H1_BND_FN), ("hi_lin", H1_BND_FN),
                ("ctx", C.c_void_p)]


class _Heat1dCfg(C.Structure):
    _fields_ = [("x_lo", C.c_double), ("x_hi", C.c_double), ("t_lo", C.c_double), ("t_hi", C.c_double),
                ("n", C.c_int64), ("nt", C.c_int64), ("accumulation_point", C.c_double),
                ("alpha_scale", C.c_double), ("nonuniform", C.c_int32), ("implicit_scheme", C.c_int32),
                ("tridiagonal", C.c_int32), ("direction", C.c_int32)]


class _Bs(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("strike", "rate", "sig", "maturity", "x_hi")]


@dataclasses.dataclass
class Heat1dCfg:
    """pde_discretization_config_1d + grid_config_hints_1d + heat_implicit_solver_config of one 1-D solve."""
    x_lo: float
    x_hi: float
    n: int
    nt: int
    t_lo: float = 0.0
    t_hi: float = 1.0
    accumulation_point: float = 0.0
    alpha_scale: float = 3.0
    nonuniform: bool = False
    implicit_scheme: int = CRANK_NICOLSON
    tridiagonal: int = THOMAS_LU
    direction: int = BACKWARD

    def c(self) -> _Heat1dCfg:
        return _Heat1dCfg(self.x_lo, self.x_hi, self.t_lo, self.t_hi, self.n, self.nt, self.accumulation_point,
                          self.alpha_scale, int(self.nonuniform), self.implicit_scheme, self.tridiagonal,
                          self.direction)


def _h1lib():
    L = lib()
    if not getattr(L, "_h1_ready", False):
        for sfx, T in (("f64", C.c_double), ("f32", C.c_float)):
            P = C.POINTER(T)
            getattr(L, f"lss_or_heat1d_solve_{sfx}").argtypes = [C.POINTER(_Heat1dPde), C.POINTER(_Heat1dCfg), P, P]
            getattr(L, f"lss_or_heat1d_grid_{sfx}").argtypes = [C.POINTER(_Heat1dCfg), P, P]
        L.lss_or_bs_heat1d.argtypes = [C.POINTER(_Bs), C.POINTER(_Heat1dPde)]
        L._h1_ready = True
    return L


def bs_heat1d_pde(strike: float, rate: float, sig: float, maturity: float, x_hi: float) -> Pde:
    """Black-Scholes call data of the reference's 1-D tests (lss_black_scholes_equation_t.hpp:66-80)."""
    p = _Bs(strike, rate, sig, maturity, x_hi)
    pde = _Heat1dPde()
    _h1lib().lss_or_bs_heat1d(C.byref(p), C.byref(pde))
    return Pde(pde, p)


def custom_heat1d_pde(a: Callable, b: Callable, c: Callable, init: Callable, lower: tuple, upper: tuple,
                      src: Optional[Callable] = None) -> Pde:
    """1-D data from Python callables: a, b, c, src take (t, x), init (x); lower/upper are
    (BC_DIRICHLET | BC_NEUMANN, value(t)) or (BC_ROBIN, lin(t), value(t))."""
    fa, fb, fc = (H1_COEF_FN(lambda _c, t, x, g=g: float(g(t, x))) for g in (a, b, c))
    fs = H1_COEF_FN(lambda _c, t, x: float(src(t, x))) if src is not None else H1_COEF_FN()
    fi = H1_INIT_FN(lambda _c, x: float(init(x)))

    def bnd(e):
        val = H1_BND_FN(lambda _c, t, g=e[-1]: float(g(t)))
        lin = H1_BND_FN(lambda _c, t, g=e[1]: float(g(t))) if len(e) == 3 else H1_BND_FN()
        return int(e[0]), val, lin

    lt, lv, ll = bnd(lower)
    ht, hv, hl = bnd(upper)
    pde = _Heat1dPde(fa, fb, fc, fs, fi, lt, ht, lv, ll, hv, hl, None)
    return Pde(pde, (fa, fb, fc, fs, fi, lv, ll, hv, hl))


def heat1d_solve(pde: Pde, cfg: Heat1dCfg, fp: str = "f64", surfaces: bool = False):
    """Final level (n,), or (final level, every level (nt, n)) with surfaces=True."""
    dt = _np_dtype(fp)
    out = np.zeros(cfg.n, dtype=dt)
    surf = np.zeros((cfg.nt, cfg.n), dtype=dt) if surfaces else None
    rc = getattr(_h1lib(), f"lss_or_heat1d_solve_{fp}")(C.byref(pde.pde), C.byref(cfg.c()), _ptr(out, fp),
                                                       _ptr(surf, fp) if surfaces else None)
    if rc != 0:
        raise ValueError(f"oracle 1-D solve failed rc={rc}")
    return (out, surf) if surfaces else out


def heat1d_grid(cfg: Heat1dCfg, fp: str = "f64"):
    """(xs, times): the physical grid X(zeta_i) and the loop times."""
    dt = _np_dtype(fp)
    xs = np.zeros(cfg.n, dtype=dt)
    times = np.zeros(cfg.nt - 1, dtype=dt)
    getattr(_h1lib(), f"lss_or_heat1d_grid_{fp}")(C.byref(cfg.c()), _ptr(xs, fp), _ptr(times, fp))
    return xs, times
