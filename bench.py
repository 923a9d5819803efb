#!/usr/bin/env python3
"""Benchmark of the MI355X Heston/SABR ADI stepper.

Headline workload (one "step" = one full solve of the book, inputs resident in HBM):
  the book of 2048 Heston calls, seed 3: NS ~ U{64..512}, Nv ~ U{32..256}, steps ~ U{50..500},
  K ~ U[50,150], T ~ U[0.1,3], kappa ~ U[0.5,5], theta_v ~ U[0.02,0.4], sigma ~ U[0.1,1],
  rho ~ U[-0.9,0], r ~ U[0,0.05], scheme ~ U{DR, CS, MCS, HV}; the call of the single-problem
  case on each problem's own sizes: S in [0, 2K], v in [0, 1.2], uniform grid, theta = zeta = 0.5;
  fp64, separable coefficients, all problems stepped together in one plan whose time loop is
  replayed from a hipGraph.
metric: ADI grid-point-steps/s (sum over problems of nx*ny*steps) of the batched book.

Also reported: the plain per-problem stepper on a sample of the same book, the batched tridiagonal
kernel on B = 65536 systems of n = 256 (fp64, both layouts), the single 100x50x100 Heston DR solve,
a parity sample of the book against the CPU oracle, the oracle itself timed on host cores, and:
  adjoint: Heston MCS 1024 x 512, 1000 steps (K = 100, r = 0.02, kappa = 3, theta_v = 0.2, sigma = 0.8,
    rho = 0.2, S in [0, 200], v in [0, 1.2], hints (100, 10, 15, Nonuniform), theta = zeta = 0.5): forward,
    checkpointed forward (every 32 steps) and backward of the bilinear price at (S, v) = (100, 0.04);
    gradient wrt kappa, theta_v, sigma, rho, r checked by central differences (h = 1e-6 relative);
  SABR book: 256 Hundsdorfer-Verwer problems 512 x 256, 500 steps, seed 4, r = 0 (beta ~ U[0.3,1],
    nu ~ U[0.2,1], rho ~ U[-0.7,0.3], F0 = K ~ U[0.01,0.05], alpha0 ~ U[0.01,0.1], T ~ U[0.25,5],
    F in [0, 5K], alpha in [0, 5 alpha0]) in fp64 and fp32, with the batched adjoint of the prices
    d/d(alpha0, nu, rho, beta).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import linear_system_solvers_amd as L  # noqa: E402

HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of the dominant kernels, from rocprofv3 FETCH_SIZE / WRITE_SIZE passes over this
# workload (scaled by the same counters read for a copy kernel of known size); see profiles/
TRAFFIC_JSON = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "traffic.json")


def measured_traffic(kind: str):
    try:
        with open(TRAFFIC_JSON) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    e = t.get("kernels", {}).get(kind)
    return None if e is None else e.get("bytes_per_launch")


def book_models(count: int = 2048, seed: int = 3, coeff_mode: int = L.COEFF_SEPARABLE,
                line_solver: int = L.LINE_PARTITIONED):
    rng = np.random.default_rng(seed)
    models = []
    for _ in range(count):
        nx = int(rng.integers(64, 513))
        ny = int(rng.integers(32, 257))
        steps = int(rng.integers(50, 501))
        K = float(rng.uniform(50, 150))
        T = float(rng.uniform(0.1, 3.0))
        kappa = float(rng.uniform(0.5, 5.0))
        theta_v = float(rng.uniform(0.02, 0.4))
        sigma = float(rng.uniform(0.1, 1.0))
        rho = float(rng.uniform(-0.9, 0.0))
        r = float(rng.uniform(0.0, 0.05))
        scheme = int(rng.integers(0, 4))
        models.append(L.Model(kind=L.MODEL_HESTON, strike=K, rate=r, sig_sig=sigma, sig_kappa=kappa,
                              sig_theta=theta_v, rho=rho, maturity=T, x_lo=0.0, x_hi=2.0 * K, y_lo=0.0, y_hi=1.2,
                              t_lo=0.0, t_hi=T, nx=nx, ny=ny, nt=steps + 1, grid=L.UNIFORM, splitting=scheme,
                              weighting=0.5, implicit_scheme=L.CRANK_NICOLSON, tridiagonal=L.THOMAS_LU,
                              direction=L.BACKWARD, coeff_mode=coeff_mode, line_solver=line_solver))
    return models


def point_steps(models) -> int:
    return int(sum(m.nx * m.ny * m.steps for m in models))


def rel_err(g, c) -> float:
    floor = 1e-8 * np.abs(c).max()
    return float(np.max(np.abs(g - c) / np.maximum(np.abs(c), max(floor, 1e-300))))


def time_stream(torch, fn, reps: int) -> float:
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(st)
    for _ in range(reps):
        fn()
    e1.record(st)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / 1e3


def bench_book(torch, models, steps: int, warmup: int, log, graph: bool = True):
    t0 = time.time()
    plan = L.Plan.from_models(models, L.FP64)
    log(f"book plan: {len(models)} problems built in {time.time() - t0:.1f}s")
    ins = [torch.from_numpy(plan.initial(p)).cuda() for p in range(plan.count)]
    outs = [torch.empty_like(u) for u in ins]
    pin, pout = [u.data_ptr() for u in ins], [u.data_ptr() for u in outs]
    stream = torch.cuda.current_stream().cuda_stream
    t0 = time.time()
    for _ in range(max(warmup, 1)):
        plan.run_dev(pin, pout, stream, graph=graph)
    torch.cuda.synchronize()
    log(f"book warmup ({max(warmup, 1)} solves incl. autotuning and graph capture) {time.time() - t0:.1f}s; "
        f"mapping (S-lines, v-lines per workgroup) = {plan.autotune(stream)}")
    secs = time_stream(torch, lambda: plan.run_dev(pin, pout, stream, graph=graph), steps)
    log(f"book timed: {steps} solves in {secs:.3f}s")
    prof = L.run_profiled(plan, pin, pout, stream)
    return plan, ins, outs, secs, prof


def bench_plain(torch, models, log, sample_every: int = 32):
    sample = models[::sample_every]
    secs = 0.0
    stream = torch.cuda.current_stream().cuda_stream
    for m in sample:
        plan = L.Plan.from_models([m], L.FP64)
        u0 = torch.from_numpy(plan.initial(0)).cuda()
        out = torch.empty_like(u0)
        plan.run_dev([u0.data_ptr()], [out.data_ptr()], stream, graph=True)  # capture + warm
        torch.cuda.synchronize()
        secs += time_stream(torch, lambda: plan.run_dev([u0.data_ptr()], [out.data_ptr()], stream, graph=True), 1)
        plan.close()
    ps = point_steps(sample)
    log(f"plain stepper: {len(sample)} problems one at a time, {ps:.3e} point-steps in {secs:.3f}s")
    return ps / secs, len(sample)


def bench_tridiag(torch, log, B=65536, n=256):
    res = {}
    rng = np.random.default_rng(1)
    a = rng.uniform(-1, 0, (B, n))
    c = rng.uniform(-1, 0, (B, n))
    b = 2.5 + np.abs(a) + np.abs(c)
    d = rng.standard_normal((B, n))
    stream = torch.cuda.current_stream().cuda_stream
    for layout, name in ((L.LAYOUT_CONTIGUOUS, "contiguous"), (L.LAYOUT_STRIDED, "strided")):
        arrs = [np.ascontiguousarray(v if layout == L.LAYOUT_CONTIGUOUS else v.T) for v in (a, b, c, d)]
        ta, tb, tc, td = (torch.from_numpy(v).cuda() for v in arrs)
        tx = torch.empty_like(td)
        fn = lambda: L.tridiag_batched(L.FP64, layout, B, n, ta.data_ptr(), tb.data_ptr(), tc.data_ptr(),
                                       td.data_ptr(), tx.data_ptr(), stream)
        fn()
        secs = time_stream(torch, fn, 20) / 20
        gbs = 5 * 8 * B * n / secs / 1e9
        res[name] = {"us": round(secs * 1e6, 1), "GB/s": round(gbs, 1), "roofline_frac": round(gbs / HBM_PEAK_GBS, 3)}
        log(f"tridiag {name}: {secs * 1e6:.1f} us, {gbs:.0f} GB/s")
    return res


def bench_config0(torch, log):
    from tests.helpers import config0
    m = config0(coeff_mode=L.COEFF_FIELDS)
    plan = L.Plan.from_models([m], L.FP64)
    u0 = torch.from_numpy(plan.initial(0)).cuda()
    out = torch.empty_like(u0)
    stream = torch.cuda.current_stream().cuda_stream
    plan.run_dev([u0.data_ptr()], [out.data_ptr()], stream, graph=True)
    secs = time_stream(torch, lambda: plan.run_dev([u0.data_ptr()], [out.data_ptr()], stream, graph=True), 10) / 10
    log(f"config0 single solve: {secs * 1e3:.3f} ms")
    return m, out.cpu().numpy(), secs


def parity_sample(models, outs, log, count=8):
    """The smallest book problems of each scheme (so the CPU check stays short) against the oracle.
    Problems whose oracle solution is not finite or grows past ten times S_max (the scheme is unstable
    there at the drawn step count) are passed over; check_all covers them."""
    from oracle import oracle as O
    from tests.helpers import oracle_solve
    order = sorted(range(len(models)), key=lambda q: models[q].nx * models[q].ny * models[q].steps)
    worst, worst_norm, seen, picked, skipped = 0.0, 0.0, {}, 0, 0
    for q in order:
        s = models[q].splitting
        if seen.get(s, 0) >= count // 4:
            continue
        c = oracle_solve(O, models[q])
        if not np.isfinite(c).all() or np.abs(c).max() > 10.0 * models[q].x_hi:
            skipped += 1
            continue
        g = outs[q]
        worst = max(worst, rel_err(g, c)) if np.isfinite(g).all() else float("inf")
        worst_norm = max(worst_norm, float(np.abs(g - c).max() / np.abs(c).max())) if np.isfinite(g).all() \
            else float("inf")
        seen[s] = seen.get(s, 0) + 1
        picked += 1
        if picked >= count:
            break
    log(f"parity sample: {picked} book problems vs oracle ({skipped} unstable ones passed over), "
        f"max per-value rel err {worst:.2e}, max normwise {worst_norm:.2e}")
    return worst, worst_norm, picked


CONDITIONING_FACTOR = 100.0


def check_all(models, outs, log, exact_outs=None, threads: int = 0, tol: float = 1e-10):
    """Every book problem against the oracle (cached coefficients; same arithmetic as the faithful
    mode), on a pool of host threads (ctypes releases the GIL during each solve).

    exact_outs, when given, are the book's surfaces from the exact GPU path (coefficient fields and
    sequential Thomas lines, the oracle's own operation order): each must equal the oracle's surface
    bit for bit (NaN where the oracle has NaN), whatever the problem's conditioning.

    outs are the fast path's surfaces (separable coefficients, partitioned lines, fused multiply-
    adds): each is held to tol normwise.  The book's parameter ranges include problems on which the
    scheme itself is unstable at the drawn step count (the explicit v = 0 row and the explicit
    mixed-derivative corrections at large steps; the solution then grows without bound): there any
    change of rounding is amplified by orders of magnitude.  For every problem over tol the oracle is
    run three more times, with the reference's other tridiagonal solver (double sweep instead of
    Thomas LU) and with the strike and the rate perturbed by a relative 1e-15; the largest change
    among them (the spread) measures what the problem does to a change of rounding, and the fast
    path is accepted when its distance to the oracle is within CONDITIONING_FACTOR times the spread.
    Where the oracle overflows (non-finite values) the fast path must overflow too.  A problem is
    counted as unstable when the oracle overflows or grows past ten times S_max."""
    import concurrent.futures as cf
    import dataclasses
    from oracle import oracle as O
    from tests.helpers import oracle_solve
    threads = threads or host_threads()
    t0 = time.time()

    def normwise(g, c):
        return float(np.abs(g - c).max() / np.abs(c).max())

    def one(q):
        m = models[q]
        c = oracle_solve(O, m)
        g = outs[q]
        exact = None if exact_outs is None else bool(np.array_equal(exact_outs[q], c, equal_nan=True))
        if not np.isfinite(c).all():
            return q, exact, True, None, None, None, bool(not np.isfinite(g).all())
        unstable = bool(np.abs(c).max() > 10.0 * m.x_hi)
        if not np.isfinite(g).all():
            return q, exact, unstable, float("inf"), float("inf"), 0.0, True
        en = normwise(g, c)
        spread = None
        if en > tol:
            variants = (dict(tridiagonal=L.DOUBLE_SWEEP), dict(strike=m.strike * (1 + 1e-15)),
                        dict(rate=m.rate * (1 + 1e-15)))
            spread = max(normwise(oracle_solve(O, dataclasses.replace(m, **v)), c) for v in variants)
        return q, exact, unstable, en, rel_err(g, c), spread, False

    done, n_exact, exact_bad = 0, 0, []
    st = {k: {"problems": 0, "within_tolerance": 0, "max_rel_err_normwise": 0.0, "max_rel_err_per_value": 0.0,
              "within_conditioning": 0, "max_err_over_spread": 0.0} for k in ("stable", "unstable")}
    overflow, failed, cond = 0, [], []
    order = sorted(range(len(models)), key=lambda q: -models[q].nx * models[q].ny * models[q].steps)
    with cf.ThreadPoolExecutor(threads) as ex:
        for q, exact, unstable, en, ev, spread, g_nonfinite in ex.map(one, order):
            done += 1
            if exact is not None:
                n_exact += exact
                if not exact:
                    exact_bad.append(q)
            d = st["unstable" if unstable else "stable"]
            d["problems"] += 1
            if en is None:  # the oracle overflowed
                overflow += 1
                if not g_nonfinite:
                    failed.append((q, float("nan"), float("nan")))
            elif spread is None:
                d["within_tolerance"] += 1
                d["max_rel_err_normwise"] = max(d["max_rel_err_normwise"], en)
                d["max_rel_err_per_value"] = max(d["max_rel_err_per_value"], ev)
            elif en <= CONDITIONING_FACTOR * spread:
                d["within_conditioning"] += 1
                d["max_err_over_spread"] = max(d["max_err_over_spread"], en / spread)
                cond.append((q, en, spread))
            else:
                failed.append((q, en, spread))
            if done % 256 == 0:
                log(f"check-all: {done}/{len(models)} problems, {len(failed)} failed"
                    + ("" if exact_outs is None else f", exact path bit-identical {n_exact}/{done}")
                    + f" ({time.time() - t0:.0f}s)")
    S, U = st["stable"], st["unstable"]
    log(f"check-all: {done} problems in {time.time() - t0:.0f}s on {threads} threads"
        + ("" if exact_outs is None else f"; exact path bit-identical to the oracle on {n_exact}/{done}")
        + f"; fast path on the {S['problems']} stable problems: {S['within_tolerance']} within {tol:.0e} normwise "
        f"(max {S['max_rel_err_normwise']:.2e}), {S['within_conditioning']} within conditioning (max "
        f"{S['max_err_over_spread']:.1f}x spread); on the {U['problems']} unstable ones: {U['within_tolerance']} "
        f"within {tol:.0e}, {U['within_conditioning']} within conditioning (max {U['max_err_over_spread']:.1f}x "
        f"spread), {overflow} overflow in the oracle and on the GPU; {len(failed)} failed")
    rnd = lambda v: float(f"{v:.3e}")  # noqa: E731
    res = {"problems": done, "threads": threads, "tolerance_normwise": tol, "conditioning_factor": CONDITIONING_FACTOR,
           "stable": S, "unstable": U, "overflow_in_oracle_and_gpu": overflow,
           "within_conditioning_sample": [[q, rnd(e), rnd(s)] for q, e, s in sorted(cond)[:12]],
           "failed": len(failed), "failed_sample": [[q, rnd(e), rnd(s)] for q, e, s in sorted(failed)[:12]],
           "seconds": round(time.time() - t0, 1)}
    if exact_outs is not None:
        res["exact_path_bit_identical"] = n_exact
        res["exact_path_mismatch_sample"] = sorted(exact_bad)[:12]
    return res


def exact_book(torch, log):
    """The book through the exact path (coefficient fields, sequential Thomas lines): its surfaces, and
    the time of one graph-replayed solve (the throughput cost of the reference's operation order)."""
    models = book_models(2048, coeff_mode=L.COEFF_FIELDS, line_solver=L.LINE_EXACT)
    t0 = time.time()
    plan = L.Plan.from_models(models, L.FP64)
    ins = [torch.from_numpy(plan.initial(p)).cuda() for p in range(plan.count)]
    outs = [torch.empty_like(u) for u in ins]
    pin, pout = [u.data_ptr() for u in ins], [u.data_ptr() for u in outs]
    stream = torch.cuda.current_stream().cuda_stream
    plan.run_dev(pin, pout, stream, graph=True)
    torch.cuda.synchronize()
    secs = time_stream(torch, lambda: plan.run_dev(pin, pout, stream, graph=True), 1)
    res = [o.cpu().numpy() for o in outs]
    plan.close()
    log(f"exact-path book: {secs * 1e3:.1f} ms per solve ({point_steps(models) / secs:.3e} grid-point-steps/s); "
        f"plan, two solves and copy back {time.time() - t0:.1f}s")
    return res, secs


def host_threads() -> int:
    """Worker threads for host-side work: the CPU share of one GPU on the box (16), or fewer cores."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(models, log, budget_s: float = 10.0, threads: int = 0):
    """The oracle in faithful mode (every coefficient evaluated through the transform chain at every use,
    as the reference does) on the smallest problems of the book until the budget: on one thread (the
    reference is single-threaded) and with the problems distributed over `threads` host threads."""
    import concurrent.futures as cf
    import threading
    from oracle import oracle as O
    from tests.helpers import oracle_cfg, oracle_pde
    order = sorted(range(len(models)), key=lambda q: models[q].nx * models[q].ny * models[q].steps)

    def solve(q):
        m = models[q]
        O.solve(oracle_pde(O, m), oracle_cfg(O, m, cached=False), "f64")
        return m.nx * m.ny * m.steps

    done, ps, t0 = 0, 0, time.time()
    for q in order:
        ps += solve(q)
        done += 1
        if time.time() - t0 > budget_s:
            break
    secs1 = time.time() - t0
    log(f"cpu baseline: {done} problems, {ps:.3e} point-steps in {secs1:.1f}s (1 thread)")
    threads = threads or host_threads()
    lock, it = threading.Lock(), iter(order)
    tot = {"ps": 0, "n": 0}
    t0 = time.time()

    def worker():
        while time.time() - t0 < budget_s:
            with lock:
                q = next(it, None)
            if q is None:
                return
            v = solve(q)
            with lock:
                tot["ps"] += v
                tot["n"] += 1

    with cf.ThreadPoolExecutor(threads) as ex:
        for f in [ex.submit(worker) for _ in range(threads)]:
            f.result()
    secsn = time.time() - t0
    log(f"cpu baseline: {tot['n']} problems, {tot['ps']:.3e} point-steps in {secsn:.1f}s ({threads} threads)")
    return {"one": (ps / secs1, done, ps), "all": (tot["ps"] / secsn, tot["n"], tot["ps"]), "threads": threads}


def adjoint_model(nx=1024, ny=512, steps=1000, **over):
    """v clustering beta_scale = 15: with 1000 steps the reference's explicit v = 0 row needs
    3 nu E <= 2, i.e. beta_scale <= ~15 (beta_scale = 20 already diverges in the reference itself)."""
    m = L.Model(kind=L.MODEL_HESTON, strike=100.0, rate=0.02, sig_sig=0.8, sig_kappa=3.0, sig_theta=0.2, rho=0.2,
                maturity=1.0, x_lo=0.0, x_hi=200.0, y_lo=0.0, y_hi=1.2, t_lo=0.0, t_hi=1.0, nx=nx, ny=ny,
                nt=steps + 1, grid=L.NONUNIFORM, accumulation_point=100.0, alpha_scale=10.0, beta_scale=15.0,
                splitting=L.MODIFIED_CRAIG_SNEYD, weighting=0.5, implicit_scheme=L.CRANK_NICOLSON,
                tridiagonal=L.THOMAS_LU, direction=L.BACKWARD, coeff_mode=L.COEFF_SEPARABLE,
                line_solver=L.LINE_PARTITIONED)
    for k, v in over.items():
        setattr(m, k, v)
    return m


HESTON_ATTR = {"kappa": "sig_kappa", "theta_v": "sig_theta", "sigma": "sig_sig", "rho": "rho", "r": "rate"}


def bench_adjoint(torch, log, reps=3, ckpt=32):
    m = adjoint_model()
    plan = L.Plan.from_models([m], L.FP64)
    xs, ys = plan.grid(0)
    w = L.bilinear_weights(xs, ys, 100.0, 0.04)
    u0 = torch.from_numpy(plan.initial(0)).cuda()
    out = torch.empty_like(u0)
    lam = torch.from_numpy(w).cuda()
    gin = torch.empty_like(u0)
    stream = torch.cuda.current_stream().cuda_stream
    fwd = lambda: plan.run_dev([u0.data_ptr()], [out.data_ptr()], stream, graph=True)  # noqa: E731
    fwd_ck = lambda: plan.forward([u0.data_ptr()], [out.data_ptr()], ckpt, stream)  # noqa: E731
    fwd()
    t_fwd = time_stream(torch, fwd, reps) / reps
    price = float(np.sum(w * out.cpu().numpy()))
    fwd_ck()
    plan.backward([lam.data_ptr()], [gin.data_ptr()], stream)
    t_ck = time_stream(torch, fwd_ck, reps) / reps
    t_bwd, dp = 0.0, None
    for _ in range(reps):
        fwd_ck()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        dp = plan.backward([lam.data_ptr()], [gin.data_ptr()], stream)
        t_bwd += time.perf_counter() - t0
    t_bwd /= reps
    grad = {name: float(dp[0, k]) for k, name in enumerate(L.HESTON_PARAMS)}
    log(f"adjoint 1024x512x1000 MCS: forward {t_fwd * 1e3:.1f} ms, checkpointed forward {t_ck * 1e3:.1f} ms, "
        f"backward {t_bwd * 1e3:.1f} ms, price {price:.8f}")
    # central differences of the GPU forward at full size: h = 1e-6 relative (as specified) and 1e-4
    # (the solve's own rounding noise, ~1e-13 of the price, limits the smaller step)
    fd, worst = {1e-6: {}, 1e-4: {}}, {1e-6: 0.0, 1e-4: 0.0}
    for hr in (1e-6, 1e-4):
        for name in L.HESTON_PARAMS:
            attr = HESTON_ATTR[name]
            v = getattr(m, attr)
            h = hr * abs(v)
            vals = []
            for sgn in (1.0, -1.0):
                pm = L.Plan.from_models([adjoint_model(**{attr: v + sgn * h})], L.FP64)
                o = torch.empty_like(u0)
                pm.run_dev([u0.data_ptr()], [o.data_ptr()], stream, graph=False)
                torch.cuda.synchronize()
                vals.append(float(np.sum(w * o.cpu().numpy())))
                pm.close()
            fd[hr][name] = (vals[0] - vals[1]) / (2 * h)
            worst[hr] = max(worst[hr], abs(grad[name] - fd[hr][name]) / max(abs(fd[hr][name]), 1e-12))
    log("adjoint gradient " + ", ".join(f"{k} {grad[k]:.8e} (fd {fd[1e-4][k]:.8e})" for k in grad) +
        f"; max rel diff {worst[1e-6]:.2e} (h 1e-6), {worst[1e-4]:.2e} (h 1e-4)")
    ps = m.nx * m.ny * m.steps
    plan.close()
    return {"workload": "heston_mcs_1024x512x1000_price_at_100_0.04", "checkpoint_every": ckpt,
            "forward_ms": round(t_fwd * 1e3, 2), "forward_ckpt_ms": round(t_ck * 1e3, 2),
            "backward_ms": round(t_bwd * 1e3, 2),
            "ratio_backward_over_forward": round(t_bwd / t_fwd, 3),
            "ratio_gradient_over_forward": round((t_ck + t_bwd) / t_fwd, 3),
            "forward_grid_point_steps_per_s": ps / t_fwd, "price": price, "gradient": grad,
            "fd_central_h1e-6": fd[1e-6], "fd_central_h1e-4": fd[1e-4],
            "max_rel_diff_vs_fd_h1e-6": worst[1e-6], "max_rel_diff_vs_fd_h1e-4": worst[1e-4]}


def explicit_models(count=1024, seed=5, coeff_mode=L.COEFF_SEPARABLE):
    """The reference's explicit Euler Heston case (50 x 30 grid, 8000 steps, S in [50, 200], v in
    [0, 1], hints (K, 3, 50), backward; unit_tests/common/lss_print_t.hpp:2599-2703) with the strike
    drawn from U[80, 120] per problem."""
    rng = np.random.default_rng(seed)
    models = []
    for _ in range(count):
        K = float(rng.uniform(80.0, 120.0))
        models.append(L.Model(kind=L.MODEL_HESTON, strike=K, rate=0.03, sig_sig=0.3, sig_kappa=2.0, sig_theta=0.2,
                              rho=0.8, maturity=1.0, x_lo=50.0, x_hi=200.0, y_lo=0.0, y_hi=1.0, t_lo=0.0, t_hi=1.0,
                              nx=50, ny=30, nt=8000, grid=L.NONUNIFORM, accumulation_point=K, alpha_scale=3.0,
                              beta_scale=50.0, splitting=L.EXPLICIT_EULER, direction=L.BACKWARD,
                              coeff_mode=coeff_mode, line_solver=L.LINE_PARTITIONED))
    return models


def bench_heat1d(torch, log, count=16384, n=256, nt=1001, oracle_check=2):
    """1-D Black-Scholes book (the reference's one-dimensional implicit solver): count calls with the
    strike, volatility and rate drawn per problem, n-point uniform grids, nt - 1 Crank-Nicolson steps,
    each problem's whole loop in one launch.  Exact order (one lane per problem) and partitioned (one
    wavefront per problem, the loop in registers); the exact path is checked against the oracle, the
    partitioned against the exact path."""
    from oracle import oracle as O
    from linear_system_solvers_amd import heat1d as H
    rng = np.random.default_rng(6)
    draws = [(float(rng.uniform(8.0, 12.0)), float(rng.uniform(0.01, 0.1)), float(rng.uniform(0.15, 0.4)))
             for _ in range(count)]
    stream = torch.cuda.current_stream().cuda_stream
    ps = count * n * (nt - 1)
    res = {"workload": f"black_scholes_1d_{count}x{n}x{nt - 1}_cn", "grid_point_steps_per_solve": ps}
    outs = {}
    for ls, name in ((L.LINE_EXACT, "exact"), (L.LINE_PARTITIONED, "partitioned")):
        models = [H.BsModel(strike=K, rate=r, sig=sg, n=n, nt=nt, line_solver=ls) for K, r, sg in draws]
        plan = H.Heat1dPlan.from_bs_models(models, L.FP64)
        u0 = torch.from_numpy(np.stack([plan.initial(p) for p in range(count)])).cuda()
        out = torch.empty_like(u0)
        stride = n * u0.element_size()
        pin = [u0.data_ptr() + p * stride for p in range(count)]
        pout = [out.data_ptr() + p * stride for p in range(count)]
        plan.run_dev(pin, pout, stream=stream)
        torch.cuda.synchronize()
        secs = time_stream(torch, lambda: plan.run_dev(pin, pout, stream=stream), 3) / 3
        outs[name] = out.cpu().numpy()
        res[name] = {"ms": round(secs * 1e3, 3), "grid_point_steps_per_s": ps / secs,
                     "us_per_step": round(secs / (nt - 1) * 1e6, 3)}
        log(f"1-D Black-Scholes book ({name}): {count} problems {n}x{nt - 1}, {secs * 1e3:.2f} ms, "
            f"{ps / secs:.3e} grid-point-steps/s")
        if name == "exact":
            same = 0
            for q in range(oracle_check):
                K, r, sg = draws[q]
                cfg = O.Heat1dCfg(x_lo=0.0, x_hi=20.0, n=n, nt=nt, t_lo=0.0, t_hi=1.0, accumulation_point=0.0,
                                  alpha_scale=3.0, nonuniform=False, implicit_scheme=O.CRANK_NICOLSON,
                                  tridiagonal=O.THOMAS_LU, direction=O.BACKWARD)
                o = O.heat1d_solve(O.bs_heat1d_pde(K, r, sg, 1.0, 20.0), cfg, "f64")
                same += int(np.array_equal(outs[name][q], o))
            res["exact_bit_identical_to_oracle"] = f"{same}/{oracle_check}"
        plan.close()
    a, b = outs["partitioned"], outs["exact"]
    res["partitioned_vs_exact_max_normwise"] = float((np.abs(a - b).max(axis=1) / np.abs(b).max(axis=1)).max())
    log(f"1-D Black-Scholes book: exact path bit-identical to the oracle on {res['exact_bit_identical_to_oracle']} "
        f"checked; partitioned vs exact max normwise {res['partitioned_vs_exact_max_normwise']:.2e}")
    return res


def bench_levels(torch, log, count=64):
    """Every time level (solve(container_3d&)): a book of count Heston DR calls on the 100 x 50 uniform grid
    with 100 steps (strike drawn per problem) written level by level into device surfaces, against the
    same plan's final-level solve."""
    from tests.helpers import config0
    rng = np.random.default_rng(8)
    models = [config0(strike=float(rng.uniform(80.0, 120.0)), coeff_mode=L.COEFF_SEPARABLE,
                      line_solver=L.LINE_PARTITIONED) for _ in range(count)]
    plan = L.Plan.from_models(models, L.FP64)
    stream = torch.cuda.current_stream().cuda_stream
    u0 = torch.from_numpy(np.stack([plan.initial(p) for p in range(count)])).cuda()
    nx, ny, st = plan.shape(0)
    out = torch.empty_like(u0)
    lv = torch.empty((count, st + 1, nx, ny), dtype=u0.dtype, device=u0.device)
    es = u0.element_size()
    pin = [u0.data_ptr() + p * nx * ny * es for p in range(count)]
    pout = [out.data_ptr() + p * nx * ny * es for p in range(count)]
    plv = [lv.data_ptr() + p * (st + 1) * nx * ny * es for p in range(count)]
    plan.run_dev(pin, pout, stream, graph=False)
    plan.run_levels_dev(pin, plv, stream)
    torch.cuda.synchronize()
    t_final = time_stream(torch, lambda: plan.run_dev(pin, pout, stream, graph=False), 3) / 3
    t_levels = time_stream(torch, lambda: plan.run_levels_dev(pin, plv, stream), 3) / 3
    same = bool(torch.equal(lv[:, -1], out))
    level_bytes = count * (st + 1) * nx * ny * es
    plan.close()
    log(f"every level: {count} problems {nx}x{ny}x{st}, levels {t_levels * 1e3:.2f} ms against final level only "
        f"{t_final * 1e3:.2f} ms (launch by launch); {level_bytes / 1e9:.2f} GB of levels at "
        f"{level_bytes / t_levels / 1e9:.0f} GB/s of solve time; last level equals the final-level solve: {same}")
    return {"workload": f"heston_dr_{count}x{nx}x{ny}x{st}_every_level", "levels_ms": round(t_levels * 1e3, 3),
            "final_level_ms": round(t_final * 1e3, 3), "level_bytes": level_bytes,
            "last_level_equals_final": same}


def bench_source(torch, log):
    """Heat-source ADI (rhs_intermed_1_source of DR/CS/MCS/HV) through the C++ drop-in: a partitioned
    batched_heston_book of 64 problems (100 x 50 points, 100 steps, the four schemes in turn) solved
    from host containers without and with the source F = 0.05 v S (1 + t) (tests/cpp/dropin_extras.cpp,
    source_bench); the source is sampled per step, so its plan reads two more surfaces per stage-1
    point."""
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "cpp", "dropin_extras")
    r = subprocess.run([exe, "source_bench"], capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        raise RuntimeError(f"dropin_extras source_bench failed: {r.stderr[-400:]}")
    res = {"workload": "heston_source_book_64x100x50x100_cpp_dropin"}
    for ln in r.stdout.splitlines():
        f = ln.split()
        if len(f) == 4 and f[0] == "source_bench":
            ms, ps = float(f[2]), float(f[3])
            res[f[1]] = {"ms": round(ms, 3), "grid_point_steps_per_s": ps / (ms / 1e3)}
    log(f"heat-source book (C++ drop-in, host containers): without source {res['without']['ms']:.2f} ms, with "
        f"source {res['with']['ms']:.2f} ms ({res['with']['grid_point_steps_per_s']:.3e} grid-point-steps/s)")
    return res


def bench_explicit(torch, log, count=1024, oracle_check=2):
    """Explicit Euler book: 1024 problems of 50 x 30 points, each problem's whole loop in one workgroup
    with its time levels resident in LDS (one launch per solve); the step-by-step kernel (every level
    output) is timed per launch beside it."""
    from oracle import oracle as O
    from tests.helpers import oracle_solve
    models = explicit_models(count)
    plan = L.Plan.from_models(models, L.FP64)
    ins = [torch.from_numpy(plan.initial(p)).cuda() for p in range(plan.count)]
    outs = [torch.empty_like(u) for u in ins]
    pin, pout = [u.data_ptr() for u in ins], [u.data_ptr() for u in outs]
    stream = torch.cuda.current_stream().cuda_stream
    plan.run_dev(pin, pout, stream, graph=True)
    torch.cuda.synchronize()
    secs = time_stream(torch, lambda: plan.run_dev(pin, pout, stream, graph=True), 2) / 2
    ps = point_steps(models)
    resident = plan.kernel_config()["explicit_resident"]
    prof = L.run_profiled(plan, pin, pout, stream)  # (launch by launch)
    e_ms, e_bytes, e_n = prof["euler"]
    gbs = e_bytes / (e_ms / 1e3) / 1e9
    # the resident loop reads 9 neighbours and writes 1 value per interior point-step in LDS (an upper
    # figure: boundary points read fewer)
    lds_gbs = 10 * 8 * ps / secs / 1e9
    worst = 0.0
    for q in range(oracle_check):
        c = oracle_solve(O, models[q])
        worst = max(worst, float(np.abs(outs[q].cpu().numpy() - c).max() / np.abs(c).max()))
    log(f"explicit Euler book: {count} problems 50x30x7999, {secs * 1e3:.1f} ms, {ps / secs:.3e} grid-point-steps/s "
        f"({'one resident launch, ' + format(lds_gbs, '.0f') + ' GB/s of LDS traffic' if resident else 'per-step launches'}); "
        f"step-by-step k_euler {e_ms / e_n * 1e3:.2f} us per launch, {gbs:.0f} GB/s; vs oracle (normwise, "
        f"{oracle_check} problems) {worst:.2e}")
    plan.close()
    return {"workload": f"heston_explicit_euler_{count}x50x30x7999", "ms": round(secs * 1e3, 2),
            "grid_point_steps_per_s": ps / secs, "resident_in_lds": resident,
            "resident_lds_GBs": round(lds_gbs, 1) if resident else None,
            "per_step_kernel_us": round(e_ms / e_n * 1e3, 2), "euler_kernel_GBs": round(gbs, 1),
            "euler_kernel_roofline_frac": round(gbs / HBM_PEAK_GBS, 4), "euler_avg_launch_us": round(e_ms / e_n * 1e3, 2),
            "vs_oracle_normwise": worst, "problems_checked": oracle_check}


def sabr_models(count=256, seed=4, nx=512, ny=256, steps=500):
    rng = np.random.default_rng(seed)
    models, a0s = [], []
    for _ in range(count):
        beta = float(rng.uniform(0.3, 1.0))
        nu = float(rng.uniform(0.2, 1.0))
        rho = float(rng.uniform(-0.7, 0.3))
        K = float(rng.uniform(0.01, 0.05))
        a0 = float(rng.uniform(0.01, 0.1))
        T = float(rng.uniform(0.25, 5.0))
        models.append(L.Model(kind=L.MODEL_SABR, strike=K, rate=0.0, sig_sig=nu, rho=rho, beta=beta, maturity=T,
                              x_lo=0.0, x_hi=5.0 * K, y_lo=0.0, y_hi=5.0 * a0, t_lo=0.0, t_hi=T, nx=nx, ny=ny,
                              nt=steps + 1, grid=L.NONUNIFORM, accumulation_point=K, alpha_scale=10.0,
                              beta_scale=50.0, splitting=L.HUNDSDORFER_VERWER, weighting=0.5,
                              implicit_scheme=L.CRANK_NICOLSON, tridiagonal=L.THOMAS_LU, direction=L.BACKWARD,
                              coeff_mode=L.COEFF_SEPARABLE, line_solver=L.LINE_PARTITIONED))
        a0s.append(a0)
    return models, a0s


def bench_sabr(torch, log, count=256, fd_check=3, oracle_check=1):
    models, a0s = sabr_models(count)
    ps = point_steps(models)
    stream = torch.cuda.current_stream().cuda_stream
    res = {"workload": f"sabr_hv_book_{count}x512x256x500_seed4", "grid_point_steps_per_solve": ps}
    outs = {}
    for fp, name in ((L.FP64, "f64"), (L.FP32, "f32")):
        plan = L.Plan.from_models(models, fp)
        ins = [torch.from_numpy(plan.initial(p)).cuda() for p in range(count)]
        o = [torch.empty_like(u) for u in ins]
        pin, pout = [u.data_ptr() for u in ins], [u.data_ptr() for u in o]
        plan.run_dev(pin, pout, stream, graph=True)
        secs = time_stream(torch, lambda: plan.run_dev(pin, pout, stream, graph=True), 2) / 2
        res[name] = {"ms": round(secs * 1e3, 2), "grid_point_steps_per_s": ps / secs,
                     "kernel_config": plan.kernel_config()}
        outs[name] = [t.cpu().numpy().astype(np.float64) for t in o]
        log(f"sabr book {name}: {secs * 1e3:.1f} ms, {ps / secs:.3e} grid-point-steps/s")
        if fp == L.FP64:
            ws = []
            for p in range(count):
                xs, ys = plan.grid(p)
                ws.append(L.bilinear_weights(xs, ys, models[p].strike, a0s[p]))
            lams = [torch.from_numpy(wq).cuda() for wq in ws]
            plan.forward(pin, pout, 32, stream)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            dp = plan.backward([t.data_ptr() for t in lams], None, stream)
            t_bwd = time.perf_counter() - t0
            res["adjoint_backward_ms"] = round(t_bwd * 1e3, 2)
            res["adjoint_ratio_backward_over_forward"] = round(t_bwd / secs, 3)
            # d price / d alpha0 from the adjoint: alpha_max = 5 alpha0 and the price point alpha0 move
            # together, so the price is read at a fixed eta and d/d alpha0 = (d/d of the alpha-axis
            # scale) / alpha0 (the plan's fourth SABR parameter term)
            da0 = [float(dp[p, 3]) / a0s[p] for p in range(count)]
            worst = 0.0
            for p in range(fd_check):
                for k, attr in enumerate(("sig_sig", "rho", "beta", "alpha0")):
                    v = a0s[p] if attr == "alpha0" else getattr(models[p], attr)
                    h = 1e-6 * abs(v)
                    vals = []
                    for sgn in (1.0, -1.0):
                        mq = sabr_models(count, nx=models[p].nx)[0][p]
                        if attr == "alpha0":
                            mq.y_hi = 5.0 * (v + sgn * h)
                        else:
                            setattr(mq, attr, v + sgn * h)
                        pq = L.Plan.from_models([mq], L.FP64)
                        wq = ws[p]
                        if attr == "alpha0":
                            gx, gy = pq.grid(0)
                            wq = L.bilinear_weights(gx, gy, mq.strike, v + sgn * h)
                        vals.append(float(np.sum(wq * pq.run_host()[0])))
                        pq.close()
                    fdv = (vals[0] - vals[1]) / (2 * h)
                    got = da0[p] if attr == "alpha0" else dp[p, k]
                    worst = max(worst, abs(got - fdv) / max(abs(fdv), 1e-12))
            res["adjoint_params"] = ["alpha0", "nu", "rho", "beta"]
            res["adjoint_sample"] = [[float(da0[p]), *map(float, dp[p, :3])] for p in range(min(count, 4))]
            res["adjoint_fd_max_rel_diff"] = worst
            res["adjoint_fd_problems_checked"] = fd_check
            log(f"sabr adjoint: backward {t_bwd * 1e3:.1f} ms, fd check on {fd_check} problems max rel diff {worst:.2e}")
        plan.close()
    d = max(float(np.abs(a - b).max() / np.abs(b).max()) for a, b in zip(outs["f32"], outs["f64"]))
    res["f32_vs_f64_max_normwise"] = d
    if oracle_check:
        from oracle import oracle as O
        from tests.helpers import oracle_solve
        worst = 0.0
        for p in range(oracle_check):
            c = oracle_solve(O, models[p])
            worst = max(worst, float(np.abs(outs["f64"][p] - c).max() / np.abs(c).max()))
            d32 = float(np.abs(outs["f32"][p] - c).max() / np.abs(c).max())
        res["f64_vs_oracle_max_normwise"] = worst
        res["f32_vs_f64_oracle_normwise"] = d32
        log(f"sabr parity: f64 vs oracle {worst:.2e}, f32 vs oracle {d32:.2e}, f32 vs f64 GPU {d:.2e}")
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--count", type=int, default=2048)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--no-graph", action="store_true", help="launch the book's steps without hipGraph replay")
    ap.add_argument("--counters", default="",
                    help="counter-collection run: S,V,X mapping (e.g. 8,4,1) set instead of autotuned; the book is "
                         "solved --steps times with no events and no timing (hipEvents are not used under --pmc)")
    ap.add_argument("--only", choices=("adjoint", "sabr", "tridiag", "explicit", "heat1d", "levels", "source"),
                    default="",
                    help="run only this secondary workload (no book, no JSON bench line)")
    ap.add_argument("--no-check-all", action="store_true",
                    help="skip checking every book problem (fast and exact GPU paths) against the oracle on host "
                         "threads (about 4 min on 16 threads, outside the timed region)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run reference kernels of known HBM traffic (a 256 MiB fill, a 256 MiB elementwise "
                         "read and write) for scaling the HBM counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; the book shards by problem with no exchange, "
              "so no multi-GPU figure is produced", file=sys.stderr)
        sys.exit(2)

    def log(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    import torch
    torch.cuda.init()
    if args.calibrate_copy:
        # reference kernels of known traffic: a fill writing 256 MiB, an elementwise kernel reading
        # 256 MiB and writing 256 MiB, and (an independent second scale) one reading 2 x 256 MiB and
        # writing 256 MiB; each reads what an earlier kernel wrote, as the ADI stages do
        x = torch.empty(32 * 1024 * 1024, dtype=torch.float64, device="cuda").uniform_()
        y = torch.empty_like(x)
        torch.mul(x, 2.0, out=y)
        a, b = torch.full_like(x, 1.5), torch.full_like(x, 2.5)
        torch.cuda.synchronize()
        c = torch.empty_like(x)
        torch.add(a, b, out=c)
        torch.cuda.synchronize()
        log("calibration: fill of 256 MiB; 256 MiB read + 256 MiB written (mul); two fills; 512 MiB read + "
            "256 MiB written (add)")
        del x, y, a, b, c
    if args.only:  # one secondary workload alone (for kernel traces of it)
        fn = {"adjoint": bench_adjoint, "sabr": bench_sabr, "tridiag": bench_tridiag,
              "explicit": bench_explicit, "heat1d": bench_heat1d, "levels": bench_levels,
              "source": bench_source}[args.only]
        print(json.dumps({"only": args.only, "result": fn(torch, log)}), flush=True)
        return
    models = book_models(args.count)
    total_ps = point_steps(models)
    if args.counters:
        x, v, z = (int(t) for t in args.counters.split(","))
        plan = L.Plan.from_models(models, L.FP64)
        plan.set_mapping(x, v, z)
        ins = [torch.from_numpy(plan.initial(p)).cuda() for p in range(plan.count)]
        outs = [torch.empty_like(u) for u in ins]
        for _ in range(args.steps):
            plan.run_dev([u.data_ptr() for u in ins], [u.data_ptr() for u in outs], 0, graph=False)
        torch.cuda.synchronize()
        log(f"counter run: {args.steps} book solves with mapping {(x, v, z)}")
        print(json.dumps({"counters_run": True, "solves": args.steps, "mapping": [x, v, z],
                          "grid_point_steps_per_solve": total_ps}), flush=True)
        return
    log(f"book: {len(models)} problems, {total_ps:.3e} grid-point-steps per solve")
    plan, ins, outs, secs, prof = bench_book(torch, models, args.steps, args.warmup, log, graph=not args.no_graph)
    value = total_ps * args.steps / secs
    ms_per_step = secs / args.steps * 1e3
    sweeps = {k: v for k, v in prof.items() if v[2] > 0}
    dom = max(sweeps, key=lambda k: sweeps[k][0])
    d_ms, d_bytes, d_n = sweeps[dom]
    achieved = d_bytes / (d_ms / 1e3) / 1e9
    total_algo = sum(v[1] for v in sweeps.values())
    total_ms = sum(v[0] for v in sweeps.values())
    log("per-kernel: " + ", ".join(f"{k} {v[0]:.1f}ms {v[1] / (v[0] / 1e3) / 1e9:.0f}GB/s x{v[2]}"
                                   for k, v in sweeps.items()))
    result = {
        "metric": "ADI grid-point-steps/s batched vs plain (rel err <=1e-10); adjoint/forward ratio",
        "value": value,
        "unit": "grid-point-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "heston_book_2048_seed3", "problems": len(models), "seed": 3,
                   "grid_point_steps_per_solve": total_ps, "coeff": "separable", "line_solver": "partitioned",
                   "graph": not args.no_graph, "kernel_config": plan.kernel_config(),
                   "assumed": {"grid": "grid_enum::Uniform (sinh map, accumulation point 0, as in the 100x50 Heston DR call)",
                               "S_range": "[0, 2K]", "v_range": "[0, 1.2]", "theta": 0.5,
                               "implicit_scheme": "CrankNicolson", "zeta_weighting": 0.5,
                               "hints": "grid_config_hints_2d defaults (point 0, alpha_scale 3, beta_scale 50)",
                               "boundaries": "Dirichlet 0 at S=0, Neumann -1 (dU/dS=1) at S_max, "
                                             "Dirichlet U=S at v_max",
                               "direction": "Backward", "payoff": "max(S-K, 0)"}},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(dom),
                     "kernel": ("k_" + dom.rstrip("12") + ("_p" if dom.startswith(("x", "v")) else "")),
                     "kernel_kind": dom, "launches": int(d_n), "avg_launch_us": round(d_ms / d_n * 1e3, 2),
                     "algo_bytes_per_launch": round(d_bytes / d_n),
                     "whole_step_algo_GBs": round(total_algo / (total_ms / 1e3) / 1e9, 1)},
    }
    failed = 0
    if not args.no_check_all:
        ex_outs, ex_secs = exact_book(torch, log)
        ca = check_all(models, [o.cpu().numpy() for o in outs], log, exact_outs=ex_outs)
        ca["exact_path_book_ms"] = round(ex_secs * 1e3, 1)
        ca["exact_path_over_fast_path_time"] = round(ex_secs / (ms_per_step / 1e3), 2)
        result["check_all"] = ca
        S, U = ca["stable"], ca["unstable"]
        within = S["within_tolerance"] + U["within_tolerance"]
        cond = S["within_conditioning"] + U["within_conditioning"]
        failed = ca["failed"]
        result["parity_qualifier"] = (
            f"{within}/{ca['problems']} book problems within 1e-10 normwise of the CPU oracle (stable: "
            f"{S['within_tolerance']}/{S['problems']}, max per-value {S['max_rel_err_per_value']:.1e}); {cond} more "
            f"(stable {S['within_conditioning']}, unstable in the reference scheme {U['within_conditioning']}) "
            f"within {CONDITIONING_FACTOR:.0f}x their own sensitivity to rounding (max {max(S['max_err_over_spread'], U['max_err_over_spread']):.1f}x); "
            f"{ca['overflow_in_oracle_and_gpu']} overflow in both; {failed} failed; exact GPU path bit-identical "
            f"on {ca.get('exact_path_bit_identical', 0)}/{ca['problems']}")
        log("parity: " + result["parity_qualifier"])
    if not args.no_extras:
        from tests.helpers import rel_err as _re  # noqa: F401
        outs_np = [o.cpu().numpy() for o in outs]
        worst, worst_norm, nchk = parity_sample(models, outs_np, log)
        plain_rate, nplain = bench_plain(torch, models, log)
        result["plain_stepper"] = {"value": plain_rate, "unit": "grid-point-steps/s",
                                   "sample": f"every 32nd book problem ({nplain}), each solved alone"}
        result["batched_over_plain"] = round(value / plain_rate, 2)
        result["parity"] = {"book_sample_max_rel_err_per_value": worst, "book_sample_max_rel_err_normwise": worst_norm,
                            "problems_checked": nchk, "tolerance_normwise": 1e-10,
                            "per_value_floor": "1e-8 * max|c|"}
        result["tridiag_B65536_n256_f64"] = bench_tridiag(torch, log)
        m0, u0, s0 = bench_config0(torch, log)
        result["single_100x50x100_dr_ms"] = round(s0 * 1e3, 3)
        plan.close()
        plan = None
        result["adjoint"] = bench_adjoint(torch, log)
        # the metric's second part: a full gradient (checkpointed forward + backward) over one forward solve
        result["adjoint_over_forward"] = result["adjoint"]["ratio_gradient_over_forward"]
        result["sabr_book"] = bench_sabr(torch, log)
        result["explicit_euler_book"] = bench_explicit(torch, log)
        result["heat1d_book"] = bench_heat1d(torch, log)
        result["every_level"] = bench_levels(torch, log)
        result["heat_source_book"] = bench_source(torch, log)
    if not args.no_cpu:
        cb = cpu_baseline(models, log)
        r1, n1, p1 = cb["one"]
        rn, nn, pn = cb["all"]
        result["cpu_baseline"] = {
            "value": rn, "unit": "grid-point-steps/s", "cores": cb["threads"], "kind": "port",
            "sample": f"the smallest book problems, oracle faithful mode (every coefficient through the closure and "
                      f"transform chain at every use, like the reference), about 10 s each way: {nn} problems "
                      f"({pn:.3e} point-steps) over {cb['threads']} threads (problems distributed, one per thread "
                      f"at a time); {n1} problems ({p1:.3e} point-steps) on 1 thread",
            "single_thread_value": r1, "cpu_model": cpu_model(), "nproc": os.cpu_count(),
            "threads_used": cb["threads"]}
    print(json.dumps(result), flush=True)
    if plan is not None:
        plan.close()
    if failed:
        log(f"{failed} book problems outside both the tolerance and their sensitivity bound")
        sys.exit(3)


if __name__ == "__main__":
    main()
