#!/usr/bin/env python3
"""Benchmark of the MI355X Heston/SABR ADI stepper.

Headline workload (one "step" = one full solve of the book, inputs resident in HBM):
  the book of 2048 Heston calls, seed 3: NS ~ U{64..512}, Nv ~ U{32..256}, steps ~ U{50..500},
  K ~ U[50,150], T ~ U[0.1,3], kappa ~ U[0.5,5], theta_v ~ U[0.02,0.4], sigma ~ U[0.1,1],
  rho ~ U[-0.9,0], r ~ U[0,0.05], scheme ~ U{DR, CS, MCS, HV}; S in [0, 2K], v in [0, 1.2],
  hints (K, 10, 50, Nonuniform), theta = 0.5, zeta = 0.5, fp64, separable coefficients, all problems
  stepped together in one plan whose time loop is replayed from a hipGraph.
metric: ADI grid-point-steps/s (sum over problems of nx*ny*steps) of the batched book.

Also reported: the plain per-problem stepper on a sample of the same book, the batched tridiagonal
kernel on B = 65536 systems of n = 256 (fp64, both layouts), the single 100x50x100 Heston DR solve,
a parity sample of the book against the CPU oracle, and the oracle itself timed on host cores.
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
                              t_lo=0.0, t_hi=T, nx=nx, ny=ny, nt=steps + 1, grid=L.NONUNIFORM,
                              accumulation_point=K, alpha_scale=10.0, beta_scale=50.0, splitting=scheme,
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


def bench_book(torch, models, steps: int, warmup: int, log):
    t0 = time.time()
    plan = L.Plan.from_models(models, L.FP64)
    log(f"book plan: {len(models)} problems built in {time.time() - t0:.1f}s")
    ins = [torch.from_numpy(plan.initial(p)).cuda() for p in range(plan.count)]
    outs = [torch.empty_like(u) for u in ins]
    pin, pout = [u.data_ptr() for u in ins], [u.data_ptr() for u in outs]
    stream = torch.cuda.current_stream().cuda_stream
    t0 = time.time()
    for _ in range(max(warmup, 1)):
        plan.run_dev(pin, pout, stream, graph=True)
    torch.cuda.synchronize()
    log(f"book warmup ({max(warmup, 1)} solves incl. graph capture) {time.time() - t0:.1f}s")
    secs = time_stream(torch, lambda: plan.run_dev(pin, pout, stream, graph=True), steps)
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
    from oracle import oracle as O
    from tests.helpers import oracle_solve
    # smallest problems of each scheme, so the CPU check stays short
    order = sorted(range(len(models)), key=lambda q: models[q].nx * models[q].ny * models[q].steps)
    picked, seen = [], {}
    for q in order:
        s = models[q].splitting
        if seen.get(s, 0) < count // 4:
            picked.append(q)
            seen[s] = seen.get(s, 0) + 1
        if len(picked) >= count:
            break
    worst, worst_norm = 0.0, 0.0
    for q in picked:
        c = oracle_solve(O, models[q])
        worst = max(worst, rel_err(outs[q], c))
        worst_norm = max(worst_norm, float(np.abs(outs[q] - c).max() / np.abs(c).max()))
    log(f"parity sample: {len(picked)} book problems vs oracle, max per-value rel err {worst:.2e}, "
        f"max normwise {worst_norm:.2e}")
    return worst, worst_norm, len(picked)


def cpu_baseline(models, log, budget_s: float = 15.0):
    """The oracle (faithful mode: every coefficient evaluated through the transform chain at every use,
    as the reference does), single thread, on the smallest problems of the book until the budget."""
    from oracle import oracle as O
    from tests.helpers import oracle_cfg, oracle_pde
    order = sorted(range(len(models)), key=lambda q: models[q].nx * models[q].ny * models[q].steps)
    done, ps, t0 = 0, 0, time.time()
    for q in order:
        m = models[q]
        O.solve(oracle_pde(O, m), oracle_cfg(O, m, cached=False), "f64")
        done += 1
        ps += m.nx * m.ny * m.steps
        if time.time() - t0 > budget_s:
            break
    secs = time.time() - t0
    log(f"cpu baseline: {done} problems, {ps:.3e} point-steps in {secs:.1f}s (1 core)")
    return ps / secs, done, ps


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--count", type=int, default=2048)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first copy 256 MiB device to device three times (reference kernel for HBM counters)")
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
        x = torch.empty(32 * 1024 * 1024, dtype=torch.float64, device="cuda").uniform_()
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        log("calibration copy: 3 x 256 MiB device-to-device")
        del x, y
    models = book_models(args.count)
    total_ps = point_steps(models)
    log(f"book: {len(models)} problems, {total_ps:.3e} grid-point-steps per solve")
    plan, ins, outs, secs, prof = bench_book(torch, models, args.steps, args.warmup, log)
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
        "config": {"workload": "heston_book_2048_seed3", "problems": len(models),
                   "grid_point_steps_per_solve": total_ps, "coeff": "separable", "line_solver": "partitioned", "graph": True},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": ("k_" + dom.rstrip("12") + ("_p" if dom.startswith(("x", "v")) else "")),
                     "kernel_kind": dom, "launches": int(d_n), "avg_launch_us": round(d_ms / d_n * 1e3, 2),
                     "algo_bytes_per_launch": round(d_bytes / d_n),
                     "whole_step_algo_GBs": round(total_algo / (total_ms / 1e3) / 1e9, 1)},
    }
    plan_sep = None
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
    if not args.no_cpu:
        rate, nprob, ps = cpu_baseline(models, log)
        result["cpu_baseline"] = {"value": rate, "unit": "grid-point-steps/s", "cores": 1, "kind": "port",
                                  "sample": f"{nprob} smallest book problems ({ps:.3e} point-steps), oracle faithful "
                                            f"mode (closure-per-coefficient evaluation like the reference), 1 thread"}
    print(json.dumps(result), flush=True)
    plan.close()


if __name__ == "__main__":
    main()
