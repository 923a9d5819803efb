"""Writes heston_mcs_1024x512x1000_oracle.json: the CPU oracle's price and central-difference gradient
for the adjoint workload (Heston MCS, 1024 x 512 non-uniform grid, 1000 steps; bench.adjoint_model),
the "CPU reference gradient" of the adjoint checks.

    python tests/golden/make_heston_mcs_1024_fd.py      (about 8 minutes on 8 cores)

Price: bilinear interpolation of the oracle's final surface at (S, v) = (100, 0.04) on the oracle's own
grid.  Gradient: (price(p + h) - price(p - h)) / 2h with h = 1e-4 |p| for kappa, theta_v, sigma, rho, r
(the solve's rounding noise, about 1e-13 of the price, limits smaller steps).  Also a few values of
the surface, which pin a fresh oracle run (the GPU test compares its forward with one).
"""
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import bench  # noqa: E402
import linear_system_solvers_amd as L  # noqa: E402
from oracle import oracle as O  # noqa: E402
from tests.helpers import oracle_cfg, oracle_solve  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "heston_mcs_1024x512x1000_oracle.json")
H_REL = 1e-4
SAMPLE = [(0, 0), (100, 37), (511, 9), (511, 10), (600, 255), (1022, 400), (1023, 511), (300, 0)]


def price(m):
    c = oracle_solve(O, m)
    _, _, xs, ys, _ = O.grid(oracle_cfg(O, m))
    return float(np.sum(L.bilinear_weights(xs, ys, 100.0, 0.04) * c)), c


def main():
    t0 = time.time()
    m = bench.adjoint_model()
    p0, c = price(m)
    print(f"price {p0:.17g} ({time.time() - t0:.0f}s)", flush=True)
    fd = {}
    for name, attr in bench.HESTON_ATTR.items():
        v = getattr(m, attr)
        h = H_REL * abs(v)
        vals = [price(bench.adjoint_model(**{attr: v + sgn * h}))[0] for sgn in (1.0, -1.0)]
        fd[name] = (vals[0] - vals[1]) / (2 * h)
        print(f"{name}: {fd[name]:.17g} ({time.time() - t0:.0f}s)", flush=True)
    res = {"workload": "heston_mcs_1024x512x1000 (bench.adjoint_model)", "point": [100.0, 0.04],
           "price": p0, "fd_central_h_rel": H_REL, "fd_gradient": fd,
           "surface_max": float(np.abs(c).max()), "surface_sum": float(c.sum()),
           "surface_sample": [[i, j, float(c[i, j])] for i, j in SAMPLE]}
    with open(OUT, "w") as f:
        json.dump(res, f, indent=1)
    print(f"wrote {OUT} in {time.time() - t0:.0f}s", flush=True)


if __name__ == "__main__":
    main()
