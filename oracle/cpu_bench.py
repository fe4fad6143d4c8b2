"""TEST INFRASTRUCTURE ONLY: one worker of bench.py's CPU baseline.

Runs the reference's own level step (oracle/_ref: courant_fine, set_unew,
godunov_fine, set_uold on a fully refined periodic level, driven by
oracle/ref_uniform.py) on one core for a time budget and prints one JSON
line {"cells": cell-updates done, "seconds": time of those steps, "steps",
"kind", "n"}.  The level is the benchmark workload's initial state generated
at side `n` (the same generator and parameters, a smaller box); set-up time
(tree building, installation) is not counted.  Where the reference build is
absent the oracle's C restatement of the same step runs instead
("kind": "port").  bench.py starts one worker per host core it uses, all at
once, and adds their rates.

    python -m oracle.cpu_bench --workload turb --n 128 --budget 15 --seed 4
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def initial_state(workload, n, seed):
    from ramses_amd import ics
    if workload == "turb":  # the 512^3 Mach-5 workload's generator
        return ics.turbulent_host(n, 5.0, gamma=5.0 / 3.0, rho_sigma=0.5, seed=seed, workers=1), \
            dict(gamma=5.0 / 3.0, riemann="hllc", slope_type=1, courant_factor=0.5)
    if workload == "turb2":  # a serving request's field: Mach 2, uniform density
        return ics.turbulent_host(n, 2.0, gamma=5.0 / 3.0, rho_sigma=0.0, seed=seed, workers=1), \
            dict(gamma=5.0 / 3.0, riemann="hllc", slope_type=1, courant_factor=0.5)
    if workload == "sedov":  # the Sedov blast of the 64^3 workload
        import numpy as np
        gamma = 1.4
        dx = 1.0 / n
        xc = (np.arange(n) + 0.5) * dx
        w = np.maximum(1.0 - np.abs(xc - 0.5) / dx, 0.0)
        r = w[:, None, None] * w[None, :, None] * w[None, None, :]
        u = np.zeros((5, n, n, n))
        u[0] = 1.0
        u[4] = (1e-5 + (gamma - 1.0) * r / dx ** 3) / (gamma - 1.0)
        return u, dict(gamma=gamma, riemann="hllc", slope_type=1, courant_factor=0.8)
    raise ValueError(workload)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="turb", choices=["turb", "turb2", "sedov"])
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--budget", type=float, default=15.0)
    ap.add_argument("--seed", type=int, default=4)
    a = ap.parse_args(argv)
    import oracle
    u, par = initial_state(a.workload, a.n, a.seed)
    steps = 0
    if oracle.reference_available():
        from oracle.ref_uniform import ReferenceUniformLevel
        lvl = ReferenceUniformLevel(a.n, par["gamma"], riemann=par["riemann"], slope_type=par["slope_type"],
                                    courant_factor=par["courant_factor"])
        lvl.load(u)
        kind = "reference"
        t0 = time.perf_counter()
        while True:
            lvl.step()
            steps += 1
            el = time.perf_counter() - t0
            if el >= a.budget:
                break
    else:
        O = oracle.Oracle()
        O.set_params(**par)
        kind = "port"
        t0 = time.perf_counter()
        while True:
            u, _ = O.uniform_step(u, 1.0 / a.n, nthreads=1)
            steps += 1
            el = time.perf_counter() - t0
            if el >= a.budget:
                break
    print(json.dumps({"cells": a.n ** 3 * steps, "seconds": el, "steps": steps, "kind": kind, "n": a.n}),
          flush=True)


if __name__ == "__main__":
    main()
