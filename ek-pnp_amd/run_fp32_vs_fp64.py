#!/usr/bin/env python3
"""FP32 deviation storage against FP64 storage at full horizon.

A workload of bench.py (default: the 4096 x 1024 channel, tau 0.6, zeta -2,
kappa h 32, E_x 5e-5, seeded cold start) is run twice on the GPU from the
same state: FP64 populations (the path that matches the FP64 oracle to
rounding in the parity tests) and FP32 deviation storage with FP64
arithmetic (the production path).  At every checkpoint the relative L2
distance of u_x, psi, c+, c- and rho between the two is reported (for a
batch: the largest over its channels), with the fluid mass drift and the
u_x residual of each (for a batch: of channel 0); the CPU oracle itself
would need hours for this horizon, so the FP64 GPU path stands in for it.

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.dirname(HERE))

import ekpnp  # noqa: E402
from bench import apply_walls, workload  # noqa: E402


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=50_000)
    ap.add_argument("--every", type=int, default=10_000)
    ap.add_argument("--config", default="c1", choices=("c1", "c2", "c4"))
    args = ap.parse_args()
    params, _, seed, desc = workload(args.config)
    runs = {}
    for prec in ("fp64", "fp32"):
        s = ekpnp.Solver(params, precision=prec)
        apply_walls(s, args.config, params, seed)
        s.initialization(ekpnp.COLD_NOISE, seed)
        s.init_equilibrium()
        s.step(1)
        runs[prec] = s
    mass_idx, res_idx = ekpnp.RED_NAMES.index("mass"), ekpnp.RED_NAMES.index("residual")
    m0 = {k: float(s.reductions()[0, mass_idx]) for k, s in runs.items()}
    out = {"experiment": "FP32 deviation storage vs FP64 storage", "workload": desc, "seed": seed,
           "checkpoints": []}
    done = 1
    t0 = time.perf_counter()
    while done < args.steps:
        n = min(args.every, args.steps - done)
        for s in runs.values():
            s.reductions()  # residual snapshot 100 steps before the checkpoint
            s.step(n - 100)
            s.reductions()
            s.step(100)
        done += n
        red = {k: s.reductions()[0] for k, s in runs.items()}
        rec = {"step": done}
        for f in ("ux", "psi", "cpos", "cneg", "rho"):
            rec[f] = max(rel_l2(runs["fp32"].get_field(c, f), runs["fp64"].get_field(c, f))
                         for c in range(len(params)))
        for k in runs:
            rec[f"mass_drift_{k}"] = (float(red[k][mass_idx]) - m0[k]) / m0[k]
            rec[f"residual_{k}"] = float(red[k][res_idx])
        out["checkpoints"].append(rec)
        log(f"step {done}: " + " ".join(f"{k}={v:.3e}" for k, v in rec.items() if k != "step"))
    out["wall_s"] = round(time.perf_counter() - t0, 1)
    for s in runs.values():
        s.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
