#!/usr/bin/env python3
"""Surrogate warm start on the heterogeneous-wall-charge channel.

The 8192 x 2048 channel with wall zeta(x) = -1.5 + sum_k a_k sin(2 pi k x/nx +
phi_k) (k = 1..4, seed 4), kappa h 64, E_x 1e-4, tau 0.8, FP32 populations,
is run three ways for the same number of steps: cold (uniform c0 + noise),
surrogate-warm and physics-baseline-warm.  A warm start broadcasts the
profiles predicted for the channel's mean parameters (zeta0, kappa h, E_x,
n_y) along x; it is then accepted or rejected as in the batched experiment: 100
settling steps, then the u_x residual over a 200-step probe must be at most
`accept`, else the channel restarts from its cold fields (the fallback).

The network comes from a --save file of run_time_to_steady.py; it was trained
on x-uniform channels of n_y <= 128, so this channel (n_y = 2048, kappa h 64,
x-varying walls) lies outside what it has seen.  Reported per arm: the probe
residual and the decision, the residual after the run, and the relative L2
distance of the x-averaged psi, c+, c-, u_x from the cold arm's; the cold arm
reports its own residual over the same probe window.

The "-local" arms predict psi, c+-, rho column by column from the local wall
potential zeta(x) (profiles at 65 levels spanning zeta(x)'s range, linearly
interpolated), so the double layer follows the wall charge; u_x stays the
profile of the mean zeta broadcast in x (an x-varying u_x would violate
continuity).  --arms picks the arms; cold always runs.

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
from ekpnp import surrogate as S  # noqa: E402
from bench import apply_walls, wall_zeta, workload  # noqa: E402


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def column_fields(p, predict, zeta, levels=65):
    """(rho, ux, c+, c-, psi) as [ny][nx] arrays: psi, c+-, rho per column from
    profiles predicted at `levels` wall potentials spanning zeta's range,
    linearly interpolated to each column's zeta(x); u_x from the mean-zeta
    profile."""
    zl = np.linspace(zeta.min(), zeta.max(), levels)
    prof = []
    for z in zl:
        q = p.copy()
        q.psi_lo = q.psi_hi = float(z)
        q.cpos_lo = q.cpos_hi = float(np.exp(-z))
        q.cneg_lo = q.cneg_hi = float(np.exp(z))
        psi, cp, cn, _ = predict(q)
        prof.append(np.stack([S.osmotic_rho(q, psi, cp, cn), cp, cn, psi]))
    prof = np.stack(prof)                                  # [levels][4][ny]
    t = (zeta - zl[0]) / (zl[1] - zl[0])
    i = np.clip(np.floor(t).astype(int), 0, levels - 2)
    w = (t - i)[None, None, :]
    cols = (1.0 - w) * prof[i].transpose(1, 2, 0) + w * prof[i + 1].transpose(1, 2, 0)  # [4][ny][nx]
    _, _, _, ux = predict(p)
    return cols[0], np.repeat(ux[:, None], p.nx, axis=1), cols[1], cols[2], cols[3]


def run_arm(params, predict, steps, settle, probe, accept, tag, local=False):
    s = ekpnp.Solver(params, precision="fp32")
    apply_walls(s, "c4", params, 4)
    s.initialization(ekpnp.COLD_NOISE, 4)
    info = {"arm": tag}
    t0 = time.perf_counter()
    if predict is not None:
        p = params[0]
        cold = {f: s.get_field(0, f) for f in ("rho", "ux", "uy", "cpos", "cneg")}
        if local:
            rho2, ux2, cp2, cn2, psi2 = column_fields(p, predict, wall_zeta(p.nx, 4))
            s.set_fields(0, rho2, ux2, np.zeros((p.ny, p.nx)), cp2, cn2, psi2)
        else:
            psi, cp, cn, ux = predict(p)
            b = lambda v: np.repeat(v[:, None], p.nx, axis=1)  # noqa: E731
            s.set_fields(0, b(S.osmotic_rho(p, psi, cp, cn)), b(ux), np.zeros((p.ny, p.nx)), b(cp), b(cn),
                         b(psi))
        s.init_equilibrium()
        s.step(settle)
        s.reductions()
        s.step(probe)
        red = s.reductions()[0]
        res = float(red[7])
        ok = bool(np.isfinite(res) and res <= accept and red[6] == 0)
        info.update(probe_residual=res, accepted=ok)
        log(f"{tag}: probe residual {res:.3e} -> {'accepted' if ok else 'rejected, cold fallback'}")
        if not ok:
            s.set_fields(0, cold["rho"], cold["ux"], cold["uy"], cold["cpos"], cold["cneg"])
            s.init_equilibrium()
            done = 0
        else:
            done = settle + probe
    else:
        # the cold start's residual over the same probe window, for comparison
        s.init_equilibrium()
        s.step(settle)
        s.reductions()
        s.step(probe)
        info["probe_residual"] = float(s.reductions()[0][7])
        done = settle + probe
    while done < steps:
        n = min(20000, steps - done)
        s.step(n)
        done += n
        log(f"{tag}: {done} steps")
    s.reductions()  # snapshot: the residual below is over the last 100 steps
    s.step(100)
    red = s.reductions()[0]
    s.synchronize()
    info["wall_s"] = round(time.perf_counter() - t0, 2)
    info["residual_after"] = float(red[7])
    info["nonfinite"] = int(red[6])
    prof = {f: s.get_field(0, f).mean(axis=1) for f in ("psi", "cpos", "cneg", "ux")}
    s.close()
    return info, prof


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--net", required=True, help="a --save file of run_time_to_steady.py")
    ap.add_argument("--steps", type=int, default=100_000)
    ap.add_argument("--settle", type=int, default=100)
    ap.add_argument("--probe", type=int, default=200)
    ap.add_argument("--accept", type=float, default=1e-4)
    ap.add_argument("--arms", default="surrogate,physics baseline",
                    help="comma-separated among: surrogate, physics baseline, surrogate-local, "
                         "physics baseline-local")
    args = ap.parse_args()
    arms = args.arms.split(",")
    params, _, _, desc = workload("c4")
    saved = np.load(args.net)
    net = S.MLP(max_batch=8)
    net.set_params(saved["params"])

    def surrogate(q):
        # the channel's parameters hold the walls' mean potential zeta0
        return S.residual_profiles(q, net.forward(S.channel_features([q]))[0])

    cold_info, cold_prof = run_arm(params, None, args.steps, args.settle, args.probe, args.accept, "cold")
    out = {"experiment": "surrogate warm start on the heterogeneous-wall-charge channel", "workload": desc,
           "steps": args.steps, "accept": args.accept, "arms": [cold_info]}
    preds = {"surrogate": surrogate, "physics baseline": S.baseline}
    for tag in arms:
        local = tag.endswith("-local")
        pred = preds[tag[:-len("-local")] if local else tag]
        info, prof = run_arm(params, pred, args.steps, args.settle, args.probe, args.accept, tag, local)
        info["rel_l2_vs_cold"] = {f: float(np.linalg.norm(prof[f] - cold_prof[f]) / np.linalg.norm(cold_prof[f]))
                                  for f in prof}
        out["arms"].append(info)
    net.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
