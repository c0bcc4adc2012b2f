#!/usr/bin/env python3
"""The learned warm start at the scale of the surrogate configuration: about
20 000 training steady states of batched-sweep channels with n_y from 64 to
1024, and 512 held-out channels from the same distribution, cold against
surrogate-warm to the u_x residual 1e-8.

Cold starts of n_y = 1024 channels need millions of steps, so the training
set is built in n_y bands, each from warm starts of the network trained on
the bands below it (a steady state does not depend on the start; a warm
start the residual probe rejects restarts cold), and the network is then
trained on every band:

  gen   --band B [--net N]   steady states of band B (cold without --net)
  train --bands ...          a network on the saved bands -> --out
  heldout --net N            the held-out comparison (one JSON line)
  sweep --net N              the 256-channel batched sweep to the residual,
                             warm-started (one JSON line)

Training channels and held-out channels are x-uniform, so training
channels run at n_x = 4 and held-out channels at n_x = 8 (the steady state
and, for warm starts, the step count do not depend on n_x); `heldout --nx
drawn` runs the held-out channels at the n_x they were drawn with instead.
Each stage prints a progress line at least every ~20 000 steps.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import ekpnp  # noqa: E402
from ekpnp import surrogate as S  # noqa: E402
import run_time_to_steady as T  # noqa: E402

BANDS = {1: (64, 128, 192, 256), 2: (320, 384, 448, 512), 3: (576, 640, 704, 768), 4: (832, 896, 960, 1024)}
TOL, EVERY = 1e-8, 100
CHUNK = 20_000  # steps between progress lines


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def band_channels(band, n, seed):
    ps, rej = S.sample_channels(n, seed + band, nx_choices=(4,), ny_choices=BANDS[band])
    return ps, rej


def gen(args):
    ps, rej = band_channels(args.band, args.n, args.seed)
    log(f"band {args.band}: {len(ps)} channels, n_y {BANDS[args.band]} (redrawn {rej})")
    t0 = time.perf_counter()
    deadline = t0 + args.time_limit if args.time_limit > 0 else None
    seed = args.seed + args.band
    if args.net:
        net = S.load_net(args.net)
        s, _, rejected = T.warm_start(ps, T.predictor(net, True), "fp32", seed, 200, 1e-4)
        net.close()
        offset = 300  # the settling and probe steps
        log(f"band {args.band}: warm starts rejected {len(rejected)}")
    else:
        s = ekpnp.Solver(ps, precision="fp32")
        s.initialization(ekpnp.COLD_NOISE, seed)
        s.init_equilibrium()
        rejected, offset = [], 0
    # the steady profiles of the channels that met the residual, harvested
    # after each slice (so that a run stopped at the deadline keeps them)
    got = {}

    def harvest(idx):
        for c in idx:
            got[int(c)] = S.residual_targets(ps[c], T.xavg(s, c, "psi"), T.xavg(s, c, "cpos"), T.xavg(s, c, "cneg"),
                                             T.xavg(s, c, "ux"))

    steps = T.run_to_residual(s, TOL, EVERY, args.max_steps, f"band {args.band}", CHUNK, on_chunk=harvest,
                              deadline=deadline)
    s.close()
    steps = np.where(steps >= 0, steps + offset, -1)
    ok = steps > 0
    X = S.channel_features(ps)
    Y = np.full((len(ps), len(next(iter(got.values())))), np.nan) if got else np.zeros((len(ps), 0))
    for c, y in got.items():
        Y[c] = y
    log(f"band {args.band}: {int(ok.sum())}/{len(ps)} converged, median steps "
        f"{np.median(steps[ok]) if ok.any() else float('nan'):.0f}, wall {time.perf_counter() - t0:.1f} s")
    # targets in FP32 (the network computes in BF16)
    np.savez(args.out, X=X, Y=Y.astype(np.float32), steps=steps, ny=np.array([p.ny for p in ps]), rejected=np.array(rejected, np.int64),
             band=args.band, warm=bool(args.net))


def train(args):
    Xs, Ys = [], []
    for path in args.bands:
        with np.load(path, allow_pickle=False) as z:
            ok = z["steps"] > 0
            Xs.append(z["X"][ok])
            Ys.append(z["Y"][ok])
            log(f"{path}: {int(ok.sum())} states")
    X, Y = np.concatenate(Xs), np.concatenate(Ys)
    t0 = time.perf_counter()
    net, hist = T.train(X, Y, args.epochs, args.batch, args.lr, args.seed, n_hidden=args.hidden)
    log(f"trained on {len(X)} states in {time.perf_counter() - t0:.1f} s, val mse {hist[-1][2]:.3e}")
    ny = sorted({int(v) for path in args.bands for v in np.load(path)["ny"]})
    S.save_net(args.out, net, command=" ".join(["run_c3_scale.py"] + sys.argv[1:]), train_channels=len(X),
               ny=ny, nx_train=4, seed=args.seed, targets="residual", epochs=args.epochs,
               val_mse=float(hist[-1][2]))
    net.close()


def heldout(args):
    ny_all = tuple(v for b in sorted(BANDS) for v in BANDS[b])
    drawn, rej = S.sample_channels(args.n, args.seed, nx_choices=(32, 64, 128, 256), ny_choices=ny_all)
    ps = drawn
    if args.nx != "drawn":
        # the drawn physics at n_x = 8 (x-uniform walls: the same steady state)
        ps = []
        for q in drawn:
            c = q.copy()
            c.nx = 8
            c.lx = c.nx * c.dx  # the x length sets the x-FFT wavenumbers
            ps.append(c)
    log(f"held out: {len(ps)} channels, n_x {args.nx}, n_y {min(p.ny for p in ps)}..{max(p.ny for p in ps)}")
    net = S.load_net(args.net)
    s, warm, warm_wall, res, rejected = T.run_warm(ps, T.predictor(net, True), "fp32", args.seed, TOL, EVERY,
                                                  args.max_steps, 200, 1e-4, chunk=CHUNK)
    s.close()
    net.close()
    log(f"warm: median {np.median(warm[warm > 0]):.0f} steps, makespan {warm.max()}, wall {warm_wall:.1f} s, "
        f"rejected {len(rejected)}")
    # the cold arm runs until every channel converged or the time limit: a
    # channel still running then has a lower bound on its cold step count
    deadline = time.perf_counter() + args.cold_time_limit if args.cold_time_limit > 0 else None
    s, cold, cold_wall = T.run_cold(ps, "fp32", args.seed, TOL, EVERY, args.max_steps, chunk=CHUNK,
                                    deadline=deadline)
    cold_run = s.steps_done
    s.close()
    censored = [c for c in range(len(ps)) if cold[c] < 0]
    log(f"cold: {len(ps) - len(censored)} converged (median {np.median(cold[cold > 0]):.0f} steps), "
        f"{len(censored)} still running after {cold_run} steps, wall {cold_wall:.1f} s")
    summ = T.summarize(cold, warm)
    ok = (cold > 0) & (warm > 0)
    lb = np.where(ok, cold, cold_run) / np.maximum(warm, 1)  # per-channel gain, a lower bound where censored
    ny = np.array([p.ny for p in ps])
    by_ny = {}
    for v in sorted(set(ny.tolist())):
        m = ny == v
        by_ny[int(v)] = {"channels": int(m.sum()), "cold_converged": int((m & ok).sum()),
                         "warm_median": float(np.median(warm[m])),
                         "gain_median_or_lower_bound": round(float(np.median(lb[m])), 2)}
    out = {"experiment": "time-to-steady at the surrogate configuration's scale, surrogate-warm vs cold",
           "tol": TOL, "check_every": EVERY, "precision": "fp32", "network": os.path.basename(args.net),
           "network_meta": S.net_meta(args.net), "heldout": {
               "channels": len(ps), "seed": args.seed, "redrawn": rej, "nx": args.nx,
               "ny": [int(p.ny) for p in ps], "nx_drawn": [int(p.nx) for p in drawn],
               "cold_steps": cold.tolist(), "warm_steps": warm.tolist(), "rejected": rejected,
               "cold_censored": censored, "cold_steps_run": int(cold_run),
               "probe_residuals": [float(v) for v in res],
               "cold_wall_s": round(cold_wall, 2), "warm_wall_s": round(warm_wall, 2),
               "summary_both_converged": summ,
               "gain_median_all_lower_bound": round(float(np.median(lb)), 2),
               "makespan_gain_lower_bound": round(float(max(cold_run, cold.max()) / warm.max()), 2),
               "by_ny": by_ny}}
    print(json.dumps(out), flush=True)


def sweep(args):
    """The batched sweep (256 channels of seed 2: n_x 32..256, n_y 64..1024,
    the bench's config c2) run to the residual from the network's warm
    starts in one context; one JSON line."""
    ps, rej = S.sample_channels(256, 2)
    nodes = sum(p.nx * p.ny for p in ps)
    log(f"sweep: {len(ps)} channels, {nodes} nodes, n_y {min(p.ny for p in ps)}..{max(p.ny for p in ps)}")
    deadline = time.perf_counter() + args.time_limit if args.time_limit > 0 else None
    net = S.load_net(args.net)
    s, warm, wall, res, rejected = T.run_warm(ps, T.predictor(net, True), "fp32", 2, TOL, EVERY, args.max_steps,
                                              200, 1e-4, chunk=CHUNK, deadline=deadline)
    run = s.steps_done
    red = s.reductions()
    s.close()
    net.close()
    ok = warm > 0
    log(f"sweep: {int(ok.sum())}/{len(ps)} at residual {TOL:g}, median {np.median(warm[ok]):.0f}, "
        f"makespan {warm.max()}, wall {wall:.1f} s")
    print(json.dumps({"experiment": "batched sweep (256 channels, seed 2) to the u_x residual, warm-started",
                      "tol": TOL, "check_every": EVERY, "precision": "fp32", "network": os.path.basename(args.net),
                      "channels": len(ps), "nodes": nodes, "converged": int(ok.sum()), "rejected": rejected,
                      "steps": warm.tolist(), "steps_run": int(run), "median_steps": float(np.median(warm[ok])),
                      "makespan": int(warm.max()), "wall_s": round(wall, 2),
                      "nonfinite_channels": int((red[:, 6] > 0).sum()),
                      "ny": [int(p.ny) for p in ps], "nx": [int(p.nx) for p in ps]}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    g = sub.add_parser("gen")
    g.add_argument("--band", type=int, required=True, choices=sorted(BANDS))
    g.add_argument("--n", type=int, default=5000)
    g.add_argument("--net", default="")
    g.add_argument("--seed", type=int, default=1000)
    g.add_argument("--max-steps", type=int, default=8_000_000)
    g.add_argument("--time-limit", type=float, default=0.0,
                   help="seconds: start no slice after this; channels not converged by then are left out")
    g.add_argument("--out", required=True)
    t = sub.add_parser("train")
    t.add_argument("--bands", nargs="+", required=True)
    t.add_argument("--epochs", type=int, default=3000)
    t.add_argument("--batch", type=int, default=256)
    t.add_argument("--lr", type=float, default=2e-3)
    t.add_argument("--hidden", type=int, default=256, help="units per hidden layer (multiple of 32)")
    t.add_argument("--seed", type=int, default=1000)
    t.add_argument("--out", required=True)
    h = sub.add_parser("heldout")
    h.add_argument("--net", required=True)
    h.add_argument("--n", type=int, default=512)
    h.add_argument("--seed", type=int, default=900000)
    h.add_argument("--nx", choices=("8", "drawn"), default="8")
    h.add_argument("--max-steps", type=int, default=20_000_000)
    h.add_argument("--cold-time-limit", type=float, default=0.0,
                   help="seconds for the cold arm; channels still running then are reported as censored")
    w = sub.add_parser("sweep")
    w.add_argument("--net", required=True)
    w.add_argument("--max-steps", type=int, default=8_000_000)
    w.add_argument("--time-limit", type=float, default=0.0)
    args = ap.parse_args()
    {"gen": gen, "train": train, "heldout": heldout, "sweep": sweep}[args.cmd](args)


if __name__ == "__main__":
    main()
