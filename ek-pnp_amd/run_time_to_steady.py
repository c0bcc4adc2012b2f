#!/usr/bin/env python3
"""Time-to-steady with surrogate warm starts versus cold starts.

1. Training data: seeded batched-sweep channels (zeta ~ U[-4,-0.4], kappa h ~
   logU[1,50], E_x ~ logU[1e-5,1e-3], draws failing the resolution/stability
   rule redrawn) are run by the batched solver from a cold start to the
   u_x residual tolerance; their x-averaged steady profiles of psi, c+, c-,
   u_x are the targets.  The steady state of an x-uniform channel does not
   depend on nx, so training channels use a narrow nx.
2. The MLP (4 x 256 GELU, BF16 MFMA kernels) is trained on the GPU, by
   default on the difference between those profiles and the physics
   baseline (Poisson-Boltzmann on the channel's nodes + Helmholtz-
   Smoluchowski u_x), so that BF16 only has to resolve a small correction.
3. Held-out channels (another seed, nx from {32, 64, 128, 256}) are run four
   times: cold (uniform c0 + noise), warm (predicted profiles broadcast in x),
   baseline-warm (the physics baseline alone) and the reference's own start
   (gpu_initialization with its 501 PB-Picard iterations, LBM.cu:95-160);
   --arms picks a subset (without "warm" nothing is trained).
   Every warm channel first settles for 100 steps and then runs a probe of
   `probe` steps; a channel whose u_x residual over the probe exceeds
   `accept` (or is not finite) is rejected
   and restarted cold.  Both arms finish with the plain solver to the same
   residual; step counts include settling and probe.

Prints one JSON line with per-channel steps and the summary.
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


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def xavg(s, ch, name):
    return s.get_field(ch, name).mean(axis=1)


def run_to_residual(s, tol, every, max_steps, tag, chunk=100_000, on_chunk=None, deadline=None):
    """run_to_residual in slices of `chunk` steps with a progress line per
    slice; a channel's count is the step at which it first met tol (channels
    that met it are stepped again briefly in later slices, until the first
    check re-freezes them).  on_chunk(indices) sees the channels that met tol
    in each slice (their fields are steady when it runs); past `deadline`
    (time.perf_counter()) no further slice starts."""
    n = len(s.channels)
    done = np.full(n, -1, dtype=np.int64)
    start = s.steps_done
    while (done < 0).any() and s.steps_done - start < max_steps:
        if deadline is not None and time.perf_counter() > deadline:
            log(f"{tag}: deadline reached")
            break
        off = s.steps_done - start
        res = s.run_to_residual(tol, every, min(chunk, max_steps - off))
        new = (done < 0) & (res >= 0)
        done[new] = off + res[new]
        log(f"{tag}: {int((done >= 0).sum())}/{n} at residual {tol:g} after {s.steps_done - start} steps")
        if on_chunk is not None and new.any():
            on_chunk(np.flatnonzero(new))
    return done


def run_cold(channels, precision, seed, tol, every, max_steps, tag="cold", chunk=100_000, deadline=None):
    s = ekpnp.Solver(channels, precision=precision)
    s.initialization(ekpnp.COLD_NOISE, seed)
    s.init_equilibrium()
    s.synchronize()
    t0 = time.perf_counter()
    steps = run_to_residual(s, tol, every, max_steps, tag, chunk, deadline=deadline)
    s.synchronize()
    return s, steps, time.perf_counter() - t0


def run_picard(channels, precision, seed, tol, every, max_steps, probe, accept=np.inf, settle=100):
    """The reference's start: uniform rho0, u = 0, psi and c+- from the
    PB-Picard iteration on the channel's nodes, equilibrium populations.
    Such a start has the double layer formed but no flow (u_x still grows
    linearly over the probe, so the warm arms' residual test would reject
    every channel) and rho uniform, with no pressure to balance the body
    force, so strongly charged channels blow up: only channels with
    non-finite values after settle + probe restart from their cold fields
    (accept = inf; a finite accept applies the warm arms' test).  Step
    counts include settle + probe."""
    t0 = time.perf_counter()
    s = ekpnp.Solver(channels, precision=precision)
    s.initialization(ekpnp.COLD_NOISE, seed)
    cold = [{f: s.get_field(c, f) for f in ("rho", "ux", "uy", "cpos", "cneg")} for c in range(len(channels))]
    s.initialization(ekpnp.PB_PICARD, seed)
    s.init_equilibrium()
    s.step(settle)
    s.reductions()  # residual snapshot
    s.step(probe)
    red = s.reductions()
    res = red[:, 7]
    rejected = [c for c in range(len(channels)) if not (np.isfinite(res[c]) and res[c] <= accept)
                or red[c, 6] > 0]
    for c in rejected:
        f = cold[c]
        s.set_fields(c, f["rho"], f["ux"], f["uy"], f["cpos"], f["cneg"])
    if rejected:
        s.init_equilibrium()
    steps = run_to_residual(s, tol, every, max_steps, "pb-picard")
    s.synchronize()
    total = np.where(steps >= 0, steps + settle + probe, -1)
    return s, total, time.perf_counter() - t0, res, rejected


def generate(channels, precision, seed, tol, every, max_steps, residual):
    s, steps, wall = run_cold(channels, precision, seed, tol, every, max_steps, tag="training data")
    X = S.channel_features(channels)
    tf = S.residual_targets if residual else S.targets
    T = np.stack([tf(p, xavg(s, c, "psi"), xavg(s, c, "cpos"), xavg(s, c, "cneg"), xavg(s, c, "ux"))
                  for c, p in enumerate(channels)])
    s.close()
    return X, T, steps, wall


def train(X, T, epochs, batch, lr0, seed, n_hidden=256):
    rng = np.random.default_rng(seed)
    n = X.shape[0]
    nval = max(1, n // 10)
    perm = rng.permutation(n)
    val, tr = perm[:nval], perm[nval:]
    net = S.MLP(n_hidden=n_hidden, max_batch=batch, seed=seed)
    steps_per_epoch = max(1, len(tr) // batch)
    total = epochs * steps_per_epoch
    it = 0
    hist = []
    for ep in range(epochs):
        order = rng.permutation(tr)
        for k in range(steps_per_epoch):
            idx = order[k * batch:(k + 1) * batch]
            lr = 1e-5 + 0.5 * (lr0 - 1e-5) * (1 + np.cos(np.pi * it / total))
            net.train_step(X[idx], T[idx], lr)
            it += 1
        if ep % max(1, epochs // 10) == 0 or ep == epochs - 1:
            yv = net.forward(X[val])
            vl = float(np.mean((yv - T[val]) ** 2))
            yt = net.forward(X[tr[:512]])
            tl = float(np.mean((yt - T[tr[:512]]) ** 2))
            hist.append((ep, tl, vl))
            log(f"epoch {ep}: train mse {tl:.3e}  val mse {vl:.3e}")
    return net, hist


def predictor(net, residual):
    """channels -> list of (psi, c+, c-, u_x) profiles; net None = physics baseline only."""
    def run(channels):
        if net is None:
            return [S.baseline(p) for p in channels]
        pred = net.forward(S.channel_features(channels))
        inv = S.residual_profiles if residual else S.profiles
        return [inv(p, pred[c]) for c, p in enumerate(channels)]
    return run


def run_warm(channels, predict, precision, seed, tol, every, max_steps, probe, accept, corrupt=None, settle=100,
             chunk=100_000, deadline=None):
    """Warm arm.  Acceptance test: after `settle` steps (the local transient
    of starting from equilibrium populations) the u_x residual over the next
    `probe` steps must be <= accept and finite; a rejected channel restarts
    from its cold fields.  Step counts include settle + probe.
    corrupt(channel, psi, cp, cn, ux) may tamper with predictions."""
    t0 = time.perf_counter()
    s, res, rejected = warm_start(channels, predict, precision, seed, probe, accept, corrupt, settle)
    steps = run_to_residual(s, tol, every, max_steps, "warm", chunk, deadline=deadline)
    s.synchronize()
    wall = time.perf_counter() - t0
    total = np.where(steps >= 0, steps + settle + probe, -1)
    return s, total, wall, res, rejected


def warm_start(channels, predict, precision, seed, probe, accept, corrupt=None, settle=100):
    """A context started from predicted profiles, after `settle` + `probe`
    steps and the acceptance test (rejected channels restarted cold):
    (solver, probe residuals, rejected channel indices)."""
    s = ekpnp.Solver(channels, precision=precision)
    s.initialization(ekpnp.COLD_NOISE, seed)
    cold = [{f: s.get_field(c, f) for f in ("rho", "ux", "uy", "cpos", "cneg")} for c in range(len(channels))]
    prof = predict(channels)
    for c, p in enumerate(channels):
        psi, cp, cn, ux = prof[c]
        if corrupt is not None:
            psi, cp, cn, ux = corrupt(c, psi, cp, cn, ux)
        ny, nx = p.ny, p.nx
        b = lambda v: np.repeat(v[:, None], nx, axis=1)  # noqa: E731
        rho = S.osmotic_rho(p, psi, cp, cn)
        s.set_fields(c, b(rho), b(ux), np.zeros((ny, nx)), b(cp), b(cn), b(psi))
    s.init_equilibrium()
    s.step(settle)
    s.reductions()  # residual snapshot
    s.step(probe)
    red = s.reductions()
    res = red[:, 7]
    rejected = [c for c in range(len(channels)) if not (np.isfinite(res[c]) and res[c] <= accept)
                or red[c, 6] > 0]
    for c in rejected:
        f = cold[c]
        s.set_fields(c, f["rho"], f["ux"], f["uy"], f["cpos"], f["cneg"])
    if rejected:
        s.init_equilibrium()
    return s, res, rejected


def summarize(cold, warm):
    ok = (cold > 0) & (warm > 0)
    ratio = cold[ok] / warm[ok]
    return {
        "channels": int(len(cold)),
        "both_converged": int(ok.sum()),
        "cold_steps_median": float(np.median(cold[ok])) if ok.any() else None,
        "warm_steps_median": float(np.median(warm[ok])) if ok.any() else None,
        "step_ratio_median": float(np.median(ratio)) if ok.any() else None,
        "step_ratio_geomean": float(np.exp(np.mean(np.log(ratio)))) if ok.any() else None,
        "makespan_cold": int(cold.max()),
        "makespan_warm": int(warm.max()),
        "makespan_ratio": float(cold.max() / max(1, warm.max())),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=8192)
    ap.add_argument("--heldout", type=int, default=64)
    ap.add_argument("--ny", default="32,48,64,96,128")
    ap.add_argument("--nx-train", type=int, default=4,
                    help="width of the training channels (the steady state does not depend on it); 8 runs "
                         "the x-FFT on many-row tiles and is much faster for large batches of tall channels")
    ap.add_argument("--nx", default="32,64")
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--every", type=int, default=100)
    ap.add_argument("--max-steps", type=int, default=1_000_000)
    ap.add_argument("--epochs", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--probe", type=int, default=200)
    ap.add_argument("--accept", type=float, default=1e-4,
                    help="largest u_x residual after the probe for a warm start to be kept")
    ap.add_argument("--precision", default="fp32")
    ap.add_argument("--seed", type=int, default=1000)
    ap.add_argument("--save", default="", help="write the trained parameters and the training set (.npz)")
    ap.add_argument("--save-net", default="",
                    help="write the trained network alone (surrogate.save_net: parameters, shape, training "
                         "command) for surrogate.load_net")
    ap.add_argument("--load-data", default="",
                    help="take the training targets and step counts from a --save file of the same "
                         "--train/--seed/--ny/--nx-train/--targets instead of generating them")
    ap.add_argument("--targets", choices=("direct", "residual"), default="residual",
                    help="learn the profiles themselves or their difference from the physics baseline")
    ap.add_argument("--arms", default="cold,warm,baseline,picard",
                    help="held-out arms to run (cold is always run: the others are measured against it)")
    args = ap.parse_args()
    ny = tuple(int(v) for v in args.ny.split(","))
    nxs = tuple(int(v) for v in args.nx.split(","))
    arms = set(args.arms.split(","))
    if "warm" not in arms:
        return held_out_only(args, ny, nxs, arms)

    tr_ch, rej_tr = S.sample_channels(args.train, args.seed, nx_choices=(args.nx_train,), ny_choices=ny)
    log(f"training channels: {len(tr_ch)} (redrawn {rej_tr})")
    residual = args.targets == "residual"
    if args.load_data:
        saved = np.load(args.load_data)
        T, tr_steps, tr_wall = saved["T"], saved["steps"], 0.0
        X = S.channel_features(tr_ch)
        if T.shape[0] != len(tr_ch):
            raise SystemExit("--load-data file holds another training set")
    else:
        X, T, tr_steps, tr_wall = generate(tr_ch, args.precision, args.seed, args.tol, args.every, args.max_steps,
                                           residual)
    good = tr_steps > 0
    log(f"data: {good.sum()} converged, steps median {np.median(tr_steps[good]):.0f}, wall {tr_wall:.1f} s")
    rms = np.sqrt(np.mean(T[good].reshape(int(good.sum()), S.N_FIELDS, S.N_PTS) ** 2, axis=(0, 2)))
    log(f"target rms per field (psi, c+, c-, u): {rms}")
    t0 = time.perf_counter()
    net, hist = train(X[good], T[good], args.epochs, args.batch, args.lr, args.seed)
    train_wall = time.perf_counter() - t0

    ho_ch, rej_ho = S.sample_channels(args.heldout, 900000 + args.seed, nx_choices=nxs, ny_choices=ny)
    s, cold, cold_wall = run_cold(ho_ch, args.precision, 900000 + args.seed, args.tol, args.every, args.max_steps)
    s.close()
    log(f"cold: median steps {np.median(cold):.0f}, wall {cold_wall:.1f} s")
    s, warm, warm_wall, probe_res, rejected = run_warm(ho_ch, predictor(net, residual), args.precision,
                                                      900000 + args.seed, args.tol, args.every, args.max_steps,
                                                      args.probe, args.accept)
    s.close()
    log(f"warm: median steps {np.median(warm):.0f}, wall {warm_wall:.1f} s, rejected {len(rejected)}")
    s, base, base_wall, base_res, base_rej = run_warm(ho_ch, predictor(None, residual), args.precision,
                                                      900000 + args.seed, args.tol, args.every, args.max_steps,
                                                      args.probe, args.accept)
    s.close()
    log(f"physics baseline: median steps {np.median(base):.0f}, wall {base_wall:.1f} s, rejected {len(base_rej)}")
    picard = None
    if "picard" in arms:
        s, pic, pic_wall, pic_res, pic_rej = run_picard(ho_ch, args.precision, 900000 + args.seed, args.tol,
                                                        args.every, args.max_steps, args.probe)
        s.close()
        log(f"pb-picard: median steps {np.median(pic):.0f}, wall {pic_wall:.1f} s, rejected {len(pic_rej)}")
        picard = {"steps": pic.tolist(), "rejected": pic_rej, "wall_s": round(pic_wall, 2),
                  "probe_residuals": [float(v) for v in pic_res], "summary": summarize(cold, pic)}
    out = {
        "experiment": "time-to-steady, surrogate-warm vs cold",
        "tol": args.tol, "check_every": args.every, "precision": args.precision,
        "train": {"channels": len(tr_ch), "redrawn": rej_tr, "converged": int(good.sum()),
                  "wall_s": round(tr_wall, 2), "train_wall_s": round(train_wall, 2), "history": hist},
        "heldout": {"channels": len(ho_ch), "redrawn": rej_ho, "nx": list(nxs), "ny": list(ny),
                    "cold_wall_s": round(cold_wall, 2), "warm_wall_s": round(warm_wall, 2),
                    "wall_ratio": round(cold_wall / warm_wall, 3), "rejected": rejected,
                    "probe_residual_median": float(np.median(probe_res)),
                    "probe_residuals": [float(v) for v in probe_res],
                    "cold_steps": cold.tolist(), "warm_steps": warm.tolist(),
                    "summary": summarize(cold, warm),
                    "baseline_steps": base.tolist(), "baseline_rejected": base_rej,
                    "baseline_probe_residual_median": float(np.median(base_res)),
                    "baseline_summary": summarize(cold, base), "pb_picard": picard},
        "targets": args.targets,
    }
    if args.save:
        np.savez(args.save, params=net.get_params(), X=X, T=T, steps=tr_steps)
    if args.save_net:
        S.save_net(args.save_net, net, command=" ".join(["run_time_to_steady.py"] + sys.argv[1:]),
                   train_channels=int(good.sum()), ny=list(ny), nx_train=args.nx_train, seed=args.seed,
                   targets=args.targets, epochs=args.epochs, val_mse=float(hist[-1][2]))
    print(json.dumps(out), flush=True)


def held_out_only(args, ny, nxs, arms):
    """The held-out set alone: cold, optionally the physics baseline and the
    reference's PB-Picard start, each against cold."""
    ho_ch, rej_ho = S.sample_channels(args.heldout, 900000 + args.seed, nx_choices=nxs, ny_choices=ny)
    seed = 900000 + args.seed
    s, cold, cold_wall = run_cold(ho_ch, args.precision, seed, args.tol, args.every, args.max_steps)
    s.close()
    log(f"cold: median steps {np.median(cold):.0f}, wall {cold_wall:.1f} s")
    out = {"experiment": "time-to-steady on the held-out channels, other starts vs cold", "tol": args.tol,
           "check_every": args.every, "precision": args.precision,
           "heldout": {"channels": len(ho_ch), "redrawn": rej_ho, "nx": list(nxs), "ny": list(ny),
                       "cold_steps": cold.tolist(), "cold_wall_s": round(cold_wall, 2)}}
    if "baseline" in arms:
        s, base, base_wall, base_res, base_rej = run_warm(ho_ch, predictor(None, True), args.precision, seed,
                                                          args.tol, args.every, args.max_steps, args.probe,
                                                          args.accept)
        s.close()
        log(f"physics baseline: median steps {np.median(base):.0f}, wall {base_wall:.1f} s")
        out["heldout"]["baseline"] = {"steps": base.tolist(), "rejected": base_rej, "wall_s": round(base_wall, 2),
                                      "summary": summarize(cold, base)}
    if "picard" in arms:
        s, pic, pic_wall, pic_res, pic_rej = run_picard(ho_ch, args.precision, seed, args.tol, args.every,
                                                        args.max_steps, args.probe)
        s.close()
        log(f"pb-picard: median steps {np.median(pic):.0f}, wall {pic_wall:.1f} s, rejected {len(pic_rej)}")
        out["heldout"]["pb_picard"] = {"steps": pic.tolist(), "rejected": pic_rej, "wall_s": round(pic_wall, 2),
                                       "probe_residuals": [float(v) for v in pic_res],
                                       "summary": summarize(cold, pic)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
