#!/usr/bin/env python3
"""Throughput of the EK-PNP coupled step on one MI355X.

A step is one coupled time step (fused LBM collide-stream of fluid + both ion
species, forward x-FFT, wall-normal tridiagonal solve, inverse x-FFT) over the
whole workload, with every input resident in HBM.  Default workload: the
single 4096 x 1024 channel (tau 0.6, zeta -2, kappa h 32, E_x 5e-5, FP32
deviation populations, FP64 Poisson and reductions), seeded cold start.

Prints one JSON line (see the keys below).  value = MLUPS, lattice updates
per second over the whole job / 1e6 (main.cu:248-250 counts all NX*NY nodes
per step; so does this).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ek-pnp_amd"))

METRIC = "MLUPS (% HBM peak); time-to-steady @res 1e-8, surrogate-warm vs cold"
HBM_PEAK_GBS = 8000.0

# algorithmic bytes per lattice update
#   populations 27 x (read + write); coupling in FP64: psi read + rho_e write
#   by the LBM kernel, rho_e read + psi write by the Poisson solve
def alg_bytes(precision):
    pop = 4 if precision == "fp32" else 8
    lbm = 27 * 2 * pop + 16
    return lbm, lbm + 16


def workload(name):
    import ekpnp

    if name == "c1":
        p = ekpnp.lattice_params(4096, 1024, tau=0.6, zeta=-2.0, kappa_h=32.0, e_x=5e-5)
        return [p], "fp32", 1, "single channel 4096x1024, tau 0.6, zeta -2, kappa_h 32, E_x 5e-5"
    if name == "c0":
        p = ekpnp.lattice_params(64, 128, tau=0.8, zeta=-1.0, kappa_h=8.0, e_x=1e-4)
        return [p], "fp64", 0, "single channel 64x128, tau 0.8, zeta -1, kappa_h 8, E_x 1e-4, FP64"
    if name == "c2":
        from ekpnp import surrogate as S

        ps, _ = S.sample_channels(256, 2)
        return ps, "fp32", 2, ("batched sweep: 256 channels, seed 2, nx ~ U{32,64,128,256}, ny ~ U{64..1024 "
                               "step 64}, zeta ~ U[-4,-0.4], kappa_h ~ logU[1,50], E_x ~ logU[1e-5,1e-3]")
    if name == "c4":
        p = ekpnp.lattice_params(8192, 2048, tau=0.8, zeta=-1.5, kappa_h=64.0, e_x=1e-4)
        return [p], "fp32", 4, ("single channel 8192x2048, wall zeta(x) = -1.5 + sum_k a_k sin(2 pi k x/nx "
                                "+ phi_k), k = 1..4, seed 4, kappa_h 64, E_x 1e-4")
    raise SystemExit(f"unknown config {name}")


def wall_zeta(nx, seed, zeta0=-1.5, modes=4, amp=0.5):
    """zeta(x) = zeta0 + sum_k a_k sin(2 pi k x / nx + phi_k), a_k ~ U[0, amp],
    phi_k ~ U[0, 2 pi), drawn from a seeded numpy stream."""
    import numpy as np

    rng = np.random.default_rng(seed)
    a = rng.uniform(0.0, amp, modes)
    phi = rng.uniform(0.0, 2 * np.pi, modes)
    x = np.arange(nx)
    return zeta0 + sum(a[k] * np.sin(2 * np.pi * (k + 1) * x / nx + phi[k]) for k in range(modes))


def apply_walls(s, name, params, seed):
    """Per-x wall potential and Boltzmann wall concentrations (the 8192 x 2048
    heterogeneous-wall channel)."""
    import numpy as np

    if name != "c4":
        return
    for ch, p in enumerate(params):
        z = wall_zeta(p.nx, seed)
        s.set_wall_profile(ch, psi_lo=z, psi_hi=z, cpos_lo=np.exp(-z), cpos_hi=np.exp(-z),
                           cneg_lo=np.exp(z), cneg_hi=np.exp(z))


def calibrate_copy():
    """The library's non-temporal float4 copy of known size (ekpnp_copy_bandwidth:
    1 GiB read + 1 GiB written per dispatch): the reference kernel for scaling
    the FETCH_SIZE / WRITE_SIZE counters, and the streaming ceiling."""
    import ekpnp

    r = ekpnp.copy_bandwidth(1 << 30, reps=5)
    print(json.dumps({"copy_bytes_read": 1 << 30, "copy_bytes_written": 1 << 30,
                      "ms_per_copy": round(r["ms_per_copy"], 4), "GBps": round(r["GBps"], 1)}), flush=True)


def measured_traffic(config, precision, wrows=0):
    """HBM bytes per launch of the row kernel's interior rows, for the tile
    shape the run used, from the committed counter summary of this workload
    (FETCH_SIZE and WRITE_SIZE scaled by the copy calibration), or None when
    no summary exists for it."""
    path = os.path.join(ROOT, "profiles", f"k_lbm_{config}_{precision}_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    by_shape = d.get("bytes_per_launch_by_tile_rows", {})
    if str(wrows) in by_shape:
        return by_shape[str(wrows)]
    return d.get("bytes_per_launch") if wrows == 0 else None


def timed_start(warmup, steps, precision):
    """First step of the timed region: at least warmup + the graph-building
    calls, and no FP32 re-centring inside the building calls or the timed
    one."""
    pre = (1 + (steps & 1)) * steps
    s0 = warmup + pre
    if precision != "fp32":
        return s0
    import ekpnp

    while True:
        hits = [r for r in ekpnp.rebase_steps(s0 + steps) if s0 - pre < r <= s0 + steps]
        if not hits:
            return s0
        s0 = hits[-1] + pre


NET_PATH = os.path.join(ROOT, "ek-pnp_amd", "ekpnp", "warmstart_net.npz")


def time_to_steady_leg(n_channels, seed, tol=1e-8, every=100, max_steps=2_000_000):
    """Batched time-to-steady on held-out channels: the same seeded channels
    run cold (uniform c0 + noise) and warm-started from the committed
    network's predicted profiles (accepted by the u_x residual test after a
    100 + 200-step probe, else restarted cold), each batch in one context to
    the u_x residual tol (FP64 residual over the interior rows, checked every
    `every` steps; warm counts include the settling and probe steps)."""
    import numpy as np

    import run_time_to_steady as T
    from ekpnp import surrogate as S

    meta = S.net_meta(NET_PATH)
    # a bounded sample: the network's n_y up to 128 (cold starts of taller
    # channels take millions of steps; the full held-out comparison over
    # n_y 64..1024 is run_c3_scale.py heldout)
    ny = tuple(int(v) for v in meta["ny"] if int(v) <= 128)
    net = S.load_net(NET_PATH)
    chans, _ = S.sample_channels(n_channels, seed, nx_choices=(32, 64), ny_choices=ny)
    s, cold, cold_wall = T.run_cold(chans, "fp32", seed, tol, every, max_steps)
    s.close()
    s, warm, warm_wall, res, rejected = T.run_warm(chans, T.predictor(net, True), "fp32", seed, tol, every,
                                                  max_steps, 200, 1e-4)
    s.close()
    net.close()
    summ = T.summarize(cold, warm)
    return {
        "tol": tol, "check_every": every, "seed": seed,
        "channels": f"{len(chans)} held-out channels, nx in {{32, 64}}, ny in {list(ny)}, zeta, kappa h, E_x "
                    f"as the batched sweep", "network": os.path.relpath(NET_PATH, ROOT),
        "network_trained_on": f"{meta.get('train_channels')} steady states (seed {meta.get('seed')})",
        "cold_steps_median": summ["cold_steps_median"], "warm_steps_median": summ["warm_steps_median"],
        "step_gain_median": round(summ["step_ratio_median"], 2) if summ["step_ratio_median"] else None,
        "makespan_cold": summ["makespan_cold"], "makespan_warm": summ["makespan_warm"],
        "makespan_gain": round(summ["makespan_ratio"], 2),
        "wall_s_cold": round(cold_wall, 2), "wall_s_warm": round(warm_wall, 2),
        "wall_gain": round(cold_wall / warm_wall, 2), "rejected": len(rejected),
        "converged": summ["both_converged"],
    }


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(params, seconds_target=15.0, single_seconds=6.0):
    """The FP64 oracle on the host cores, bounded sample of the same workload;
    then the same on one thread, a shorter sample."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    p = params[0]
    o = O.Oracle(p)
    o.initialization(0, 1, 0)
    o.init_equilibrium()
    threads = O.max_threads()
    t1 = o.time_steps(1)
    n = max(1, min(200, int(seconds_target / max(t1, 1e-6))))
    t = o.time_steps(n)
    nodes = p.nx * p.ny
    O.set_threads(1)
    t1s = o.time_steps(1)
    n1 = max(1, min(50, int(single_seconds / max(t1s, 1e-6))))
    ts = o.time_steps(n1)
    O.set_threads(threads)
    return {
        "value": round(n * nodes / t / 1e6, 3),
        "unit": "MLUPS",
        "cores": threads,
        "kind": "port",
        "sample": f"FP64 oracle, {p.nx}x{p.ny} channel, {n} coupled steps after 1 warm-up step "
                  f"({t:.1f} s), OpenMP over rows",
        "cpu": cpu_model(),
        "host_cpus": os.cpu_count(),
        "single_thread": {"value": round(n1 * nodes / ts / 1e6, 3), "unit": "MLUPS",
                          "sample": f"{n1} coupled steps after 1 warm-up step ({ts:.1f} s), 1 thread"},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--config", default="c4", choices=("c0", "c1", "c2", "c4"))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--tts-channels", type=int, default=32,
                    help="held-out channels of the time-to-steady leg (0: skip it)")
    ap.add_argument("--tts-seed", type=int, default=905000)
    ap.add_argument("--stage-steps", type=int, default=20)
    ap.add_argument("--compute", choices=("fp32", "fp64"), default="fp64",
                    help="collision arithmetic for FP32 population storage")
    ap.add_argument("--tune-reps", type=int, default=10,
                    help="Poisson solves per candidate plan in the autotuner (0: default plan)")
    ap.add_argument("--plan", default=None, metavar="FFT_ELEMS,TRI_ROWS,FAMILY[,LBM_WROWS]",
                    help="use this Poisson plan (family 0 one-tile, 1 persistent, 2 row-pair) and, if given, "
                         "this LBM tile shape (0, 2, 4 rows per row-kernel tile) instead of autotuning, e.g. "
                         "to profile the kernels the autotuners picked")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 1 GiB device copy three times (the reference kernel "
                         "for scaling the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU: channels and y-slabs would shard "
                                   "across GPUs, but no multi-GPU path is built, so --gpus must be 1"}))
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy()
        return

    import ekpnp

    params, precision, seed, desc = workload(args.config)
    nodes = sum(p.nx * p.ny for p in params)
    print(f"bench: {args.config} ({desc}), {precision}, warmup {args.warmup}, steps {args.steps}",
          file=sys.stderr, flush=True)
    s = ekpnp.Solver(params, precision=precision, compute_fp64=(args.compute == "fp64"))
    plan_vals = [int(v) for v in args.plan.split(",")] if args.plan else []
    if plan_vals:
        fe, tr, fam = plan_vals[:3]
        s.set_poisson_plan(fe, tr, fam)
        plan = s.poisson_plan()
    else:
        plan = s.autotune(args.tune_reps) if args.tune_reps > 0 else None
    apply_walls(s, args.config, params, seed)
    s.initialization(ekpnp.COLD_NOISE, seed)
    s.init_equilibrium()
    s.step(args.warmup)
    # the row kernel's tile shape, timed on the warmed-up state (not advanced)
    if len(plan_vals) > 3:
        s.set_lbm_tiles(plan_vals[3])
        lbm_tiles = {"wrows": s.lbm_tiles()}
    else:
        lbm_tiles = s.autotune_lbm(args.tune_reps) if args.tune_reps > 0 else {"wrows": s.lbm_tiles()}
    # The timed call replays cached hipGraphs: an untimed call of the same
    # length captures them first (two for odd counts, which end on the other
    # lattice parity).  FP32 storage re-centres its references on a schedule
    # that starts at 8 steps and doubles to 4096 (and cuts a call's graph
    # chunks there), so both calls are placed between two re-centrings, after
    # the cold start's early ones.
    start = timed_start(args.warmup, args.steps, precision)
    pre = 1 + (args.steps & 1)
    s.step(start - pre * args.steps - args.warmup)
    for _ in range(pre):
        s.step(args.steps)
    assert s.steps_done == start
    s.synchronize()
    t0 = time.perf_counter()
    s.step(args.steps)
    s.synchronize()
    t1 = time.perf_counter()
    dt = t1 - t0
    mlups = args.steps * nodes / dt / 1e6
    print(f"bench: timed {args.steps} steps in {dt:.3f} s", file=sys.stderr, flush=True)

    # per-kernel durations, HIP events on the library's stream
    st = s.time_stages(args.stage_steps)
    # the streaming ceiling on this part: the library's non-temporal copy,
    # of 1 GiB and of one lattice's populations (the LBM stage reads one
    # lattice and writes the other; the copy rate falls with the footprint)
    copy = ekpnp.copy_bandwidth(1 << 30, reps=10)
    lattice_bytes = 27 * nodes * (4 if precision == "fp32" else 8)
    copy_fp = ekpnp.copy_bandwidth(lattice_bytes // 16 * 16, reps=10)
    red = s.reductions()
    walls_in_tri = s.poisson_plan()["walls_in_tri"]
    if plan is not None:
        plan["walls_in_tri"] = walls_in_tri
    s.close()
    lbm_b, coupled_b = alg_bytes(precision)
    if walls_in_tri:
        # the LBM stage is the interior-row launch; rows 0 and ny-1 are
        # updated inside the tridiagonal launch
        stage_nodes = sum(p.nx * (p.ny - 2) if p.nx % 64 == 0 else p.nx * p.ny for p in params)
        kernel = ("k_lbm_rows<false>: fused collide-stream of the interior rows "
                  "(the two wall rows ride on the tridiagonal launch)")
        if any(p.nx % 64 for p in params):
            kernel = ("LBM stage: k_lbm_rows<false> (interior rows of the channels with nx a multiple of 64; "
                      "their wall rows ride on the tridiagonal launch) + k_lbm (every row of the others)")
    else:
        stage_nodes = nodes
        kernel = ("LBM stage: k_lbm_rows<false> (interior rows) + k_lbm_rows<true> (wall rows), "
                  "fused collide-stream")
    achieved = lbm_b * stage_nodes / (st["lbm"] * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": round(mlups, 2),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "untimed_steps": start,
        "ms_per_step": round(dt / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 populations, f64 Poisson" if precision == "fp32" else "f64",
        "data": "synthetic (seeded cold start: c0 = 1 with U(-1e-3, 1e-3) noise)",
        "config": {"workload": desc, "nodes": nodes, "precision": precision,
                   "collision_arithmetic": "f64" if (precision == "fp64" or args.compute == "fp64") else "f32",
                   "poisson_plan": plan, "lbm_tiles": lbm_tiles},
        "hbm_frac_coupled": round(mlups * 1e6 * coupled_b / (HBM_PEAK_GBS * 1e9), 4),
        "stages_ms": {k: round(v, 5) for k, v in st.items()},
        "roofline": {
            "bound": "hbm",
            "kernel": kernel,
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic(args.config, precision, lbm_tiles["wrows"]),
        },
        "finite": bool(red[:, 6].sum() == 0),
        "copy_ceiling_gbps": round(copy["GBps"], 1),
        "lbm_frac_of_copy": round(achieved / copy["GBps"], 4),
        "copy_same_footprint": {"bytes": lattice_bytes // 16 * 16, "gbps": round(copy_fp["GBps"], 1),
                                "lbm_frac": round(achieved / copy_fp["GBps"], 4)},
    }
    if args.tts_channels > 0 and os.path.exists(NET_PATH):
        print("bench: time to steady, cold vs surrogate-warm", file=sys.stderr, flush=True)
        out["time_to_steady"] = time_to_steady_leg(args.tts_channels, args.tts_seed)
    if not args.no_cpu_baseline:
        print("bench: cpu baseline (oracle)", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(params)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
