"""The benchmark workloads at their full sizes against the FP64 oracle.

A few coupled steps of the 4096 x 1024 channel (row-pair x-FFTs) and of the
8192 x 2048 channel with zeta(x) walls (the bench workload: wide row-pair
x-FFTs, the plan the autotuner picks for it),
FP32 deviation storage with FP64 arithmetic (the row LBM kernel, its wall rows
on the tridiagonal launch, the tridiagonal at its real column height),
compared field by field with the oracle started from the
same seeded state; and fluid mass conservation over a few hundred steps of the
full-size channel, a property that does not depend on the oracle.
"""
import numpy as np
import pytest

import ekpnp
import oracle as O
from helpers import assert_parity, field_errors

pytestmark = pytest.mark.gpu


def wall_zeta(nx, seed, zeta0=-1.5, modes=4, amp=0.5):
    """zeta(x) = zeta0 + sum_k a_k sin(2 pi k x / nx + phi_k), seeded numpy stream."""
    rng = np.random.default_rng(seed)
    a = rng.uniform(0.0, amp, modes)
    phi = rng.uniform(0.0, 2 * np.pi, modes)
    x = np.arange(nx)
    return zeta0 + sum(a[k] * np.sin(2 * np.pi * (k + 1) * x / nx + phi[k]) for k in range(modes))


def _run_pair(p, seed, nsteps, walls=None, plan=None):
    s = ekpnp.Solver([p], precision="fp32")
    if plan is not None:
        s.set_poisson_plan(*plan)
    o = O.Oracle(p, poisson_method=O.POISSON_XDFT_THOMAS)
    if walls is not None:
        s.set_wall_profile(0, **walls)
        o.set_wall_profile(**walls)
    s.initialization(ekpnp.COLD_NOISE, seed)
    o.initialization(ekpnp.COLD_NOISE, seed, 0)
    s.init_equilibrium()
    o.init_equilibrium()
    s.step(nsteps)
    o.step(nsteps)
    return s, o


def test_bench_channel_4096x1024_against_the_oracle():
    print("case 4096x1024 tau 0.6 zeta -2 kappa_h 32 E_x 5e-5 seed 1, 12 steps", flush=True)
    p = ekpnp.lattice_params(4096, 1024, tau=0.6, zeta=-2.0, kappa_h=32.0, e_x=5e-5)
    s, o = _run_pair(p, 1, 12)
    # the default plan for this width: row-pair x-FFTs, wall rows on the
    # tridiagonal launch
    plan = s.poisson_plan()
    print("plan", plan, flush=True)
    assert plan["fft_kernels"] == "row-pair" and plan["walls_in_tri"]
    errs = field_errors(s, 0, o)
    print(" ".join(f"{k}={v:.2e}" for k, v in errs.items()), flush=True)
    for f in ("rho", "ux", "cpos", "cneg", "psi"):
        assert errs[f] < 1e-5, (f, errs[f])


def test_wall_profile_channel_8192x2048_against_the_oracle():
    print("case 8192x2048 tau 0.8 zeta(x) seed 4 kappa_h 64 E_x 1e-4, 4 steps", flush=True)
    p = ekpnp.lattice_params(8192, 2048, tau=0.8, zeta=-1.5, kappa_h=64.0, e_x=1e-4)
    z = wall_zeta(p.nx, 4)
    walls = dict(psi_lo=z, psi_hi=z, cpos_lo=np.exp(-z), cpos_hi=np.exp(-z), cneg_lo=np.exp(z), cneg_hi=np.exp(z))
    s, o = _run_pair(p, 4, 4, walls, plan=(4096, 16, 2))
    plan = s.poisson_plan()
    print("plan", plan, flush=True)
    assert plan["fft_kernels"] == "row-pair" and plan["walls_in_tri"]
    errs = field_errors(s, 0, o)
    print(" ".join(f"{k}={v:.2e}" for k, v in errs.items()), flush=True)
    for f in ("rho", "ux", "cpos", "cneg", "psi"):
        assert errs[f] < 1e-5, (f, errs[f])


@pytest.mark.timeout(900)
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_batched_sweep_against_the_oracle(precision):
    # the 256-channel batched sweep (seed 2: 18.2 M nodes, nx 32..256, ny
    # 64..1024; row-kernel and general-kernel channels, the wall rows on the
    # tridiagonal launch) 32 steps from the seeded cold start (FP32
    # re-centrings at 8, 16, 32 included), every channel against the oracle
    # started from the same state.  FP64 storage: within 1e-11 (max-norm with
    # a cancellation floor).  FP32 storage: rho, c+-, psi within 1e-5 relative
    # L2 on every channel; u_x, which starts at rest and is still ~1e-6 here
    # on channels with a weak applied field, is quantised by the FP32 storage
    # at ~1e-5 of itself through the first ~1000 steps of a cold start
    # (channel 0, E_x 1.5e-5: 1.5e-5 from step 16 to 512, 8.7e-6 at 1024,
    # 1.6e-6 at 8192; the 5000-step comparison of this sweep against FP64
    # storage stays within 1e-5 on every channel), so here it is bounded at
    # 3e-5 and within 1e-5 on at least 90 % of the channels
    from ekpnp import surrogate as S

    print("case batched sweep vs oracle", precision, flush=True)
    ps, _ = S.sample_channels(256, 2)
    s = ekpnp.Solver(ps, precision=precision)
    s.initialization(ekpnp.COLD_NOISE, 2)
    s.init_equilibrium()
    s.step(32)
    errs = []
    for ch, p in enumerate(ps):
        o = O.Oracle(p, poisson_method=O.POISSON_XDFT_THOMAS)
        o.initialization(ekpnp.COLD_NOISE, 2, ch)
        o.init_equilibrium()
        o.step(32)
        if precision == "fp64":
            assert_parity(s, ch, o, 1e-11, tag=f"sweep ch{ch} {p.nx}x{p.ny}")
        else:
            errs.append(field_errors(s, ch, o, fields=("rho", "ux", "cpos", "cneg", "psi")))
        o.close()
    s.close()
    if precision == "fp32":
        for f in ("rho", "ux", "cpos", "cneg", "psi"):
            e = np.array([r[f] for r in errs])
            print(f"{f}: largest {e.max():.2e} (channel {int(e.argmax())}), median {np.median(e):.2e}, "
                  f"{int((e >= 1e-5).sum())} of {len(e)} at or above 1e-5", flush=True)
        for f in ("rho", "cpos", "cneg", "psi"):
            assert max(r[f] for r in errs) < 1e-5, f
        ux = np.array([r["ux"] for r in errs])
        assert ux.max() < 3e-5 and (ux < 1e-5).mean() >= 0.9


def test_bench_channel_conserves_fluid_mass():
    # full-way bounce-back walls and periodic x: the fluid mass is an exact
    # invariant of the scheme; FP32 storage rounds each stored deviation, so
    # it may drift only at that level
    print("case 4096x1024 mass, 400 steps", flush=True)
    p = ekpnp.lattice_params(4096, 1024, tau=0.6, zeta=-2.0, kappa_h=32.0, e_x=5e-5)
    s = ekpnp.Solver([p], precision="fp32")
    s.initialization(ekpnp.COLD_NOISE, 1)
    s.init_equilibrium()
    s.step(2)
    m0 = s.reductions()[0, ekpnp.RED_NAMES.index("mass")]
    s.step(400)
    red = s.reductions()[0]
    m1 = red[ekpnp.RED_NAMES.index("mass")]
    print("mass", m0, m1, (m1 - m0) / m0, flush=True)
    assert red[ekpnp.RED_NAMES.index("nonfinite")] == 0
    assert abs(m1 - m0) <= 1e-7 * m0


def test_bench_channel_fp32_tracks_fp64_storage_across_a_recentring():
    # 6000 steps in one call (the storage reference is re-centred at step
    # 4096 inside it): FP32 deviation storage against FP64 storage, the path
    # that matches the oracle to rounding, within the 1e-5 relative-L2 bar
    print("case 4096x1024 fp32 vs fp64 storage, 6000 steps", flush=True)
    p = ekpnp.lattice_params(4096, 1024, tau=0.6, zeta=-2.0, kappa_h=32.0, e_x=5e-5)
    out = {}
    for prec in ("fp64", "fp32"):
        s = ekpnp.Solver([p], precision=prec)
        s.initialization(ekpnp.COLD_NOISE, 1)
        s.init_equilibrium()
        s.step(6000)
        out[prec] = {f: s.get_field(0, f) for f in ("ux", "psi", "cpos", "cneg")}
        s.close()
    errs = {f: float(np.linalg.norm(out["fp32"][f] - out["fp64"][f]) / np.linalg.norm(out["fp64"][f]))
            for f in out["fp64"]}
    print(" ".join(f"{k}={v:.2e}" for k, v in errs.items()), flush=True)
    for f, e in errs.items():
        assert e < 1e-5, (f, e)



def test_batched_sweep_fp32_tracks_fp64_storage():
    # the 256-channel batched sweep (seed 2: nx in {32, 64, 128, 256}, ny in
    # 64..1024, mixed zeta, kappa h, E_x; row kernel and k_lbm channels, wall
    # rows on the tridiagonal launch) after 5000 steps: FP32 deviation
    # storage against FP64 storage, every channel within the 1e-5 bar
    from ekpnp import surrogate as S

    print("case batched sweep, 256 channels seed 2, 5000 steps", flush=True)
    ps, _ = S.sample_channels(256, 2)
    out = {}
    for prec in ("fp64", "fp32"):
        s = ekpnp.Solver(ps, precision=prec)
        s.initialization(ekpnp.COLD_NOISE, 2)
        s.init_equilibrium()
        s.step(5000)
        out[prec] = [{f: s.get_field(c, f) for f in ("ux", "psi", "cpos", "cneg")} for c in range(len(ps))]
        s.close()
    worst = {f: 0.0 for f in ("ux", "psi", "cpos", "cneg")}
    for a, b in zip(out["fp32"], out["fp64"]):
        for f in worst:
            worst[f] = max(worst[f], float(np.linalg.norm(a[f] - b[f]) / np.linalg.norm(b[f])))
    print(" ".join(f"{k}={v:.2e}" for k, v in worst.items()), flush=True)
    for f, e in worst.items():
        assert e < 1e-5, (f, e)


@pytest.mark.timeout(900)
def test_wall_profile_channel_fp32_tracks_fp64_storage_20000_steps():
    # The 8192 x 2048 channel with zeta(x) walls: inside the double layer the
    # concentrations vary along x by a factor of up to ~50, which a reference
    # per row cannot follow (its stored ion deviations were of the order of
    # the concentration, and u_x drifted past 1e-5 after ~12 000 steps); with
    # a reference per 64-node block FP32 storage must stay within the 1e-5
    # relative-L2 bar of FP64 storage (the path that matches the oracle to
    # rounding) at every checkpoint through 20 000 steps.
    print("case 8192x2048 zeta(x) seed 4, fp32 vs fp64 storage, 20000 steps", flush=True)
    p = ekpnp.lattice_params(8192, 2048, tau=0.8, zeta=-1.5, kappa_h=64.0, e_x=1e-4)
    z = wall_zeta(p.nx, 4)
    walls = dict(psi_lo=z, psi_hi=z, cpos_lo=np.exp(-z), cpos_hi=np.exp(-z), cneg_lo=np.exp(z), cneg_hi=np.exp(z))
    checkpoints = (5001, 10001, 15001, 20000)
    fields = ("ux", "uy", "psi", "cpos", "cneg")
    out = {}
    for prec in ("fp64", "fp32"):
        s = ekpnp.Solver([p], precision=prec)
        s.set_wall_profile(0, **walls)
        s.initialization(ekpnp.COLD_NOISE, 4)
        s.init_equilibrium()
        out[prec] = []
        for n in checkpoints:
            s.step(n - s.steps_done)
            out[prec].append({f: s.get_field(0, f) for f in fields})
            print(prec, "step", n, flush=True)
        s.close()
    worst = {f: 0.0 for f in fields}
    for n, a, b in zip(checkpoints, out["fp32"], out["fp64"]):
        errs = {f: float(np.linalg.norm(a[f] - b[f]) / np.linalg.norm(b[f])) for f in fields}
        print("step", n, " ".join(f"{k}={v:.2e}" for k, v in errs.items()), flush=True)
        for f in fields:
            worst[f] = max(worst[f], errs[f])
    for f, e in worst.items():
        assert e < 1e-5, (f, e)
