"""GPU path (libekpnp.so through the C ABI) against the FP64 CPU oracle.

Every comparison runs both sides from the same seeded input.  FP64 storage
must agree to rounding (1e-11 relative L2 per field); FP32 deviation storage
to the 1e-5 relative-L2 bar on u_x, psi, c+ and c-.
"""
import numpy as np
import pytest

import ekpnp
import oracle as O
from helpers import FIELDS, assert_parity, field_errors, rel_l2

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-11


def _pair(params, precision="fp64", seed=0, mode=ekpnp.COLD_NOISE, method=O.POISSON_XDFT_THOMAS):
    o = O.Oracle(params, poisson_method=method)
    o.initialization(mode, seed, 0)
    o.init_equilibrium()
    s = ekpnp.Solver([params], precision=precision)
    s.initialization(mode, seed)
    s.init_equilibrium()
    return s, o


def _report(tag, errs):
    print(tag, " ".join(f"{k}={v:.2e}" for k, v in errs.items()), flush=True)


def test_initial_fields_match():
    p = ekpnp.lattice_params(64, 128, tau=0.8, zeta=-1.0, kappa_h=8.0, e_x=1e-4)
    s, o = _pair(p)
    assert_parity(s, 0, o, 1e-13, fields=("rho", "cpos", "cneg", "psi", "ex", "ey"), tag="init")


@pytest.mark.parametrize("nsteps", [1, 10, 100, 1000])
def test_fp64_64x128_channel_steps(nsteps):
    # the 64 x 128 channel: tau 0.8, zeta -1, kappa h 8, E_x 1e-4, seed 0
    p = ekpnp.lattice_params(64, 128, tau=0.8, zeta=-1.0, kappa_h=8.0, e_x=1e-4)
    s, o = _pair(p)
    s.step(nsteps)
    o.step(nsteps)
    assert_parity(s, 0, o, FP64_TOL, tag=f"fp64 n={nsteps}")


def test_fp64_populations_match():
    p = ekpnp.lattice_params(32, 40, tau=0.7, zeta=-1.5, kappa_h=6.0, e_x=2e-4)
    s, o = _pair(p, seed=3)
    s.step(25)
    o.step(25)
    for sp in ("f", "h", "g"):
        e = rel_l2(s.get_populations(0, sp), o.get_populations(sp))
        print("pops", sp, e, flush=True)
        assert e < FP64_TOL


def test_reference_si_case_dft_path():
    # the reference's own compiled-in case: NX = 10 (non-power-of-two x rows
    # use the direct DFT), NY = 101, SI units, g_eq0 quirk on, zero force,
    # gpu_initialization start (c = 0); 300 steps
    p = ekpnp.reference_params()
    s, o = _pair(p, mode=ekpnp.REFERENCE)
    s.step(300)
    o.step(300)
    errs = field_errors(s, 0, o)
    _report("reference SI", errs)
    for f in ("rho", "cpos", "cneg", "psi"):
        assert errs[f] < FP64_TOL, (f, errs[f])
    # SI velocities are O(1e-12) m/s here: compare on their own scale
    for f in ("ux", "uy"):
        a, b = s.get_field(0, f), o.get_field(f)
        assert np.max(np.abs(a - b)) <= 1e-11 * max(np.max(np.abs(b)), 1e-30) + 1e-30


def test_fp64_force_and_flags():
    # body force on, ions drifting in the applied field, moving top wall
    p = ekpnp.lattice_params(16, 24, tau=0.9, zeta=-2.0, kappa_h=4.0, e_x=1e-3)
    p.flags = ekpnp.FLAG_ION_DRIFT_APPLIED | ekpnp.FLAG_G0_FLUID_U
    p.uw = 1e-3
    p.exf = 1e-6
    s, o = _pair(p, seed=7)
    s.step(200)
    o.step(200)
    assert_parity(s, 0, o, 1e-10, tag="flags")


def test_stream_collide_save_then_fast_poisson_matches_step():
    p = ekpnp.lattice_params(32, 48, tau=0.8, zeta=-1.0, kappa_h=5.0, e_x=1e-4)
    s, o = _pair(p, seed=11)
    for _ in range(5):
        s.stream_collide_save()
        s.fast_poisson()
    o.step(5)
    assert_parity(s, 0, o, FP64_TOL, tag="split ops")


@pytest.mark.parametrize("nx,ny", [(64, 130), (48, 66), (256, 1026), (32, 2050), (8192, 6),
                                   (4, 300), (33, 70), (2047, 5), (2046, 9)])
def test_poisson_only(nx, ny):
    # psi from random c+- with per-x wall potentials, both oracle methods;
    # the last four widths take the direct DFT with several rows per tile
    # (up to 298 rows of nx = 4, ragged last tiles under the smaller plans;
    # odd widths; two rows of the widest non-power-of-two rows per tile)
    rng = np.random.default_rng(nx * 7919 + ny)
    p = ekpnp.lattice_params(nx, ny, tau=0.8, zeta=-1.0, kappa_h=8.0, e_x=1e-4)
    shp = (ny, nx)
    cp = 1 + 0.1 * rng.standard_normal(shp)
    cn = 1 + 0.1 * rng.standard_normal(shp)
    lo = -1 + 0.3 * rng.standard_normal(nx)
    hi = -0.5 + 0.3 * rng.standard_normal(nx)
    s = ekpnp.Solver([p], precision="fp64")
    s.set_wall_profile(0, psi_lo=lo, psi_hi=hi)
    s.set_fields(0, np.ones(shp), np.zeros(shp), np.zeros(shp), cp, cn)
    o = O.Oracle(p, poisson_method=O.POISSON_XDFT_THOMAS)
    o.set_wall_profile(psi_lo=lo, psi_hi=hi)
    o.set_fields(np.ones(shp), np.zeros(shp), np.zeros(shp), cp, cn)
    want = o.get_field("psi")
    e = rel_l2(s.get_field(0, "psi"), want)
    print("poisson", nx, ny, e, flush=True)
    assert e < 1e-12
    # every work decomposition solves the same system
    for fft_elems, tri_rows, persistent in ((1024, 8, True), (4096, 16, True), (256, 4, True),
                                            (2048, 16, False), (512, 8, False)):
        s.set_poisson_plan(fft_elems, tri_rows, persistent)
        s.fast_poisson()
        e = rel_l2(s.get_field(0, "psi"), want)
        print("poisson plan", fft_elems, tri_rows, persistent, e, flush=True)
        assert e < 1e-12
    # the autotuner keeps psi and leaves a valid plan behind
    before = s.get_field(0, "psi")
    info = s.autotune(3)
    print("autotune", info, flush=True)
    assert info["fft_elems"] in (1024, 2048, 4096) and info["tri_rows"] in (8, 16)
    assert np.array_equal(s.get_field(0, "psi"), before)
    s.fast_poisson()
    assert rel_l2(s.get_field(0, "psi"), want) < 1e-12


@pytest.mark.parametrize("nx,ny", [(2048, 7), (4096, 6), (4096, 4), (2048, 34), (8192, 7), (8192, 4), (8192, 10)])
def test_row_pair_fft_kernels(nx, ny):
    # the row-pair x-FFTs (two rows per work-group, the second row's loads in
    # flight during the first's transform): odd interior row counts leave a
    # one-row last tile (holding the upper wall lift), ny = 4 puts both wall
    # lifts in one tile.  nx = 8192 takes the wide kernels (one row buffer,
    # the second row in registers, the two-level twiddle table), asked for
    # with fft_elems = 4096, against the one-tile kernels with one row per
    # tile and the same table
    rng = np.random.default_rng(nx + ny)
    p = ekpnp.lattice_params(nx, ny, tau=0.8, zeta=-1.0, kappa_h=8.0, e_x=1e-4)
    shp = (ny, nx)
    cp = 1 + 0.1 * rng.standard_normal(shp)
    cn = 1 + 0.1 * rng.standard_normal(shp)
    lo = -1 + 0.3 * rng.standard_normal(nx)
    hi = -0.5 + 0.3 * rng.standard_normal(nx)
    s = ekpnp.Solver([p], precision="fp64")
    s.set_wall_profile(0, psi_lo=lo, psi_hi=hi)
    o = O.Oracle(p, poisson_method=O.POISSON_XDFT_THOMAS)
    o.set_wall_profile(psi_lo=lo, psi_hi=hi)
    o.set_fields(np.ones(shp), np.zeros(shp), np.zeros(shp), cp, cn)
    want = o.get_field("psi")
    elems = min(nx, 4096)
    s.set_poisson_plan(elems, 16, 0)
    s.set_fields(0, np.ones(shp), np.zeros(shp), np.zeros(shp), cp, cn)
    one_tile = s.get_field(0, "psi")
    s.set_poisson_plan(elems, 16, 2)
    assert s.poisson_plan()["fft_kernels"] == "row-pair"
    s.fast_poisson()
    pair = s.get_field(0, "psi")
    e = rel_l2(pair, want)
    print("row-pair", nx, ny, e, np.abs(pair - one_tile).max(), flush=True)
    assert e < 1e-12
    assert rel_l2(pair, one_tile) < 1e-14
    # a plan the channels do not allow falls back to the one-tile kernels
    s.set_poisson_plan(elems // 2, 16, 2)
    assert s.poisson_plan()["fft_kernels"] == "one-tile"
    s.fast_poisson()
    assert rel_l2(s.get_field(0, "psi"), want) < 1e-12


def test_plan_arguments_are_validated():
    s = ekpnp.Solver([ekpnp.lattice_params(64, 20)], precision="fp64")
    for args in ((0, 16, 0), (8192, 16, 0), (2048, 0, 0), (2048, 17, 0), (2048, 16, 3), (2048, 16, -1)):
        with pytest.raises(ekpnp.EkpnpError):
            s.set_poisson_plan(*args)
    s.set_poisson_plan(2048, 8, 1)
    plan = s.poisson_plan()
    assert plan["fft_elems"] == 2048 and plan["tri_rows"] == 8 and plan["fft_kernels"] == "persistent"
    assert not plan["walls_in_tri"]  # FP64 storage: no row kernel


@pytest.mark.parametrize("nx", [2048, 8192])
def test_row_pair_fft_kernels_in_a_coupled_batch(nx):
    # two nx-wide channels on the row-pair kernels (nx = 8192: the wide ones)
    # beside a 10-wide one on the direct DFT, FP64 storage, against the oracle
    # after 40 steps
    ps = [ekpnp.lattice_params(nx, 20, tau=0.7, zeta=-1.5, kappa_h=6.0, e_x=2e-4),
          ekpnp.lattice_params(nx, 11, tau=0.8, zeta=-1.0, kappa_h=3.0, e_x=1e-4),
          ekpnp.lattice_params(10, 16, tau=0.9, zeta=-0.5, kappa_h=3.0, e_x=1e-5)]
    s = ekpnp.Solver(ps, precision="fp64")
    s.set_poisson_plan(min(nx, 4096), 16, 2)
    assert s.poisson_plan()["fft_kernels"] == "row-pair"
    s.initialization(ekpnp.COLD_NOISE, 4)
    s.init_equilibrium()
    s.step(40)
    for ch, p in enumerate(ps):
        o = O.Oracle(p, poisson_method=O.POISSON_XDFT_THOMAS)
        o.initialization(ekpnp.COLD_NOISE, 4, ch)
        o.init_equilibrium()
        o.step(40)
        assert_parity(s, ch, o, FP64_TOL, tag=f"row-pair batch nx={nx} ch={ch}")


def test_batched_channels_match_single():
    # three channels of different size and physics in one context
    ps = [ekpnp.lattice_params(64, 128, tau=0.8, zeta=-1.0, kappa_h=8.0, e_x=1e-4),
          ekpnp.lattice_params(32, 64, tau=0.7, zeta=-2.0, kappa_h=4.0, e_x=5e-4),
          ekpnp.lattice_params(10, 37, tau=0.9, zeta=-0.5, kappa_h=3.0, e_x=1e-5)]
    s = ekpnp.Solver(ps, precision="fp64")
    s.initialization(ekpnp.COLD_NOISE, 5)
    s.init_equilibrium()
    s.step(60)
    for ch, p in enumerate(ps):
        o = O.Oracle(p)
        o.initialization(ekpnp.COLD_NOISE, 5, ch)
        o.init_equilibrium()
        o.step(60)
        assert_parity(s, ch, o, FP64_TOL, tag=f"batched ch{ch}")


def test_many_small_channels_match_single_runs():
    # 1200 small channels of mixed widths (power-of-two FFT rows, direct-DFT
    # rows, the row kernel and the general kernel) in one context: every tile
    # table and per-channel offset at a large channel count.  Each channel
    # starts from its own seeded fields (independent of its index), and a
    # sample of channels run alone must match the batch bit for bit after
    # 40 steps (FP32 storage, re-centrings at 8, 16, 32 included)
    rng = np.random.default_rng(12)
    widths = (4, 6, 8, 10, 16, 24, 32, 64, 128)
    ps, fields = [], []
    while len(ps) < 1200:
        nx, ny = int(rng.choice(widths)), int(rng.integers(4, 49))
        p = ekpnp.lattice_params(nx, ny, tau=float(rng.uniform(0.6, 1.0)), zeta=float(rng.uniform(-2.0, -0.3)),
                                 kappa_h=float(rng.uniform(0.5, 2.0)), e_x=float(rng.uniform(1e-5, 5e-4)))
        if ekpnp.check_params([p])[0]:
            continue  # flagged by the stability rules: redrawn
        shp = (ny, nx)
        cp = 1.0 + 0.01 * rng.standard_normal(shp)
        cn = 1.0 + 0.01 * rng.standard_normal(shp)
        ps.append(p)
        fields.append((np.ones(shp), np.zeros(shp), np.zeros(shp), cp, cn))
    print("case 1200 small channels", flush=True)

    def run(idx):
        s = ekpnp.Solver([ps[i] for i in idx], precision="fp32")
        for c, i in enumerate(idx):
            s.set_fields(c, *fields[i])
        s.init_equilibrium()
        s.step(40)
        return s

    big = run(range(len(ps)))
    for i in (0, 1, 2, 3, 377, 600, 911, 1198, 1199):
        one = run([i])
        for f in ("rho", "ux", "uy", "cpos", "cneg", "psi"):
            assert np.array_equal(big.get_field(i, f), one.get_field(0, f)), (i, ps[i].nx, ps[i].ny, f)
        one.close()
    red = big.reductions()
    assert (red[:, 6] == 0).all()
    big.close()


def _fp32_errors_64x128(compute64):
    p = ekpnp.lattice_params(64, 128, tau=0.8, zeta=-1.0, kappa_h=8.0, e_x=1e-4)
    o = O.Oracle(p)
    o.initialization(ekpnp.COLD_NOISE, 0, 0)
    o.init_equilibrium()
    s = ekpnp.Solver([p], precision="fp32", compute_fp64=compute64)
    s.initialization(ekpnp.COLD_NOISE, 0)
    s.init_equilibrium()
    for n in (3000, 17000):
        s.step(n)
        o.step(n)
        errs = field_errors(s, 0, o)
        _report(f"fp32 compute64={compute64} n={o.steps}", errs)
    return errs


def test_fp32_deviation_storage_64x128():
    # FP32 storage with FP64 collision arithmetic (the default): the 1e-5
    # relative-L2 bar after 20000 steps (past the 3 rebases it triggers)
    errs = _fp32_errors_64x128(True)
    for f in ("ux", "cpos", "cneg", "psi"):
        assert errs[f] < 1e-5, (f, errs[f])


def test_fp32_arithmetic_variant_is_bounded():
    # FP32 collision arithmetic (compute_fp64=False) is an option for
    # throughput comparisons, not the default, and does NOT meet the 1e-5 bar:
    # its u_x error grows to ~6e-5 over 20000 steps (the small drift velocity
    # is a difference of O(1) population sums), bounded here a few times
    # above what it measures at step 20000 (u_x 6.0e-5, psi 2.8e-7, c+-
    # 3.4e-8), so that a change in its error shows
    errs = _fp32_errors_64x128(False)
    for f in ("cpos", "cneg"):
        assert errs[f] < 2e-7, (f, errs[f])
    assert errs["psi"] < 3e-6
    assert errs["ux"] < 2e-4


@pytest.mark.parametrize("nx,ny", [(256, 40), (512, 24), (64, 37), (128, 20), (192, 9), (320, 5)])
def test_row_kernel_is_bit_identical_to_the_general_kernel(nx, ny, monkeypatch):
    # channels with nx a multiple of 64 run on k_lbm_rows (FP32 storage, FP64
    # arithmetic; a wave's nodes share a row, tiles may hold several rows or
    # parts of two); EKPNP_LBM_ROWS=0 sends every tile to k_lbm.  Same
    # operations in the same order: identical bits, also with the macroscopic
    # output of the last step and across a rebase.
    p = ekpnp.lattice_params(nx, ny, tau=0.7, zeta=-1.5, kappa_h=6.0, e_x=2e-4)
    runs = []
    for env in ("1", "0"):
        monkeypatch.setenv("EKPNP_LBM_ROWS", env)
        s = ekpnp.Solver([p], precision="fp32")
        s.initialization(ekpnp.COLD_NOISE, 5)
        s.init_equilibrium()
        s.step(150)
        s.rebase()
        s.step(151)
        runs.append({sp: s.get_populations(0, sp) for sp in ("f", "h", "g")} |
                    {f: s.get_field(0, f) for f in ("rho", "ux", "cpos", "cneg", "psi")})
    for k in runs[0]:
        np.testing.assert_array_equal(runs[0][k], runs[1][k], err_msg=k)


def test_wall_rows_on_the_tridiagonal_launch_are_bit_identical(monkeypatch):
    # The coupled step may update the row kernel's wall rows inside the
    # tridiagonal launch (k_tridiag_walls) instead of a launch of their own;
    # EKPNP_WALLS_IN_TRI=0 at plan time keeps the separate launch.  A batch of
    # row-kernel channels (one tile spanning several rows, one wide) and a
    # k_lbm channel, stepped through the graphs (macroscopic output on the
    # last step), the per-operator calls and a time_stages pass: identical
    # bits both ways.
    ps = [ekpnp.lattice_params(64, 30, tau=0.7, zeta=-1.5, kappa_h=6.0, e_x=2e-4),
          ekpnp.lattice_params(512, 20, tau=0.8, zeta=-1.0, kappa_h=4.0, e_x=1e-4),
          ekpnp.lattice_params(48, 26, tau=0.9, zeta=-2.0, kappa_h=5.0, e_x=3e-4)]
    runs = []
    for env in ("1", "0"):
        monkeypatch.setenv("EKPNP_WALLS_IN_TRI", env)
        s = ekpnp.Solver(ps, precision="fp32")
        assert s.poisson_plan()["walls_in_tri"] == (env == "1")
        s.initialization(ekpnp.COLD_NOISE, 9)
        s.init_equilibrium()
        s.step(97)
        s.stream_collide_save()
        s.fast_poisson()
        s.time_stages(3)
        s.step(64)
        out = {}
        for ch in range(len(ps)):
            for sp in ("f", "h", "g"):
                out[(ch, sp)] = s.get_populations(ch, sp)
            for f in ("rho", "ux", "uy", "cpos", "cneg", "psi"):
                out[(ch, f)] = s.get_field(ch, f)
        out["red"] = s.reductions()
        runs.append(out)
        s.close()
    for k in runs[0]:
        np.testing.assert_array_equal(runs[0][k], runs[1][k], err_msg=str(k))


def test_recentring_inside_a_long_step_call():
    # FP32 storage re-centres every 4096 steps, also inside one long call:
    # step(5000) and step(4096) + step(904) re-centre at the same step and
    # give identical bits
    p = ekpnp.lattice_params(256, 40, tau=0.7, zeta=-1.5, kappa_h=6.0, e_x=2e-4)
    runs = []
    for split in ((5000,), (4096, 904)):
        s = ekpnp.Solver([p], precision="fp32")
        s.initialization(ekpnp.COLD_NOISE, 6)
        s.init_equilibrium()
        for n in split:
            s.step(n)
        runs.append({sp: s.get_populations(0, sp) for sp in ("f", "h", "g")} | {"psi": s.get_field(0, "psi")})
        s.close()
    for k in runs[0]:
        np.testing.assert_array_equal(runs[0][k], runs[1][k], err_msg=k)


def test_fp32_row_kernel_channel_against_the_oracle():
    # a 256-wide FP32 channel (row kernel) and a 96-wide one (k_lbm) in one
    # batch, each against its FP64 oracle after 2000 steps
    pa = ekpnp.lattice_params(256, 34, tau=0.8, zeta=-1.0, kappa_h=6.0, e_x=1e-4)
    pb = ekpnp.lattice_params(96, 40, tau=0.7, zeta=-1.5, kappa_h=5.0, e_x=2e-4)
    s = ekpnp.Solver([pa, pb], precision="fp32")
    s.initialization(ekpnp.COLD_NOISE, 7)
    s.init_equilibrium()
    s.step(2000)
    for ch, p in enumerate((pa, pb)):
        o = O.Oracle(p, poisson_method=O.POISSON_XDFT_THOMAS)
        o.initialization(ekpnp.COLD_NOISE, 7, ch)
        o.init_equilibrium()
        o.step(2000)
        errs = field_errors(s, ch, o)
        _report(f"fp32 batched ch={ch}", errs)
        for f in ("ux", "cpos", "cneg", "psi"):
            assert errs[f] < 1e-5, (ch, f, errs[f])


def test_rebase_preserves_the_state():
    p = ekpnp.lattice_params(32, 48, tau=0.8, zeta=-2.0, kappa_h=8.0, e_x=1e-4)
    a = ekpnp.Solver([p], precision="fp32")
    b = ekpnp.Solver([p], precision="fp32")
    for s in (a, b):
        s.initialization(ekpnp.COLD_NOISE, 2)
        s.init_equilibrium()
        s.step(500)
    before = {sp: a.get_populations(0, sp) for sp in ("f", "h", "g")}
    a.rebase()
    for sp in ("f", "h", "g"):
        after = a.get_populations(0, sp)
        scale = np.abs(before[sp]).max()
        assert np.abs(after - before[sp]).max() <= 4e-7 * scale, sp
    a.step(300)
    b.step(300)
    for f in ("ux", "cpos", "cneg", "psi"):
        e = rel_l2(a.get_field(0, f), b.get_field(0, f))
        print("rebase vs none", f, e, flush=True)
        assert e < 1e-5, (f, e)


def test_fp32_reaches_tight_residual_like_fp64():
    # thin, strong double layer (kappa dy 0.38, zeta -2): the osmotic pressure
    # makes rho(y) vary across the channel far beyond FP32 resolution of the
    # residual; with per-row storage references FP32 storage still reaches a
    # 1e-8 residual, in about the FP64 step count.  The states are compared
    # at the same step (at residual 1e-8 each run is still ~1e-8 x the slow
    # time constant away from the steady state, so runs stopped at different
    # steps differ by that much whatever the precision)
    p = ekpnp.lattice_params(4, 64, tau=0.8, zeta=-2.0, kappa_h=12.0, e_x=1e-4)
    steps, runs = {}, {}
    for prec in ("fp64", "fp32"):
        s = ekpnp.Solver([p], precision=prec)
        s.initialization(ekpnp.COLD_NOISE, 7)
        s.init_equilibrium()
        steps[prec] = int(s.run_to_residual(1e-8, 100, 600000)[0])
        runs[prec] = s
    print("steps", steps["fp64"], steps["fp32"], flush=True)
    assert steps["fp64"] > 0 and steps["fp32"] > 0
    assert abs(steps["fp32"] - steps["fp64"]) <= 0.1 * steps["fp64"]
    # bring both to the same step
    n = max(steps.values()) + 1000
    for prec, s in runs.items():
        s.step(n - s.steps_done)
        assert s.steps_done == n
    for f in ("ux", "cpos", "cneg", "psi"):
        e = rel_l2(runs["fp32"].get_field(0, f), runs["fp64"].get_field(0, f))
        print("fp32 vs fp64 at step", n, f, e, flush=True)
        assert e < 1e-5, (f, e)


@pytest.mark.parametrize("which", ["reference", "lattice"])
def test_pb_picard_start_parity(which):
    # LBM.cu:106-127 on the device: 501 x {Boltzmann fields, Poisson, relax}
    if which == "reference":
        p = ekpnp.reference_params()
    else:
        p = ekpnp.lattice_params(32, 66, tau=0.8, zeta=-1.0, kappa_h=6.0, e_x=1e-4)
    s, o = _pair(p, mode=ekpnp.PB_PICARD)
    errs = field_errors(s, 0, o, fields=("cpos", "cneg", "psi"))
    _report(f"pb start {which}", errs)
    assert max(errs.values()) < 1e-11
    # E: the reference keeps the field of the last un-relaxed solve; the device
    # differentiates the stored (relaxed) psi.  They differ by the Picard
    # residual only; make the oracle use the relaxed psi and step on
    o.efield()
    s.step(100)
    o.step(100)
    errs = field_errors(s, 0, o, fields=("rho", "cpos", "cneg", "psi"))
    _report(f"pb + 100 steps {which}", errs)
    assert max(errs.values()) < 1e-11


def test_run_to_residual_freezes_converged_channels_exactly():
    # a fast (short) and a slow (tall) channel: the batched driver freezes the
    # fast one when it converges; its final state, step count, populations
    # and later evolution must equal a run of that channel alone
    pa = ekpnp.lattice_params(8, 12, tau=1.0, zeta=-0.5, kappa_h=1.5, e_x=1e-4)
    pb = ekpnp.lattice_params(8, 40, tau=1.0, zeta=-0.5, kappa_h=1.5, e_x=1e-4)
    tol, every = 1e-7, 20
    both = ekpnp.Solver([pa, pb], precision="fp64")
    both.initialization(ekpnp.COLD_NOISE, 9)
    both.init_equilibrium()
    steps = both.run_to_residual(tol, every, 200000)
    print("steps", steps, flush=True)
    assert steps[0] > 0 and steps[1] > steps[0]
    single = ekpnp.Solver([pa], precision="fp64")
    single.initialization(ekpnp.COLD_NOISE, 9)
    single.init_equilibrium()
    s1 = single.run_to_residual(tol, every, 200000)
    assert s1[0] == steps[0]
    for f in ("rho", "ux", "uy", "cpos", "cneg", "psi"):
        assert np.array_equal(both.get_field(0, f), single.get_field(0, f)), f
    for sp in ("f", "h", "g"):
        assert np.array_equal(both.get_populations(0, sp), single.get_populations(0, sp)), sp
    # after the driver returns every channel steps again, in lock-step
    both.step(7)
    single.step(7)
    assert np.array_equal(both.get_field(0, "ux"), single.get_field(0, "ux"))


def test_heterogeneous_wall_charge_parity():
    # heterogeneous-wall physics on a small grid: zeta(x) with 4 Fourier modes on both
    # walls, Boltzmann wall concentrations, per-x lifts in the Poisson solve
    nx, ny = 64, 48
    p = ekpnp.lattice_params(nx, ny, tau=0.8, zeta=-1.5, kappa_h=6.0, e_x=1e-4)
    rng = np.random.default_rng(4)
    a, phi = rng.uniform(0, 0.5, 4), rng.uniform(0, 2 * np.pi, 4)
    x = np.arange(nx)
    z = -1.5 + sum(a[k] * np.sin(2 * np.pi * (k + 1) * x / nx + phi[k]) for k in range(4))
    walls = dict(psi_lo=z, psi_hi=z[::-1].copy(), cpos_lo=np.exp(-z), cpos_hi=np.exp(-z[::-1]),
                 cneg_lo=np.exp(z), cneg_hi=np.exp(z[::-1]))
    for prec, tol in (("fp64", FP64_TOL), ("fp32", 1e-5)):
        s = ekpnp.Solver([p], precision=prec)
        s.set_wall_profile(0, **walls)
        s.initialization(ekpnp.COLD_NOISE, 4)
        s.init_equilibrium()
        o = O.Oracle(p)
        o.set_wall_profile(**walls)
        o.initialization(ekpnp.COLD_NOISE, 4, 0)
        o.init_equilibrium()
        s.step(400)
        o.step(400)
        if prec == "fp64":
            assert_parity(s, 0, o, tol, tag="zeta(x) fp64")
        else:
            errs = field_errors(s, 0, o)
            _report("zeta(x) fp32", errs)
            for f in ("ux", "uy", "cpos", "cneg", "psi"):
                assert errs[f] < tol, (f, errs[f])


def _zeta_walls(nx, seed=4):
    rng = np.random.default_rng(seed)
    a, phi = rng.uniform(0, 0.5, 4), rng.uniform(0, 2 * np.pi, 4)
    x = np.arange(nx)
    z = -1.5 + sum(a[k] * np.sin(2 * np.pi * (k + 1) * x / nx + phi[k]) for k in range(4))
    return dict(psi_lo=z, psi_hi=z[::-1].copy(), cpos_lo=np.exp(-z), cpos_hi=np.exp(-z[::-1]),
                cneg_lo=np.exp(z), cneg_hi=np.exp(z[::-1]))


@pytest.mark.timeout(600)
@pytest.mark.parametrize("nx,ny,nsteps", [(64, 48, 16000), (256, 32, 8000)])
def test_heterogeneous_wall_charge_fp32_long_horizon(nx, ny, nsteps):
    # FP32 storage on zeta(x) walls against the FP64 oracle over many steps,
    # through the early re-centrings of the cold start and several at the
    # 4096-step interval; 256 columns hold four 64-node reference blocks per
    # row (64 columns: one)
    print(f"case zeta(x) fp32 {nx}x{ny}, {nsteps} steps", flush=True)
    p = ekpnp.lattice_params(nx, ny, tau=0.8, zeta=-1.5, kappa_h=6.0, e_x=1e-4)
    walls = _zeta_walls(nx)
    s = ekpnp.Solver([p], precision="fp32")
    s.set_wall_profile(0, **walls)
    s.initialization(ekpnp.COLD_NOISE, 4)
    s.init_equilibrium()
    o = O.Oracle(p, poisson_method=O.POISSON_XDFT_THOMAS)
    o.set_wall_profile(**walls)
    o.initialization(ekpnp.COLD_NOISE, 4, 0)
    o.init_equilibrium()
    for n in (nsteps // 4, nsteps):
        s.step(n - s.steps_done)
        o.step(n - o.steps)
        errs = field_errors(s, 0, o)
        _report(f"zeta(x) fp32 {nx}x{ny} n={n}", errs)
        for f in ("ux", "uy", "cpos", "cneg", "psi"):
            assert errs[f] < 1e-5, (n, f, errs[f])


def test_run_to_residual_frozen_fp32_channel_keeps_its_references():
    # FP32 storage: a channel frozen by run_to_residual keeps its lattice and
    # its storage references through the re-centrings the other channel's
    # steps bring; here the run starts at an odd step (so the frozen state
    # sits in the other lattice half of the time) and re-centres while the
    # fast channel is frozen.  Its state at the end must equal a run of that
    # channel alone, bit for bit.
    pa = ekpnp.lattice_params(8, 12, tau=1.0, zeta=-0.5, kappa_h=1.5, e_x=1e-4)
    pb = ekpnp.lattice_params(64, 40, tau=1.0, zeta=-0.5, kappa_h=1.5, e_x=1e-4)
    tol, every = 1e-7, 20
    both = ekpnp.Solver([pa, pb], precision="fp32")
    single = ekpnp.Solver([pa], precision="fp32")
    for s in (both, single):
        s.initialization(ekpnp.COLD_NOISE, 9)
        s.init_equilibrium()
        s.step(1)
    steps = both.run_to_residual(tol, every, 400000)
    s1 = single.run_to_residual(tol, every, 400000)
    print("steps", steps, s1, flush=True)
    assert steps[0] > 0 and steps[1] > steps[0] and s1[0] == steps[0]
    # re-centrings fell between the freeze and the end of the call
    between = [r for r in ekpnp.rebase_steps(1 + steps[1]) if 1 + steps[0] < r <= 1 + steps[1]]
    assert between, "no re-centring while the channel was frozen"
    for f in ("rho", "ux", "uy", "cpos", "cneg", "psi"):
        assert np.array_equal(both.get_field(0, f), single.get_field(0, f)), f
    for sp in ("f", "h", "g"):
        assert np.array_equal(both.get_populations(0, sp), single.get_populations(0, sp)), sp


def test_nonfinite_population_is_reported_for_its_channel():
    # failure injection from the host: a NaN planted in channel 1's populations
    # must stop run_to_residual with ERR_UNSTABLE naming channel 1, while
    # channel 0's fields stay finite
    ps = [ekpnp.lattice_params(16, 20, tau=0.8, zeta=-1.0, kappa_h=3.0, e_x=1e-4),
          ekpnp.lattice_params(16, 24, tau=0.8, zeta=-1.0, kappa_h=3.0, e_x=1e-4)]
    s = ekpnp.Solver(ps, precision="fp32")
    s.initialization(ekpnp.COLD_NOISE, 1)
    s.init_equilibrium()
    s.step(10)
    h = s.get_populations(1, "h")
    h[3, 10, 5] = np.nan
    s.set_populations(1, "h", h)
    with pytest.raises(ekpnp.EkpnpError) as ei:
        s.run_to_residual(1e-12, 20, 2000)
    assert ei.value.code == ekpnp.ERR_UNSTABLE
    assert "channel 1" in str(ei.value)
    assert np.isfinite(s.get_field(0, "ux")).all()
    red = s.reductions()
    assert red[1, 6] > 0 and red[0, 6] == 0


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_restart_channel_recovers_only_that_channel(precision):
    # recovery after a non-finite report: ekpnp_restart_channel restarts
    # channel 1 alone.  A twin context that was never corrupted, and that
    # restarts its channel 1 at the same step, must match the recovered one
    # bit for bit in every channel (the corruption leaves no trace; channels
    # 0 and 2 go on untouched); and channel 1 must follow a fresh context's
    # channel 1 from the same seed: bit for bit with FP64 storage, within
    # 1e-5 with FP32 storage (its re-centring steps are the context's, not a
    # fresh start's)
    ps = [ekpnp.lattice_params(64, 40, tau=0.8, zeta=-1.0, kappa_h=5.0, e_x=1e-4),
          ekpnp.lattice_params(96, 36, tau=0.7, zeta=-2.0, kappa_h=4.0, e_x=2e-4),
          ekpnp.lattice_params(128, 24, tau=0.9, zeta=-0.5, kappa_h=3.0, e_x=5e-5)]
    print("case restart_channel", precision, flush=True)

    def start():
        s = ekpnp.Solver(ps, precision=precision)
        s.initialization(ekpnp.COLD_NOISE, 7)
        s.init_equilibrium()
        return s

    a, b = start(), start()
    a.step(50)
    b.step(50)
    h = a.get_populations(1, "h")
    h[2, 10, 5] = np.nan
    a.set_populations(1, "h", h)
    b.set_populations(1, "h", b.get_populations(1, "h"))
    with pytest.raises(ekpnp.EkpnpError) as ei:
        a.run_to_residual(1e-12, 20, 200)
    assert ei.value.code == ekpnp.ERR_UNSTABLE and "channel 1" in str(ei.value)
    assert a.steps_done == 70
    b.step(20)
    a.restart_channel(1, ekpnp.COLD_NOISE, 7)
    b.restart_channel(1, ekpnp.COLD_NOISE, 7)
    a.step(100)
    b.step(100)
    fields = ("rho", "ux", "uy", "cpos", "cneg", "psi")
    for ch in range(3):
        for f in fields:
            assert np.array_equal(a.get_field(ch, f), b.get_field(ch, f)), (ch, f)
        for sp in ("f", "h", "g"):
            assert np.array_equal(a.get_populations(ch, sp), b.get_populations(ch, sp)), (ch, sp)
    red = a.reductions()
    assert (red[:, 6] == 0).all()
    c = start()
    c.step(100)
    for f in fields:
        e = rel_l2(a.get_field(1, f), c.get_field(1, f))
        if precision == "fp64":
            assert np.array_equal(a.get_field(1, f), c.get_field(1, f)), (f, e)
        else:
            assert e < 1e-5, (f, e)
    for s in (a, b, c):
        s.close()
    # arguments and state
    d = ekpnp.Solver(ps[:1], precision=precision)
    d.initialization(ekpnp.COLD_NOISE, 1)
    with pytest.raises(ekpnp.EkpnpError) as ei:
        d.restart_channel(0)
    assert ei.value.code == ekpnp.ERR_STATE
    d.init_equilibrium()
    with pytest.raises(ekpnp.EkpnpError) as ei:
        d.restart_channel(0, ekpnp.PB_PICARD, 0)
    assert ei.value.code == ekpnp.ERR_UNSUPPORTED
    with pytest.raises(ekpnp.EkpnpError):
        d.restart_channel(1)
    d.close()


@pytest.mark.parametrize("nx,ny,precision", [(8, 4, "fp64"), (2, 5, "fp64"), (3, 4, "fp64"), (8192, 6, "fp64"),
                                             (512, 2050, "fp32"), (256, 4, "fp32"), (24, 7, "fp32")])
def test_extreme_shapes(nx, ny, precision):
    # smallest grids (two interior rows, nx = 2 and 3 on the direct DFT, the
    # smallest power-of-two FFT), the widest row and the tallest column the
    # plan supports, and the row kernel on a channel with a single interior
    # row pair; 30 steps against the oracle
    print(f"case nx={nx} ny={ny} {precision}", flush=True)
    p = ekpnp.lattice_params(nx, ny, tau=0.8, zeta=-1.0, kappa_h=2.0, e_x=1e-4)
    s, o = _pair(p, precision=precision, seed=nx + ny)
    s.step(30)
    o.step(30)
    if precision == "fp64":
        assert_parity(s, 0, o, FP64_TOL, tag=f"{nx}x{ny}")
    else:
        # FP32 deviation storage: on two interior rows u_x (~3e-4 here) is a
        # small difference of O(1) populations rounded to FP32; the
        # re-centrings at steps 8 and 16 keep it within the bar (measured
        # 6.9e-6 at 256 x 4, 3.3e-6 at 24 x 7); u_y of these x-uniform
        # channels is noise-level after 30 steps, its relative error meaningless
        errs = field_errors(s, 0, o)
        _report(f"fp32 {nx}x{ny}", errs)
        for f in ("ux", "cpos", "cneg", "psi"):
            assert errs[f] < 1e-5, (f, errs[f])
