"""The FP64 global reductions (k_reduce_partial / k_reduce_final) against
brute-force numpy FP64 sums of the fields the solver hands out, and the
discrete Gauss law tying the charge reduction to the Poisson solve.

The reference has no global reductions (only host loops over copied fields,
LBM.cu:1394-1451); these sums decide every residual-driven freeze and every
time-to-steady count, so each is checked here on a single channel and on a
batch of channels of different sizes, including the residual over a known
interval.
"""
import numpy as np
import pytest

import ekpnp

pytestmark = pytest.mark.gpu

IDX = {k: i for i, k in enumerate(ekpnp.RED_NAMES)}


def _sums(s, ch):
    rho, ux, cp, cn = (s.get_field(ch, f) for f in ("rho", "ux", "cpos", "cneg"))
    inner = slice(1, -1)
    return {"mass": rho.sum(), "cpos": cp.sum(), "cneg": cn.sum(), "charge": (cp - cn)[inner].sum(),
            "ux2": (ux[inner] ** 2).sum()}, ux


def _close(a, b, tol=1e-12):
    return abs(a - b) <= tol * max(abs(b), 1e-300)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_reductions_match_numpy_sums(precision):
    ps = [ekpnp.lattice_params(64, 128, tau=0.8, zeta=-1.0, kappa_h=8.0, e_x=1e-4),
          ekpnp.lattice_params(256, 34, tau=0.7, zeta=-2.0, kappa_h=5.0, e_x=3e-4),
          ekpnp.lattice_params(10, 37, tau=0.9, zeta=-0.5, kappa_h=3.0, e_x=1e-5)]
    for chans in ([ps[0]], ps):
        print(f"case {precision} channels={len(chans)}", flush=True)
        s = ekpnp.Solver(chans, precision=precision)
        s.initialization(ekpnp.COLD_NOISE, 3)
        s.init_equilibrium()
        s.step(300)
        red0 = s.reductions()  # snapshot of ux
        before = [_sums(s, c)[1] for c in range(len(chans))]
        interval = 100
        s.step(interval)
        red = s.reductions()
        for c in range(len(chans)):
            want, ux = _sums(s, c)
            inner = slice(1, -1)
            dux2 = ((ux - before[c])[inner] ** 2).sum()
            got = {k: red[c, IDX[k]] for k in ("mass", "cpos", "cneg", "charge", "ux2")}
            print(c, {k: (got[k], want[k]) for k in got}, flush=True)
            for k in got:
                assert _close(got[k], want[k]), (c, k, got[k], want[k])
            assert _close(red[c, IDX["dux2"]], dux2, 1e-10), (c, red[c, IDX["dux2"]], dux2)
            assert red[c, IDX["nonfinite"]] == 0
            res = np.sqrt(dux2 / want["ux2"]) / interval
            assert _close(red[c, IDX["residual"]], res, 1e-10), (c, red[c, IDX["residual"]], res)
            # the first call has no earlier snapshot: residual 0 by definition
            assert red0[c, IDX["residual"]] == 0.0 and red[c, IDX["residual"]] > 0.0
        s.close()


def test_nonfinite_count_covers_every_field():
    # a NaN planted in one positive-ion population of channel 1: the
    # collide-stream makes that node's c+, its charge and (through the body
    # force) u non-finite (one node counted); the Poisson solve then spreads
    # the charge through psi to every interior node of the channel (the wall
    # rows keep their wall potentials); channel 0 stays clean
    ps = [ekpnp.lattice_params(16, 20, tau=0.8, zeta=-1.0, kappa_h=3.0, e_x=1e-4),
          ekpnp.lattice_params(64, 24, tau=0.8, zeta=-1.0, kappa_h=3.0, e_x=1e-4)]
    s = ekpnp.Solver(ps, precision="fp64")
    s.initialization(ekpnp.COLD_NOISE, 1)
    s.init_equilibrium()
    s.step(4)
    h = s.get_populations(1, "h")
    h[0, 10, 7] = np.nan
    s.set_populations(1, "h", h)
    s.stream_collide_save()

    def bad_nodes():
        bad = sum((~np.isfinite(s.get_field(1, k))) for k in ("rho", "ux", "uy", "cpos", "cneg", "psi"))
        return int((bad > 0).sum())

    red = s.reductions()
    print("non-finite after collide", bad_nodes(), red[:, IDX["nonfinite"]], flush=True)
    assert bad_nodes() == 1 and red[1, IDX["nonfinite"]] == 1 and red[0, IDX["nonfinite"]] == 0
    s.fast_poisson()
    red = s.reductions()
    print("non-finite after Poisson", bad_nodes(), red[:, IDX["nonfinite"]], flush=True)
    assert bad_nodes() == 64 * 22 and red[1, IDX["nonfinite"]] == 64 * 22 and red[0, IDX["nonfinite"]] == 0


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_discrete_gauss_law_on_the_device(precision):
    # The solver's psi and its own charge reduction obey the summed discrete
    # Poisson equation (kx = 0 mode, telescoped in y):
    #   sum_x [(psi(ny-1) - psi(ny-2)) - (psi(1) - psi(0))] / dy^2
    #     = -poisson_coef * (charge reduction over the interior rows)
    # on zeta(x) walls, after steps of the coupled solver
    nx, ny = 256, 66
    p = ekpnp.lattice_params(nx, ny, tau=0.8, zeta=-1.5, kappa_h=6.0, e_x=1e-4)
    rng = np.random.default_rng(3)
    x = np.arange(nx)
    z = -1.5 + 0.4 * np.sin(2 * np.pi * x / nx + rng.uniform(0, 6.28))
    s = ekpnp.Solver([p], precision=precision)
    s.set_wall_profile(0, psi_lo=z, psi_hi=z, cpos_lo=np.exp(-z), cpos_hi=np.exp(-z), cneg_lo=np.exp(z),
                       cneg_hi=np.exp(z))
    s.initialization(ekpnp.COLD_NOISE, 5)
    s.init_equilibrium()
    s.step(200)
    # psi is solved from the charge of the last step's collide, which the
    # macroscopic planes hold
    psi = s.get_field(0, "psi")
    charge = s.reductions()[0, IDX["charge"]]
    flux = ((psi[-1] - psi[-2]) - (psi[1] - psi[0])).sum() / p.dy ** 2
    print("flux", flux, "-pc*charge", -p.poisson_coef * charge, flush=True)
    assert abs(flux + p.poisson_coef * charge) <= 1e-10 * abs(flux)
