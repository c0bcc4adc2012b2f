"""Iterative Poisson solve on the GPU (ekpnp_poisson_iterative, the
reference's poisson_phi, poisson.cu:35-130) through the C ABI, against the
oracle's red-black restatement and the sparse direct solve of the 5-point
system, and against the spectral solver where the two discretisations
coincide (x-invariant charge)."""
import numpy as np
import pytest

import ekpnp
import oracle as O
from test_poisson_iterative_oracle import direct_5pt

pytestmark = pytest.mark.gpu


def _solver_with_charge(nx, ny, seed, dy=None, walls=True, x_invariant=False):
    rng = np.random.default_rng(seed)
    p = ekpnp.lattice_params(nx, ny, kappa_h=4.0)
    if dy is not None:
        p.dy = dy
    shp = (ny, nx)
    if x_invariant:
        cp = np.repeat((1 + 0.3 * rng.standard_normal(ny))[:, None], nx, axis=1)
        cn = np.repeat((1 + 0.3 * rng.standard_normal(ny))[:, None], nx, axis=1)
    else:
        cp = 1 + 0.3 * rng.standard_normal(shp)
        cn = 1 + 0.3 * rng.standard_normal(shp)
    lo = np.full(nx, p.psi_lo)
    hi = np.full(nx, p.psi_hi)
    s = ekpnp.Solver([p], precision="fp64")
    if walls and not x_invariant:
        lo = lo + 0.2 * rng.standard_normal(nx)
        hi = hi + 0.2 * rng.standard_normal(nx)
        s.set_wall_profile(0, psi_lo=lo, psi_hi=hi)
    return s, p, cp, cn, lo, hi


@pytest.mark.parametrize("nx,ny,dy,omega", [(64, 48, None, 1.8), (32, 40, 0.6, 1.7), (128, 34, None, 1.0)])
def test_iterative_matches_oracle_and_direct_solve(nx, ny, dy, omega):
    print(f"case nx={nx} ny={ny} dy={dy} omega={omega}", flush=True)
    s, p, cp, cn, lo, hi = _solver_with_charge(nx, ny, nx + ny, dy=dy)
    shp = (ny, nx)
    s.set_fields(0, np.ones(shp), np.zeros(shp), np.zeros(shp), cp, cn, psi=np.zeros(shp))
    it, ch = s.poisson_iterative(0, omega=omega, tol=1e-13, max_iter=200000, check_every=10)
    assert ch < 1e-13 and it < 200000
    psi = s.get_field(0, "psi")
    ref_it, it_o, ch_o = O.poisson_sor(cp - cn, np.zeros(shp), lo, hi, p.dx, p.dy, p.poisson_coef, omega=omega,
                                       tol=1e-13, max_iter=200000, check_every=10)
    direct = direct_5pt(cp - cn, lo, hi, p.dx, p.dy, p.poisson_coef)
    scale = np.abs(direct).max()
    print("sweeps gpu", it, "oracle", it_o, "max|gpu-oracle|", np.abs(psi - ref_it).max(),
          "max|gpu-direct|", np.abs(psi - direct).max(), flush=True)
    # same sweep order and stop rule: the sweep counts may differ only by
    # the rounding of the last change test
    assert abs(it - it_o) <= 10
    np.testing.assert_allclose(psi, ref_it, rtol=0, atol=1e-10 * scale)
    np.testing.assert_allclose(psi, direct, rtol=0, atol=1e-10 * scale)


def test_iterative_equals_spectral_for_x_invariant_charge():
    nx, ny = 64, 70
    s, p, cp, cn, lo, hi = _solver_with_charge(nx, ny, 11, x_invariant=True)
    shp = (ny, nx)
    s.set_fields(0, np.ones(shp), np.zeros(shp), np.zeros(shp), cp, cn)  # psi by the spectral solver
    spectral = s.get_field(0, "psi")
    s.set_fields(0, np.ones(shp), np.zeros(shp), np.zeros(shp), cp, cn, psi=np.zeros(shp))
    it, ch = s.poisson_iterative(0, omega=1.9, tol=1e-14, max_iter=200000, check_every=20)
    psi = s.get_field(0, "psi")
    print("sweeps", it, "max|iter-spectral|", np.abs(psi - spectral).max(), flush=True)
    np.testing.assert_allclose(psi, spectral, rtol=0, atol=1e-10 * np.abs(spectral).max())


def test_reference_defaults_gauss_seidel():
    # poisson_phi's own settings: tolerance 1e-9 on the largest change of a
    # sweep, tested after every sweep, no over-relaxation
    nx, ny = 16, 24
    s, p, cp, cn, lo, hi = _solver_with_charge(nx, ny, 5)
    shp = (ny, nx)
    s.set_fields(0, np.ones(shp), np.zeros(shp), np.zeros(shp), cp, cn, psi=np.zeros(shp))
    it, ch = s.poisson_iterative(0)
    psi = s.get_field(0, "psi")
    ref, it_o, _ = O.poisson_sor(cp - cn, np.zeros(shp), lo, hi, p.dx, p.dy, p.poisson_coef)
    direct = direct_5pt(cp - cn, lo, hi, p.dx, p.dy, p.poisson_coef)
    print("sweeps gpu", it, "oracle", it_o, "change", ch, "max|gpu-direct|", np.abs(psi - direct).max(), flush=True)
    assert ch < 1e-9 and abs(it - it_o) <= 1
    np.testing.assert_allclose(psi, ref, rtol=0, atol=1e-11 * np.abs(ref).max())
    # Gauss-Seidel stopped on the change, not the error: within change/(1 - rho)
    np.testing.assert_allclose(psi, direct, rtol=0, atol=1e-6)


def test_sweep_limit_and_untouched_channels():
    pa = ekpnp.lattice_params(32, 20, kappa_h=4.0)
    pb = ekpnp.lattice_params(16, 24, kappa_h=2.0)
    s = ekpnp.Solver([pa, pb], precision="fp64")
    s.initialization(ekpnp.COLD_NOISE, 3)
    before = s.get_field(0, "psi")
    it, ch = s.poisson_iterative(1, omega=1.5, tol=1e-12, max_iter=3)
    assert it == 3 and ch >= 1e-12
    np.testing.assert_array_equal(s.get_field(0, "psi"), before)


def test_rejects_odd_nx_and_bad_arguments():
    s = ekpnp.Solver([ekpnp.lattice_params(15, 20)], precision="fp64")
    s.initialization(ekpnp.COLD_NOISE, 1)
    with pytest.raises(ekpnp.EkpnpError):
        s.poisson_iterative(0)
    s2 = ekpnp.Solver([ekpnp.lattice_params(16, 20)], precision="fp64")
    s2.initialization(ekpnp.COLD_NOISE, 1)
    for kw in ({"omega": 2.0}, {"omega": 0.0}, {"max_iter": 0}, {"check_every": 0}, {"tol": -1.0}):
        with pytest.raises(ekpnp.EkpnpError):
            s2.poisson_iterative(0, **kw)
    with pytest.raises(ekpnp.EkpnpError):
        s2.poisson_iterative(1)
