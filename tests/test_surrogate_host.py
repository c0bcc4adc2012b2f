"""Host-side pieces of the warm-start surrogate (no GPU): the physics
baseline and the target transforms."""
import numpy as np
import pytest

from ekpnp import surrogate as S


def test_pb_profile_linear_limit_is_the_discrete_debye_huckel_solution():
    # tiny zeta: sinh(psi) = psi to 1e-12, and the discrete equation
    # psi[i-1] - (2 + kappa^2 dy^2) psi[i] + psi[i+1] = 0 is solved exactly by
    # zeta cosh(theta (i - c)) / cosh(theta c), 2 cosh(theta) = 2 + kappa^2 dy^2
    p = S.channel_of(-1e-6, 5.0, 1e-4, 4, 101)
    psi = S.pb_profile(p)
    k2 = 2.0 * p.poisson_coef * p.c_bulk * p.inv_thermal_voltage
    th = np.arccosh(1.0 + 0.5 * k2 * p.dy ** 2)
    i = np.arange(p.ny)
    c = 0.5 * (p.ny - 1)
    want = p.psi_lo * np.cosh(th * (i - c)) / np.cosh(th * c)
    assert np.abs(psi - want).max() <= 1e-12 * abs(p.psi_lo)


@pytest.mark.parametrize("zeta", [-1.0, -4.0])
def test_pb_profile_thin_layer_matches_gouy_chapman(zeta):
    # kappa h = 300 on 2001 nodes (kappa dy = 0.3): near each wall the channel
    # is a semi-infinite Gouy-Chapman layer, tanh(psi/4) = tanh(zeta/4) e^{-kappa y};
    # second-order discretisation error ~ (kappa dy)^2 / 12
    p = S.channel_of(zeta, 300.0, 1e-4, 4, 2001)
    psi = S.pb_profile(p)
    kappa = np.sqrt(2.0 * p.poisson_coef)
    y = np.arange(40) * p.dy
    gc = 4.0 * np.arctanh(np.tanh(zeta / 4.0) * np.exp(-kappa * y))
    err = np.abs(psi[:40] - gc).max() / abs(zeta)
    assert err < 0.02
    # discrete equation residual at the Newton solution
    a = psi[1:-1]
    f = psi[:-2] - 2 * a + psi[2:] - 2 * p.dy ** 2 * p.poisson_coef * np.sinh(a)
    assert np.abs(f).max() < 1e-12


def test_baseline_velocity_is_helmholtz_smoluchowski_in_the_bulk():
    p = S.channel_of(-2.0, 40.0, 1e-4, 4, 256)
    psi, cp, cn, ux = S.baseline(p)
    mid = p.ny // 2
    assert abs(psi[mid]) < 1e-6
    # u = hs_scale (psi - psi_wall) / (-zeta) -> hs_scale (1 - ...) in the bulk
    assert ux[mid] == pytest.approx(S.hs_scale(p) * (0.0 - 0.5 * (psi[0] + psi[1])) / 2.0, rel=1e-5)
    assert np.allclose(cp[1:-1] * cn[1:-1], 1.0, rtol=1e-12)


def test_residual_targets_round_trip():
    # a steady state = baseline x smooth correction: the correction is what
    # gets resampled, so the round trip is exact up to linear interpolation
    # of a smooth function (~ (pi/255)^2 / 8 of its amplitude per resampling)
    p = S.channel_of(-1.5, 6.0, 3e-4, 4, 96)
    psi, cp, cn, ux = S.baseline(p)
    y = np.arange(p.ny) / (p.ny - 1.0)
    bump = np.sin(np.pi * y)
    psi2 = psi + 0.01 * p.psi_lo * bump
    cp2 = cp * np.exp(0.01 * bump)
    cn2 = cn * np.exp(-0.02 * bump)
    ux2 = ux + 0.01 * S.hs_scale(p) * bump
    t = S.residual_targets(p, psi2, cp2, cn2, ux2)
    assert t.dtype == np.float32 and t.shape == (S.N_OUT,)
    back = S.residual_profiles(p, t)
    for got, want in zip(back, (psi2, cp2, cn2, ux2)):
        assert got.shape == (p.ny,)
        assert np.abs(got - want).max() <= 5e-6 * np.abs(want).max()
    # the baseline itself maps to zero targets
    assert np.abs(S.residual_targets(p, psi, cp, cn, ux)).max() == 0.0


def test_features_stay_in_range_over_the_sampling_ranges():
    ps, _ = S.sample_channels(400, 5)
    X = S.channel_features(ps)
    assert X.shape == (400, S.N_IN) and X.dtype == np.float32
    assert np.all(np.abs(X) <= 1.6), np.abs(X).max(axis=0)
    # kappa dy feature is the Debye resolution the stability rule bounds
    kdy = np.array([np.sqrt(2 * p.poisson_coef) * p.dy for p in ps])
    assert np.allclose((X[:, 4] + 1.0) * 0.2, kdy, rtol=1e-5)


def test_osmotic_density_matches_the_oracle_steady_state():
    # the warm start's rho(y) from the wall-normal momentum balance against
    # the FP64 oracle's steady state (interior rows, up to the constant the
    # unmodelled wall-row mass leaves, < 10 % of the profile's variation)
    import oracle as O
    import ekpnp
    p = ekpnp.lattice_params(4, 32, tau=0.8, zeta=-2.0, kappa_h=6.0, e_x=1e-4)
    o = O.Oracle(p)
    o.initialization(ekpnp.COLD_NOISE, 3, 0)
    o.init_equilibrium()
    o.step(60000)
    rho = o.get_field("rho").mean(axis=1)[1:-1]
    pred = S.osmotic_rho(p, *(o.get_field(f).mean(axis=1) for f in ("psi", "cpos", "cneg")))[1:-1]
    dev = rho - rho.mean()
    assert np.linalg.norm((pred - pred.mean()) - dev) < 0.02 * np.linalg.norm(dev)
    assert np.linalg.norm(pred - rho) < 0.1 * np.linalg.norm(dev)


def test_column_fields_follow_the_local_wall_potential():
    # the column-local warm start of the heterogeneous-wall channel: psi, c+-
    # and rho interpolated per column between profiles at wall-potential
    # levels; at a level's own zeta the column is that level's profile
    # exactly, the wall row carries zeta(x) itself (psi is linear in the
    # level there), and u_x is the mean-zeta profile in every column
    import run_wall_profile_warmstart as R

    p = S.channel_of(-1.5, 8.0, 1e-4, 16, 40)
    zeta = np.linspace(-2.0, -1.0, 16)
    rho, ux, cp, cn, psi = R.column_fields(p, S.baseline, zeta, levels=16)
    assert rho.shape == ux.shape == cp.shape == cn.shape == psi.shape == (40, 16)
    np.testing.assert_allclose(psi[0], zeta, rtol=0, atol=1e-14)
    np.testing.assert_allclose(psi[-1], zeta, rtol=0, atol=1e-14)
    for x in (0, 7, 15):
        q = p.copy()
        q.psi_lo = q.psi_hi = float(zeta[x])
        q.cpos_lo = q.cpos_hi = float(np.exp(-zeta[x]))
        q.cneg_lo = q.cneg_hi = float(np.exp(zeta[x]))
        bpsi, bcp, bcn, _ = S.baseline(q)
        np.testing.assert_allclose(psi[:, x], bpsi, rtol=1e-12, atol=1e-14)
        np.testing.assert_allclose(cp[:, x], bcp, rtol=1e-12)
        np.testing.assert_allclose(rho[:, x], S.osmotic_rho(q, bpsi, bcp, bcn), rtol=1e-12)
    u_mean = S.baseline(p)[3]
    for x in range(16):
        np.testing.assert_array_equal(ux[:, x], u_mean)
