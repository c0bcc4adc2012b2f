"""Pins the FP64 CPU oracle (CPU only).

The reference ships no tests, fixtures or golden outputs, so every check here
comes from the mathematics: exact conservation and decay laws of the TRT
collision, exact transport of single populations by streaming and bounce-back,
the anti-bounce-back map, closed-form Poisson solutions computed with numpy /
scipy routines, the D2Q9 equilibrium moments, and force-driven Poiseuille flow
(exact for TRT with Lambda = 3/16 and bounce-back walls).  Each is chosen so a
plausible slip (wrong sign, swapped direction, dropped term) fails it.
"""
import numpy as np
import pytest
import scipy.linalg

import ekpnp
import oracle as O

EX = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1])
EY = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1])
OPP = np.array([0, 3, 4, 1, 2, 7, 8, 5, 6])
W = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)


def lattice(nx=8, ny=10, **kw):
    return ekpnp.lattice_params(nx, ny, **kw)


def uniform_state(o, rho=1.0, ux=0.0, uy=0.0, cp=1.0, cn=1.0, psi=None):
    shp = o.shape
    f = lambda v: np.full(shp, v, dtype=np.float64)  # noqa: E731
    o.set_fields(f(rho), f(ux), f(uy), f(cp), f(cn), f(0.0) if psi is None else psi)
    o.init_equilibrium()


# ---------------------------------------------------------------------------
# equilibrium moments (LBM.cu:195-305)
# ---------------------------------------------------------------------------
def test_equilibrium_moments():
    p = lattice(6, 7, tau=0.8)
    o = O.Oracle(p)
    rng = np.random.default_rng(1)
    shp = o.shape
    rho = 1 + 0.01 * rng.standard_normal(shp)
    ux = 0.01 * rng.standard_normal(shp)
    uy = 0.01 * rng.standard_normal(shp)
    cp = 1 + 0.1 * rng.standard_normal(shp)
    cn = 1 + 0.1 * rng.standard_normal(shp)
    psi = 0.1 * rng.standard_normal(shp)
    o.set_fields(rho, ux, uy, cp, cn, psi)
    o.init_equilibrium()
    ex, ey = o.get_field("ex"), o.get_field("ey")
    cs2 = 1 / 3
    for sp, X, vx, vy in (("f", rho, ux, uy), ("h", cp, ux + p.mob_pos * ex, uy + p.mob_pos * ey),
                          ("g", cn, ux + p.mob_neg * ex, uy + p.mob_neg * ey)):
        q = o.get_populations(sp)
        np.testing.assert_allclose(q.sum(0), X, rtol=1e-14)
        np.testing.assert_allclose(np.tensordot(EX, q, 1), X * vx, atol=1e-15)
        np.testing.assert_allclose(np.tensordot(EY, q, 1), X * vy, atol=1e-15)
        # second moments: X (cs^2 delta_ab + v_a v_b)
        np.testing.assert_allclose(np.tensordot(EX * EX, q, 1), X * (cs2 + vx * vx), rtol=1e-13)
        np.testing.assert_allclose(np.tensordot(EX * EY, q, 1), X * vx * vy, atol=1e-15)


# ---------------------------------------------------------------------------
# TRT collision (LBM.cu:331-804)
# ---------------------------------------------------------------------------
def _collide_once(p, pre):
    """Run the collide stage on given pre-collision populations {sp: [9,ny,nx]}."""
    o = O.Oracle(p)
    uniform_state(o)
    for sp, a in pre.items():
        o.set_populations(sp, a)
    o.stage("collide")
    post = {}
    for sp in ("f", "h", "g"):
        rest = o.get_populations(sp)[0]
        post[sp] = np.concatenate([rest[None], o.get_post(sp)], 0)
    return post, o


def _equilibrium_pops(p, rho=1.0, ux=0.0, uy=0.0, cp=1.0, cn=1.0):
    o = O.Oracle(p)
    uniform_state(o, rho, ux, uy, cp, cn)
    return {sp: o.get_populations(sp) for sp in ("f", "h", "g")}


def test_collision_conserves_mass_and_equilibrium_is_fixed_point():
    p = lattice(6, 8, tau=0.7, e_x=0.0)  # no body force: F = 0
    eq = _equilibrium_pops(p, rho=1.02, ux=0.01, uy=-0.005, cp=1.1, cn=0.9)
    post, _ = _collide_once(p, eq)
    rows = slice(1, -1)  # the y = 0 row takes the mirrored velocity
    for sp in ("f", "h", "g"):
        np.testing.assert_allclose(post[sp][:, rows], eq[sp][:, rows], rtol=1e-14, atol=1e-16)
    rng = np.random.default_rng(2)
    pert = {sp: eq[sp] * (1 + 0.01 * rng.standard_normal(eq[sp].shape)) for sp in eq}
    post, _ = _collide_once(p, pert)
    for sp in ("f", "h", "g"):
        np.testing.assert_allclose(post[sp].sum(0), pert[sp].sum(0), rtol=1e-14)


@pytest.mark.parametrize("parity", ["even", "odd"])
def test_trt_nonequilibrium_decay_rates(parity):
    p = lattice(4, 6, tau=0.73)
    eq = _equilibrium_pops(p)
    eps = 1e-4
    d = np.zeros(9)
    if parity == "even":  # moment-free symmetric perturbation
        d[[1, 3]] = eps
        d[[2, 4]] = -eps
    else:  # moment-free antisymmetric perturbation
        d[5], d[7], d[6], d[8] = eps / 2, -eps / 2, eps / 2, -eps / 2
        d[2], d[4] = -eps, eps
    pre = {sp: eq[sp] + d[:, None, None] for sp in eq}
    post, _ = _collide_once(p, pre)
    cs2, dt = 1 / 3, 1.0
    lam_f = p.nu / cs2 / dt
    om_f = {"even": 1 / (lam_f + 0.5), "odd": 1 / (p.magic_f / lam_f + 0.5)}[parity]
    lam_c = p.diff_pos / cs2 / dt
    om_c = {"even": 1 / (p.magic_pos / lam_c + 0.5), "odd": 1 / (lam_c + 0.5)}[parity]
    rows = slice(1, -1)
    for sp, om in (("f", om_f), ("h", om_c), ("g", om_c)):
        got = (post[sp] - eq[sp])[:, rows]
        np.testing.assert_allclose(got, (1 - om) * d[:, None, None] * np.ones_like(got), rtol=1e-9,
                                   atol=1e-18)


def test_guo_force_momentum_increment():
    # uniform charge density and field: F = force_coef (c+ - c-)(Ex + E_app) + exf
    p = lattice(4, 6, tau=0.8, e_x=3e-3)
    p.exf = 2e-5
    o = O.Oracle(p)
    shp = o.shape
    f = lambda v: np.full(shp, v)  # noqa: E731
    psi = np.tile(-0.02 * np.arange(p.nx, dtype=np.float64), (p.ny, 1))  # Ex = 0.02 except the wrap
    o.set_fields(f(1.0), f(0.0), f(0.0), f(1.2), f(0.8), psi)
    o.init_equilibrium()
    pre = o.get_populations("f")
    o.stage("collide")
    post = np.concatenate([o.get_populations("f")[0][None], o.get_post("f")], 0)
    ex = o.get_field("ex")
    Fx = p.force_coef * 0.4 * (ex + p.e_applied) + p.exf
    rows = slice(1, -1)
    dj = np.tensordot(EX, post - pre, 1)[rows]
    np.testing.assert_allclose(dj, Fx[rows], rtol=1e-10)
    np.testing.assert_allclose(np.tensordot(EY, post - pre, 1)[rows], 0, atol=1e-16)
    # velocity written by collide includes half the force impulse
    np.testing.assert_allclose(o.get_field("ux")[rows], 0.5 * Fx[rows], rtol=1e-12)


# ---------------------------------------------------------------------------
# streaming, bounce-back and anti-bounce-back (LBM.cu:806-1024)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("d", range(1, 9))
def test_stream_moves_a_single_population_along_its_velocity(d):
    p = lattice(7, 9)
    o = O.Oracle(p)
    uniform_state(o)
    post = np.zeros((8,) + o.shape)
    x0, y0 = 3, 4
    post[d - 1, y0, x0] = 1.0
    o.set_post("f", post)
    o.stage("stream")
    pre = o.get_populations("f")[1:]
    hit = np.argwhere(pre != 0)
    assert hit.tolist() == [[d - 1, (y0 + EY[d]) % p.ny, (x0 + EX[d]) % p.nx]]


def test_full_way_bounce_back_returns_population_to_fluid():
    # a population arriving at the bottom wall node moving down is reversed by
    # the boundary stage and streamed back into (x,1) moving up: compare a run
    # with an extra population against the same run without it
    p = lattice(5, 8, e_x=0.0)
    runs = []
    for extra in (0.0, 0.25):
        o = O.Oracle(p)
        uniform_state(o)
        pre = o.get_populations("f")
        pre[4, 0, 2] += extra
        o.set_populations("f", pre)
        o.stage("collide")
        o.stage("boundary")
        o.stage("stream")
        runs.append(o.get_populations("f"))
    diff = runs[1] - runs[0]
    assert abs(diff[2, 1, 2] - 0.25) < 1e-15
    diff[2, 1, 2] = 0.0
    assert np.abs(diff).max() < 1e-15


def test_anti_bounce_back_wall_concentration_map():
    # c_wall(n+1) = 2 c_w - c_wall(n) (bc_charge on a mass-conserving collision)
    p = lattice(6, 8)
    o = O.Oracle(p)
    shp = o.shape
    cp = np.ones(shp)
    cp[0, :] = p.cpos_lo + 0.03
    cp[-1, :] = p.cpos_hi - 0.02
    o.set_fields(np.ones(shp), np.zeros(shp), np.zeros(shp), cp, np.ones(shp), np.zeros(shp))
    o.init_equilibrium()
    o.stream_collide_save()
    o.stream_collide_save()  # the second collide writes the moments of the first result
    c = o.get_field("cpos")
    np.testing.assert_allclose(c[0], p.cpos_lo - 0.03, rtol=1e-12)
    np.testing.assert_allclose(c[-1], p.cpos_hi + 0.02, rtol=1e-12)


def test_poiseuille_exact_with_magic_three_sixteenths():
    # body force exf only (no charge coupling): steady TRT profile with
    # Lambda = 3/16 and walls half-way between nodes 0|1 and ny-2|ny-1
    ny, F = 14, 1e-6
    p = lattice(4, ny, tau=1.1, zeta=0.0, kappa_h=1.0, e_x=0.0)
    p.magic_f = 3.0 / 16.0
    p.force_coef = 0.0
    p.exf = F
    o = O.Oracle(p)
    uniform_state(o)
    o.step(12000)
    y = np.arange(ny, dtype=np.float64)
    u_exact = F / (2 * p.rho0 * p.nu) * (y - 0.5) * (ny - 1.5 - y)
    ux = o.get_field("ux")[:, 0]
    np.testing.assert_allclose(ux[1:-1], u_exact[1:-1], rtol=1e-9)


# ---------------------------------------------------------------------------
# Poisson (poisson.cu:180-306) and E field (poisson.cu:145-174)
# ---------------------------------------------------------------------------
def _numpy_poisson(p, cp, cn, lo, hi):
    """Independent solve: numpy FFT in x, scipy banded solve in y."""
    ny, nx = cp.shape
    M = ny - 2
    rhs = -p.poisson_coef * (cp - cn)[1:-1] * p.dy ** 2
    rhs[0] -= lo
    rhs[-1] -= hi
    R = np.fft.fft(rhs, axis=1)
    k = np.fft.fftfreq(nx, d=1.0 / nx)
    kx = 2 * np.pi * k / p.lx
    out = np.empty_like(R)
    for j in range(nx):
        ab = np.zeros((3, M), dtype=np.complex128)
        ab[0, 1:] = 1.0
        ab[1, :] = -(2.0 + (kx[j] * p.dy) ** 2)
        ab[2, :-1] = 1.0
        out[:, j] = scipy.linalg.solve_banded((1, 1), ab, R[:, j])
    psi = np.empty((ny, nx))
    psi[1:-1] = np.fft.ifft(out, axis=1).real
    psi[0], psi[-1] = lo, hi
    return psi


@pytest.mark.parametrize("nx,ny,method", [(16, 21, 0), (16, 21, 1), (12, 15, 0), (12, 15, 1), (64, 66, 1)])
def test_poisson_matches_independent_solve(nx, ny, method):
    rng = np.random.default_rng(nx + 100 * ny)
    p = lattice(nx, ny, kappa_h=3.0)
    shp = (ny, nx)
    cp = 1 + 0.2 * rng.standard_normal(shp)
    cn = 1 + 0.2 * rng.standard_normal(shp)
    lo = -1 + 0.2 * rng.standard_normal(nx)
    hi = 0.5 + 0.2 * rng.standard_normal(nx)
    o = O.Oracle(p, poisson_method=method)
    o.set_wall_profile(psi_lo=lo, psi_hi=hi)
    o.set_fields(np.ones(shp), np.zeros(shp), np.zeros(shp), cp, cn)
    ref = _numpy_poisson(p, cp, cn, lo, hi)
    np.testing.assert_allclose(o.get_field("psi"), ref, rtol=0, atol=1e-11 * np.abs(ref).max())


def test_poisson_manufactured_single_mode_closed_form():
    # psi = A sin(2 pi m x/nx) sinh(q y)/sinh(q (ny-1)) is discretely harmonic
    # when 2 cosh(q) - 2 = (2 sin(pi m/nx))^2 ... for the spectral x operator:
    # cosh(q) = 1 + kx^2/2, so a zero right-hand side with that top-wall
    # profile must reproduce it exactly
    nx, ny, m = 32, 40, 3
    p = lattice(nx, ny)
    kx = 2 * np.pi * m / nx
    q = np.arccosh(1 + kx * kx / 2)
    x = np.arange(nx)
    y = np.arange(ny)[:, None]
    psi = np.sin(kx * x)[None, :] * np.sinh(q * y) / np.sinh(q * (ny - 1))
    o = O.Oracle(p)
    o.set_wall_profile(psi_lo=np.zeros(nx), psi_hi=psi[-1])
    shp = (ny, nx)
    o.set_fields(np.ones(shp), np.zeros(shp), np.zeros(shp), np.ones(shp), np.ones(shp))
    np.testing.assert_allclose(o.get_field("psi"), psi, atol=1e-13)


def test_poisson_methods_agree_on_reference_size():
    # odd extension + 2-D DFT (the reference's method) vs x-DFT + Thomas on the
    # reference's own grid (NX = 10, NY = 101) and SI coefficients
    p = ekpnp.reference_params()
    rng = np.random.default_rng(4)
    shp = (p.ny, p.nx)
    cp = 0.01 * (1 + 0.1 * rng.standard_normal(shp))
    cn = 0.01 * (1 + 0.1 * rng.standard_normal(shp))
    res = []
    for method in (0, 1):
        o = O.Oracle(p, poisson_method=method)
        o.set_fields(np.full(shp, 1000.0), np.zeros(shp), np.zeros(shp), cp, cn)
        res.append(o.get_field("psi"))
    np.testing.assert_allclose(res[0], res[1], rtol=0, atol=1e-12 * np.abs(res[1]).max())


def test_efield_central_differences():
    nx, ny = 16, 9
    p = lattice(nx, ny)
    o = O.Oracle(p)
    x = np.arange(nx)
    y = np.arange(ny)[:, None]
    k = 2 * np.pi / nx
    psi = np.sin(k * x)[None, :] + 0.3 * y * np.ones((1, nx))
    shp = (ny, nx)
    o.set_fields(np.ones(shp), np.zeros(shp), np.zeros(shp), np.ones(shp), np.ones(shp), psi)
    ex = o.get_field("ex")
    ey = o.get_field("ey")
    # (psi(x-1) - psi(x+1))/2 = -cos(kx) sin(k)
    np.testing.assert_allclose(ex, np.broadcast_to(-np.cos(k * x) * np.sin(k), shp), atol=1e-14)
    np.testing.assert_allclose(ey, -0.3, atol=1e-14)


# ---------------------------------------------------------------------------
# coupled physics: Poisson-Boltzmann + Helmholtz-Smoluchowski steady state
# ---------------------------------------------------------------------------
def _pb_newton(ny, zeta, pc):
    """psi'' = 2 pc sinh(psi) on nodes 0..ny-1, psi = zeta at both ends."""
    psi = np.full(ny, zeta) * 0.0
    psi[0] = psi[-1] = zeta
    for _ in range(100):
        r = psi[:-2] - 2 * psi[1:-1] + psi[2:] - 2 * pc * np.sinh(psi[1:-1])
        J = np.zeros((3, ny - 2))
        J[0, 1:] = 1
        J[1] = -2 - 2 * pc * np.cosh(psi[1:-1])
        J[2, :-1] = 1
        dpsi = scipy.linalg.solve_banded((1, 1), J, -r)
        psi[1:-1] += dpsi
        if np.abs(dpsi).max() < 1e-15:
            break
    return psi


def _steady_channel(ny, steps):
    p = lattice(4, ny, tau=1.0, zeta=-1.0, kappa_h=3.0, e_x=1e-4, diff=0.1)
    o = O.Oracle(p)
    o.initialization(ekpnp.COLD_NOISE, 0, 0)
    o.init_equilibrium()
    o.step(steps)
    return p, o


def test_coupled_steady_state_converges_to_pb_and_hs():
    # the ghost-node anti-bounce-back wall of the reference (LBM.cu:915-1023)
    # is first-order in dy, so the steady state approaches the
    # Poisson-Boltzmann solution on the same nodes at first order: the error
    # must halve when the grid is refined at fixed kappa h
    errs, hs, boltz = [], [], []
    for ny, steps in ((18, 4000), (34, 14000)):
        p, o = _steady_channel(ny, steps)
        psi = o.get_field("psi")[:, 0]
        pb = _pb_newton(ny, -1.0, p.poisson_coef)
        errs.append(np.abs(psi - pb).max())
        # Helmholtz-Smoluchowski from the solver's own psi, wall plane at y = 1/2
        psi_w = 0.5 * (psi[0] + psi[1])
        u_hs = p.e_applied * (psi - psi_w) / (p.rho0 * p.nu) * (p.force_coef / p.poisson_coef)
        ux = o.get_field("ux")[:, 0]
        mid = slice(2, -2)
        hs.append(np.abs(ux[mid] - u_hs[mid]).max() / np.abs(u_hs).max())
        # Boltzmann relation c+ c- = 1 holds to the same order
        cpcn = o.get_field("cpos")[1:-1, 0] * o.get_field("cneg")[1:-1, 0]
        boltz.append(np.abs(cpcn - 1).max())
    ratio = errs[1] / errs[0]
    print("psi errors", errs, "HS errors", hs, "c+c- errors", boltz)
    assert boltz[1] < 0.6 * boltz[0]
    assert 0.4 < ratio < 0.6, (errs, ratio)
    assert hs[1] < 0.03 and hs[1] < 0.7 * hs[0], hs


def test_pb_picard_start_reaches_discrete_pb_solution():
    # the reference's start (LBM.cu:106-127): under-relaxed Picard on the
    # Boltzmann concentrations converges to the discrete Poisson-Boltzmann
    # solution on the same nodes (Newton restatement above)
    ny = 34
    p = lattice(4, ny, zeta=-1.0, kappa_h=3.0)
    o = O.Oracle(p)
    o.initialization(ekpnp.PB_PICARD, 0, 0)
    psi = o.get_field("psi")[:, 0]
    np.testing.assert_allclose(psi, _pb_newton(ny, -1.0, p.poisson_coef), atol=1e-10)
    # concentrations are the Boltzmann fields of the iterate before the last solve
    np.testing.assert_allclose(o.get_field("cpos")[:, 0], np.exp(-psi), rtol=1e-9)
    np.testing.assert_allclose(o.get_field("cneg")[:, 0], np.exp(psi), rtol=1e-9)
    np.testing.assert_array_equal(o.get_field("ux"), 0.0)


# ---------------------------------------------------------------------------
# the flag and moving-wall branches of the collision and boundary stages
# ---------------------------------------------------------------------------
def test_couette_flow_with_moving_top_wall_is_exact():
    # The top wall moves at uw: its bounce-back carries the momentum terms
    # -+2 rho0 uw w_i/(cs^2 CFL) (LBM.cu:812-846); the bottom wall is at rest.
    # Uncharged fluid (zeta = 0, no field): the steady profile is linear,
    # with the walls half-way between the wall and the first fluid nodes,
    # u(y) = uw (y - 1/2)/(ny - 2), exact for TRT with Lambda = 3/16.  A wrong
    # sign, weight or direction pairing of the diagonal wall terms changes the
    # slope (the axis terms act on populations that stream along the wall row
    # and never reach the fluid, so no fluid observable depends on them).
    ny, uw = 12, 1e-3
    p = lattice(4, ny, tau=1.0, zeta=0.0, kappa_h=1.0, e_x=0.0)
    p.magic_f = 3.0 / 16.0
    p.uw = uw
    o = O.Oracle(p)
    uniform_state(o)
    o.step(30000)
    y = np.arange(ny, dtype=np.float64)
    u_exact = uw * (y - 0.5) / (ny - 2)
    ux = o.get_field("ux")
    np.testing.assert_allclose(ux[1:-1], np.broadcast_to(u_exact[1:-1, None], ux[1:-1].shape), rtol=1e-9)
    np.testing.assert_allclose(o.get_field("uy")[1:-1], 0.0, atol=1e-15)


@pytest.mark.parametrize("applied", [False, True])
def test_ion_drift_in_the_applied_field(applied):
    # The ions drift at v = u + K E; with EKPNP_FLAG_ION_DRIFT_APPLIED the
    # applied field joins the solved one, v = u + K (E + E_app) (the
    # reference drifts in E only, LBM.cu:479-488).  Uniform ions, no solved
    # field, fluid at rest and uncoupled (force_coef = 0): the equilibria's
    # x-flux is c K E_app with the flag and 0 without, and that equilibrium
    # is a fixed point of the collision.
    p = lattice(6, 8, tau=0.8, zeta=0.0, e_x=2e-3)
    p.force_coef = 0.0
    if applied:
        p.flags = ekpnp.FLAG_ION_DRIFT_APPLIED
    o = O.Oracle(p)
    uniform_state(o, cp=1.3, cn=0.7)
    rows = slice(1, -1)
    for sp, c, k in (("h", 1.3, p.mob_pos), ("g", 0.7, p.mob_neg)):
        q = o.get_populations(sp)
        want = c * k * p.e_applied if applied else 0.0
        # (sums of O(1) populations: rounding ~1e-16; the flux itself is 1e-4)
        np.testing.assert_allclose(np.tensordot(EX, q, 1), want, rtol=1e-12, atol=1e-15)
        np.testing.assert_allclose(np.tensordot(EY, q, 1), 0.0, atol=1e-15)
    pre = {sp: o.get_populations(sp) for sp in ("h", "g")}
    o.stage("collide")
    for sp in ("h", "g"):
        post = np.concatenate([o.get_populations(sp)[0][None], o.get_post(sp)], 0)
        np.testing.assert_allclose(post[:, rows], pre[sp][:, rows], rtol=1e-13, atol=1e-16)


@pytest.mark.parametrize("quirk", [False, True])
def test_negative_ion_rest_equilibrium_quirk_mass_defect(quirk):
    # LBM.cu:493 builds g_eq,0 from the fluid |u|^2 instead of the drift
    # velocity's |v|^2 (v = u + Kn E), as LBM.cu:242 does at initialisation.
    # Starting from the drift-velocity equilibrium (init_equilibrium), one
    # collision then changes the negative-ion mass of every interior node by
    #   omega+ w0 c- (|v|^2 - |u|^2) / (2 cs^2)
    # (its even rate, the rest weight), and leaves it unchanged without the
    # quirk.  Here u = 0 and E = (0, Ey) uniform from a linear potential.
    p = lattice(6, 9, tau=0.8, zeta=0.0, e_x=0.0)
    p.force_coef = 0.0
    if quirk:
        p.flags = ekpnp.FLAG_G0_FLUID_U
    o = O.Oracle(p)
    shp = o.shape
    cn = 1.4
    psi = np.broadcast_to(-0.05 * np.arange(p.ny, dtype=np.float64)[:, None], shp).copy()
    o.set_fields(np.ones(shp), np.zeros(shp), np.zeros(shp), np.full(shp, 1.1), np.full(shp, cn), psi)
    o.init_equilibrium()
    ey = o.get_field("ey")
    mass0 = o.get_populations("g").sum(0)
    o.stage("collide")
    post = np.concatenate([o.get_populations("g")[0][None], o.get_post("g")], 0)
    rows = slice(1, -1)
    cs2 = 1.0 / 3.0
    lam = p.diff_neg / cs2
    om_plus = 1.0 / (p.magic_neg / lam + 0.5)
    v2 = (p.mob_neg * ey) ** 2
    defect = om_plus * (4.0 / 9.0) * cn * v2 / (2 * cs2) if quirk else 0.0 * v2
    np.testing.assert_allclose((post.sum(0) - mass0)[rows], defect[rows], rtol=1e-9, atol=1e-15)
    assert quirk is False or defect[rows].min() > 1e-7


@pytest.mark.parametrize("method", [0, 1])
def test_poisson_discrete_gauss_law(method):
    # Summing the discrete Poisson equation over the interior rows and every
    # column removes the x-derivative (periodic: the kx = 0 mode only) and
    # telescopes the y-differences to the wall rows:
    #   sum_x [(psi(ny-1) - psi(ny-2)) - (psi(1) - psi(0))] / dy^2
    #     = -poisson_coef sum_{x, 0<y<ny-1} (c+ - c-)
    # the total charge balances the wall flux of E, for random charge and
    # per-x wall potentials
    nx, ny = 24, 31
    rng = np.random.default_rng(11)
    p = lattice(nx, ny, kappa_h=4.0)
    shp = (ny, nx)
    cp = 1 + 0.3 * rng.standard_normal(shp)
    cn = 1 + 0.3 * rng.standard_normal(shp)
    lo = -1 + 0.2 * rng.standard_normal(nx)
    hi = -0.3 + 0.2 * rng.standard_normal(nx)
    o = O.Oracle(p, poisson_method=method)
    o.set_wall_profile(psi_lo=lo, psi_hi=hi)
    o.set_fields(np.ones(shp), np.zeros(shp), np.zeros(shp), cp, cn)
    psi = o.get_field("psi")
    flux = ((psi[-1] - psi[-2]) - (psi[1] - psi[0])).sum() / p.dy ** 2
    charge = (cp - cn)[1:-1].sum()
    np.testing.assert_allclose(flux, -p.poisson_coef * charge, rtol=1e-11)
