"""Pins the oracle's iterative Poisson solve (poisson_phi / gpu_poisson,
poisson.cu:35-130) on the CPU.

The reference holds no fixtures for this dead-code path, so the checks come
from the mathematics: the fixed point is the solution of the 5-point system,
computed independently by a sparse direct solve; x-invariant charge reduces
the 5-point system to exactly the tridiagonal system of the spectral solver
(kx = 0 only), and a constant charge between walls to a parabola that the
second difference reproduces exactly.  A planted sign slip in the oracle's
source term must fail the direct-solve check.
"""
import numpy as np
import pytest
import scipy.sparse
import scipy.sparse.linalg

import ekpnp
import oracle as O


def direct_5pt(rho_e, lo, hi, dx, dy, pc):
    """Sparse direct solve of the 5-point system (periodic x, Dirichlet rows)."""
    ny, nx = rho_e.shape
    m = ny - 2
    n = m * nx
    idx = lambda y, x: (y - 1) * nx + (x % nx)  # noqa: E731
    ax, ay = 1.0 / dx ** 2, 1.0 / dy ** 2
    rows, cols, vals = [], [], []
    b = -pc * rho_e[1:-1].reshape(-1).copy()
    for y in range(1, ny - 1):
        for x in range(nx):
            i = idx(y, x)
            rows += [i, i, i]
            cols += [i, idx(y, x - 1), idx(y, x + 1)]
            vals += [-2 * ax - 2 * ay, ax, ax]
            for yy, wall in ((y - 1, lo), (y + 1, hi)):
                if yy == 0 or yy == ny - 1:
                    b[i] -= ay * wall[x]
                else:
                    rows.append(i)
                    cols.append(idx(yy, x))
                    vals.append(ay)
    A = scipy.sparse.csr_matrix((vals, (rows, cols)), shape=(n, n))
    psi = np.empty((ny, nx))
    psi[1:-1] = scipy.sparse.linalg.spsolve(A, b).reshape(m, nx)
    psi[0], psi[-1] = lo, hi
    return psi


def random_case(nx, ny, seed):
    rng = np.random.default_rng(seed)
    rho = rng.uniform(-1, 1, (ny, nx))
    lo = -1 + 0.2 * rng.standard_normal(nx)
    hi = 0.5 + 0.2 * rng.standard_normal(nx)
    return rho, lo, hi


@pytest.mark.parametrize("nx,ny,dx,dy,omega", [(32, 20, 1.0, 1.0, 1.0), (32, 20, 1.0, 0.7, 1.7),
                                               (16, 33, 0.5, 1.0, 1.85)])
def test_sor_fixed_point_is_the_5pt_solution(nx, ny, dx, dy, omega):
    rho, lo, hi = random_case(nx, ny, nx * ny)
    pc = 0.3
    ref = direct_5pt(rho, lo, hi, dx, dy, pc)
    psi, it, ch = O.poisson_sor(rho, np.zeros((ny, nx)), lo, hi, dx, dy, pc, omega=omega, tol=1e-14,
                                max_iter=200000)
    assert ch < 1e-14 and it < 200000
    np.testing.assert_allclose(psi, ref, rtol=0, atol=1e-10 * np.abs(ref).max())


def test_planted_source_sign_slip_fails_the_direct_check():
    nx, ny = 16, 12
    rho, lo, hi = random_case(nx, ny, 7)
    ref = direct_5pt(rho, lo, hi, 1.0, 1.0, 0.5)
    bad, _, _ = O.poisson_sor(-rho, np.zeros((ny, nx)), lo, hi, 1.0, 1.0, 0.5, omega=1.6, tol=1e-13)
    assert np.max(np.abs(bad - ref)) > 1e-3


def test_sor_constant_charge_gives_the_exact_parabola():
    # psi'' = -pc rho between psi(0) = a, psi(L) = b: the second difference of
    # a quadratic is exact, so the discrete solution is the parabola itself
    nx, ny, dy, pc, rho0, a, b = 8, 30, 0.5, 0.2, 1.3, -0.7, 0.4
    y = np.arange(ny) * dy
    L = (ny - 1) * dy
    exact = a + (b - a) * y / L + 0.5 * pc * rho0 * y * (L - y)
    psi, _, _ = O.poisson_sor(np.full((ny, nx), rho0), np.zeros((ny, nx)), np.full(nx, a), np.full(nx, b),
                              1.0, dy, pc, omega=1.8, tol=1e-15, max_iter=100000)
    np.testing.assert_allclose(psi, np.repeat(exact[:, None], nx, axis=1), rtol=0, atol=1e-12)


def test_sor_agrees_with_the_spectral_oracle_for_x_invariant_charge():
    nx, ny = 12, 26
    rng = np.random.default_rng(3)
    p = ekpnp.lattice_params(nx, ny, kappa_h=3.0)
    prof_p = 1 + 0.3 * rng.standard_normal(ny)
    prof_n = 1 + 0.3 * rng.standard_normal(ny)
    cp = np.repeat(prof_p[:, None], nx, axis=1)
    cn = np.repeat(prof_n[:, None], nx, axis=1)
    shp = (ny, nx)
    o = O.Oracle(p, poisson_method=O.POISSON_XDFT_THOMAS)
    o.set_fields(np.ones(shp), np.zeros(shp), np.zeros(shp), cp, cn)
    spectral = o.get_field("psi")
    psi, _, _ = O.poisson_sor(cp - cn, np.zeros(shp), np.full(nx, p.psi_lo), np.full(nx, p.psi_hi), p.dx, p.dy,
                              p.poisson_coef, omega=1.8, tol=1e-15, max_iter=100000)
    np.testing.assert_allclose(psi, spectral, rtol=0, atol=1e-11 * np.abs(spectral).max())


def test_sor_stop_rules():
    rho, lo, hi = random_case(16, 12, 1)
    _, it, ch = O.poisson_sor(rho, np.zeros((12, 16)), lo, hi, 1.0, 1.0, 0.5, tol=1e-9, max_iter=3)
    assert it == 3 and ch >= 1e-9
    _, it, ch = O.poisson_sor(rho, np.zeros((12, 16)), lo, hi, 1.0, 1.0, 0.5, tol=1e-9, max_iter=100000,
                              check_every=7)
    assert it % 7 == 0 and ch < 1e-9
