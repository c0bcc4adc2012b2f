"""Restart, file outputs and diagnostics through the C ABI (GPU).

The text formats are those of LBM.cu (save_data_end 1272-1312, read_data
1350-1391, save_data_tecplot 1100-1269, save_data_dmd 1314-1348, current
1394-1417, record_umax 1419-1451); the expected values below are computed
with numpy from the fields the solver returns.
"""
import re

import numpy as np
import pytest

import ekpnp

pytestmark = pytest.mark.gpu


def _solver(prec, nsteps, seed=3):
    ps = [ekpnp.lattice_params(32, 24, tau=0.8, zeta=-1.5, kappa_h=4.0, e_x=2e-4),
          ekpnp.lattice_params(16, 40, tau=0.7, zeta=-0.8, kappa_h=6.0, e_x=1e-4)]
    s = ekpnp.Solver(ps, precision=prec)
    s.initialization(ekpnp.COLD_NOISE, seed)
    s.init_equilibrium()
    s.step(nsteps)
    return s, ps


FIELD_NAMES = ("rho", "ux", "uy", "cpos", "cneg", "psi", "ex", "ey")


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_checkpoint_resumes_bit_exactly(prec, tmp_path):
    a, ps = _solver(prec, 137)
    a.reductions()  # residual snapshot is part of the state
    path = str(tmp_path / "state.ckpt")
    a.save_checkpoint(path)
    a.step(150)
    ra = a.reductions()
    b = ekpnp.Solver(ps, precision=prec)
    b.load_checkpoint(path)
    assert b.steps_done == 137
    b.step(150)
    rb = b.reductions()
    assert np.array_equal(ra, rb)
    for ch in range(2):
        for f in FIELD_NAMES:
            assert np.array_equal(a.get_field(ch, f), b.get_field(ch, f)), (ch, f)
        for sp in ("f", "h", "g"):
            assert np.array_equal(a.get_populations(ch, sp), b.get_populations(ch, sp)), (ch, sp)


def test_checkpoint_rejects_a_different_context(tmp_path):
    a, ps = _solver("fp32", 10)
    path = str(tmp_path / "state.ckpt")
    a.save_checkpoint(path)
    other = ekpnp.Solver(ps, precision="fp64")
    with pytest.raises(ekpnp.EkpnpError) as ei:
        other.load_checkpoint(path)
    assert ei.value.code == ekpnp.ERR_INVALID
    swapped = ekpnp.Solver(ps[::-1], precision="fp32")
    with pytest.raises(ekpnp.EkpnpError):
        swapped.load_checkpoint(path)
    with pytest.raises(ekpnp.EkpnpError) as ei:
        a.load_checkpoint(str(tmp_path / "missing.ckpt"))
    assert ei.value.code == ekpnp.ERR_IO


def test_data_end_format_and_restart(tmp_path):
    s, ps = _solver("fp64", 200)
    ch = 1
    ny, nx = s.shape(ch)
    path = tmp_path / "data_end.dat"
    s.save_data_end(ch, 12.5, str(path))
    lines = path.read_text().splitlines()
    assert len(lines) == nx * ny
    num = r"\s*-?\d+\.\d{6}"
    assert all(re.fullmatch(" ".join([num] * 9), ln) for ln in lines[:50])
    tab = np.loadtxt(str(path))
    assert tab.shape == (nx * ny, 9)
    assert np.all(tab[:, 0] == 12.5)
    cols = ("ux", "uy", "rho", "cpos", "psi", "ex", "ey", "cneg")
    for k, f in enumerate(cols, start=1):
        v = s.get_field(ch, f).ravel()
        assert np.abs(tab[:, k] - v).max() <= 5.0000001e-7, f
    # restart: the fields come back as the file's 6-decimal values
    r = ekpnp.Solver(ps, precision="fp64")
    r.initialization(ekpnp.COLD_NOISE, 0)
    t = r.read_data(ch, str(path))
    assert t == 12.5
    for k, f in ((1, "ux"), (2, "uy"), (3, "rho"), (4, "cpos"), (5, "psi"), (8, "cneg")):
        assert np.array_equal(r.get_field(ch, f).ravel(), tab[:, k]), f
    r.init_equilibrium()
    r.step(5)
    assert np.isfinite(r.get_field(ch, "ux")).all()


def test_current_umax_and_charge_data(tmp_path):
    s, ps = _solver("fp64", 300)
    for ch, p in enumerate(ps):
        cp = s.get_field(ch, "cpos")
        ey = s.get_field(ch, "ey")
        uy = s.get_field(ch, "uy")
        top = 2.0 * cp[-2] - cp[-3]
        want_i = np.sum(top * ey[-1]) * p.mob_pos * p.dy
        # the device adds the terms in another order: a bound on the
        # rounding of a sum of nx terms, relative to the sum of their sizes
        scale = np.sum(np.abs(top * ey[-1])) * p.mob_pos * p.dy
        assert abs(s.current(ch) - want_i) <= 1e-13 * scale, (s.current(ch), want_i, scale)
        uy2 = uy.copy()
        uy2[-1] = 2.0 * uy[-2] - uy[-3]
        want_u = max(0.0, uy2.max())
        path = tmp_path / f"umax{ch}.dat"
        assert s.record_umax(ch, 1.0, str(path)) == want_u
        assert s.record_umax(ch, 2.0, str(path)) == want_u
        rows = path.read_text().splitlines()
        assert len(rows) == 2 and rows[1].split()[0] == "2.000000"
        cpath = tmp_path / f"charge{ch}.dat"
        s.save_charge_data(ch, str(cpath))
        got = np.loadtxt(str(cpath))
        cp2 = cp.copy()
        cp2[-1] = top
        assert np.abs(got - cp2.ravel()).max() <= 5.0000001e-7


def test_current_and_umax_reduce_on_the_device_in_a_batch():
    # larger channels than above, behind another channel in the context: the
    # device reductions run over many work-groups and the top row spans
    # several of them; same expectations as the reference's host loops
    ps = [ekpnp.lattice_params(32, 24, tau=0.8, zeta=-1.5, kappa_h=4.0, e_x=2e-4),
          ekpnp.lattice_params(512, 260, tau=0.7, zeta=-2.0, kappa_h=8.0, e_x=3e-4),
          ekpnp.lattice_params(1024, 66, tau=0.9, zeta=-1.0, kappa_h=5.0, e_x=1e-4)]
    s = ekpnp.Solver(ps, precision="fp32")
    z = -2.0 + 0.5 * np.sin(2 * np.pi * np.arange(512) / 512)
    s.set_wall_profile(1, psi_lo=z, psi_hi=z, cpos_lo=np.exp(-z), cpos_hi=np.exp(-z), cneg_lo=np.exp(z),
                       cneg_hi=np.exp(z))
    s.initialization(ekpnp.COLD_NOISE, 5)
    s.init_equilibrium()
    s.step(150)
    for ch, p in enumerate(ps):
        print("case channel", ch, p.nx, "x", p.ny, flush=True)
        cp = s.get_field(ch, "cpos")
        ey = s.get_field(ch, "ey")
        uy = s.get_field(ch, "uy")
        top = 2.0 * cp[-2] - cp[-3]
        want_i = np.sum(top * ey[-1]) * p.mob_pos * p.dy
        scale = np.sum(np.abs(top * ey[-1])) * p.mob_pos * p.dy
        got_i = s.current(ch)
        assert abs(got_i - want_i) <= 1e-13 * scale, (ch, got_i, want_i, scale)
        uy2 = uy.copy()
        uy2[-1] = 2.0 * uy[-2] - uy[-3]
        assert s.record_umax(ch, 0.0, None) == max(0.0, uy2.max())
    s.close()


def test_tecplot_zone(tmp_path):
    s, ps = _solver("fp64", 120)
    ch = 0
    p = ps[ch]
    ny, nx = s.shape(ch)
    path = tmp_path / "data.dat"
    s.save_tecplot(ch, 0.0, str(path), True)
    s.step(10)
    s.save_tecplot(ch, 7.0, str(path), False)
    text = path.read_text().splitlines()
    assert text[0].startswith('VARIABLES="x","y","u","v","p","Positive charge"')
    zones = [i for i, ln in enumerate(text) if ln.startswith("ZONE")]
    assert len(zones) == 2
    assert text[zones[1]] == f'ZONE T="t=7", F=POINT, I = {nx}, J = {ny}'
    tab = np.array([[float(v) for v in ln.split()] for ln in text[zones[1] + 1:]])
    assert tab.shape == (nx * ny, 14)
    F = {f: s.get_field(ch, f) for f in FIELD_NAMES}
    ux, uy = F["ux"], F["uy"]
    # interior vorticity (central differences, periodic x), scaled by 1/0.25
    vor = ((np.roll(uy, -1, 1) - np.roll(uy, 1, 1)) / p.dx / 2
           - (np.roll(ux, -1, 0) - np.roll(ux, 1, 0)) / p.dy / 2) / 0.25
    got = tab[:, 9].reshape(ny, nx)
    assert np.abs(got[1:-1] - vor[1:-1]).max() <= 1e-11 * max(1.0, np.abs(vor).max())
    # x, y, u at interior rows and the extrapolated wall rows
    assert np.allclose(tab[:, 0].reshape(ny, nx)[0], p.dx * np.arange(nx))
    u = tab[:, 2].reshape(ny, nx)
    assert np.abs(u[1:-1] - ux[1:-1]).max() <= 1e-12
    assert np.abs(u[0] - (2 * ux[1] - ux[2])).max() <= 1e-12
