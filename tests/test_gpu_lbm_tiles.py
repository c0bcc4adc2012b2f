"""Tile shapes of the row collide-stream kernel (ekpnp_set_lbm_tiles,
ekpnp_autotune_lbm): every shape gives the same fields and populations bit
for bit, a shape can change between steps, and the shape autotuner leaves
the state as it found it."""
import numpy as np
import pytest

import ekpnp

pytestmark = pytest.mark.gpu

FIELDS = ("rho", "ux", "uy", "cpos", "cneg", "psi")


def _channels():
    # rows of 1, 2, 3, 5 and 128 reference blocks (tiles of 2 or 4 rows leave
    # ragged ends in x), odd and even interior row counts (ragged ends in y),
    # and a 48-wide channel on the general kernel in the same context; none
    # flagged by the stability rules
    return [ekpnp.lattice_params(64, 37, tau=0.8, zeta=-1.0, kappa_h=6.0, e_x=1e-4),
            ekpnp.lattice_params(128, 22, tau=0.7, zeta=-2.0, kappa_h=4.0, e_x=3e-4),
            ekpnp.lattice_params(192, 9, tau=0.9, zeta=-0.5, kappa_h=1.2, e_x=5e-5),
            ekpnp.lattice_params(320, 14, tau=0.6, zeta=-1.5, kappa_h=2.0, e_x=2e-4),
            ekpnp.lattice_params(8192, 12, tau=0.8, zeta=-1.5, kappa_h=1.5, e_x=1e-4),
            ekpnp.lattice_params(48, 30, tau=0.75, zeta=-1.2, kappa_h=5.0, e_x=1e-4)]


def _state(s, n):
    out = {}
    for c in range(n):
        for f in FIELDS:
            out[(c, f)] = s.get_field(c, f)
        for sp in ("f", "h", "g"):
            out[(c, sp)] = s.get_populations(c, sp)
    return out


def _run(shapes, seed=3):
    """shapes: [(wrows, steps), ...] applied in order from one cold start."""
    ps = _channels()
    s = ekpnp.Solver(ps, precision="fp32")
    s.initialization(ekpnp.COLD_NOISE, seed)
    s.init_equilibrium()
    for wrows, steps in shapes:
        s.set_lbm_tiles(wrows)
        assert s.lbm_tiles() == wrows
        s.step(steps)
    st = _state(s, len(ps))
    red = s.reductions()
    s.close()
    return st, red


def _assert_same(a, b, tag):
    for key in a:
        assert np.array_equal(a[key], b[key]), (tag, key)


def test_every_tile_shape_gives_the_same_bits():
    # 70 steps: re-centrings at 8, 16, 32 and 64 inside the calls
    print("case lbm tile shapes 0 / 2 / 4, 70 steps", flush=True)
    base, red = _run([(0, 70)])
    assert np.isfinite(red).all() and (red[:, ekpnp.RED_NAMES.index("nonfinite")] == 0).all()
    for wrows in (2, 4):
        other, _ = _run([(wrows, 70)])
        _assert_same(base, other, f"wrows={wrows}")


def test_tile_shape_changes_between_steps():
    print("case lbm tile shape 0 -> 4 -> 2 between steps", flush=True)
    base, _ = _run([(0, 70)])
    mixed, _ = _run([(0, 25), (4, 20), (2, 25)])
    _assert_same(base, mixed, "mixed")


def test_shape_autotuner_leaves_the_state_unchanged():
    print("case lbm tile autotuner between steps", flush=True)
    ps = _channels()
    twins = []
    for tune in (False, True):
        s = ekpnp.Solver(ps, precision="fp32")
        s.initialization(ekpnp.COLD_NOISE, 5)
        s.init_equilibrium()
        s.step(30)
        if tune:
            psi_before = s.get_field(4, "psi")
            info = s.autotune_lbm(3)
            print("autotune_lbm", info, flush=True)
            assert info["wrows"] in (0, 2, 4) and info["ms"] > 0.0
            assert s.lbm_tiles() == info["wrows"]
            # the charge density the next Poisson solve reads is put back
            s.fast_poisson()
            assert np.array_equal(s.get_field(4, "psi"), psi_before)
        s.step(40)
        twins.append(_state(s, len(ps)))
        s.close()
    _assert_same(twins[0], twins[1], "autotuned")


def test_shape_arguments_and_contexts_without_the_row_kernel():
    s = ekpnp.Solver([ekpnp.lattice_params(64, 20)], precision="fp64")
    for bad in (-1, 1, 3, 8):
        with pytest.raises(ekpnp.EkpnpError):
            s.set_lbm_tiles(bad)
    with pytest.raises(ekpnp.EkpnpError):
        s.autotune_lbm(2)  # no populations yet
    s.initialization(ekpnp.COLD_NOISE, 1)
    s.init_equilibrium()
    # FP64 storage has no row kernel: nothing to choose, the shape is kept
    assert s.autotune_lbm(2) == {"wrows": 0, "ms": 0.0}
    s.close()
