"""Pins the FP64 MLP reference (CPU only): BF16 rounding and gradients."""
import numpy as np

import mlp_ref as R


def test_bf16_round_to_nearest_even():
    one = 1.0
    ulp = 2.0 ** -7  # BF16 spacing at 1
    vals = np.array([one, one + ulp / 2, one + 3 * ulp / 2, one + ulp / 4, -2.5, 1e-30, 3.0e38])
    got = R.bf16(vals)
    assert got[0] == 1.0
    assert got[1] == 1.0                 # tie -> even
    assert got[2] == one + 2 * ulp       # tie -> even (up)
    assert got[3] == 1.0
    assert got[4] == -2.5
    assert abs(got[5] - 1e-30) / 1e-30 < 2 ** -8
    assert np.isfinite(got[6])


def test_reference_gradient_matches_finite_differences():
    shp = R.shapes(4, 32, 2, 32)
    p = R.init_params(shp, 4, seed=3)
    rng = np.random.default_rng(0)
    x = rng.uniform(-1, 1, (8, 4))
    t = rng.standard_normal((8, 32)) * 0.3
    _, g = R.loss_grad(p, x, t, shp, emulate=False)
    # central differences on a sample of coordinates, including biases and
    # the first and last layers
    idx = np.unique(np.concatenate([rng.integers(0, p.size, 60), [0, 1, p.size - 1, p.size - 33]]))
    h = 1e-6
    for i in idx:
        if i < 32 * 32 and (i % 32) >= 4:  # padding columns of layer 1: no effect
            continue
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (R.loss_grad(pp, x, t, shp, False)[0] - R.loss_grad(pm, x, t, shp, False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (i, fd, g[i])


def test_padding_columns_have_zero_gradient():
    shp = R.shapes(4, 32, 1, 32)
    p = R.init_params(shp, 4, seed=1)
    x = np.random.default_rng(1).uniform(-1, 1, (5, 4))
    _, g = R.loss_grad(p, x, np.zeros((5, 32)), shp, emulate=True)
    W1 = g[:32 * 32].reshape(32, 32)
    assert np.all(W1[:, 4:] == 0.0)
