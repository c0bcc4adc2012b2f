"""The learned warm start end to end on the GPU (small scale): training on
steady states the batched solver produces, the gain over cold starts, and
the acceptance test's fallback when a prediction is corrupted."""
import os
import sys

import numpy as np
import pytest

import ekpnp
from ekpnp import surrogate as S

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "ek-pnp_amd"))
import run_time_to_steady as T  # noqa: E402

pytestmark = pytest.mark.gpu

NY = (32, 48)
# a kept warm start changes u_x by < 1e-4 per step over the 200-step probe
# that follows 100 settling steps (accurate predictions sit near 1e-6, a
# sign-flipped state near 1e-3)
ACCEPT = 1e-4


@pytest.fixture(scope="module")
def trained():
    tr, _ = S.sample_channels(768, 1000, nx_choices=(4,), ny_choices=NY)
    X, Y, steps, _ = T.generate(tr, "fp32", 1000, 1e-8, 100, 400000, residual=True)
    good = steps > 0
    print("training channels converged", int(good.sum()), "of", len(tr), flush=True)
    assert good.mean() > 0.95
    net, hist = T.train(X[good], Y[good], 400, 128, 2e-3, 1000)
    print("val mse", hist[-1][2], flush=True)
    return net


def test_warm_start_cuts_steps_to_steady(trained):
    ho, _ = S.sample_channels(12, 901000, nx_choices=(32,), ny_choices=NY)
    s, cold, _ = T.run_cold(ho, "fp32", 901000, 1e-8, 100, 400000)
    s.close()
    s, warm, _, res, rejected = T.run_warm(ho, T.predictor(trained, True), "fp32", 901000, 1e-8, 100, 400000,
                                           200, ACCEPT)
    s.close()
    print("cold", cold.tolist(), "warm", warm.tolist(), "rejected", rejected, flush=True)
    assert (cold > 0).all() and (warm > 0).all()
    assert not rejected
    # the network's prediction beats the cold start on every channel and by
    # a wide margin overall
    assert np.median(cold / warm) > 2.0
    assert (warm < cold).all()


def test_corrupted_prediction_is_rejected_and_falls_back_to_cold(trained):
    ho, _ = S.sample_channels(6, 902000, nx_choices=(32,), ny_choices=NY)
    s, cold, _ = T.run_cold(ho, "fp32", 902000, 1e-8, 100, 400000)
    s.close()

    def flip(c, psi, cp, cn, ux):
        # channels 1 and 4: sign-flipped potential, swapped ions, reversed flow
        if c in (1, 4):
            return -psi, cn, cp, -ux
        return psi, cp, cn, ux

    s, warm, _, res, rejected = T.run_warm(ho, T.predictor(trained, True), "fp32", 902000, 1e-8, 100, 400000,
                                           200, ACCEPT, corrupt=flip)
    s.close()
    print("probe residuals", res.tolist(), "rejected", rejected, flush=True)
    assert sorted(rejected) == [1, 4]
    # a rejected channel restarts from the identical cold state: same steps
    # as the cold arm, plus settling (100) and probe (200)
    for c in (1, 4):
        assert warm[c] == cold[c] + 300, (c, warm[c], cold[c])
    for c in (0, 2, 3, 5):
        assert warm[c] < cold[c]


NET_PATH = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "ek-pnp_amd", "ekpnp",
                        "warmstart_net.npz")


@pytest.mark.skipif(not os.path.exists(NET_PATH), reason="no committed warm-start network")
@pytest.mark.timeout(600)
def test_committed_network_meets_the_fivefold_target():
    # the network committed with the library (trained on the batched
    # solver's steady states over n_y 64..1024) on held-out channels of
    # another seed and width, n_y 64 and 128 (cold runs of taller channels
    # take millions of steps): every warm start kept, faster than cold on
    # every channel, and a median per-channel gain of at least 5
    net = S.load_net(NET_PATH)
    ho, _ = S.sample_channels(16, 903000, nx_choices=(32,), ny_choices=(64, 128))
    s, cold, _ = T.run_cold(ho, "fp32", 903000, 1e-8, 100, 3_000_000)
    s.close()
    s, warm, _, res, rejected = T.run_warm(ho, T.predictor(net, True), "fp32", 903000, 1e-8, 100, 3_000_000,
                                           200, ACCEPT)
    s.close()
    net.close()
    gain = cold / np.maximum(warm, 1)
    print("cold", cold.tolist(), "warm", warm.tolist(), "rejected", rejected, "median gain",
          float(np.median(gain)), flush=True)
    assert (cold > 0).all() and (warm > 0).all()
    assert not rejected
    assert (warm < cold).all()
    assert np.median(gain) >= 5.0


@pytest.mark.skipif(not os.path.exists(NET_PATH), reason="no committed warm-start network")
@pytest.mark.timeout(900)
def test_batched_sweep_runs_to_the_residual_from_warm_starts():
    # the batched sweep (256 channels of seed 2: n_x 32..256, n_y 64..1024,
    # 18.2 M nodes) in one context, warm-started from the committed network
    # and finished by the plain method to the 1e-8 u_x residual: every
    # prediction kept, every channel converged (measured makespan 158 500
    # steps; cold starts of its n_y >= 896 channels take millions), fields
    # finite and the fluid mass kept
    net = S.load_net(NET_PATH)
    ps, _ = S.sample_channels(256, 2)
    print("case batched sweep to the residual, warm-started", flush=True)
    predict = T.predictor(net, True)
    # the warm start's density: the osmotic profile of the predicted columns
    prof = predict(ps)
    mass0 = np.array([p.nx * S.osmotic_rho(p, *prof[c][:3]).sum() for c, p in enumerate(ps)])
    s, warm, wall, res, rejected = T.run_warm(ps, predict, "fp32", 2, 1e-8, 100, 400_000, 200, ACCEPT,
                                              chunk=20_000)
    red = s.reductions()
    mass = red[:, ekpnp.RED_NAMES.index("mass")]
    s.close()
    net.close()
    print("converged", int((warm > 0).sum()), "median", float(np.median(warm[warm > 0])), "makespan",
          int(warm.max()), "wall", round(wall, 1), "rejected", rejected, flush=True)
    assert not rejected
    assert (warm > 0).all()
    assert (red[:, ekpnp.RED_NAMES.index("nonfinite")] == 0).all()
    # every channel keeps the fluid mass it was started with
    drift = np.abs(mass / mass0 - 1.0)
    print("largest mass drift", float(drift.max()), flush=True)
    assert drift.max() < 1e-6
