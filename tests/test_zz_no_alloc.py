"""ekpnp_reductions_noalloc: the allocation-free variant of ekpnp_reductions
(partial sums and per-channel records in the library's fixed device scratch
array) against ekpnp_reductions on identical contexts, bit for bit, on two
calls in a row (the second on a smaller context, after the first left its
larger data in the scratch array), and the size limit."""
import numpy as np
import pytest

import ekpnp

pytestmark = pytest.mark.gpu


def _pair(ps, steps, precision="fp32"):
    out = []
    for _ in range(2):
        s = ekpnp.Solver(ps, precision=precision)
        s.initialization(ekpnp.COLD_NOISE, 6)
        s.init_equilibrium()
        s.step(steps)
        out.append(s)
    return out


def _same(a, b):
    # bit-identical records, NaN-free
    assert np.isfinite(a).all() and np.isfinite(b).all()
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), (a, b)


def test_noalloc_reductions_equal_the_allocating_ones():
    big = [ekpnp.lattice_params(256, 130, tau=0.8, zeta=-1.0, kappa_h=8.0, e_x=1e-4),
           ekpnp.lattice_params(64, 66, tau=0.7, zeta=-2.0, kappa_h=4.0, e_x=3e-4),
           ekpnp.lattice_params(48, 40, tau=0.9, zeta=-0.5, kappa_h=3.0, e_x=1e-5)]
    small = [ekpnp.lattice_params(32, 20, tau=0.8, zeta=-1.5, kappa_h=3.0, e_x=2e-4)]
    for name, ps in (("three channels", big), ("one small channel", small)):
        print("case", name, flush=True)
        a, b = _pair(ps, 150)
        # first call (snapshot), then a residual over 60 more steps
        _same(a.reductions(), b.reductions_noalloc())
        a.step(60)
        b.step(60)
        ra, rb = a.reductions(), b.reductions_noalloc()
        _same(ra, rb)
        assert (ra[:, 7] > 0).all()  # a real residual over the interval
        a.close()
        b.close()


def test_noalloc_reductions_reject_contexts_beyond_the_scratch():
    # two 8192 x 2050 channels: 2 x 65600 work-groups of 256 nodes, 8 doubles
    # each, exceed the 2^20-double scratch
    ps = [ekpnp.lattice_params(8192, 2050, tau=0.8, zeta=-1.0, kappa_h=64.0, e_x=1e-4)] * 2
    s = ekpnp.Solver(ps, precision="fp32")
    with pytest.raises(ekpnp.EkpnpError) as ei:
        s.reductions_noalloc()
    assert ei.value.code == ekpnp.ERR_UNSUPPORTED
    assert "scratch" in str(ei.value)
    s.close()
