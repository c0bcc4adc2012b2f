"""The main.cu loop on the C ABI (ek-pnp_amd/ekpnp_run): periodic outputs at
the reference's iteration test (i % K == 1, main.cu:206-231) and the final
data_end.dat, against the Python binding stepping the same channel."""
import os
import subprocess

import numpy as np
import pytest

import ekpnp

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "ek-pnp_amd", "ekpnp_run")


def test_driver_outputs_match_the_binding(tmp_path):
    if not os.path.exists(EXE):
        pytest.fail("ekpnp_run is not built (make -C ek-pnp_amd)")
    n = 120
    r = subprocess.run([EXE, "--config", "c0", "--steps", str(n), "--nsave", "50", "--ndmd", "60",
                        "--print-current", "40", "--dir", str(tmp_path)],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    # iterations 0..n; outputs after iterations 1, 51, 101 (nsave) ...
    tec = (tmp_path / "data.dat").read_text()
    assert tec.count("ZONE") == 1 + 3
    assert (tmp_path / "umax.dat").read_text().count("\n") == 3      # 1, 41, 81
    cd = np.loadtxt(str(tmp_path / "charge_data"))
    assert cd.size == 2 * 64 * 128                                    # 1, 61
    assert r.stdout.count("Current =") == 3
    # data_end.dat after n + 1 iterations, as the binding computes them
    p = ekpnp.lattice_params(64, 128, tau=0.8, zeta=-1.0, kappa_h=8.0, e_x=1e-4)
    s = ekpnp.Solver([p], precision="fp64")
    s.initialization(ekpnp.COLD_NOISE, 0)
    s.init_equilibrium()
    s.step(n + 1)
    tab = np.loadtxt(str(tmp_path / "data_end.dat"))
    assert tab.shape == (64 * 128, 9)
    assert tab[0, 0] == pytest.approx(p.dt * (n + 1), abs=5e-7)
    for k, f in enumerate(("ux", "uy", "rho", "cpos", "psi", "ex", "ey", "cneg"), start=1):
        assert np.abs(tab[:, k] - s.get_field(0, f).ravel()).max() <= 5.0000001e-7, f
