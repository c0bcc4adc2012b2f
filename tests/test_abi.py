"""The C ABI boundary (CPU only: no compute calls)."""
import ctypes
import os
import re
import subprocess

import pytest

import ekpnp

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ekpnp.h")
HEADERS = [HEADER, os.path.join(ROOT, "include", "ekpnp_mlp.h")]


def declared_functions():
    names = set()
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(ekpnp_[a-z_0-9]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_entry_point():
    lib = ekpnp.lib()
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the Python mirror binds each of them with a signature
    from ekpnp import surrogate
    bound = {n for n, _, _ in ekpnp.SIGNATURES + surrogate.MLP_SIGNATURES}
    assert set(names) == bound, set(names) ^ bound


def test_nm_lists_only_the_c_abi():
    out = subprocess.run(["nm", "-D", "--defined-only", ekpnp.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T ekpnp_" in l}
    assert exported == set(declared_functions())


def test_struct_layout_matches_header(tmp_path):
    fields = [f for f, _ in ekpnp.ChannelParams._fields_]
    src = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void){",
           'printf("%zu %zu\\n", sizeof(ekpnp_channel_params), sizeof(ekpnp_options));']
    for f in fields:
        src.append(f'printf("%zu\\n", offsetof(ekpnp_channel_params, {f}));')
    for f, _ in ekpnp.Options._fields_:
        src.append(f'printf("%zu\\n", offsetof(ekpnp_options, {f}));')
    src.append("return 0;}")
    c = tmp_path / "layout.c"
    c.write_text("\n".join(src))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-o", str(exe), str(c)], check=True)
    vals = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True)
            .stdout.split()]
    assert vals[0] == ctypes.sizeof(ekpnp.ChannelParams)
    assert vals[1] == ctypes.sizeof(ekpnp.Options)
    offs = vals[2:2 + len(fields)]
    assert offs == [getattr(ekpnp.ChannelParams, f).offset for f in fields]
    oo = vals[2 + len(fields):]
    assert oo == [getattr(ekpnp.Options, f).offset for f, _ in ekpnp.Options._fields_]


def test_parameter_helpers_match_c():
    lib = ekpnp.lib()
    a = ekpnp.ChannelParams()
    lib.ekpnp_reference_params(ctypes.byref(a))
    assert a.as_dict() == ekpnp.reference_params().as_dict()
    for args in ((64, 128, 0.8, -1.0, 8.0, 1e-4, 0.05), (4096, 1024, 0.6, -2.0, 32.0, 5e-5, 0.05),
                 (33, 70, 0.55, -3.7, 41.0, 1e-3, 0.02)):
        c = ekpnp.ChannelParams()
        assert lib.ekpnp_lattice_params(ctypes.byref(c), *args) == 0
        py = ekpnp.lattice_params(args[0], args[1], tau=args[2], zeta=args[3], kappa_h=args[4],
                                  e_x=args[5], diff=args[6])
        for k, v in c.as_dict().items():
            assert v == pytest.approx(py.as_dict()[k], rel=1e-15, abs=0), k
    # invalid inputs are rejected without a device
    assert lib.ekpnp_lattice_params(ctypes.byref(ekpnp.ChannelParams()), 64, 3, 0.8, -1, 8, 1e-4, 0.05) != 0
    assert lib.ekpnp_lattice_params(ctypes.byref(ekpnp.ChannelParams()), 64, 64, 0.5, -1, 8, 1e-4, 0.05) != 0


def test_create_rejects_bad_parameters_before_touching_a_device():
    lib = ekpnp.lib()
    ctx = ctypes.c_void_p()
    p = ekpnp.lattice_params(16, 16)
    p.nu = -1.0
    arr = (ekpnp.ChannelParams * 1)(p)
    assert lib.ekpnp_create(arr, 1, None, ctypes.byref(ctx)) == ekpnp.ERR_INVALID
    assert b"viscosity" in lib.ekpnp_create_error()
    # a width changed without the x length (the x-FFT wavenumbers follow lx)
    q = ekpnp.lattice_params(256, 16)
    q.nx = 8
    arr = (ekpnp.ChannelParams * 1)(q)
    assert lib.ekpnp_create(arr, 1, None, ctypes.byref(ctx)) == ekpnp.ERR_INVALID
    assert b"lx differs from nx * dx" in lib.ekpnp_create_error()
    big = ekpnp.lattice_params(16384, 16)
    arr = (ekpnp.ChannelParams * 1)(big)
    assert lib.ekpnp_create(arr, 1, None, ctypes.byref(ctx)) == ekpnp.ERR_UNSUPPORTED
    assert lib.ekpnp_create(arr, 0, None, ctypes.byref(ctx)) == ekpnp.ERR_INVALID


def test_python_defaults_match_the_c_defaults():
    # the binding's Solver must not silently pick other options than
    # ekpnp_default_options (FP64 collision arithmetic, graphs on)
    import inspect
    o = ekpnp.Options()
    ekpnp.lib().ekpnp_default_options(ctypes.byref(o))
    sig = inspect.signature(ekpnp.Solver.__init__).parameters
    assert o.compute_fp64 == 1 and sig["compute_fp64"].default is True
    assert o.use_graph == 1 and sig["use_graph"].default is True


def test_check_params_flags_unresolved_and_fast_drift_channels():
    # kappa = sqrt(2 pc c_bulk inv_vt) = kappa_h / h in lattice units; the
    # drift is K (2 kappa) sinh(|zeta|/2) for psi in kT/e
    import math
    ok = ekpnp.lattice_params(64, 128, zeta=-1.0, kappa_h=8.0)            # kappa dy 0.126
    debye = ekpnp.lattice_params(64, 40, zeta=-0.5, kappa_h=9.0)          # kappa dy 0.46, drift 0.012
    drift = ekpnp.lattice_params(64, 200, zeta=-4.0, kappa_h=30.0)        # kappa dy 0.30, drift 0.11
    ref = ekpnp.reference_params()                                        # kappa dy 0.11 (SI units)
    # tau -> 1/2 with a strong field: the Helmholtz-Smoluchowski velocity
    # |zeta| E_x / nu (force_coef = poisson_coef, rho0 = 1) = 2e-2 / (0.01/3) = 6
    flow = ekpnp.lattice_params(64, 128, tau=0.51, zeta=-2.0, kappa_h=8.0, e_x=1e-2)
    # just inside the flow limit: 0.09 at tau 0.8
    slow = ekpnp.lattice_params(64, 128, tau=0.8, zeta=-3.0, kappa_h=8.0, e_x=3e-3)
    flags = ekpnp.check_params([ok, debye, drift, ref, flow, slow])
    assert flags.tolist() == [0, ekpnp.CHECK_DEBYE_UNRESOLVED, ekpnp.CHECK_WALL_DRIFT, 0, ekpnp.CHECK_FLOW_FAST, 0]
    kappa = 30.0 / (0.5 * 199)
    assert 0.05 < 0.05 * 2 * kappa * math.sinh(2.0) and kappa < 0.4
    assert abs(2.0 * 1e-2 / ((0.51 - 0.5) / 3) - 6.0) < 1e-9 and 3.0 * 3e-3 / 0.1 < 0.1
    # the count and argument checks of the C entry point
    arr = (ekpnp.ChannelParams * 6)(ok, debye, drift, ref, flow, slow)
    assert ekpnp.lib().ekpnp_check_params(arr, 6, None) == 3
    assert ekpnp.lib().ekpnp_check_params(arr, 0, None) == ekpnp.ERR_INVALID


def test_create_rejects_flagged_channels_on_request_before_touching_a_device():
    lib = ekpnp.lib()
    ctx = ctypes.c_void_p()
    good = ekpnp.lattice_params(64, 128, zeta=-1.0, kappa_h=8.0)
    bad = ekpnp.lattice_params(64, 40, zeta=-0.5, kappa_h=9.0)
    arr = (ekpnp.ChannelParams * 2)(good, bad)
    o = ekpnp.Options()
    lib.ekpnp_default_options(ctypes.byref(o))
    assert o.reject_unresolved == 0
    o.reject_unresolved = 1
    assert lib.ekpnp_create(arr, 2, ctypes.byref(o), ctypes.byref(ctx)) == ekpnp.ERR_UNSTABLE
    msg = lib.ekpnp_create_error()
    assert b"channel 1" in msg and b"kappa dy" in msg, msg
    assert not ctx.value
