"""ctypes wrapper of the FP64 CPU oracle (oracle/libekpnp_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product (ek-pnp_amd/) never
imports it.  The oracle restates the reference's four-stage step
(LBM.cu:331-1024), its odd-extension FFT Poisson solve (poisson.cu:180-306)
and an equivalent x-DFT + Thomas form; see ekpnp_oracle.cpp for the line map.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libekpnp_oracle.so")
sys.path.insert(0, os.path.join(os.path.dirname(_HERE), "ek-pnp_amd"))
from ekpnp import ChannelParams, FIELDS, SPECIES  # noqa: E402  (struct layout only)

POISSON_ODD_EXTENSION = 0   # the reference's method (poisson.cu:180-306)
POISSON_XDFT_THOMAS = 1     # x-DFT + Thomas along y

_lib = None
_DP = ctypes.POINTER(ctypes.c_double)
_VP = ctypes.c_void_p


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        h = ctypes.CDLL(LIB_PATH)
        sig = {
            "ekpnp_oracle_set_threads": (None, [ctypes.c_int]),
            "ekpnp_oracle_max_threads": (ctypes.c_int, []),
            "ekpnp_oracle_create": (_VP, [ctypes.POINTER(ChannelParams)]),
            "ekpnp_oracle_destroy": (None, [_VP]),
            "ekpnp_oracle_set_wall_profile": (None, [_VP, _DP, _DP, _DP, _DP, _DP, _DP]),
            "ekpnp_oracle_initialization": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_uint64, ctypes.c_int,
                                                           ctypes.c_int]),
            "ekpnp_oracle_set_fields": (None, [_VP, _DP, _DP, _DP, _DP, _DP, _DP, ctypes.c_int]),
            "ekpnp_oracle_init_equilibrium": (None, [_VP]),
            "ekpnp_oracle_stage_collide": (None, [_VP]),
            "ekpnp_oracle_stage_boundary": (None, [_VP]),
            "ekpnp_oracle_stage_stream": (None, [_VP]),
            "ekpnp_oracle_stage_bc_charge": (None, [_VP]),
            "ekpnp_oracle_efield": (None, [_VP]),
            "ekpnp_oracle_stream_collide_save": (None, [_VP]),
            "ekpnp_oracle_fast_poisson": (None, [_VP, ctypes.c_int]),
            "ekpnp_oracle_step": (None, [_VP, ctypes.c_int, ctypes.c_int]),
            "ekpnp_oracle_time_steps": (ctypes.c_double, [_VP, ctypes.c_int, ctypes.c_int]),
            "ekpnp_oracle_get_field": (None, [_VP, ctypes.c_int, _DP]),
            "ekpnp_oracle_set_field": (None, [_VP, ctypes.c_int, _DP]),
            "ekpnp_oracle_get_populations": (None, [_VP, ctypes.c_int, _DP]),
            "ekpnp_oracle_set_populations": (None, [_VP, ctypes.c_int, _DP]),
            "ekpnp_oracle_get_post": (None, [_VP, ctypes.c_int, _DP]),
            "ekpnp_oracle_set_post": (None, [_VP, ctypes.c_int, _DP]),
            "ekpnp_oracle_steps": (ctypes.c_longlong, [_VP]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _lib = h
    return _lib


def set_threads(n: int) -> None:
    lib().ekpnp_oracle_set_threads(int(n))


def max_threads() -> int:
    return int(lib().ekpnp_oracle_max_threads())


def _p(a):
    return None if a is None else a.ctypes.data_as(_DP)


def _f64(a, shape):
    arr = np.ascontiguousarray(a, dtype=np.float64).reshape(shape)
    return arr


class Oracle:
    """One channel advanced by the FP64 CPU restatement of the reference."""

    def __init__(self, params: ChannelParams, poisson_method: int = POISSON_XDFT_THOMAS):
        self.p = params.copy()
        self.nx, self.ny = params.nx, params.ny
        self.method = poisson_method
        self._h = lib().ekpnp_oracle_create(ctypes.byref(self.p))
        if not self._h:
            raise ValueError("oracle rejected the parameters")

    def close(self):
        if self._h:
            lib().ekpnp_oracle_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def shape(self):
        return (self.ny, self.nx)

    def set_wall_profile(self, psi_lo=None, psi_hi=None, cpos_lo=None, cpos_hi=None, cneg_lo=None,
                         cneg_hi=None):
        arrs = [None if a is None else _f64(a, (self.nx,)) for a in
                (psi_lo, psi_hi, cpos_lo, cpos_hi, cneg_lo, cneg_hi)]
        lib().ekpnp_oracle_set_wall_profile(self._h, *[_p(a) for a in arrs])

    def initialization(self, mode: int = 0, seed: int = 0, channel: int = 0):
        rc = lib().ekpnp_oracle_initialization(self._h, mode, seed, channel, self.method)
        if rc != 0:
            raise ValueError("unknown initialization mode")

    def set_fields(self, rho, ux, uy, cpos, cneg, psi=None):
        arrs = [_f64(a, self.shape) for a in (rho, ux, uy, cpos, cneg)]
        ps = None if psi is None else _f64(psi, self.shape)
        lib().ekpnp_oracle_set_fields(self._h, *[_p(a) for a in arrs], _p(ps), self.method)

    def init_equilibrium(self):
        lib().ekpnp_oracle_init_equilibrium(self._h)

    def stream_collide_save(self):
        lib().ekpnp_oracle_stream_collide_save(self._h)

    def fast_poisson(self, method: int | None = None):
        lib().ekpnp_oracle_fast_poisson(self._h, self.method if method is None else method)

    def efield(self):
        lib().ekpnp_oracle_efield(self._h)

    def stage(self, name: str):
        getattr(lib(), f"ekpnp_oracle_stage_{name}")(self._h)

    def step(self, n: int = 1):
        lib().ekpnp_oracle_step(self._h, int(n), self.method)

    def time_steps(self, n: int) -> float:
        return float(lib().ekpnp_oracle_time_steps(self._h, int(n), self.method))

    def get_field(self, name: str) -> np.ndarray:
        out = np.empty(self.shape, dtype=np.float64)
        lib().ekpnp_oracle_get_field(self._h, FIELDS[name], _p(out))
        return out

    def set_field(self, name: str, a) -> None:
        arr = _f64(a, self.shape)
        lib().ekpnp_oracle_set_field(self._h, FIELDS[name], _p(arr))

    def get_populations(self, species: str) -> np.ndarray:
        out = np.empty((9,) + self.shape, dtype=np.float64)
        lib().ekpnp_oracle_get_populations(self._h, SPECIES[species], _p(out))
        return out

    def set_populations(self, species: str, a) -> None:
        arr = _f64(a, (9,) + self.shape)
        lib().ekpnp_oracle_set_populations(self._h, SPECIES[species], _p(arr))

    def get_post(self, species: str) -> np.ndarray:
        out = np.empty((8,) + self.shape, dtype=np.float64)
        lib().ekpnp_oracle_get_post(self._h, SPECIES[species], _p(out))
        return out

    def set_post(self, species: str, a) -> None:
        arr = _f64(a, (8,) + self.shape)
        lib().ekpnp_oracle_set_post(self._h, SPECIES[species], _p(arr))

    @property
    def steps(self) -> int:
        return int(lib().ekpnp_oracle_steps(self._h))


def poisson_sor(rho_e, psi, psi_lo, psi_hi, dx: float, dy: float, poisson_coef: float, omega: float = 1.0,
                tol: float = 1e-9, max_iter: int = 10 ** 6, check_every: int = 1):
    """Iterative 5-point Poisson solve: restatement of poisson_phi and
    gpu_poisson (poisson.cu:35-130) in red-black order.

    Solves (psi_L + psi_R - 2 psi)/dx^2 + (psi_U + psi_D - 2 psi)/dy^2
    = -poisson_coef rho_e on rows 1..ny-2, periodic in x, with rows 0 and
    ny-1 held at psi_lo(x), psi_hi(x) (poisson.cu:96-103).  gpu_poisson's
    update psi = 0.25 (psi_L + psi_R + psi_U + psi_D + source), source =
    charge/eps dx^2 (poisson.cu:118-124), is written for dx != dy as
      psi_gs = (r (psi_L + psi_R) + psi_U + psi_D + poisson_coef dy^2 rho_e)
               / (2 (1 + r)),  r = (dy/dx)^2,
    and relaxed with omega (psi_gs itself for omega = 1).  The reference
    updates every node at once, in place; here one sweep updates the nodes
    with (x + y) even, then those with (x + y) odd (a Gauss-Seidel order with
    the same fixed point).  Stops when the largest |change| of a sweep is
    below tol (poisson.cu:50-57), tested every check_every sweeps, or after
    max_iter sweeps.  Returns (psi, sweeps, change).
    """
    rho_e = np.asarray(rho_e, dtype=np.float64)
    ny, nx = rho_e.shape
    p = np.array(psi, dtype=np.float64, copy=True).reshape(ny, nx)
    p[0] = psi_lo
    p[ny - 1] = psi_hi
    r = (dy / dx) ** 2
    inv = 1.0 / (2.0 * (1.0 + r))
    src = poisson_coef * dy * dy * rho_e
    yy, xx = np.mgrid[0:ny, 0:nx]
    interior = (yy >= 1) & (yy <= ny - 2)
    masks = [interior & ((xx + yy) % 2 == c) for c in (0, 1)]
    it, change = 0, np.inf
    while it < max_iter:
        ch = 0.0
        for m in masks:
            gs = (r * (np.roll(p, 1, axis=1) + np.roll(p, -1, axis=1)) + np.roll(p, -1, axis=0)
                  + np.roll(p, 1, axis=0) + src) * inv
            new = gs if omega == 1.0 else p + omega * (gs - p)
            ch = max(ch, float(np.max(np.abs(new[m] - p[m]))))
            p[m] = new[m]
        it += 1
        if it % check_every == 0 or it == max_iter:
            change = ch
            if change < tol:
                break
    return p, it, change
