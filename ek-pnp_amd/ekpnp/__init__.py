"""Python host mirror of the EK-PNP C ABI (include/ekpnp.h).

The reference drives its coupled step from main.cu through the host
operators of LBM.h:149-169 (``initialization``, ``init_equilibrium``,
``stream_collide_save``, ``fast_Poisson``, the ``save_data_*`` writers).  This
module exposes the same operators, with the same names, on top of
``libekpnp.so`` (hand-written gfx950 kernels):

    s = Solver([lattice_params(64, 128, tau=0.8, zeta=-1, kappa_h=8, e_x=1e-4)])
    s.initialization(COLD_NOISE, seed=0)      # LBM.cu:95
    s.init_equilibrium()                      # LBM.cu:183
    for _ in range(n):                        # main.cu:190-199
        s.stream_collide_save()               # LBM.cu:307
        s.fast_poisson()                      # poisson.cu:180
    ux = s.get_field(0, "ux")                 # save_data_end columns, LBM.cu:1298

There is no CPU fallback: every call goes to the GPU library, and importing
this module raises if the library has not been built.
"""
from __future__ import annotations

import ctypes
import os
from typing import Iterable, Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# EKPNP_LIB selects another build of the same library (e.g. a tuning variant)
LIB_PATH = os.environ.get("EKPNP_LIB", os.path.join(os.path.dirname(_HERE), "libekpnp.so"))

# status codes (include/ekpnp.h)
OK, ERR_INVALID, ERR_HIP, ERR_NOMEM, ERR_STATE, ERR_UNSUPPORTED, ERR_UNSTABLE, ERR_IO = 0, -1, -2, -3, -4, -5, -6, -7
POP_FP32, POP_FP64 = 0, 1
FLAG_ION_DRIFT_APPLIED = 0x1
FLAG_G0_FLUID_U = 0x2
COLD_NOISE, REFERENCE, PB_PICARD = 0, 1, 2
FIELDS = {"rho": 0, "ux": 1, "uy": 2, "cpos": 3, "cneg": 4, "psi": 5, "ex": 6, "ey": 7}
SPECIES = {"f": 0, "h": 1, "g": 2}
RED_NAMES = ("mass", "cpos", "cneg", "charge", "dux2", "ux2", "nonfinite", "residual")

_DOUBLE_FIELDS = (
    "dx", "dy", "dt", "cfl", "cs2", "rho0", "nu", "diff_pos", "diff_neg", "mob_pos", "mob_neg",
    "magic_f", "magic_pos", "magic_neg", "psi_lo", "psi_hi", "cpos_lo", "cpos_hi", "cneg_lo",
    "cneg_hi", "c_bulk", "poisson_coef", "force_coef", "e_applied", "exf", "uw", "lx", "init_noise",
    "inv_thermal_voltage", "pb_omega", "pb_iterations",
)


class ChannelParams(ctypes.Structure):
    """ekpnp_channel_params: one channel's size and physics (LBM.h:31-112)."""

    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("flags", ctypes.c_uint32),
                ("reserved0", ctypes.c_int32)] + [(n, ctypes.c_double) for n in _DOUBLE_FIELDS]

    def copy(self) -> "ChannelParams":
        c = ChannelParams()
        ctypes.pointer(c)[0] = self
        return c

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class Options(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int32), ("device", ctypes.c_int32),
                ("use_graph", ctypes.c_int32), ("compute_fp64", ctypes.c_int32),
                ("reject_unresolved", ctypes.c_int32)]


class EkpnpError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"ekpnp error {code}: {message}")
        self.code = code


_lib_handle = None

# (name, restype, argtypes) for every entry point of include/ekpnp.h
_P = ctypes.POINTER
# x-FFT kernel families of a Poisson plan (ekpnp_set_poisson_plan's fft_persistent)
FFT_KERNELS = ("one-tile", "persistent", "row-pair")

_VP = ctypes.c_void_p
_DP = _P(ctypes.c_double)
SIGNATURES = [
    ("ekpnp_reference_params", None, [_P(ChannelParams)]),
    ("ekpnp_lattice_params", ctypes.c_int, [_P(ChannelParams), ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                            ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double]),
    ("ekpnp_default_options", None, [_P(Options)]),
    ("ekpnp_check_params", ctypes.c_int, [_P(ChannelParams), ctypes.c_int, _P(ctypes.c_uint32)]),
    ("ekpnp_create", ctypes.c_int, [_P(ChannelParams), ctypes.c_int, _P(Options), _P(_VP)]),
    ("ekpnp_destroy", None, [_VP]),
    ("ekpnp_last_error", ctypes.c_char_p, [_VP]),
    ("ekpnp_create_error", ctypes.c_char_p, []),
    ("ekpnp_initialization", ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_uint64]),
    ("ekpnp_set_wall_profile", ctypes.c_int, [_VP, ctypes.c_int, _DP, _DP, _DP, _DP, _DP, _DP]),
    ("ekpnp_set_fields", ctypes.c_int, [_VP, ctypes.c_int, _DP, _DP, _DP, _DP, _DP, _DP]),
    ("ekpnp_init_equilibrium", ctypes.c_int, [_VP]),
    ("ekpnp_restart_channel", ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_int, ctypes.c_uint64]),
    ("ekpnp_stream_collide_save", ctypes.c_int, [_VP]),
    ("ekpnp_fast_poisson", ctypes.c_int, [_VP]),
    ("ekpnp_poisson_iterative", ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                               ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                               ctypes.POINTER(ctypes.c_double)]),
    ("ekpnp_step", ctypes.c_int, [_VP, ctypes.c_int]),
    ("ekpnp_rebase", ctypes.c_int, [_VP]),
    ("ekpnp_run_to_residual", ctypes.c_int, [_VP, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                             _P(ctypes.c_int)]),
    ("ekpnp_get_field", ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_int, _DP]),
    ("ekpnp_get_populations", ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_int, _DP]),
    ("ekpnp_set_populations", ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_int, _DP]),
    ("ekpnp_reductions", ctypes.c_int, [_VP, _DP]),
    ("ekpnp_reductions_noalloc", ctypes.c_int, [_VP, _DP]),
    ("ekpnp_synchronize", ctypes.c_int, [_VP]),
    ("ekpnp_save_data_end", ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_double, ctypes.c_char_p]),
    ("ekpnp_read_data", ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_char_p, _DP]),
    ("ekpnp_save_tecplot", ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_double, ctypes.c_char_p, ctypes.c_int]),
    ("ekpnp_save_charge_data", ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_char_p]),
    ("ekpnp_current", ctypes.c_int, [_VP, ctypes.c_int, _DP]),
    ("ekpnp_record_umax", ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_double, ctypes.c_char_p, _DP]),
    ("ekpnp_save_checkpoint", ctypes.c_int, [_VP, ctypes.c_char_p]),
    ("ekpnp_load_checkpoint", ctypes.c_int, [_VP, ctypes.c_char_p]),
    ("ekpnp_get_params", ctypes.c_int, [_VP, ctypes.c_int, _P(ChannelParams)]),
    ("ekpnp_set_poisson_plan", ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    ("ekpnp_autotune", ctypes.c_int, [_VP, ctypes.c_int, _DP]),
    ("ekpnp_get_poisson_plan", ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_int)]),
    ("ekpnp_set_lbm_tiles", ctypes.c_int, [_VP, ctypes.c_int]),
    ("ekpnp_get_lbm_tiles", ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_int)]),
    ("ekpnp_autotune_lbm", ctypes.c_int, [_VP, ctypes.c_int, _DP]),
    ("ekpnp_num_channels", ctypes.c_int, [_VP]),
    ("ekpnp_steps_done", ctypes.c_longlong, [_VP]),
    ("ekpnp_device_bytes", ctypes.c_size_t, [_VP]),
    ("ekpnp_time_stages", ctypes.c_int, [_VP, ctypes.c_int, _DP]),
    ("ekpnp_copy_bandwidth", ctypes.c_int, [ctypes.c_int, ctypes.c_size_t, ctypes.c_int, ctypes.c_int, _DP, _DP]),
]


def lib() -> ctypes.CDLL:
    """Load libekpnp.so (raises if it was not built)."""
    global _lib_handle
    if _lib_handle is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C ek-pnp_amd` "
                               "(or __graft_entry__.build()); there is no CPU fallback")
        h = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _lib_handle = h
    return _lib_handle


def reference_params() -> ChannelParams:
    """The reference's compiled-in SI case (LBM.h:31-112, main.cu:38-44).

    Same values as the C ABI's ekpnp_reference_params (a CPU test checks
    that the two agree field by field)."""
    p = ChannelParams()
    p.nx, p.ny = 10, 101                         # LBM.h:35-36
    p.flags = FLAG_G0_FLUID_U                    # LBM.cu:493 as shipped
    p.dx = p.dy = 1.0e-6 / 100.0                 # LBM.h:42-43
    p.cfl = 0.01                                 # LBM.h:50
    p.dt = 0.01 * 1.0e-6 / 100.0                 # LBM.h:51
    p.cs2 = 1.0 / 3.0 / (0.01 * 0.01)            # LBM.h:52
    p.rho0 = 1000.0                              # LBM.h:53
    p.nu = 0.889e-6                              # LBM.h:65
    p.diff_pos, p.diff_neg = 1.0e-8, 1.0e-8      # LBM.h:64, 71
    p.mob_pos, p.mob_neg = 4.245e-7, -4.245e-7   # LBM.h:67, 74
    p.magic_f = p.magic_pos = p.magic_neg = 1.0e-6  # LBM.h:110-112
    p.psi_lo = p.psi_hi = -5.2574e-3             # LBM.h:58, 60
    p.cpos_lo = p.cpos_hi = 1.2364549e-2         # LBM.h:55-56
    p.cneg_lo = p.cneg_hi = 8.087639e-3          # LBM.h:81, 83
    p.c_bulk = 0.01                              # chargeinf, LBM.h:54
    p.poisson_coef = 9.64e4 / 6.95e-10           # convertCtoCharge/eps, LBM.h:63, 90
    p.force_coef = 0.0                           # LBM.cu:388-389 as shipped
    p.e_applied = 1.0e3                          # LBM.h:62
    p.exf = p.uw = 0.0                           # LBM.h:46-47
    p.lx = 1.0e-7                                # LBM.h:40
    p.init_noise = 0.0
    p.inv_thermal_voltage = 1.6e-19 / 1.38e-23 / 273.0  # electron/kB/roomT, LBM.h:87-89
    p.pb_omega = 0.05                            # LBM.h:91
    p.pb_iterations = 501.0                      # LBM.cu:116
    return p


def lattice_params(nx: int, ny: int, tau: float = 0.8, zeta: float = -1.0, kappa_h: float = 8.0,
                   e_x: float = 1e-4, diff: float = 0.05) -> ChannelParams:
    """Lattice-unit channel: dx = dt = 1, cs^2 = 1/3, psi in kT/e, c in units of c0.

    nu = (tau - 1/2)/3; TRT magic 1/4 for all species; D+- = diff and
    K+- = +-diff (Einstein relation with psi in kT/e); wall psi = zeta,
    wall c+ = exp(-zeta), wall c- = exp(zeta); kappa = kappa_h/h with
    h = (ny-1)/2; the Poisson right-hand side is -(kappa^2/2)(c+ - c-) and the
    body force (kappa^2/2)(c+ - c-)(E + E_x), so the Helmholtz-Smoluchowski
    velocity is E_x (psi - zeta)/(rho0 nu).  Same values as the C ABI's
    ekpnp_lattice_params."""
    if nx < 2 or ny < 4 or not tau > 0.5 or not diff > 0:
        raise EkpnpError(ERR_INVALID, "invalid lattice parameters")
    p = ChannelParams()
    p.nx, p.ny, p.flags = nx, ny, 0
    p.dx = p.dy = p.dt = p.cfl = 1.0
    p.cs2 = 1.0 / 3.0
    p.rho0 = 1.0
    p.nu = (tau - 0.5) / 3.0
    p.diff_pos = p.diff_neg = diff
    p.mob_pos, p.mob_neg = diff, -diff
    p.magic_f = p.magic_pos = p.magic_neg = 0.25
    p.psi_lo = p.psi_hi = zeta
    p.cpos_lo = p.cpos_hi = float(np.exp(-zeta))
    p.cneg_lo = p.cneg_hi = float(np.exp(zeta))
    p.c_bulk = 1.0
    kappa = kappa_h / (0.5 * (ny - 1))
    p.poisson_coef = 0.5 * kappa * kappa
    p.force_coef = p.poisson_coef
    p.e_applied = e_x
    p.exf = p.uw = 0.0
    p.lx = float(nx)
    p.init_noise = 1.0e-3
    p.inv_thermal_voltage = 1.0
    p.pb_omega = 0.05
    p.pb_iterations = 501.0
    return p


REBASE_INTERVAL = 4096  # EKPNP_REBASE_INTERVAL


def rebase_steps(upto: int):
    """Steps after init_equilibrium at which FP32 storage re-centres its
    references inside ekpnp_step (include/ekpnp.h): 8, 16, ..., 4096, then
    every 4096, up to and including `upto`."""
    out, last = [], 0
    while True:
        last = max(8, 2 * last) if last < REBASE_INTERVAL else last + REBASE_INTERVAL
        if last > upto:
            return out
        out.append(last)


def copy_bandwidth(nbytes: int = 1 << 30, reps: int = 10, grid: int = 1024, device: int = 0) -> dict:
    """Non-temporal float4 device copy of nbytes (ekpnp_copy_bandwidth):
    {"GBps": read + write bytes per second / 1e9, "ms_per_copy": ...}."""
    g = ctypes.c_double(0.0)
    ms = ctypes.c_double(0.0)
    rc = lib().ekpnp_copy_bandwidth(int(device), int(nbytes), int(reps), int(grid), ctypes.byref(g),
                                    ctypes.byref(ms))
    if rc != OK:
        raise EkpnpError(rc, "copy bandwidth measurement failed")
    return {"GBps": g.value, "ms_per_copy": ms.value, "bytes": int(nbytes), "grid": int(grid), "reps": int(reps)}


CHECK_DEBYE_UNRESOLVED, CHECK_WALL_DRIFT, CHECK_FLOW_FAST = 0x1, 0x2, 0x4


def check_params(channels: Sequence[ChannelParams]) -> np.ndarray:
    """EKPNP_CHECK_* flags per channel (ekpnp_check_params; no device needed):
    Debye length unresolved (kappa dy > 0.4), ions' wall drift above 0.05,
    electro-osmotic (Helmholtz-Smoluchowski) flow above 0.1 dx/dt."""
    if isinstance(channels, ChannelParams):
        channels = [channels]
    arr = (ChannelParams * len(channels))(*channels)
    out = (ctypes.c_uint32 * len(channels))()
    rc = lib().ekpnp_check_params(arr, len(channels), out)
    if rc < 0:
        raise EkpnpError(rc, "ekpnp_check_params: no channels")
    return np.array(out[:], dtype=np.uint32)


def _dptr(a: Optional[np.ndarray]):
    if a is None:
        return None
    return a.ctypes.data_as(_DP)


def _as_f64(a, shape) -> np.ndarray:
    arr = np.ascontiguousarray(a, dtype=np.float64)
    if arr.shape != shape and arr.size == int(np.prod(shape)):
        arr = arr.reshape(shape)
    if arr.shape != shape:
        raise ValueError(f"expected shape {shape}, got {arr.shape}")
    return arr


class Solver:
    """One EK-PNP context: every channel of `channels` advances in lock-step."""

    def __init__(self, channels: Sequence[ChannelParams], precision: str = "fp32", device: int = 0,
                 use_graph: bool = True, compute_fp64: bool = True, reject_unresolved: bool = False):
        L = lib()
        if isinstance(channels, ChannelParams):
            channels = [channels]
        self.channels = [c.copy() for c in channels]
        arr = (ChannelParams * len(self.channels))(*self.channels)
        opt = Options()
        L.ekpnp_default_options(ctypes.byref(opt))
        opt.precision = {"fp32": POP_FP32, "fp64": POP_FP64}[precision]
        opt.device = device
        opt.use_graph = 1 if use_graph else 0
        opt.compute_fp64 = 1 if compute_fp64 else 0
        opt.reject_unresolved = 1 if reject_unresolved else 0
        self.precision = precision
        ctx = _VP()
        rc = L.ekpnp_create(arr, len(self.channels), ctypes.byref(opt), ctypes.byref(ctx))
        if rc != OK:
            raise EkpnpError(rc, L.ekpnp_create_error().decode())
        self._ctx = ctx

    # -- lifetime ---------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "_ctx", None):
            lib().ekpnp_destroy(self._ctx)
            self._ctx = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int) -> None:
        if rc != OK:
            raise EkpnpError(rc, lib().ekpnp_last_error(self._ctx).decode())

    def shape(self, ch: int):
        return (self.channels[ch].ny, self.channels[ch].nx)

    # -- state ------------------------------------------------------------
    def initialization(self, mode: int = COLD_NOISE, seed: int = 0) -> None:
        self._check(lib().ekpnp_initialization(self._ctx, mode, seed))

    def set_wall_profile(self, ch: int, psi_lo=None, psi_hi=None, cpos_lo=None, cpos_hi=None,
                         cneg_lo=None, cneg_hi=None) -> None:
        nx = self.channels[ch].nx
        arrs = [None if a is None else _as_f64(a, (nx,)) for a in
                (psi_lo, psi_hi, cpos_lo, cpos_hi, cneg_lo, cneg_hi)]
        self._check(lib().ekpnp_set_wall_profile(self._ctx, ch, *[_dptr(a) for a in arrs]))

    def set_fields(self, ch: int, rho, ux, uy, cpos, cneg, psi=None) -> None:
        shp = self.shape(ch)
        arrs = [_as_f64(a, shp) for a in (rho, ux, uy, cpos, cneg)]
        ps = None if psi is None else _as_f64(psi, shp)
        self._check(lib().ekpnp_set_fields(self._ctx, ch, *[_dptr(a) for a in arrs], _dptr(ps)))

    def init_equilibrium(self) -> None:
        self._check(lib().ekpnp_init_equilibrium(self._ctx))

    def restart_channel(self, ch: int, mode: int = COLD_NOISE, seed: int = 0) -> None:
        """Restart channel ch alone from its starting fields at equilibrium
        (ekpnp_restart_channel); the other channels keep their state."""
        self._check(lib().ekpnp_restart_channel(self._ctx, ch, mode, seed))

    # -- stepping ---------------------------------------------------------
    def stream_collide_save(self) -> None:
        self._check(lib().ekpnp_stream_collide_save(self._ctx))

    def fast_poisson(self) -> None:
        self._check(lib().ekpnp_fast_poisson(self._ctx))

    def poisson_iterative(self, channel: int = 0, omega: float = 1.0, tol: float = 1e-9,
                          max_iter: int = 10 ** 6, check_every: int = 1) -> tuple:
        """Red-black SOR solve of the 5-point Poisson system of one channel
        (poisson_phi, poisson.cu:35-130; its defaults: tol 1e-9, 1e6 sweeps,
        a convergence check every sweep).  Overwrites that channel's psi;
        returns (sweeps, largest |change| of the last checked sweep)."""
        it = ctypes.c_int(0)
        ch = ctypes.c_double(0.0)
        self._check(lib().ekpnp_poisson_iterative(self._ctx, int(channel), float(omega), float(tol), int(max_iter),
                                                  int(check_every), ctypes.byref(it), ctypes.byref(ch)))
        return it.value, ch.value

    def step(self, n: int = 1) -> None:
        self._check(lib().ekpnp_step(self._ctx, int(n)))

    def rebase(self) -> None:
        """Re-centre FP32 storage on the current row means (no-op for FP64)."""
        self._check(lib().ekpnp_rebase(self._ctx))

    def run_to_residual(self, tol: float = 1e-8, check_every: int = 100, max_steps: int = 10 ** 6):
        out = (ctypes.c_int * len(self.channels))()
        self._check(lib().ekpnp_run_to_residual(self._ctx, tol, check_every, max_steps, out))
        return np.array(out[:], dtype=np.int64)

    def synchronize(self) -> None:
        self._check(lib().ekpnp_synchronize(self._ctx))

    # -- output -----------------------------------------------------------
    def get_field(self, ch: int, name: str) -> np.ndarray:
        out = np.empty(self.shape(ch), dtype=np.float64)
        self._check(lib().ekpnp_get_field(self._ctx, ch, FIELDS[name], _dptr(out)))
        return out

    def get_populations(self, ch: int, species: str) -> np.ndarray:
        ny, nx = self.shape(ch)
        out = np.empty((9, ny, nx), dtype=np.float64)
        self._check(lib().ekpnp_get_populations(self._ctx, ch, SPECIES[species], _dptr(out)))
        return out

    def set_populations(self, ch: int, species: str, pops) -> None:
        ny, nx = self.shape(ch)
        arr = _as_f64(pops, (9, ny, nx))
        self._check(lib().ekpnp_set_populations(self._ctx, ch, SPECIES[species], _dptr(arr)))

    def reductions(self) -> np.ndarray:
        out = np.empty((len(self.channels), 8), dtype=np.float64)
        self._check(lib().ekpnp_reductions(self._ctx, _dptr(out)))
        return out

    def reductions_noalloc(self) -> np.ndarray:
        """reductions() through ekpnp_reductions_noalloc (no device
        allocation; the library's fixed scratch array)."""
        out = np.empty((len(self.channels), 8), dtype=np.float64)
        self._check(lib().ekpnp_reductions_noalloc(self._ctx, _dptr(out)))
        return out

    # -- files and diagnostics ----------------------------------------------
    def save_data_end(self, ch: int, time: float, path: str) -> None:
        self._check(lib().ekpnp_save_data_end(self._ctx, ch, float(time), os.fsencode(path)))

    def read_data(self, ch: int, path: str) -> float:
        t = ctypes.c_double(0.0)
        self._check(lib().ekpnp_read_data(self._ctx, ch, os.fsencode(path), ctypes.byref(t)))
        return t.value

    def save_tecplot(self, ch: int, time: float, path: str, first: bool) -> None:
        self._check(lib().ekpnp_save_tecplot(self._ctx, ch, float(time), os.fsencode(path), int(bool(first))))

    def save_charge_data(self, ch: int, path: str) -> None:
        self._check(lib().ekpnp_save_charge_data(self._ctx, ch, os.fsencode(path)))

    def current(self, ch: int) -> float:
        out = ctypes.c_double(0.0)
        self._check(lib().ekpnp_current(self._ctx, ch, ctypes.byref(out)))
        return out.value

    def record_umax(self, ch: int, time: float = 0.0, path: str = None) -> float:
        out = ctypes.c_double(0.0)
        p = None if path is None else os.fsencode(path)
        self._check(lib().ekpnp_record_umax(self._ctx, ch, float(time), p, ctypes.byref(out)))
        return out.value

    def save_checkpoint(self, path: str) -> None:
        self._check(lib().ekpnp_save_checkpoint(self._ctx, os.fsencode(path)))

    def load_checkpoint(self, path: str) -> None:
        self._check(lib().ekpnp_load_checkpoint(self._ctx, os.fsencode(path)))

    def set_poisson_plan(self, fft_elems: int = 2048, tri_rows: int = 16, persistent=True) -> None:
        """persistent: False/0 one-tile x-FFT kernels, True/1 persistent
        pipelined ones, 2 row-pair ones (channels permitting; otherwise the
        one-tile kernels)."""
        self._check(lib().ekpnp_set_poisson_plan(self._ctx, int(fft_elems), int(tri_rows), int(persistent)))

    def autotune(self, reps: int = 20) -> dict:
        out = np.zeros(4, dtype=np.float64)
        self._check(lib().ekpnp_autotune(self._ctx, int(reps), _dptr(out)))
        return {"fft_elems": int(out[0]), "tri_rows": int(out[1]), "poisson_ms": float(out[2]),
                "persistent": int(out[3]) == 1, "fft_kernels": FFT_KERNELS[int(out[3])]}

    def set_lbm_tiles(self, wrows: int) -> None:
        """Tile shape of the row collide-stream kernel: 0 four consecutive
        64-node blocks per work-group (default), 2 or 4 that many interior rows
        x 4 / wrows blocks.  Results are identical for every shape."""
        self._check(lib().ekpnp_set_lbm_tiles(self._ctx, int(wrows)))

    def lbm_tiles(self) -> int:
        out = ctypes.c_int(0)
        self._check(lib().ekpnp_get_lbm_tiles(self._ctx, ctypes.byref(out)))
        return int(out.value)

    def autotune_lbm(self, reps: int = 10) -> dict:
        """Times the row kernel's interior launch per tile shape on the current
        state (not advanced) and keeps the fastest (needs populations)."""
        out = np.zeros(2, dtype=np.float64)
        self._check(lib().ekpnp_autotune_lbm(self._ctx, int(reps), _dptr(out)))
        return {"wrows": int(out[0]), "ms": float(out[1])}

    def poisson_plan(self) -> dict:
        """The plan in force (ekpnp_get_poisson_plan)."""
        out = (ctypes.c_int * 4)()
        self._check(lib().ekpnp_get_poisson_plan(self._ctx, out))
        return {"fft_elems": out[0], "tri_rows": out[1], "persistent": out[2] == 1, "fft_kernels": FFT_KERNELS[out[2]],
                "walls_in_tri": bool(out[3])}

    def get_params(self, ch: int) -> ChannelParams:
        out = ChannelParams()
        self._check(lib().ekpnp_get_params(self._ctx, ch, ctypes.byref(out)))
        return out

    @property
    def steps_done(self) -> int:
        return int(lib().ekpnp_steps_done(self._ctx))

    @property
    def device_bytes(self) -> int:
        return int(lib().ekpnp_device_bytes(self._ctx))

    def time_stages(self, nsteps: int) -> dict:
        out = np.zeros(5, dtype=np.float64)
        self._check(lib().ekpnp_time_stages(self._ctx, int(nsteps), _dptr(out)))
        return {"lbm": out[0], "fft_fwd": out[1], "tridiag": out[2], "fft_inv": out[3], "step": out[4]}


__all__ = [
    "OK", "ERR_INVALID", "ERR_HIP", "ERR_NOMEM", "ERR_STATE", "ERR_UNSUPPORTED", "ERR_UNSTABLE", "ERR_IO",
    "ChannelParams", "Options", "Solver", "EkpnpError", "lib", "reference_params", "lattice_params",
    "FIELDS", "SPECIES", "RED_NAMES", "COLD_NOISE", "REFERENCE", "PB_PICARD", "POP_FP32", "POP_FP64",
    "FLAG_ION_DRIFT_APPLIED", "FLAG_G0_FLUID_U", "SIGNATURES", "LIB_PATH", "REBASE_INTERVAL", "rebase_steps", "copy_bandwidth", "check_params",
    "CHECK_DEBYE_UNRESOLVED", "CHECK_WALL_DRIFT", "CHECK_FLOW_FAST",
]
