"""Warm-start surrogate: channel parameters -> wall-normal steady profiles.

The network (include/ekpnp_mlp.h) maps a channel's (zeta, kappa h, E_x,
n_y), presented as the six derived features of features(), to psi, c+, c-
and u_x at N_PTS points spread uniformly over the wall-normal coordinate y
in [0, ny-1].  It is trained and evaluated by the library's BF16 MFMA
kernels; this module only prepares features/targets, resamples profiles,
supplies the physics baseline and drives training.

Targets are normalised so that every output is O(1) across the parameter
range:
    psi / zeta,   log(c+) / (-zeta),   log(c-) / zeta,
    u_x nu (poisson_coef / force_coef) / (E_x (-zeta))
(for Boltzmann ions and a Helmholtz-Smoluchowski flow all four are close to
psi/zeta shifted to vanish at the wall for u_x).  The reference has no
counterpart; its only warm start is the PB-Picard iteration (LBM.cu:106-127).
"""
from __future__ import annotations

import ctypes
from typing import Sequence

import numpy as np
from scipy.linalg import solve_banded

from . import ChannelParams, EkpnpError, lib

N_PTS = 256
N_FIELDS = 4          # psi, c+, c-, u_x
N_OUT = N_PTS * N_FIELDS
N_IN = 6              # from (zeta, kappa h, E_x, n_y); see features()


class MLPConfig(ctypes.Structure):
    _fields_ = [("n_in", ctypes.c_int32), ("n_hidden", ctypes.c_int32), ("n_layers", ctypes.c_int32),
                ("n_out", ctypes.c_int32), ("max_batch", ctypes.c_int32), ("device", ctypes.c_int32),
                ("seed", ctypes.c_uint64)]


_VP = ctypes.c_void_p
_FP = ctypes.POINTER(ctypes.c_float)
MLP_SIGNATURES = [
    ("ekpnp_mlp_create", ctypes.c_int, [ctypes.POINTER(MLPConfig), ctypes.POINTER(_VP)]),
    ("ekpnp_mlp_destroy", None, [_VP]),
    ("ekpnp_mlp_last_error", ctypes.c_char_p, [_VP]),
    ("ekpnp_mlp_num_params", ctypes.c_longlong, [_VP]),
    ("ekpnp_mlp_get_params", ctypes.c_int, [_VP, _FP]),
    ("ekpnp_mlp_set_params", ctypes.c_int, [_VP, _FP]),
    ("ekpnp_mlp_forward", ctypes.c_int, [_VP, _FP, ctypes.c_int, _FP]),
    ("ekpnp_mlp_loss_grad", ctypes.c_int, [_VP, _FP, _FP, ctypes.c_int, _FP, _FP]),
    ("ekpnp_mlp_train_step", ctypes.c_int, [_VP, _FP, _FP, ctypes.c_int, ctypes.c_float, _FP]),
    ("ekpnp_mlp_set_adam", ctypes.c_int, [_VP, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float]),
]

_bound = False


def _lib():
    global _bound
    h = lib()
    if not _bound:
        for name, res, args in MLP_SIGNATURES:
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _bound = True
    return h


def _f32(a, shape=None):
    arr = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None:
        arr = arr.reshape(shape)
    return arr


def _fp(a):
    return a.ctypes.data_as(_FP)


def _batch(a, width: int, what: str) -> np.ndarray:
    """float32 [B, width] copy of a; raises on any other shape (the C side
    reads exactly B x width values)."""
    arr = np.ascontiguousarray(a, dtype=np.float32)
    if arr.ndim != 2 or arr.shape[1] != width:
        raise ValueError(f"{what} must have shape [batch, {width}], got {arr.shape}")
    return arr


class MLP:
    """GPU multilayer perceptron (BF16 MFMA, FP32 accumulation, Adam)."""

    def __init__(self, n_in=N_IN, n_hidden=256, n_layers=4, n_out=N_OUT, max_batch=512, seed=0, device=0):
        L = _lib()
        self.cfg = MLPConfig(n_in, n_hidden, n_layers, n_out, max_batch, device, seed)
        h = _VP()
        rc = L.ekpnp_mlp_create(ctypes.byref(self.cfg), ctypes.byref(h))
        if rc != 0:
            raise EkpnpError(rc, "ekpnp_mlp_create failed")
        self._h = h
        self.n_params = int(L.ekpnp_mlp_num_params(h))

    def close(self):
        if getattr(self, "_h", None):
            _lib().ekpnp_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise EkpnpError(rc, _lib().ekpnp_mlp_last_error(self._h).decode())

    def layer_shapes(self):
        """[(out, in_padded), ...] per layer, in parameter-vector order."""
        c = self.cfg
        dims = [((c.n_in + 31) // 32) * 32] + [c.n_hidden] * c.n_layers + [c.n_out]
        return [(dims[i + 1], dims[i]) for i in range(len(dims) - 1)]

    def get_params(self) -> np.ndarray:
        out = np.empty(self.n_params, dtype=np.float32)
        self._check(_lib().ekpnp_mlp_get_params(self._h, _fp(out)))
        return out

    def set_params(self, p) -> None:
        arr = _f32(p, (self.n_params,))
        self._check(_lib().ekpnp_mlp_set_params(self._h, _fp(arr)))

    def forward(self, x) -> np.ndarray:
        x = _batch(x, self.cfg.n_in, "x")
        B = x.shape[0]
        y = np.empty((B, self.cfg.n_out), dtype=np.float32)
        for s in range(0, B, self.cfg.max_batch):
            xb = np.ascontiguousarray(x[s:s + self.cfg.max_batch])
            yb = np.empty((xb.shape[0], self.cfg.n_out), dtype=np.float32)
            self._check(_lib().ekpnp_mlp_forward(self._h, _fp(xb), xb.shape[0], _fp(yb)))
            y[s:s + xb.shape[0]] = yb
        return y

    def loss_grad(self, x, t):
        x, t = _batch(x, self.cfg.n_in, "x"), _batch(t, self.cfg.n_out, "t")
        if x.shape[0] != t.shape[0] or x.shape[0] > self.cfg.max_batch:
            raise ValueError("x and t need the same batch size, at most max_batch")
        loss = ctypes.c_float()
        g = np.empty(self.n_params, dtype=np.float32)
        self._check(_lib().ekpnp_mlp_loss_grad(self._h, _fp(x), _fp(t), x.shape[0], ctypes.byref(loss), _fp(g)))
        return float(loss.value), g

    def train_step(self, x, t, lr: float) -> float:
        x, t = _batch(x, self.cfg.n_in, "x"), _batch(t, self.cfg.n_out, "t")
        if x.shape[0] != t.shape[0] or x.shape[0] > self.cfg.max_batch:
            raise ValueError("x and t need the same batch size, at most max_batch")
        loss = ctypes.c_float()
        self._check(_lib().ekpnp_mlp_train_step(self._h, _fp(x), _fp(t), x.shape[0], lr, ctypes.byref(loss)))
        return float(loss.value)

    def set_adam(self, beta1=0.9, beta2=0.999, eps=1e-8, weight_decay=0.0):
        self._check(_lib().ekpnp_mlp_set_adam(self._h, beta1, beta2, eps, weight_decay))


# ---------------------------------------------------------------------------
# features, targets, resampling
# ---------------------------------------------------------------------------
def features(zeta, kappa_h, e_x, ny) -> np.ndarray:
    """[n, 6] network inputs, each roughly in [-1, 1], all functions of
    (zeta, kappa h, E_x, n_y): zeta, log kappa h, log E_x, log n_y, the Debye
    resolution kappa dy = kappa h / ((n_y - 1) / 2) (the steady state's
    departure from Poisson-Boltzmann is a wall-discretisation effect), and
    the Gouy-Chapman charge scale sinh(|zeta| / 2)."""
    zeta, kappa_h, e_x, ny = (np.atleast_1d(np.asarray(v, dtype=np.float64)) for v in (zeta, kappa_h, e_x, ny))
    kdy = kappa_h / (0.5 * (ny - 1.0))
    return np.stack([zeta / 2.0 + 1.0,
                     np.log(kappa_h) / np.log(50.0) * 2.0 - 1.0,
                     (np.log10(e_x) + 4.0) / 1.0,
                     np.log2(ny / 128.0) / 2.0,
                     kdy / 0.2 - 1.0,
                     np.sinh(0.5 * np.abs(zeta)) / np.sinh(2.0) * 2.0 - 1.0], axis=1).astype(np.float32)


def resample(profile: np.ndarray, n: int) -> np.ndarray:
    """Linear resampling of a profile on nodes 0..len-1 to n points on the same span."""
    src = np.linspace(0.0, 1.0, profile.shape[-1])
    dst = np.linspace(0.0, 1.0, n)
    return np.interp(dst, src, profile)


def hs_scale(p: ChannelParams) -> float:
    """u_x scale of a channel: E_x (-zeta) force_coef / (poisson_coef rho0 nu)."""
    return p.e_applied * (-p.psi_lo) * (p.force_coef / p.poisson_coef) / (p.rho0 * p.nu)


def _normalized(p: ChannelParams, psi, cpos, cneg, ux):
    """The four profiles on the channel's own rows, each O(1)."""
    z = p.psi_lo
    return [np.asarray(psi, dtype=np.float64) / z,
            np.log(np.maximum(cpos, 1e-300)) / (-z),
            np.log(np.maximum(cneg, 1e-300)) / z,
            np.asarray(ux, dtype=np.float64) / hs_scale(p)]


def _denormalized(p: ChannelParams, parts):
    z = p.psi_lo
    psi = parts[0] * z
    cpos = np.exp(parts[1] * (-z))
    cneg = np.exp(parts[2] * z)
    ux = parts[3] * hs_scale(p)
    # walls exactly as the boundary conditions hold them
    psi[0], psi[-1] = p.psi_lo, p.psi_hi
    cpos[0], cpos[-1] = p.cpos_lo, p.cpos_hi
    cneg[0], cneg[-1] = p.cneg_lo, p.cneg_hi
    return psi, cpos, cneg, ux


def _split(y_norm):
    return [np.asarray(y_norm[i * N_PTS:(i + 1) * N_PTS], dtype=np.float64) for i in range(N_FIELDS)]


def targets(p: ChannelParams, psi, cpos, cneg, ux) -> np.ndarray:
    """Normalised 1024-vector from x-averaged wall-normal profiles (length ny)."""
    return np.concatenate([resample(a, N_PTS) for a in _normalized(p, psi, cpos, cneg, ux)]).astype(np.float32)


def profiles(p: ChannelParams, y_norm: np.ndarray):
    """Inverse of targets(): (psi, c+, c-, u_x) on the channel's ny rows."""
    return _denormalized(p, [resample(a, p.ny) for a in _split(y_norm)])


# ---------------------------------------------------------------------------
# physics baseline: Poisson-Boltzmann on the channel's own nodes plus the
# Helmholtz-Smoluchowski velocity it implies
# ---------------------------------------------------------------------------
def pb_profile(p: ChannelParams, tol: float = 1e-13, max_iter: int = 200) -> np.ndarray:
    """psi on nodes 0..ny-1 solving psi[i-1] - 2 psi[i] + psi[i+1] =
    -dy^2 poisson_coef (c+ - c-), c+- = c_bulk exp(-+ psi inv_vt), with the
    wall potentials as Dirichlet values (damped Newton, tridiagonal)."""
    ny = p.ny
    s2 = p.dy * p.dy * p.poisson_coef * p.c_bulk
    ivt = p.inv_thermal_voltage
    psi = np.linspace(p.psi_lo, p.psi_hi, ny)
    psi[1:-1] *= 0.5
    for _ in range(max_iter):
        a = psi[1:-1]
        f = psi[:-2] - 2.0 * a + psi[2:] - 2.0 * s2 * np.sinh(a * ivt)
        diag = -2.0 - 2.0 * s2 * ivt * np.cosh(a * ivt)
        step = _thomas(np.ones_like(a), diag, np.ones_like(a), -f)
        lim = np.max(np.abs(step))
        damp = 1.0 if lim <= 1.0 else 1.0 / lim
        psi[1:-1] = a + damp * step
        if lim < tol * max(1.0, np.max(np.abs(psi))):
            break
    return psi


def _thomas(lo, diag, up, rhs):
    """Tridiagonal solve (LAPACK's banded solver): lo[i] x[i-1] + diag[i] x[i]
    + up[i] x[i+1] = rhs[i]; lo[0] and up[-1] are not used."""
    ab = np.zeros((3, diag.size))
    ab[0, 1:] = up[:-1]
    ab[1] = diag
    ab[2, :-1] = lo[1:]
    return solve_banded((1, 1), ab, rhs)


def baseline(p: ChannelParams):
    """(psi, c+, c-, u_x) on the channel's ny rows: Poisson-Boltzmann potential,
    Boltzmann ions, and u_x = hs_scale (psi - psi_w) / (-zeta) with the no-slip
    plane half a node inside each wall (full-way bounce-back walls)."""
    psi = pb_profile(p)
    ivt = p.inv_thermal_voltage
    cpos = p.c_bulk * np.exp(-psi * ivt)
    cneg = p.c_bulk * np.exp(psi * ivt)
    cpos[0], cpos[-1] = p.cpos_lo, p.cpos_hi
    cneg[0], cneg[-1] = p.cneg_lo, p.cneg_hi
    lo = 0.5 * (psi[0] + psi[1])
    hi = 0.5 * (psi[-1] + psi[-2])
    y = np.arange(p.ny, dtype=np.float64)
    wall = lo + (hi - lo) * np.clip((y - 0.5) / max(p.ny - 2.0, 1.0), 0.0, 1.0)
    ux = hs_scale(p) * (psi - wall) / (-p.psi_lo)
    return psi, cpos, cneg, ux


def osmotic_rho(p: ChannelParams, psi, cpos, cneg) -> np.ndarray:
    """Density profile of the steady x-uniform channel implied by psi and
    c+-: the wall-normal momentum balance cs^2 d(rho)/dy = F_y with the
    body force F_y = force_coef (c+ - c-) E_y, E_y = -dpsi/dy by central
    differences, integrated with the trapezoidal rule over the interior rows
    and shifted to mean rho0 there (the mass is conserved; the bounce-back
    wall rows, whose density the balance does not describe, are set to
    rho0)."""
    psi = np.asarray(psi, dtype=np.float64)
    ey = (psi[:-2] - psi[2:]) / (2.0 * p.dy)                     # rows 1 .. ny-2
    fy = p.force_coef * (np.asarray(cpos)[1:-1] - np.asarray(cneg)[1:-1]) * ey
    rho_i = np.concatenate([[0.0], np.cumsum(0.5 * (fy[1:] + fy[:-1]) * p.dy / p.cs2)])
    rho = np.full(psi.shape, p.rho0)
    rho[1:-1] = rho_i - rho_i.mean() + p.rho0
    return rho


def residual_targets(p: ChannelParams, psi, cpos, cneg, ux) -> np.ndarray:
    """Normalised profiles of a steady state minus those of the physics
    baseline, subtracted on the channel's own rows and then resampled (only
    the smooth difference is interpolated)."""
    have = _normalized(p, psi, cpos, cneg, ux)
    base = _normalized(p, *baseline(p))
    return np.concatenate([resample(a - b, N_PTS) for a, b in zip(have, base)]).astype(np.float32)


def residual_profiles(p: ChannelParams, y_norm: np.ndarray):
    """Inverse of residual_targets(): baseline on the rows plus the resampled
    correction."""
    base = _normalized(p, *baseline(p))
    return _denormalized(p, [b + resample(a, p.ny) for a, b in zip(_split(y_norm), base)])


def channel_of(zeta, kappa_h, e_x, nx, ny, tau=0.8, diff=0.05) -> ChannelParams:
    from . import lattice_params
    return lattice_params(int(nx), int(ny), tau=tau, zeta=float(zeta), kappa_h=float(kappa_h), e_x=float(e_x),
                          diff=diff)


def stable(p: ChannelParams) -> bool:
    """Acceptance rule for sampled channels, the library's ekpnp_check_params:
    the Debye length is resolved (kappa dy <= 0.4) and the Gouy-Chapman wall
    drift of the ions, K * 2 kappa sinh(|zeta|/2) in lattice units, stays far
    below the lattice sound speed (<= 0.05)."""
    from . import check_params
    return bool(check_params([p])[0] == 0)


def sample_channels(n: int, seed: int, nx_choices=(32, 64, 128, 256), ny_choices=None,
                    zeta_range=(-4.0, -0.4), kappa_h_range=(1.0, 50.0), e_x_range=(1e-5, 1e-3)):
    """Batched-sweep channels (nx, ny, zeta, kappa h, E_x from the ranges
    below) drawn from a seeded stream; draws that fail
    stable() are redrawn.  Returns (list of ChannelParams, number rejected)."""
    if ny_choices is None:
        ny_choices = tuple(range(64, 1025, 64))
    rng = np.random.default_rng(seed)
    out, rejected = [], 0
    while len(out) < n:
        nx = int(rng.choice(nx_choices))
        ny = int(rng.choice(ny_choices))
        zeta = float(rng.uniform(*zeta_range))
        kh = float(np.exp(rng.uniform(np.log(kappa_h_range[0]), np.log(kappa_h_range[1]))))
        ex = float(np.exp(rng.uniform(np.log(e_x_range[0]), np.log(e_x_range[1]))))
        p = channel_of(zeta, kh, ex, nx, ny)
        if stable(p):
            out.append(p)
        else:
            rejected += 1
    return out, rejected


def save_net(path: str, net: "MLP", **meta) -> None:
    """Write a trained network's FP32 master parameters and shape (numpy .npz,
    no pickled objects) with optional metadata (strings and numbers)."""
    c = net.cfg
    np.savez(path, params=net.get_params(), shape=np.array([c.n_in, c.n_hidden, c.n_layers, c.n_out], np.int64),
             **{k: np.asarray(v) for k, v in meta.items()})


def load_net(path: str, max_batch: int = 512, device: int = 0) -> "MLP":
    """A network written by save_net, on the GPU (numpy load without pickle)."""
    with np.load(path, allow_pickle=False) as z:
        n_in, n_hidden, n_layers, n_out = (int(v) for v in z["shape"])
        params = z["params"]
    net = MLP(n_in=n_in, n_hidden=n_hidden, n_layers=n_layers, n_out=n_out, max_batch=max_batch, device=device)
    net.set_params(params)
    return net


def net_meta(path: str) -> dict:
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k].tolist() for k in z.files if k not in ("params",)}


def channel_features(ps: Sequence[ChannelParams]) -> np.ndarray:
    kh = [np.sqrt(2.0 * p.poisson_coef) * 0.5 * (p.ny - 1) for p in ps]
    return features([p.psi_lo for p in ps], kh, [p.e_applied for p in ps], [p.ny for p in ps])
