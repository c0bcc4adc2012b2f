"""FP64 CPU reference of the surrogate MLP (include/ekpnp_mlp.h).

TEST INFRASTRUCTURE ONLY (imported by tests/ as the checker of the GPU BF16
MFMA kernels).  With emulate=True the values are rounded to BF16 exactly where
the GPU rounds them (weights, inputs, hidden activations, back-propagated
gradients), while every sum runs in FP64; with emulate=False the whole network
is FP64, which is what the finite-difference checks pin.

Parameter vector layout: for each layer the weights [out][in_padded]
row-major, then the bias [out]; the first layer's inputs are padded to a
multiple of 32 with zero columns.
"""
import numpy as np
from scipy.special import erf


def bf16(a):
    """Round-to-nearest-even to BF16, returned as float64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def dgelu(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def shapes(n_in, n_hidden, n_layers, n_out):
    dims = [((n_in + 31) // 32) * 32] + [n_hidden] * n_layers + [n_out]
    return [(dims[i + 1], dims[i]) for i in range(len(dims) - 1)]


def unpack(p, shp):
    out, off = [], 0
    for o, i in shp:
        W = p[off:off + o * i].reshape(o, i)
        off += o * i
        b = p[off:off + o]
        off += o
        out.append((np.asarray(W, dtype=np.float64), np.asarray(b, dtype=np.float64)))
    assert off == p.size
    return out


def pack(grads):
    return np.concatenate([np.concatenate([gW.ravel(), gb.ravel()]) for gW, gb in grads])


def forward(p, x, shp, emulate=True):
    layers = unpack(np.asarray(p, dtype=np.float64), shp)
    r = bf16 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    B = x.shape[0]
    a = np.zeros((B, shp[0][1]))
    a[:, :x.shape[1]] = x
    a = r(a)
    acts, zs = [a], []
    for li, (W, b) in enumerate(layers):
        z = a @ r(W).T + b
        zs.append(z)
        if li < len(layers) - 1:
            a = r(gelu(z))
            acts.append(a)
    return zs[-1], (acts, zs, layers)


def loss_grad(p, x, t, shp, emulate=True):
    y, (acts, zs, layers) = forward(p, x, shp, emulate)
    r = bf16 if emulate else (lambda a: a)
    d = y - t
    n = d.size
    loss = float(np.mean(d * d))
    g = 2.0 * d / n
    grads = []
    for li in range(len(layers) - 1, -1, -1):
        W, _ = layers[li]
        gr = r(g)
        gW = gr.T @ acts[li]
        gb = g.sum(axis=0)
        grads.append((gW, gb))
        if li > 0:
            g = (gr @ r(W)) * dgelu(zs[li - 1])
    return loss, pack(grads[::-1])


def init_params(shp, n_in, seed=0):
    rng = np.random.default_rng(seed)
    parts = []
    for li, (o, i) in enumerate(shp):
        fan = n_in if li == 0 else i
        W = rng.standard_normal((o, i)) * np.sqrt(2.0 / fan)
        if li == 0:
            W[:, n_in:] = 0.0
        parts.append(W.ravel())
        parts.append(np.zeros(o))
    return np.concatenate(parts)
