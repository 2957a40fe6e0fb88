"""ORACLE (test infrastructure only): numpy restatement of the warm-start
U-Net that ``poissonrecon_gpu_amd/csrc/unet.hip`` trains and runs on the GPU.

Only tests/ use this module; the product never imports it.  The network has
no counterpart in the reference (it is the build's learned warm start): this
file is its specification, with the GPU's rounding points mirrored so the
GPU's forward pass and gradients can be compared tightly:

  - every 3x3x3 convolution rounds its input activations and its weights to
    bfloat16 (round to nearest even) and accumulates in float64 here (float32
    on the GPU's matrix cores);
  - the data gradient of a convolution rounds the incoming gradient and the
    weights to bfloat16; the weight gradient rounds the incoming gradient and
    the layer input to bfloat16;
  - everything else (bias, ReLU, pooling, upsampling, concatenation, the 1x1
    head, the loss) is exact float64 arithmetic of float32 data.

Layout: activations are (G, G, G, C) arrays, voxel (x, y, z) x-major,
channels last.  Convolution weights are (27, Cout, Cin) with tap
t = 9 (dx + 1) + 3 (dy + 1) + (dz + 1), zero padding outside the grid.

Architecture (4 levels, 16-32-64-64 channels): input channels are the
normalised divergence and the node mask, padded to 16.
  e0: conv 16->16, conv 16->16        (G)
  e1: pool, conv 16->32, conv 32->32  (G/2)
  e2: pool, conv 32->64, conv 64->64  (G/4)
  b:  pool, conv 64->64, conv 64->64  (G/8)
  d2: up, concat e2 (128), conv 128->64, conv 64->64
  d1: up, concat e1 (96),  conv 96->32,  conv 32->32
  d0: up, concat e0 (48),  conv 48->16,  conv 16->16
  head: 1x1 conv 16->1 (no ReLU); all 3x3x3 convolutions are followed by ReLU.
Loss: sum_v mask_v (pred_v - target_v)^2 / sum_v mask_v.
"""
from __future__ import annotations

import numpy as np

CIN0 = 16
CONVS = [("e0a", 16, 16), ("e0b", 16, 16), ("e1a", 16, 32), ("e1b", 32, 32), ("e2a", 32, 64), ("e2b", 64, 64),
         ("ba", 64, 64), ("bb", 64, 64), ("d2a", 128, 64), ("d2b", 64, 64), ("d1a", 96, 32), ("d1b", 32, 32),
         ("d0a", 48, 16), ("d0b", 16, 16)]
HEAD_C = 16


def param_layout():
    """[(name, shape)] in the flat parameter order shared with the GPU."""
    out = []
    for name, cin, cout in CONVS:
        out.append((name + ".w", (27, cout, cin)))
        out.append((name + ".b", (cout,)))
    out.append(("head.w", (HEAD_C,)))
    out.append(("head.b", (1,)))
    return out


def num_params():
    return sum(int(np.prod(s)) for _, s in param_layout())


def unflatten(flat):
    d, o = {}, 0
    for name, shape in param_layout():
        n = int(np.prod(shape))
        d[name] = np.asarray(flat[o:o + n], np.float32).reshape(shape)
        o += n
    return d


def flatten(d):
    return np.concatenate([d[name].reshape(-1).astype(np.float32) for name, _ in param_layout()])


# False: no bfloat16 rounding anywhere (the exact function whose gradient
# finite differences check); True: the GPU's rounding points.
ROUND = True


def bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    if not ROUND:
        return np.asarray(a, np.float64)
    u = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def _f32(a):
    return a.astype(np.float32) if ROUND else a


def _taps():
    return [(t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1) for t in range(27)]


def _shift(a, dx, dy, dz):
    """s[x, y, z] = a[x + dx, y + dy, z + dz], zero outside."""
    G = a.shape[0]
    s = np.zeros_like(a)
    xs = slice(max(0, -dx), G - max(0, dx))
    ys = slice(max(0, -dy), G - max(0, dy))
    zs = slice(max(0, -dz), G - max(0, dz))
    xd = slice(max(0, dx), G - max(0, -dx))
    yd = slice(max(0, dy), G - max(0, -dy))
    zd = slice(max(0, dz), G - max(0, -dz))
    s[xs, ys, zs] = a[xd, yd, zd]
    return s


def conv3(a, w):
    """out[v, co] = sum_t sum_ci bf16(a)[v + off_t, ci] * bf16(w)[t, co, ci] (float64)."""
    ab = bf16(a).astype(np.float64)
    wb = bf16(w).astype(np.float64)
    out = np.zeros(a.shape[:3] + (w.shape[1],), np.float64)
    for t, (dx, dy, dz) in enumerate(_taps()):
        out += _shift(ab, dx, dy, dz) @ wb[t].T
    return out


def conv3_backward(a, w, g):
    """(d input, d weight) of conv3 for incoming gradient g."""
    gb = bf16(g).astype(np.float64)
    wb = bf16(w).astype(np.float64)
    ab = bf16(a).astype(np.float64)
    da = np.zeros(a.shape, np.float64)
    dw = np.zeros(w.shape, np.float64)
    G = a.shape[0]
    for t, (dx, dy, dz) in enumerate(_taps()):
        da += _shift(gb @ wb[t], -dx, -dy, -dz)
        dw[t] = (gb.reshape(-1, gb.shape[3]).T @ _shift(ab, dx, dy, dz).reshape(-1, a.shape[3]))
    return da, dw


def pool(a):
    G = a.shape[0] // 2
    return a.reshape(G, 2, G, 2, G, 2, -1).mean(axis=(1, 3, 5))


def pool_backward(g):
    return np.repeat(np.repeat(np.repeat(g, 2, 0), 2, 1), 2, 2) / 8.0


def up(a):
    return np.repeat(np.repeat(np.repeat(a, 2, 0), 2, 1), 2, 2)


def up_backward(g):
    G = g.shape[0] // 2
    return g.reshape(G, 2, G, 2, G, 2, -1).sum(axis=(1, 3, 5))


def forward(params, x_in):
    """x_in: (G, G, G, 16) float32.  Returns (prediction (G, G, G), cache)."""
    p = params
    cache = {}

    def cr(name, a):
        z = conv3(a, p[name + ".w"]) + p[name + ".b"].astype(np.float64)
        y = _f32(np.maximum(z, 0.0))
        cache[name] = (a, y)
        return y

    e0 = cr("e0b", cr("e0a", x_in))
    e1 = cr("e1b", cr("e1a", _f32(pool(e0))))
    e2 = cr("e2b", cr("e2a", _f32(pool(e1))))
    bt = cr("bb", cr("ba", _f32(pool(e2))))
    d2 = cr("d2b", cr("d2a", np.concatenate([up(bt), e2], axis=3)))
    d1 = cr("d1b", cr("d1a", np.concatenate([up(d2), e1], axis=3)))
    d0 = cr("d0b", cr("d0a", np.concatenate([up(d1), e0], axis=3)))
    pred = _f32(d0.astype(np.float64) @ p["head.w"].astype(np.float64) + float(p["head.b"][0]))
    cache["head"] = d0
    return pred, cache


def loss_and_grad(params, x_in, target, mask):
    """Masked MSE and its gradient with respect to every parameter (float64)."""
    p = params
    pred, cache = forward(params, x_in)
    m = mask.astype(np.float64)
    msum = m.sum()
    diff = pred.astype(np.float64) - target.astype(np.float64)
    loss = float((m * diff * diff).sum() / msum)
    gpred = 2.0 * m * diff / msum
    grads = {}
    d0 = cache["head"].astype(np.float64)
    grads["head.w"] = np.tensordot(gpred, d0, axes=([0, 1, 2], [0, 1, 2]))
    grads["head.b"] = np.array([gpred.sum()])
    g = gpred[..., None] * p["head.w"].astype(np.float64)

    def back(name, g):
        a, y = cache[name]
        g = g * (y > 0)
        grads[name + ".b"] = g.sum(axis=(0, 1, 2))
        da, dw = conv3_backward(a, p[name + ".w"], _f32(g))
        grads[name + ".w"] = dw
        return da

    g = back("d0a", back("d0b", g))
    g_up, g_e0 = g[..., :32], g[..., 32:]
    g = back("d1a", back("d1b", up_backward(g_up)))
    g_up, g_e1 = g[..., :64], g[..., 64:]
    g = back("d2a", back("d2b", up_backward(g_up)))
    g_up, g_e2 = g[..., :64], g[..., 64:]
    g = back("ba", back("bb", up_backward(g_up)))
    g = back("e2a", back("e2b", pool_backward(g) + g_e2))
    g = back("e1a", back("e1b", pool_backward(g) + g_e1))
    back("e0a", back("e0b", pool_backward(g) + g_e0))
    return loss, grads, pred


def init_params(seed=0):
    """He-normal convolution weights, zero biases (the GPU's initialiser is its own;
    tests pass explicit parameters to both sides)."""
    rng = np.random.default_rng(seed)
    d = {}
    for name, cin, cout in CONVS:
        d[name + ".w"] = (rng.standard_normal((27, cout, cin)) * np.sqrt(2.0 / (27 * cin))).astype(np.float32)
        d[name + ".b"] = np.zeros(cout, np.float32)
    d["head.w"] = (rng.standard_normal(HEAD_C) * 0.25).astype(np.float32)
    d["head.b"] = np.zeros(1, np.float32)
    return d
