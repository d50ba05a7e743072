"""CPU restatement of the convolutional preconditioner (test infrastructure:
imported only by tests/ and bench.py's CPU baseline, never by the product).

The reference has no learned component (it factorises the sparse system,
ceviche/solvers.py:37-62), so this oracle follows the network definition of
ceviche_amd/csrc/cnn.hip in float64 numpy:

  activations x[b, i, j, c] (NHWC), weights w[tap, ci, co] with tap = 3 dy + dx,
  conv3x3(x)[i, j, co] = sum_{dy, dx, ci} xpad[i + dy, j + dx, ci] w[3 dy + dx, ci, co]

with zero padding of one cell, and the U-Net of cv_unet_apply.  Pinned by
tests/test_oracle_cnn.py against scipy.ndimage.correlate (forward), the
adjoint identities <conv(x), g> = <x, conv^T(g)> and <dW, V> = <conv_V(x), g>
(backward), and central finite differences of the whole network (float64).
`quant` arguments apply a rounding (bf16_round) wherever the GPU stores an
activation, so GPU results can be compared at their own storage precision.
"""
import numpy as np

TAPS = [(t // 3, t % 3) for t in range(9)]


def bf16_round(x):
    """Round to bfloat16 (nearest even) and return float64."""
    u = np.ascontiguousarray(np.asarray(x, dtype=np.float32)).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _pad(x):
    return np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))


def conv3x3(x, w, b=None, relu=False):
    B, H, W, _ = x.shape
    xp = _pad(x)
    out = np.zeros((B, H, W, w.shape[2]))
    for t, (dy, dx) in enumerate(TAPS):
        out += xp[:, dy:dy + H, dx:dx + W, :] @ w[t]
    if b is not None:
        out = out + b
    return np.maximum(out, 0.0) if relu else out


def conv3x3_bwd_data(g, w):
    """Adjoint of conv3x3 in its input: x-gradient for output cotangent g."""
    B, H, W, _ = g.shape
    gp = np.zeros((B, H + 2, W + 2, w.shape[1]))
    for t, (dy, dx) in enumerate(TAPS):
        gp[:, dy:dy + H, dx:dx + W, :] += g @ w[t].T
    return gp[:, 1:-1, 1:-1, :]


def conv3x3_bwd_weight(x, g):
    """(dW, db) for output cotangent g."""
    B, H, W, Ci = x.shape
    Co = g.shape[3]
    xp = _pad(x)
    g2 = g.reshape(-1, Co)
    dw = np.stack([xp[:, dy:dy + H, dx:dx + W, :].reshape(-1, Ci).T @ g2 for dy, dx in TAPS])
    return dw, g2.sum(axis=0)


def down2(x, scale):
    """scale * sum of each 2x2 block; odd sizes are padded with zeros."""
    B, H, W, C = x.shape
    Hc, Wc = (H + 1) // 2, (W + 1) // 2
    xp = np.zeros((B, 2 * Hc, 2 * Wc, C))
    xp[:, :H, :W, :] = x
    return scale * xp.reshape(B, Hc, 2, Wc, 2, C).sum(axis=(2, 4))


def up2(x, H, W, scale, add=None):
    """scale * nearest-neighbour upsampling to (H, W), plus add."""
    out = scale * np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)[:, :H, :W, :]
    return out if add is None else out + add


def level_dims(H, W, levels):
    dims = [(H, W)]
    for _ in range(1, levels):
        h, w = dims[-1]
        dims.append(((h + 1) // 2, (w + 1) // 2))
    return dims


def unet_forward(x0, Ws, bs, levels, quant=None):
    """Forward pass of cv_unet_apply; returns (y, cache)."""
    q = quant or (lambda a: a)
    L = levels
    dims = level_dims(x0.shape[1], x0.shape[2], L)
    c = {'x0': x0}
    e = [None] * L
    h = [None] * L
    u = [None] * L
    bb = [None] * L
    e[0] = q(conv3x3(x0, Ws[0], bs[0], relu=True))
    for l in range(L - 1):
        h[l] = q(conv3x3(e[l], Ws[1 + l], bs[1 + l], relu=True))
        e[l + 1] = q(down2(h[l], 0.25))
    bb[L - 1] = q(conv3x3(e[L - 1], Ws[L], bs[L], relu=True))
    for l in range(L - 2, -1, -1):
        u[l] = q(up2(bb[l + 1], dims[l][0], dims[l][1], 1.0, h[l]))
        layer = L + 1 + (L - 2 - l)
        bb[l] = q(conv3x3(u[l], Ws[layer], bs[layer], relu=True))
    y = q(conv3x3(bb[0], Ws[2 * L], bs[2 * L]))
    c.update(e=e, h=h, u=u, b=bb, dims=dims)
    return y, c


def unet_backward(c, gy, Ws, levels):
    """(dWs, dbs) of <y, gy> for the cache of unet_forward (float64)."""
    L = levels
    dims = c['dims']
    e, h, u, bb = c['e'], c['h'], c['u'], c['b']
    dW = [None] * (2 * L + 1)
    db = [None] * (2 * L + 1)
    dW[2 * L], db[2 * L] = conv3x3_bwd_weight(bb[0], gy)
    g_b = conv3x3_bwd_data(gy, Ws[2 * L])          # cotangent of b_0
    g_u = [None] * L
    for l in range(0, L - 1):
        layer = L + 1 + (L - 2 - l)
        gp = g_b * (bb[l] > 0)
        dW[layer], db[layer] = conv3x3_bwd_weight(u[l], gp)
        g_u[l] = conv3x3_bwd_data(gp, Ws[layer])
        g_b = down2(g_u[l], 1.0)                     # cotangent of b_{l+1}
    gp = g_b * (bb[L - 1] > 0)
    dW[L], db[L] = conv3x3_bwd_weight(e[L - 1], gp)
    g_e = conv3x3_bwd_data(gp, Ws[L])               # cotangent of e_{L-1}
    for l in range(L - 2, -1, -1):
        g_h = up2(g_e, dims[l][0], dims[l][1], 0.25, g_u[l])
        gp = g_h * (h[l] > 0)
        dW[1 + l], db[1 + l] = conv3x3_bwd_weight(e[l], gp)
        g_e = conv3x3_bwd_data(gp, Ws[1 + l])
    gp = g_e * (e[0] > 0)
    dW[0], db[0] = conv3x3_bwd_weight(c['x0'], gp)
    return dW, db


def pack_input(r, rs, feat):
    """(H, W) complex field and (H, W, nf) features -> (1, H, W, 32)."""
    H, W = r.shape
    x = np.zeros((1, H, W, 32))
    x[0, :, :, 0] = rs * r.real
    x[0, :, :, 1] = rs * r.imag
    if feat is not None:
        x[0, :, :, 2:2 + feat.shape[2]] = feat
    return x


def unet_apply(r, rs, os_, feat, Ws, bs, levels, quant=None):
    """os * net(rs * r) as cv_unet_apply (complex (H, W) field)."""
    x0 = pack_input(r, rs, feat)
    if quant is not None:
        x0 = quant(x0)
    y, _ = unet_forward(x0, Ws, bs, levels, quant)
    return os_ * (y[0, :, :, 0] + 1j * y[0, :, :, 1])
