"""Pins the CPU oracle of the convolutional preconditioner (oracle/cnn_ref.py)
with checks that do not restate its formulas: scipy.ndimage.correlate for the
forward convolution, adjoint identities for both backward passes and for the
resampling pair, torch's bfloat16 conversion for the rounding, and central
finite differences of the whole network in float64."""
import numpy as np
import pytest
import torch
from scipy.ndimage import correlate

from oracle import cnn_ref as K


def _rand(rng, *shape):
    return rng.standard_normal(shape)


def test_conv_matches_scipy_correlate():
    rng = np.random.default_rng(0)
    x = _rand(rng, 2, 7, 9, 3)
    w = _rand(rng, 9, 3, 4)
    b = _rand(rng, 4)
    y = K.conv3x3(x, w, b)
    ref = np.zeros_like(y)
    for bi in range(2):
        for co in range(4):
            acc = np.full((7, 9), b[co])
            for ci in range(3):
                acc += correlate(x[bi, :, :, ci], w[:, ci, co].reshape(3, 3), mode='constant', cval=0.0)
            ref[bi, :, :, co] = acc
    assert np.allclose(y, ref, rtol=1e-13, atol=1e-13)
    assert np.allclose(K.conv3x3(x, w, b, relu=True), np.maximum(ref, 0), rtol=1e-13, atol=1e-13)


def test_conv_backward_adjoint_identities():
    rng = np.random.default_rng(1)
    x = _rand(rng, 2, 6, 5, 4)
    w = _rand(rng, 9, 4, 3)
    g = _rand(rng, 2, 6, 5, 3)
    # <conv(x), g> = <x, conv^T(g)>
    lhs = np.sum(K.conv3x3(x, w) * g)
    rhs = np.sum(x * K.conv3x3_bwd_data(g, w))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    # d/dW <conv_W(x) + b, g> along V equals <dW, V>; along db: sum(g) . v
    V = _rand(rng, 9, 4, 3)
    v = _rand(rng, 3)
    dw, db = K.conv3x3_bwd_weight(x, g)
    lhs = np.sum(K.conv3x3(x, V, v) * g)
    assert abs(lhs - (np.sum(dw * V) + np.sum(db * v))) <= 1e-12 * abs(lhs)


@pytest.mark.parametrize('H,W', [(8, 6), (7, 5), (1, 3)])
def test_down_up_adjoint(H, W):
    rng = np.random.default_rng(2)
    x = _rand(rng, 2, H, W, 3)
    yc = _rand(rng, 2, (H + 1) // 2, (W + 1) // 2, 3)
    lhs = np.sum(K.down2(x, 0.25) * yc)
    rhs = np.sum(x * K.up2(yc, H, W, 0.25))
    assert abs(lhs - rhs) <= 1e-12 * max(1.0, abs(lhs))
    # a constant field stays constant through down2 / 4 where blocks are full
    c = np.ones((1, 8, 8, 1))
    assert np.allclose(K.down2(c, 0.25), 1.0)


def test_bf16_round_matches_torch():
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.standard_normal(4000) * 10.0 ** rng.integers(-6, 6, 4000),
                        [0.0, -0.0, 1.0, 1.00390625, 1.005859375, 65504.0]])
    ref = torch.from_numpy(x.astype(np.float32)).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(K.bf16_round(x), ref)


def test_unet_backward_matches_finite_differences():
    rng = np.random.default_rng(4)
    L = 3
    n_layers = 2 * L + 1
    Ws = [0.4 * _rand(rng, 9, 4, 4) for _ in range(n_layers)]
    bs = [0.1 * _rand(rng, 4) for _ in range(n_layers)]
    x0 = _rand(rng, 1, 10, 9, 4)
    G = _rand(rng, 1, 10, 9, 4)

    def f(Ws_, bs_):
        y, _ = K.unet_forward(x0, Ws_, bs_, L)
        return np.sum(y * G)

    y, c = K.unet_forward(x0, Ws, bs, L)
    dW, db = K.unet_backward(c, G, Ws, L)
    h = 1e-6
    for layer in range(n_layers):
        V = _rand(rng, *Ws[layer].shape)
        Wp = [w.copy() for w in Ws]
        Wm = [w.copy() for w in Ws]
        Wp[layer] += h * V
        Wm[layer] -= h * V
        fd = (f(Wp, bs) - f(Wm, bs)) / (2 * h)
        an = np.sum(dW[layer] * V)
        assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)), (layer, fd, an)
        v = _rand(rng, 4)
        bp = [b.copy() for b in bs]
        bm = [b.copy() for b in bs]
        bp[layer] += h * v
        bm[layer] -= h * v
        fd = (f(Ws, bp) - f(Ws, bm)) / (2 * h)
        an = np.sum(db[layer] * v)
        assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)), (layer, fd, an)
