"""Pins of the FP64 oracle of the learned-RCMC network: finite differences of its loss, the closed form of a
bias-only network, and the bfloat16 rounding used for the device weights."""
import numpy as np

import pyoracle as po


def random_pairs(n, seed):
    rng = np.random.default_rng(seed)
    p = rng.standard_normal((n, 20)).astype(np.float32)
    p[:, 0] = rng.uniform(-1, 1, n)
    p[:, 1] = rng.uniform(-1, 1, n)
    return p


def random_params(seed, scale=0.3):
    rng = np.random.default_rng(seed)
    return rng.uniform(-scale, scale, po.lib().or_mlp_num_params())


def test_num_params():
    assert po.lib().or_mlp_num_params() == 2 * 64 + 64 + 64 * 64 + 64 + 16 * 64 + 16


def test_gradient_matches_central_differences():
    w = random_params(1)
    pairs = random_pairs(5, 2)
    _, g = po.mlp_loss_grad(w, pairs)
    rng = np.random.default_rng(3)
    # a sample of parameters from every block (W1, b1, W2, b2, W3, b3)
    idx = np.concatenate([rng.choice(np.arange(a, b), 6, replace=False) for a, b in
                          [(0, 128), (128, 192), (192, 4288), (4288, 4352), (4352, 5376), (5376, 5392)]])
    h = 1e-6
    for i in idx:
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        fd = (po.mlp_loss_grad(wp, pairs, False)[0] - po.mlp_loss_grad(wm, pairs, False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_bias_only_network_closed_form():
    # only output biases: taps are constant, prediction = sum_j (b_re_j + i b_im_j) s_j
    w = np.zeros(po.lib().or_mlp_num_params())
    rng = np.random.default_rng(4)
    taps = rng.standard_normal(16)
    w[5376:] = taps
    pairs = random_pairs(7, 5)
    loss, g = po.mlp_loss_grad(w, pairs)
    s = pairs[:, 2:18:2].astype(np.float64) + 1j * pairs[:, 3:18:2]
    t = pairs[:, 18].astype(np.float64) + 1j * pairs[:, 19]
    c = taps[0::2] + 1j * taps[1::2]
    base = np.stack([po.sinc8_weights(0.5 * float(u)) for u in pairs[:, 0]])
    e = np.sum(s * (c[None, :] + base), 1) - t
    assert abs(loss - np.sum(np.abs(e) ** 2)) < 1e-9 * loss
    # d loss / d b_re_j = sum 2 Re(conj(e) s_j); d / d b_im_j = sum 2 Re(conj(e) i s_j)
    gre = 2 * np.real(np.conj(e)[:, None] * s).sum(0)
    gim = 2 * np.real(np.conj(e)[:, None] * 1j * s).sum(0)
    assert np.allclose(g[5376::2], gre, rtol=1e-10, atol=1e-10)
    assert np.allclose(g[5377::2], gim, rtol=1e-10, atol=1e-10)
    # hidden-layer gradients vanish when W3 = 0
    assert np.all(g[:5376][np.r_[0:4352]] == 0)


def test_forward_matches_loss_of_constant_taps():
    w = random_params(6)
    out = po.mlp_forward(w, 0.25, -0.125)
    pair = np.zeros((1, 20), np.float32)
    pair[0, 0], pair[0, 1] = 0.25, -0.125
    pair[0, 2] = 1.0  # s_0 = 1: prediction = tap 0 = sinc weight 0 + network output 0
    loss, _ = po.mlp_loss_grad(w, pair, False)
    b0 = po.sinc8_weights(0.125)[0]
    assert abs(loss - ((out[0] + b0) ** 2 + out[1] ** 2)) < 1e-12 * max(1.0, loss)


def test_sinc8_weights():
    assert np.array_equal(po.sinc8_weights(0.0), np.eye(8)[4])
    win = np.blackman(8).astype(np.float32).astype(np.float64)
    for d in (-0.5, -0.3, 0.01, 0.25, 0.5):
        w = po.sinc8_weights(d)
        assert abs(w.sum() - 1.0) < 1e-14
        ref = win * np.sinc(d + 4 - np.arange(8))
        assert np.allclose(w, ref / ref.sum(), rtol=1e-12, atol=1e-14)


def test_bf16_round_known_values():
    x = np.array([1.0, 1.00390625, 1.005859375, 1.0 + 2 ** -8 + 2 ** -9, -3.14159265, 65504.0, 1e-20],
                 np.float32)
    y = po.bf16_round(x)
    # bfloat16 keeps 8 significant bits; ties go to even
    assert y[0] == 1.0
    assert y[1] == 1.0  # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: even (1.0)
    assert y[2] == np.float32(1.0078125)  # above the tie
    assert y[3] == np.float32(1.0078125)
    assert y[4] == np.float32(-3.140625)
    assert y[5] == np.float32(65536.0)
    assert (y.view(np.uint32) & 0xFFFF).max() == 0
    assert np.all(np.abs(y - x) <= np.abs(x) * 2 ** -8)
