"""Pins the FP64 oracle of the Poisson warm-start network (oracle/pnet_oracle.cpp).

gradient    -> central finite differences of the loss along random directions
               and single coordinates
equivariance-> a periodic shift of the input by a multiple of 4 voxels (the
               pooling factor) shifts the output the same way
linearity   -> the last layer is linear: scaling its weights and bias scales y
"""
import numpy as np
import pytest

from oracle import oracle as om


def _init(rng):
    P = om.pnet_param_count()
    params = np.zeros(P)
    for cin, cout, level, relu, woff, boff in om.pnet_layers():
        fan_in = 27 * cin
        params[woff:boff] = rng.standard_normal(boff - woff) * np.sqrt(2.0 / fan_in)
        params[boff:boff + cout] = rng.uniform(0.0, 0.1, cout)
    return params


def test_param_count_and_layers():
    assert om.pnet_param_count() == 182083
    layers = om.pnet_layers()
    assert [(l[0], l[1]) for l in layers] == [(2, 16), (16, 16), (16, 32), (32, 32), (32, 32), (32, 32),
                                               (64, 32), (48, 16), (16, 3)]


def _problem(n, seed):
    rng = np.random.default_rng(seed)
    params = _init(rng)
    inp = rng.uniform(0, 1, (n ** 3, 2))
    e = rng.standard_normal(3)
    e /= np.linalg.norm(e)
    t = rng.standard_normal(n ** 3) * 0.1
    return rng, params, inp, e, t


def test_gradient_directional_finite_difference():
    n = 8
    rng, params, inp, e, t = _problem(n, 1)
    loss, g = om.pnet_loss_grad(n, params, inp, e, t)
    for _ in range(3):
        d = rng.standard_normal(params.size)
        d /= np.linalg.norm(d)  # unit direction: a small step crosses no ReLU kink
        h = 1e-6
        lp, _ = om.pnet_loss_grad(n, params + h * d, inp, e, t)
        lm, _ = om.pnet_loss_grad(n, params - h * d, inp, e, t)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g @ d) < 1e-6 * max(1.0, abs(fd))


def test_gradient_coordinates_each_layer():
    n = 8
    rng, params, inp, e, t = _problem(n, 2)
    _, g = om.pnet_loss_grad(n, params, inp, e, t)
    for cin, cout, level, relu, woff, boff in om.pnet_layers():
        for i in [woff + 7, boff]:
            h = 1e-6
            p1, p2 = params.copy(), params.copy()
            p1[i] += h
            p2[i] -= h
            fd = (om.pnet_loss_grad(n, p1, inp, e, t)[0] - om.pnet_loss_grad(n, p2, inp, e, t)[0]) / (2 * h)
            assert abs(fd - g[i]) < 1e-6 * max(1e-3, abs(fd)), (cin, cout, i)


def test_translation_equivariance():
    n = 8
    rng, params, inp, e, t = _problem(n, 3)
    y = om.pnet_forward(n, params, inp).reshape(n, n, n, 3)
    xs = np.roll(inp.reshape(n, n, n, 2), shift=(4, 0, 4), axis=(0, 1, 2))
    ys = om.pnet_forward(n, params, xs.reshape(-1, 2)).reshape(n, n, n, 3)
    assert np.abs(np.roll(y, shift=(4, 0, 4), axis=(0, 1, 2)) - ys).max() < 1e-12


def test_last_layer_linear():
    n = 8
    rng, params, inp, e, t = _problem(n, 4)
    y = om.pnet_forward(n, params, inp)
    woff = om.pnet_layers()[-1][4]
    p2 = params.copy()
    p2[woff:] *= 2.0
    assert np.abs(om.pnet_forward(n, p2, inp) - 2.0 * y).max() < 1e-12
