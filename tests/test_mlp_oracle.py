"""CPU checks of the learned warm start's oracle (oracle/mlp_oracle.py) and of the host-side network
API without GPU compute:
  * the FP64 forward pass, loss and parameter gradient equal torch FP64 autograd of the same network;
  * central finite differences of the loss agree with the oracle's gradient;
  * the bfloat16 rounding helper rounds to nearest even (known bit patterns);
  * closed forms: with zero hidden weights the output is the last bias and the last bias's gradient
    is the mean residual;
  * feature rows of a hand-built one-ring case equal values worked out by hand.
"""
import os
import sys

import numpy as np
import pytest
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import mlp_oracle as mo  # noqa: E402

from spg_amd import native, scenes  # noqa: E402


def _random_params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    layers = []
    for o, i in mo.SHAPES:
        a = scale * np.sqrt(6.0 / (o + i))
        layers.append((rng.uniform(-a, a, size=(o, i)), rng.uniform(-0.1, 0.1, size=o)))
    return mo.flatten(layers)


def _data(seed, n=40):
    rng = np.random.default_rng(seed)
    return rng.normal(size=(n, 24)), rng.normal(size=(n, 3)), rng.normal(size=24) * 0.3, rng.uniform(0.5, 2.0, 24)


def test_num_params():
    assert mo.NUM_PARAMS == native.MLP_NUM_PARAMS == 36611


def test_forward_and_gradient_match_torch_autograd():
    p = _random_params(1)
    f, t, mean, std = _data(2)
    loss, g = mo.loss_grad(p, f, t, mean, std)
    pt = torch.tensor(p, dtype=torch.float64, requires_grad=True)
    x = (torch.tensor(f) - torch.tensor(mean)) / torch.tensor(std)
    k = 0
    h = x
    for l, (o, i) in enumerate(mo.SHAPES):
        w = pt[k:k + o * i].reshape(o, i)
        k += o * i
        b = pt[k:k + o]
        k += o
        h = torch.nn.functional.linear(h, w, b)
        if l < 3:
            h = torch.nn.functional.silu(h)
    y_ref = h
    lt = ((y_ref - torch.tensor(t)) ** 2).sum() / f.shape[0]
    lt.backward()
    y = mo.forward(p, f, mean, std)
    assert np.allclose(y, y_ref.detach().numpy(), rtol=1e-12, atol=1e-12)
    assert abs(loss - lt.item()) <= 1e-12 * abs(lt.item())
    assert np.allclose(g, pt.grad.numpy(), rtol=1e-10, atol=1e-13)


def test_gradient_matches_finite_differences():
    p = _random_params(3)
    f, t, mean, std = _data(4, n=12)
    _, g = mo.loss_grad(p, f, t, mean, std)
    rng = np.random.default_rng(5)
    # every layer's weights and biases are probed
    idx = [0, 24 * 128 - 1, 24 * 128 + 5, 3200 + 77, 3200 + 128 * 128 + 3, 36611 - 3 - 1, 36611 - 1]
    idx += list(rng.integers(0, mo.NUM_PARAMS, size=20))
    for i in idx:
        e = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[i] += e
        pm[i] -= e
        fd = (mo.loss_grad(pp, f, t, mean, std)[0] - mo.loss_grad(pm, f, t, mean, std)[0]) / (2 * e)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])) + 1e-9, (i, fd, g[i])


def test_bf16_rounding_known_patterns():
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: ties go to the even mantissa (1.0)
    assert mo.bf16(np.float32(1.0 + 2.0 ** -8)) == 1.0
    # 1 + 3 * 2^-8 is a tie between 1 + 2^-7 and 1 + 2^-6: even is 1 + 2^-6
    assert mo.bf16(np.float32(1.0 + 3 * 2.0 ** -8)) == 1.0 + 2.0 ** -6
    # just above the tie rounds up
    assert mo.bf16(np.float32(1.0 + 2.0 ** -8 + 2.0 ** -20)) == 1.0 + 2.0 ** -7
    assert mo.bf16(np.float32(-3.0)) == -3.0
    x = np.random.default_rng(1).normal(size=1000).astype(np.float32)
    assert np.max(np.abs(mo.bf16(x) - x) / np.abs(x)) <= 2.0 ** -8


def test_zero_hidden_weights_closed_form():
    layers = mo.unflatten(_random_params(6))
    layers = [(np.zeros_like(w), b) for w, b in layers]
    p = mo.flatten(layers)
    f, t, _, _ = _data(7)
    y = mo.forward(p, f)
    assert np.allclose(y, layers[3][1][None, :])
    loss, g = mo.loss_grad(p, f, t)
    assert np.allclose(g[-3:], 2.0 * (y - t).mean(axis=0))
    assert np.isclose(loss, ((y - t) ** 2).sum() / f.shape[0])


def test_emulated_forward_close_to_exact():
    p = _random_params(8)
    f, t, mean, std = _data(9, n=200)
    y = mo.forward(p, f, mean, std)
    ye = mo.forward(p, f, mean, std, emulate=True)
    assert np.linalg.norm(ye - y) / np.linalg.norm(y) < 3e-2
    _, g = mo.loss_grad(p, f, t, mean, std)
    _, ge = mo.loss_grad(p, f, t, mean, std, emulate=True)
    assert np.linalg.norm(ge - g) / np.linalg.norm(g) < 5e-2


def test_features_hand_case():
    """Three vertices in one triangle plus an anchor on vertex 0; values worked out by hand."""
    x0 = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 2.0, 0.0]])
    s = scenes.Scene(x=x0.copy(), x0=x0, m=np.ones(3),
                     energies=[("bending", np.zeros((0, 4), np.int32), np.zeros((0, 2))),
                               ("stvk", np.array([[0, 1, 2]], np.int32), np.array([[1e3, 0.3]])),
                               ("anchor", np.array([[0]], np.int32), np.array([[0.0, 0.0, 0.0, 1e5]]))])
    tables = mo.static_tables(s)
    nb, rest, stat = tables
    assert [list(j) for j in nb] == [[1, 2], [0, 2], [0, 1]]
    assert np.allclose(rest[0], [1.0, 2.0]) and np.allclose(rest[1], [1.0, np.sqrt(5.0)])
    assert np.allclose(stat, [[3.0, 0.0, 1.0], [3.0, 0.0, 0.0], [3.0, 0.0, 0.0]])
    h, g = 0.1, np.array([0.0, -10.0, 0.0])
    xt = np.array([[0.0, 0.0, 0.0], [1.1, 0.0, 0.0], [0.0, 2.0, 0.0]])
    vk = np.array([[0.0, -1.0, 0.0], [0.0, -1.0, 0.0], [0.0, 0.0, 0.0]])  # velocities after the kick
    c = np.array([[0.0, 0.1, 0.0], [0.0, 0.0, 0.0], [0.0, 0.0, 0.2]])
    F = mo.features(xt, vk, c, tables, h, g)
    s_, u_ = 0.1 ** 2 * 10.0, 0.1 * 10.0
    assert np.allclose(F[0, 0:3], [0.0, 0.1 / s_, 0.0])
    assert np.allclose(F[2, 3:6], np.array([0.0, 1.0, 0.0]) / u_)      # v - h g = (0, 1, 0)
    assert np.allclose(F[0, 6:9], np.array([0.0, 0.0, 0.1]) / s_)      # mean of c1, c2
    assert np.allclose(F[0, 9:12], np.array([0.0, 0.1, 0.2]) / s_)
    assert np.allclose(F[0, 21:24], [0.05, 0.1, 0.0])                   # strains 0.1 and 0
    assert np.allclose(F[1, 21:24], [0.05 + 0.5 * (np.sqrt(1.21 + 4) / np.sqrt(5) - 1),
                                     0.1, np.sqrt(1.21 + 4) / np.sqrt(5) - 1])
    assert np.allclose(F[:, 18:21], stat)
