"""Pins the learned scorer's float64 oracle (oracle/scorer.py), on the CPU:

  * the convolution against explicit sums (every edge position, every tap);
  * the gradient of the plain float64 network against central finite differences, along
    random directions restricted to each parameter tensor (a wrong sign, index or transposed
    operand in any layer's backward pass fails here);
  * bfloat16 rounding against torch's round-to-nearest-even cast;
  * the Adam step against torch.optim.Adam on the same gradient sequence;
  * the loss against softplus/BCE computed by torch.
"""

import numpy as np
import pytest

from oracle import scorer as sc


def _tiles(rng, n):
    return rng.standard_normal((n, sc.C0, sc.T))


def test_conv_matches_explicit_sums():
    rng = np.random.default_rng(0)
    A = rng.standard_normal((2, 5, sc.T))
    W = rng.standard_normal((3, 5, sc.K))
    b = rng.standard_normal(3)
    Y = sc.conv(W, b, A)
    for n in range(2):
        for co in range(3):
            for t in (0, 1, 3, 4, 100, 251, 252, 255):
                s = b[co]
                for ci in range(5):
                    for k in range(sc.K):
                        tt = t + k - 4
                        if 0 <= tt < sc.T:
                            s += W[co, ci, k] * A[n, ci, tt]
                assert abs(Y[n, co, t] - s) < 1e-12


def test_conv_grads_are_adjoint():
    # <conv(W, 0, A), dY> == <W, dW(dY, A)> == <A, dA(dY, W)>
    rng = np.random.default_rng(1)
    A = rng.standard_normal((3, 4, sc.T))
    W = rng.standard_normal((6, 4, sc.K))
    dY = rng.standard_normal((3, 6, sc.T))
    dW, dA = sc.conv_grads(dY, A, W)
    lhs = np.sum(sc.conv(W, np.zeros(6), A) * dY)
    assert abs(lhs - np.sum(W * dW)) < 1e-9 * abs(lhs) + 1e-9
    assert abs(lhs - np.sum(A * dA)) < 1e-9 * abs(lhs) + 1e-9


@pytest.mark.parametrize('seed', [2, 3])
def test_gradient_matches_finite_differences(seed):
    rng = np.random.default_rng(seed)
    theta = sc.init_params(rng)
    theta[-1] = 0.3
    X = _tiles(rng, 3)
    y = np.array([1.0, 0.0, 1.0])
    loss, g, _ = sc.loss_and_grad(theta, X, y)
    o = 0
    for name, shape in sc.SHAPES:
        n = int(np.prod(shape))
        for trial in range(2):
            d = np.zeros_like(theta)
            d[o:o + n] = rng.standard_normal(n)
            eps = 1e-6
            lp, _, _ = sc.loss_and_grad(theta + eps * d, X, y)
            lm, _, _ = sc.loss_and_grad(theta - eps * d, X, y)
            fd = (lp - lm) / (2 * eps)
            an = float(g @ d)
            assert abs(fd - an) <= 1e-6 * max(abs(an), 1e-3), (name, trial, fd, an)
        o += n
    assert o == sc.NPARAMS


def test_bf16_rounding_matches_torch():
    import torch
    rng = np.random.default_rng(4)
    x = np.concatenate([rng.standard_normal(10000) * 10.0 ** rng.integers(-30, 30, 10000),
                        np.array([0.0, -0.0, 1.0, 1.00390625, 1.01171875, 3.0e38, -7.0e-39])]).astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(sc.bf16(x), ref)


def test_adam_matches_torch():
    import torch
    rng = np.random.default_rng(5)
    theta = rng.standard_normal(50)
    p = torch.nn.Parameter(torch.from_numpy(theta.copy()))
    opt = torch.optim.Adam([p], lr=1e-2, betas=(0.9, 0.999), eps=1e-8)
    m = np.zeros(50)
    v = np.zeros(50)
    th = theta.copy()
    for step in range(1, 6):
        g = rng.standard_normal(50)
        p.grad = torch.from_numpy(g.copy())
        opt.step()
        th, m, v = sc.adam_step(th, m, v, g, step, 1e-2)
    np.testing.assert_allclose(th, p.detach().numpy(), rtol=1e-12, atol=1e-14)


def test_loss_matches_torch_bce():
    import torch
    rng = np.random.default_rng(6)
    theta = sc.init_params(rng)
    X = _tiles(rng, 4)
    y = np.array([0.0, 1.0, 1.0, 0.0])
    loss, _, logits = sc.loss_and_grad(theta, X, y)
    ref = torch.nn.functional.binary_cross_entropy_with_logits(torch.from_numpy(logits), torch.from_numpy(y))
    assert abs(loss - float(ref)) < 1e-12
