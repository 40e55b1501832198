"""ORACLE — test infrastructure only.

float64 numpy restatement of the learned peak scorer (a small 1-D conv net over (DM, time)
tiles of the dedispersed output): forward pass, mean binary cross-entropy loss and its exact
gradient. The reference has no learned scorer; this network stands beside its boxcar peak
statistic (src_lib/PeakFindingKernel.cu:252-456), which stays the baseline of completeness.
Only tests/ may use this module, and only as the checker.

Network (one tile X: 64 DM rows x 256 samples):
    Y_l[co, t] = b_l[co] + sum_{ci, k} W_l[co, ci, k] A_{l-1}[ci, t + k - 4]   (zero outside 0..255)
    A_0 = X, A_l = relu(Y_l), l = 1, 2, 3; 16 channels, 9 taps, Cin = 64 / 16 / 16
    pmax[c] = max_t A_3[c, t] (first argmax tau[c]), pmean[c] = mean_t A_3[c, t]
    logit = c0 + sum_c u[c] pmax[c] + v[c] pmean[c]
    loss = mean_i softplus(logit_i) - y_i logit_i
Parameter vector (float32 in the library), in this order:
    W1 (16, 64, 9) | W2 (16, 16, 9) | W3 (16, 16, 9) | b1 (16) | b2 (16) | b3 (16) | u (16) | v (16) | c0

rounding='bf16' rounds (round-to-nearest-even) exactly where the GPU kernels hold bfloat16
values: the input tiles, the conv weights, A_1 and A_2 (after relu), and the backward
signals dY_3, dY_2, dY_1 (after the relu masks) that feed the matrix products; everything
else stays float64. rounding=None is the plain float64 network, whose gradient the tests pin
with central finite differences, so that a wrong sign, index or transposition in the
backward pass fails there.
"""

import numpy as np

C0, C, K, T, PAD = 64, 16, 9, 256, 4
SHAPES = [('W1', (C, C0, K)), ('W2', (C, C, K)), ('W3', (C, C, K)), ('b1', (C,)), ('b2', (C,)), ('b3', (C,)),
          ('u', (C,)), ('v', (C,)), ('c0', (1,))]
NPARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(theta):
    theta = np.asarray(theta, dtype=np.float64)
    assert theta.shape == (NPARAMS,), theta.shape
    out, o = {}, 0
    for name, s in SHAPES:
        n = int(np.prod(s))
        out[name] = theta[o:o + n].reshape(s)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).reshape(-1) for name, _ in SHAPES])


def bf16(x):
    """Round-to-nearest-even to bfloat16, returned as float64 (via float32: every value here
    is a float32 or a float64 product that the kernels hold in float32 first)."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    nan = np.isnan(a)
    out = r.astype(np.uint32).view(np.float32).astype(np.float64)
    out[nan] = np.nan
    return out


def _rnd(x, rounding):
    return bf16(x) if rounding == 'bf16' else np.asarray(x, dtype=np.float64)


def conv(W, b, A):
    """Y[n, co, t] = b[co] + sum_{ci,k} W[co, ci, k] A[n, ci, t + k - 4]."""
    n, cin, t = A.shape
    Ap = np.zeros((n, cin, t + 2 * PAD))
    Ap[:, :, PAD:PAD + t] = A
    Y = np.broadcast_to(b[None, :, None], (n, W.shape[0], t)).copy()
    for k in range(K):
        Y += np.einsum('oc,nct->not', W[:, :, k], Ap[:, :, k:k + t], optimize=True)
    return Y


def conv_grads(dY, A, W):
    """Gradients of conv: dW[co, ci, k] = sum_{n,t} dY[n, co, t] A[n, ci, t + k - 4];
    dA[n, ci, t] = sum_{co, k} W[co, ci, k] dY[n, co, t - k + 4]."""
    n, cin, t = A.shape
    Ap = np.zeros((n, cin, t + 2 * PAD))
    Ap[:, :, PAD:PAD + t] = A
    dW = np.zeros(W.shape)
    for k in range(K):
        dW[:, :, k] = np.einsum('not,nct->oc', dY, Ap[:, :, k:k + t], optimize=True)
    dYp = np.zeros((n, dY.shape[1], t + 2 * PAD))
    dYp[:, :, PAD:PAD + t] = dY
    dA = np.zeros((n, cin, t))
    for k in range(K):
        # dY index t - k + 4 lives at padded position t - k + 8
        dA += np.einsum('oc,not->nct', W[:, :, k], dYp[:, :, 2 * PAD - k:2 * PAD - k + t], optimize=True)
    return dW, dA


def forward(theta, X, rounding=None):
    """X: (n, 64, 256). Returns (logits (n,), cache)."""
    p = unpack(theta)
    W = [_rnd(p['W1'], rounding), _rnd(p['W2'], rounding), _rnd(p['W3'], rounding)]
    A0 = _rnd(X, rounding)
    Y1 = conv(W[0], p['b1'], A0)
    A1 = _rnd(np.maximum(Y1, 0), rounding)
    Y2 = conv(W[1], p['b2'], A1)
    A2 = _rnd(np.maximum(Y2, 0), rounding)
    Y3 = conv(W[2], p['b3'], A2)
    A3 = np.maximum(Y3, 0)
    tau = np.argmax(A3, axis=2)                       # first maximum
    pmax = np.take_along_axis(A3, tau[:, :, None], axis=2)[:, :, 0]
    pmean = A3.mean(axis=2)
    logits = p['c0'][0] + pmax @ p['u'] + pmean @ p['v']
    cache = dict(p=p, W=W, A0=A0, A1=A1, A2=A2, Y3=Y3, tau=tau, pmax=pmax, pmean=pmean)
    return logits, cache


def loss_and_grad(theta, X, y, rounding=None):
    """Mean binary cross-entropy over the tiles and its gradient (NPARAMS,)."""
    y = np.asarray(y, dtype=np.float64)
    logits, c = forward(theta, X, rounding)
    n = len(y)
    loss = float(np.mean(np.logaddexp(0.0, logits) - y * logits))
    g = (1.0 / (1.0 + np.exp(-logits)) - y) / n               # d loss / d logit_i
    p = c['p']
    grads = dict(u=g @ c['pmax'], v=g @ c['pmean'], c0=np.array([g.sum()]))
    dA3 = np.broadcast_to((g[:, None] * p['v'][None, :] / T)[:, :, None], c['Y3'].shape).copy()
    np.put_along_axis(dA3, c['tau'][:, :, None],
                      np.take_along_axis(dA3, c['tau'][:, :, None], axis=2) + (g[:, None] * p['u'][None, :])[:, :, None],
                      axis=2)
    dY3 = _rnd(np.where(c['Y3'] > 0, dA3, 0.0), rounding)
    grads['b3'] = dY3.sum(axis=(0, 2))
    grads['W3'], dA2 = conv_grads(dY3, c['A2'], c['W'][2])
    dY2 = _rnd(np.where(c['A2'] > 0, dA2, 0.0), rounding)
    grads['b2'] = dY2.sum(axis=(0, 2))
    grads['W2'], dA1 = conv_grads(dY2, c['A1'], c['W'][1])
    dY1 = _rnd(np.where(c['A1'] > 0, dA1, 0.0), rounding)
    grads['b1'] = dY1.sum(axis=(0, 2))
    grads['W1'], _ = conv_grads(dY1, c['A0'], c['W'][0])
    return loss, pack(grads), logits


def init_params(rng):
    """He-normal conv weights, zero biases, small head (the library's initialization)."""
    d = {}
    d['W1'] = rng.standard_normal((C, C0, K)) * np.sqrt(2.0 / (C0 * K))
    d['W2'] = rng.standard_normal((C, C, K)) * np.sqrt(2.0 / (C * K))
    d['W3'] = rng.standard_normal((C, C, K)) * np.sqrt(2.0 / (C * K))
    for name in ('b1', 'b2', 'b3'):
        d[name] = np.zeros(C)
    d['u'] = rng.standard_normal(C) * 0.25
    d['v'] = rng.standard_normal(C) * 0.25
    d['c0'] = np.zeros(1)
    return pack(d)


def adam_step(theta, m, v, grad, step, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015, Algorithm 1), step counted from 1."""
    m = beta1 * m + (1 - beta1) * grad
    v = beta2 * v + (1 - beta2) * grad * grad
    mhat = m / (1 - beta1**step)
    vhat = v / (1 - beta2**step)
    return theta - lr * mhat / (np.sqrt(vhat) + eps), m, v
