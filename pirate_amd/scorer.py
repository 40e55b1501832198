"""Learned peak scorer: a small bf16 1-D convolutional network over (DM, time) tiles of the
dedispersed output, run and trained on gfx950 MFMA by the library's own kernels
(pirate_amd/csrc/gpu/k_scorer.hip, C-ABI pdd_scorer_*).

The reference has no learned scorer. The network stands beside the boxcar statistic of the
peak finder (src_lib/PeakFindingKernel.cu:252-456): the peak finder stays the fallback and the
baseline of the scorer's completeness (completeness at equal false-alarm rate, measured by
pirate_amd.scorer_data.completeness_at_far).

Network on one tile X (64 DM rows x 256 samples), "same" convolutions with zero padding:
    A1 = bf16(relu(conv9(W1, X) + b1))     64 -> 16 channels, 9 taps
    A2 = bf16(relu(conv9(W2, A1) + b2))    16 -> 16
    A3 = relu(conv9(W3, A2) + b3)          16 -> 16 (float32)
    logit = c0 + sum_c u[c] max_t A3[c, t] + v[c] mean_t A3[c, t]
Inputs and conv weights are rounded to bf16 on load; products accumulate in float32. The
parameters are one float32 vector of 13905 elements:
    W1 (16,64,9) | W2 (16,16,9) | W3 (16,16,9) | b1 (16) | b2 (16) | b3 (16) | u (16) | v (16) | c0
"""

import ctypes

import numpy as np

from . import _lib as L

C0, C, K, T = 64, 16, 9, 256
SHAPES = [('W1', (C, C0, K)), ('W2', (C, C, K)), ('W3', (C, C, K)), ('b1', (C,)), ('b2', (C,)), ('b3', (C,)),
          ('u', (C,)), ('v', (C,)), ('c0', (1,))]
NPARAMS = sum(int(np.prod(s)) for _, s in SHAPES)

_DTYPES = {'float32': 0, 'float16': 1, 'bfloat16': 2}


def _torch():
    import torch
    return torch


def _stream_ptr(stream):
    torch = _torch()
    s = torch.cuda.current_stream() if stream is None else stream
    return ctypes.c_void_p(s.cuda_stream)


def init_params(seed=0):
    """He-normal conv weights, zero biases, small random head (float64 numpy vector)."""
    rng = np.random.default_rng(seed)
    parts = [rng.standard_normal((C, C0, K)) * np.sqrt(2.0 / (C0 * K)),
             rng.standard_normal((C, C, K)) * np.sqrt(2.0 / (C * K)),
             rng.standard_normal((C, C, K)) * np.sqrt(2.0 / (C * K)),
             np.zeros(C), np.zeros(C), np.zeros(C),
             rng.standard_normal(C) * 0.25, rng.standard_normal(C) * 0.25, np.zeros(1)]
    return np.concatenate([p.reshape(-1) for p in parts])


class TileSource:
    """Where tiles are read from (see pdd_tile_source in include/pirate_amd.h).

    Either a contiguous (n, 64, 256) device tensor of tiles (float32, float16 or bfloat16), or
    rows of a dedispersed output (GpuDedisperser.tree_rows) with an (n, 3) int32 device tensor
    of tile origins (beam, row0, t0). The objects whose memory the source points at are kept
    referenced here."""

    def __init__(self, struct, keep=(), ntiles=None):
        self.struct = struct
        self._keep = list(keep)
        self.ntiles = ntiles

    @classmethod
    def from_tiles(cls, tiles):
        torch = _torch()
        if tiles.dim() != 3 or tuple(tiles.shape[1:]) != (C0, T) or not tiles.is_cuda or not tiles.is_contiguous():
            raise ValueError('TileSource.from_tiles: need a contiguous (n, 64, 256) device tensor')
        name = {torch.float32: 'float32', torch.float16: 'float16', torch.bfloat16: 'bfloat16'}.get(tiles.dtype)
        if name is None:
            raise ValueError(f'TileSource.from_tiles: unsupported dtype {tiles.dtype}')
        s = L.TileSource(ctypes.c_void_p(tiles.data_ptr()), _DTYPES[name], tiles.shape[0], C0 * T, C0, T, 0, T, None)
        return cls(s, [tiles], tiles.shape[0])

    def with_origins(self, origins):
        """The same rows with tile origins: an (n, 3) int32 tensor (beam, row0, t0), moved to the device."""
        torch = _torch()
        o = torch.as_tensor(origins, dtype=torch.int32)
        if o.dim() != 2 or o.shape[1] != 3:
            raise ValueError('TileSource.with_origins: origins must have shape (n, 3)')
        o = o.to('cuda').contiguous()
        s = L.TileSource(self.struct.base, self.struct.dtype, self.struct.nbeams, self.struct.beam_stride,
                         self.struct.nrows, self.struct.row_stride, self.struct.t_lo, self.struct.t_hi,
                         ctypes.c_void_p(o.data_ptr()))
        return TileSource(s, self._keep + [o], o.shape[0])


class LearnedScorer:
    """Parameters (float32 device vector), Adam state and the forward / gradient / training
    calls of the learned peak scorer. Calls run on the current torch stream (or 'stream')."""

    def __init__(self, params=None, seed=0, device='cuda'):
        torch = _torch()
        if L.lib().pdd_scorer_nparams() != NPARAMS:
            raise RuntimeError('LearnedScorer: parameter count mismatch with libpirate_amd.so')
        p = init_params(seed) if params is None else np.asarray(params, dtype=np.float64)
        if p.shape != (NPARAMS,):
            raise ValueError(f'LearnedScorer: params must have shape ({NPARAMS},)')
        self.params = torch.tensor(p, dtype=torch.float32, device=device)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.grad_buf = torch.zeros_like(self.params)
        self.step = 0

    def state_dict(self):
        return {'params': self.params.cpu().numpy(), 'm': self.m.cpu().numpy(), 'v': self.v.cpu().numpy(),
                'step': self.step}

    def load_state_dict(self, d):
        torch = _torch()
        for name in ('params', 'm', 'v'):
            getattr(self, name).copy_(torch.as_tensor(np.asarray(d[name], dtype=np.float32)))
        self.step = int(d['step'])

    @staticmethod
    def _source(src):
        return src if isinstance(src, TileSource) else TileSource.from_tiles(src)

    def forward(self, src, stream=None):
        """Logits (n,) float32 device tensor of the tiles of src (a TileSource or an (n, 64, 256) tensor)."""
        torch = _torch()
        s = self._source(src)
        out = torch.empty(s.ntiles, dtype=torch.float32, device=self.params.device)
        L.check(L.lib().pdd_scorer_forward(ctypes.c_void_p(self.params.data_ptr()), ctypes.byref(s.struct), s.ntiles,
                                           ctypes.c_void_p(out.data_ptr()), _stream_ptr(stream)))
        return out

    def grad(self, src, labels, stream=None):
        """(grad (NPARAMS,), logits (n,), loss_terms (n,)) of the mean binary cross-entropy."""
        torch = _torch()
        s = self._source(src)
        y = torch.as_tensor(labels, dtype=torch.float32).to(self.params.device).contiguous()
        if y.shape != (s.ntiles,):
            raise ValueError('LearnedScorer.grad: one label per tile')
        logits = torch.empty(s.ntiles, dtype=torch.float32, device=self.params.device)
        loss = torch.empty_like(logits)
        L.check(L.lib().pdd_scorer_grad(ctypes.c_void_p(self.params.data_ptr()), ctypes.byref(s.struct),
                                        ctypes.c_void_p(y.data_ptr()), s.ntiles, ctypes.c_void_p(self.grad_buf.data_ptr()),
                                        ctypes.c_void_p(logits.data_ptr()), ctypes.c_void_p(loss.data_ptr()),
                                        _stream_ptr(stream)))
        return self.grad_buf, logits, loss

    def grad_noalloc(self, src, labels, grad, logits, loss_terms, stream=None):
        """grad() through pdd_scorer_grad_noalloc, which allocates nothing: the caller passes
        the label vector and the outputs (float32 device tensors of NPARAMS, n, n and n
        elements); the partial gradients live in the library's fixed device array (calls that
        overlap on one GPU race on it). Results equal grad()'s bit for bit."""
        s = self._source(src)
        for t, n in ((labels, s.ntiles), (grad, NPARAMS), (logits, s.ntiles), (loss_terms, s.ntiles)):
            if not (t.is_cuda and t.is_contiguous() and t.dtype == self.params.dtype and t.numel() == n):
                raise ValueError('LearnedScorer.grad_noalloc: labels and outputs must be contiguous float32 '
                                 'device tensors of the right sizes')
        L.check(L.lib().pdd_scorer_grad_noalloc(ctypes.c_void_p(self.params.data_ptr()), ctypes.byref(s.struct),
                                                ctypes.c_void_p(labels.data_ptr()), s.ntiles,
                                                ctypes.c_void_p(grad.data_ptr()), ctypes.c_void_p(logits.data_ptr()),
                                                ctypes.c_void_p(loss_terms.data_ptr()), _stream_ptr(stream)))

    def adam_step(self, grad=None, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8, stream=None):
        g = self.grad_buf if grad is None else grad
        self.step += 1
        L.check(L.lib().pdd_scorer_adam(ctypes.c_void_p(self.params.data_ptr()), ctypes.c_void_p(self.m.data_ptr()),
                                        ctypes.c_void_p(self.v.data_ptr()), ctypes.c_void_p(g.data_ptr()), self.step,
                                        lr, beta1, beta2, eps, _stream_ptr(stream)))

    def train_step(self, src, labels, lr=1e-3, stream=None):
        """One gradient step on the tiles of src; returns the mean loss as a device scalar."""
        _, _, loss = self.grad(src, labels, stream)
        self.adam_step(lr=lr, stream=stream)
        return loss.mean()


class ScoreHead:
    """Calibrated head over the network's logit z and the peak finder's boxcar statistic p of
    the same tile: score = w . f(z, p), f = (1, z, p, z p), standardized on the training set,
    with w the logistic-regression fit of the training labels (Newton / iteratively reweighted
    least squares in float64, a small ridge term for conditioning). The network ranks tiles
    by shape (it learns to reject RFI); the boxcar statistic carries the matched-filter SNR,
    near optimal on Gaussian noise; the head weighs both. score is a log-odds."""

    NF = 4

    def __init__(self, w=None, mean=None, std=None):
        self.w = None if w is None else np.asarray(w, dtype=np.float64)
        self.mean = None if mean is None else np.asarray(mean, dtype=np.float64)
        self.std = None if std is None else np.asarray(std, dtype=np.float64)

    @staticmethod
    def _raw(z, p):
        z = np.asarray(z, dtype=np.float64)
        p = np.asarray(p, dtype=np.float64)
        return np.stack([z, p, z * p], axis=1)

    def _features(self, z, p):
        f = (self._raw(z, p) - self.mean) / self.std
        return np.concatenate([np.ones((f.shape[0], 1)), f], axis=1)

    def fit(self, z, p, labels, iters=30, ridge=1e-6):
        """Fits w; returns the per-iteration mean log-loss."""
        r = self._raw(z, p)
        self.mean = r.mean(axis=0)
        self.std = np.maximum(r.std(axis=0), 1e-12)
        X = self._features(z, p)
        y = np.asarray(labels, dtype=np.float64)
        n = X.shape[0]
        w = np.zeros(self.NF)
        hist = []
        for _ in range(iters):
            s = X @ w
            q = 0.5 * (1.0 + np.tanh(0.5 * s))                       # sigmoid
            hist.append(float(np.mean(np.logaddexp(0.0, s) - y * s)))
            g = X.T @ (y - q) - ridge * n * w
            H = (X * (q * (1.0 - q))[:, None]).T @ X + ridge * n * np.eye(self.NF)
            step = np.linalg.solve(H, g)
            w = w + step
            if np.max(np.abs(step)) < 1e-10:
                break
        self.w = w
        return hist

    def __call__(self, z, p):
        return self._features(z, p) @ self.w

    def state_dict(self):
        return {'w': self.w.tolist(), 'mean': self.mean.tolist(), 'std': self.std.tolist()}
