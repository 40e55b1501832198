"""pdd_scorer_grad_noalloc (the gradient without allocation) against pdd_scorer_grad, through
the C-ABI: identical gradients, logits and loss terms, bit for bit, on two calls in a row, the
second with fewer tiles (fewer partial gradients than the first call left in the array)."""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _call(fn, scr, src, y, g, lo, ls):
    from pirate_amd import _lib as L
    L.check(fn(ctypes.c_void_p(scr.params.data_ptr()), ctypes.byref(src.struct), ctypes.c_void_p(y.data_ptr()),
               src.ntiles, ctypes.c_void_p(g.data_ptr()), ctypes.c_void_p(lo.data_ptr()),
               ctypes.c_void_p(ls.data_ptr()), None))


def test_noalloc_gradient_equals_allocating_gradient():
    import torch
    import pirate_amd as pa
    from pirate_amd import _lib as L
    rng = np.random.default_rng(70)
    scr = pa.LearnedScorer(seed=71)
    for n in (700, 300):
        print(f'[no-alloc] {n} tiles', flush=True)
        X = torch.from_numpy(rng.standard_normal((n, 64, 256)).astype(np.float32)).cuda().to(torch.bfloat16)
        y = torch.from_numpy((rng.random(n) < 0.5).astype(np.float32)).cuda()
        src = pa.TileSource.from_tiles(X)
        outs = {}
        for name in ('pdd_scorer_grad', 'pdd_scorer_grad_noalloc'):
            g = torch.full((pa.scorer.NPARAMS,), np.nan, dtype=torch.float32, device='cuda')
            lo = torch.full((n,), np.nan, dtype=torch.float32, device='cuda')
            ls = torch.full((n,), np.nan, dtype=torch.float32, device='cuda')
            torch.cuda.synchronize()
            _call(getattr(L.lib(), name), scr, src, y, g, lo, ls)
            torch.cuda.synchronize()
            outs[name] = (g, lo, ls)
        for a, b in zip(outs['pdd_scorer_grad'], outs['pdd_scorer_grad_noalloc']):
            assert torch.isfinite(a).all()
            assert torch.equal(a, b)
    # the Python wrapper of the entry point, on caller-provided outputs
    g = torch.empty(pa.scorer.NPARAMS, dtype=torch.float32, device='cuda')
    lo = torch.empty(300, dtype=torch.float32, device='cuda')
    ls = torch.empty(300, dtype=torch.float32, device='cuda')
    scr.grad_noalloc(src, y, g, lo, ls)
    ref = scr.grad(src, y)[0]
    assert torch.equal(g, ref)
