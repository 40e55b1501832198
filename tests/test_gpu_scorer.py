"""GPU tests of the learned scorer's MFMA kernels against the float64 oracle (oracle/scorer.py,
with rounding='bf16': bf16 at exactly the places the kernels hold bf16 values).

  * forward logits of random tiles;
  * the gradient of every parameter tensor (W1, W2, W3, b1..b3, u, v, c0), per-tile logits and
    loss terms, for a batch of 6 tiles and for 600 tiles (more tiles than work groups, so that
    work groups accumulate several tiles before their partial gradient);
  * Adam against the oracle's Adam;
  * bit-identical gradients on repeated calls (fixed-order reduction);
  * tiles read in place from strided float32 / float16 rows with origins (and an out-of-range
    origin -> NaN) against the same tiles gathered into a contiguous array;
  * tiles read from a dedisperser's tree rows (rank-10, 1024-channel shapes) against the rows copied to the host.

Tolerances: the kernels accumulate in float32 and the oracle in float64; a float32 rounding can
move a value across a bf16 rounding boundary (one bf16 ulp, 2^-8 relative, on rare elements), so
comparisons use a relative bound of 2e-3 of each tensor's scale."""

import sys

import numpy as np
import pytest

from oracle import scorer as sc

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def _params(seed):
    rng = np.random.default_rng(seed)
    theta = sc.init_params(rng)
    p = sc.unpack(theta)
    # biases and head away from zero so that relus switch and every gradient is non-trivial
    for name in ('b1', 'b2', 'b3'):
        p[name][:] = rng.standard_normal(sc.C) * 0.1
    p['u'][:] = rng.standard_normal(sc.C) * 0.5
    p['v'][:] = rng.standard_normal(sc.C) * 0.5
    p['c0'][:] = 0.2
    return sc.pack(p)


def _tiles(rng, n):
    x = rng.standard_normal((n, sc.C0, sc.T))
    # a few bright features so that max-pooling picks distinct samples
    for i in range(n):
        x[i, rng.integers(0, sc.C0, 8), rng.integers(0, sc.T)] += 4.0
    return x.astype(np.float32)


def _close(got, want, rtol, what):
    got = np.asarray(got, dtype=np.float64)
    want = np.asarray(want, dtype=np.float64)
    scale = max(np.max(np.abs(want)), 1e-30)
    err = np.max(np.abs(got - want))
    assert err <= rtol * scale, f'{what}: max error {err:.3g} vs scale {scale:.3g}'
    return err / scale


@pytest.mark.parametrize('n', [5])
def test_forward_matches_oracle(n):
    import torch
    import pirate_amd as pa
    rng = np.random.default_rng(10)
    theta = _params(11)
    X = _tiles(rng, n)
    scr = pa.LearnedScorer(theta)
    logits = scr.forward(torch.from_numpy(X).cuda()).cpu().numpy()
    want, _ = sc.forward(theta, X.astype(np.float64), rounding='bf16')
    log(f'[scorer fwd] gpu {logits} oracle {want}')
    _close(logits, want, 2e-3, 'logits')


@pytest.mark.parametrize('n', [6, 600])
def test_gradient_matches_oracle(n):
    import torch
    import pirate_amd as pa
    rng = np.random.default_rng(20 + n)
    theta = _params(21 + n)
    X = _tiles(rng, n)
    y = (rng.random(n) < 0.5).astype(np.float32)
    scr = pa.LearnedScorer(theta)
    g, logits, loss = scr.grad(torch.from_numpy(X).cuda(), torch.from_numpy(y))
    g = g.cpu().numpy().astype(np.float64)
    loss_o, g_o, logits_o = sc.loss_and_grad(theta, X.astype(np.float64), y, rounding='bf16')
    _close(logits.cpu().numpy(), logits_o, 2e-3, 'logits')
    assert abs(float(loss.mean()) - loss_o) <= 2e-3 * max(1.0, abs(loss_o))
    gu, gou = sc.unpack(g), sc.unpack(g_o)
    for name, _ in sc.SHAPES:
        rel = _close(gu[name], gou[name], 2e-3, f'grad {name} (n={n})')
        log(f'[scorer grad n={n}] {name}: max rel error {rel:.2e}')


def test_gradient_is_deterministic():
    import torch
    import pirate_amd as pa
    rng = np.random.default_rng(30)
    X = torch.from_numpy(_tiles(rng, 700)).cuda()
    y = torch.from_numpy((rng.random(700) < 0.5).astype(np.float32))
    scr = pa.LearnedScorer(_params(31))
    g1 = scr.grad(X, y)[0].clone()
    g2 = scr.grad(X, y)[0].clone()
    assert torch.equal(g1, g2)


def test_adam_matches_oracle():
    import torch
    import pirate_amd as pa
    rng = np.random.default_rng(40)
    theta = _params(41)
    scr = pa.LearnedScorer(theta)
    th = theta.copy()
    m = np.zeros_like(th)
    v = np.zeros_like(th)
    for step in range(1, 4):
        g = rng.standard_normal(sc.NPARAMS).astype(np.float32)
        scr.adam_step(torch.from_numpy(g).cuda(), lr=1e-2)
        th, m, v = sc.adam_step(th, m, v, g.astype(np.float64), step, 1e-2)
    got = scr.params.cpu().numpy().astype(np.float64)
    # float32 parameters: bound by a few float32 ulps of the parameter plus of the step size
    assert np.max(np.abs(got - th) / (np.abs(th) * 1e-6 + 1e-2 * 1e-4)) <= 1.0


@pytest.mark.parametrize('dtype', ['float32', 'float16'])
def test_strided_rows_with_origins(dtype):
    import torch
    import pirate_amd as pa
    from pirate_amd import _lib as L
    import ctypes
    tdt = getattr(torch, dtype)
    rng = np.random.default_rng(50)
    nbeams, nrows, rowlen, lead = 3, 200, 700, 40          # t ranges over [-lead, rowlen - lead)
    rows = torch.from_numpy(rng.standard_normal((nbeams, nrows, rowlen)).astype(np.float32)).to('cuda', tdt)
    origins = np.array([[0, 0, -lead], [2, 136, 700 - lead - 256], [1, 17, 3], [2, 100, 250], [1, 0, 0]],
                       dtype=np.int32)
    bad = np.array([[3, 0, 0], [0, 137, 0], [0, 0, -lead - 1], [0, 0, 700 - lead - 255]], dtype=np.int32)
    allo = np.concatenate([origins, bad])
    s = L.TileSource(ctypes.c_void_p(rows.data_ptr() + lead * rows.element_size()), {'float32': 0, 'float16': 1}[dtype],
                     nbeams, nrows * rowlen, nrows, rowlen, -lead, rowlen - lead, None)
    src = pa.TileSource(s, [rows]).with_origins(allo)
    scr = pa.LearnedScorer(_params(51))
    got = scr.forward(src).cpu().numpy()
    tiles = torch.stack([rows[b, r:r + 64, t + lead:t + lead + 256] for b, r, t in origins]).contiguous()
    want = scr.forward(tiles).cpu().numpy()
    np.testing.assert_array_equal(got[:len(origins)], want)
    assert np.all(np.isnan(got[len(origins):]))


def test_tree_rows_of_a_dedisperser():
    import torch
    import pirate_amd as pa
    from gpu_cases import CASES
    d = dict(CASES['r10_1024ch'], beams_per_gpu=2, beams_per_batch=2)
    plan = pa.DedispersionPlan(pa.DedispersionConfig.from_dict(d))
    dd = pa.GpuDedisperser(plan)
    dd.allocate()
    dd.fill_all_weights(0, np.ones(dd.wt_shape(0), dtype=np.float32))
    gen = torch.Generator(device='cuda')
    gen.manual_seed(60)
    scr = pa.LearnedScorer(_params(61))
    for c in range(2):
        with dd.get_input(c) as arr:
            arr.normal_(generator=gen)
        with dd.get_output(c):
            pass
    src = dd.tree_rows(0)
    st = src.struct
    assert st.nrows == 1024 and st.t_hi == 2048 and st.t_lo == 0 and st.nbeams == 2
    origins = np.array([[0, 0, 0], [1, 960, 2048 - 256], [1, 64, 5], [0, 512, 1000]], dtype=np.int32)
    got = scr.forward(src.with_origins(origins)).cpu().numpy()
    host = dd.debug_copy('sb0').view(np.float32)
    tpad = 32                                              # 2 * max_width of the tree
    rowlen = st.row_stride
    lead = host.size - st.nbeams * st.beam_stride          # element offset of the rows in their buffer
    rows = host[lead:lead + st.nbeams * st.beam_stride].reshape(st.nbeams, st.nrows, rowlen)
    tiles = np.stack([rows[b, r:r + 64, tpad + t:tpad + t + 256] for b, r, t in origins])
    want = scr.forward(torch.from_numpy(np.ascontiguousarray(tiles)).cuda()).cpu().numpy()
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize('dtype', ['float32', 'float16'])
def test_scorer_in_chain(dtype):
    """GpuDedisperser.set_scorer: the chain scores every chunk's tiles of tree 0 right after its
    stage 2; the logits equal LearnedScorer.forward on the same tiles of the tree rows."""
    import torch
    import pirate_amd as pa
    from gpu_cases import CASES
    d = dict(CASES['r10_1024ch'], beams_per_gpu=4, beams_per_batch=2, dtype=dtype)
    plan = pa.DedispersionPlan(pa.DedispersionConfig.from_dict(d))
    dd = pa.GpuDedisperser(plan)
    dd.allocate()
    dd.fill_host_weights(0)
    scr = pa.LearnedScorer(_params(71))
    dd.set_scorer(scr, 0)
    gen = torch.Generator(device='cuda')
    gen.manual_seed(72)
    grid = np.array([[b, 64 * i, 256 * j] for b in range(2) for i in range(16) for j in range(8)], dtype=np.int32)
    for seq in range(6):
        with dd.get_input(seq) as arr:
            arr.normal_(generator=gen)
        with dd.get_output(seq) as out:
            assert tuple(out.out_score.shape) == (2, 16, 8)
            got = out.out_score.cpu().numpy().reshape(-1)
            src = dd.tree_rows(0)
            esize = 4 if dtype == 'float32' else 2
            src.struct.base = src.struct.base + out.ibeam * src.struct.beam_stride * esize
            want = scr.forward(src.with_origins(grid)).cpu().numpy()
        np.testing.assert_array_equal(got, want)
        assert np.all(np.isfinite(got))
