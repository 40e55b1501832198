"""Bit-exact parity on the exact CHORD-like float32 shapes (16384 channels over 300-1500 MHz,
1 ms samples, toplevel tree rank 14, 4 primary trees at 1/2/4/8 ms, chunks of 2048 samples),
2 beams in one batch, against the CPU oracle over 66 chunks.

66 chunks cover every lag of every ring row and wrap every ring: the longest stage-2 lag is
that of primary tree 3 (stage-2 ranks amb 7, dd 6): (2^6 - 1) * (2^7 - 1 + 2^7) = 16065
samples of 256 per chunk, i.e. 63 chunks, plus one chunk of ring capacity. out_max and the
argmax tokens of all four trees are compared with == at every chunk. Each beam runs in an
oracle instance of its own on a worker thread (about 0.6 s per beam-chunk on one core).
"""

import sys

import numpy as np
import pytest

from gpu_cases import chord16k_with_beams
from gpu_harness import OracleWorker, random_weights

pytestmark = pytest.mark.gpu

NCHUNKS = 66


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def longest_lag_chunks(plan):
    n = 0
    for t in plan.trees:
        p = t.primary_tree_index
        lag = (2**t.dd_rank - 1) * (2**t.amb_rank - 1 + (2**t.amb_rank if p > 0 else 0))
        n = max(n, -(-lag // t.nt_ds) + 1)
    return n


@pytest.mark.timeout(900)
def test_chord16k_shapes_bit_exact_vs_oracle():
    import torch
    import pirate_amd as pa
    d = chord16k_with_beams(2)
    plan = pa.DedispersionPlan(pa.DedispersionConfig.from_dict(d))
    assert plan.ntrees == 4 and plan.nfreq == 16384 and plan.nt_in == 2048
    assert longest_lag_chunks(plan) <= NCHUNKS, longest_lag_chunks(plan)
    log(f'[gpu chord16k] 2 beams x 16384 ch, rank 14, 4 trees, T=2048, float32: {NCHUNKS} chunks vs oracle')
    dd = pa.GpuDedisperser(plan)
    dd.allocate()
    rng = np.random.default_rng(2)
    W = [random_weights(rng, dd, it) for it in range(plan.ntrees)]
    for it in range(plan.ntrees):
        dd.fill_all_weights(it, W[it])
    assert dd.nbatches == 1 and dd.nbatches_wt == 1   # weight slot 0 at every chunk
    workers = []
    for b in range(2):
        w = OracleWorker(dict(d), [W[it][0, b:b + 1] for it in range(plan.ntrees)], 'float32', f'b{b}')
        w.start()
        workers.append(w)
    gen = torch.Generator(device='cuda')
    gen.manual_seed(2)
    for ichunk in range(NCHUNKS):
        with dd.get_input(ichunk) as arr:
            arr.uniform_(-1, 1, generator=gen)
            x = arr.cpu().numpy()
        with dd.get_output(ichunk) as out:
            for b, w in enumerate(workers):
                gm = [m[b:b + 1].cpu().numpy() for m in out.out_max]
                ga = [a[b:b + 1].cpu().numpy().view(np.uint32) for a in out.out_argmax]
                w.q.put((ichunk, x[b:b + 1], gm, ga))
        if ichunk % 8 == 0:
            log(f'[gpu chord16k] chunk {ichunk} launched; oracle checked {[w.nchecked for w in workers]}, '
                f'mismatches {sum(len(w.bad) for w in workers)}')
    for w in workers:
        w.q.put(None)
    for w in workers:
        w.join()
        assert w.err is None, w.err
        assert w.nchecked == NCHUNKS
    bad = [r for w in workers for r in w.bad]
    log(f'[gpu chord16k] done: {NCHUNKS} chunks x 2 beams x 4 trees, {len(bad)} mismatching (chunk, tree) records')
    assert not bad, bad[:5]
