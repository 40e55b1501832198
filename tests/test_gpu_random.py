"""Randomized GPU parity, as the reference's GpuDedisperser::test_random() does it
(src_lib/GpuDedisperser.cpp:1531-1563): each seed draws a dedispersion config (dtype, tree
rank, zones, primary trees, early triggers, subband counts, peak-finder widths, output and
weight downsampling, time chunk and beam batching; tests/random_configs.py), a chunk count
that covers the longest tree delay, and output / weight ring depths; the GPU chain must then
match the oracle bit for bit (float32), or the oracle that rounds to the storage type where
the chain stores (float16, bfloat16), on out_max and out_argmax of every tree and sequence.
"""

import sys

import numpy as np
import pytest

from gpu_harness import run_and_compare
from random_configs import draw_supported

pytestmark = pytest.mark.gpu

NSEEDS = 300


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def run_shape(cfg, seed):
    """(nchunks, nbatches_out, nbatches_wt) drawn like test_random(): enough chunks for the
    longest delay to pass through, ring depths between the active batch count and the
    number of sequences."""
    rng = np.random.default_rng(1000 + seed)
    ntree = 2 ** cfg['toplevel_tree_rank']
    nt = cfg['time_samples_per_chunk']
    npri = len(cfg['primary_trees'])
    min_nchunks = (ntree << (npri - 1)) // nt + 2
    nchunks = int(rng.integers(min_nchunks, min_nchunks + 3))
    nb = cfg['beams_per_gpu'] // cfg['beams_per_batch']
    lo = cfg['num_active_batches']
    hi = max(lo, min(nchunks * nb, 16))
    nbo = int(rng.integers(lo, hi + 1))
    nbw = int(rng.integers(lo, 2 * nbo + 1))
    return nchunks, nbo, nbw


@pytest.mark.parametrize('seed', range(NSEEDS))
def test_gpu_random_config_bit_exact(seed):
    cfg, nrej = draw_supported(seed)
    nchunks, nbo, nbw = run_shape(cfg, seed)
    trees = [(p['num_early_triggers'], p['max_width'], p['time_downsampling'], p['wt_dm_downsampling'],
              p['wt_time_downsampling']) for p in cfg['primary_trees']]
    log(f'[gpu random] seed={seed} dtype={cfg["dtype"]} rank={cfg["toplevel_tree_rank"]} nt={cfg["time_samples_per_chunk"]} '
        f'zones={cfg["zone_nfreq"]} beams={cfg["beams_per_gpu"]}/{cfg["beams_per_batch"]}/{cfg["num_active_batches"]} '
        f'sc={cfg["frequency_subband_counts"]} trees(et,W,Dout,wdm,wt)={trees} nchunks={nchunks} '
        f'nbatches_out={nbo} nbatches_wt={nbw}')
    bad, dd, plan = run_and_compare(cfg, nchunks, seed=seed, log=log, nbatches_out=nbo, nbatches_wt=nbw)
    assert not bad, f'{len(bad)} mismatching (seq, tree) outputs, first: {bad[0]}'
