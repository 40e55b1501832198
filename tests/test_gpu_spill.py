"""Lag-ring rows spilled to pinned host memory (config max_gpu_clag), against the oracle.

The reference keeps ring segments whose chunk lag exceeds max_gpu_clag in host memory zones
and copies them back before the stage-2 kernel needs them (MegaRingbuf.hpp:45-58,
GpuDedisperser.cpp:628-707, 1108-1148). Here a spilled row travels through a per-row host
delay of k whole chunks (one 2-D copy per k-group and direction per batch) and a scatter
kernel; the outputs must not change: bit-exact against the oracle (which has no ring at
all), with the run long enough for the largest lag to pass through the host, for float32
and float16 storage, several primary trees, early triggers and beam batches.
"""

import sys

import pytest

from gpu_cases import BASE, pt
from gpu_harness import run_and_compare

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


SPILL = {
    # rank 10, chunks of 128 samples: stage-2 lags up to ~15 chunks
    'clag0_fp32': (dict(BASE, zone_nfreq=[700, 300], zone_freq_edges=[400, 600, 800], toplevel_tree_rank=10,
                        time_samples_per_chunk=128, beams_per_gpu=2, beams_per_batch=1, num_active_batches=1,
                        frequency_subband_counts=[1], primary_trees=[pt(0, 8, 32, 32)], max_gpu_clag=0), 18),
    'clag2_et_fp32': (dict(BASE, zone_nfreq=[1024], zone_freq_edges=[400, 800], toplevel_tree_rank=10,
                           time_samples_per_chunk=256, beams_per_gpu=4, beams_per_batch=2, num_active_batches=2,
                           frequency_subband_counts=[2, 2, 1], primary_trees=[pt(1, 8, 32, 32), pt(1, 4, 32, 32)],
                           max_gpu_clag=2), 12),
    'clag1_fp16': (dict(BASE, dtype='float16', zone_nfreq=[1024], zone_freq_edges=[400, 800], toplevel_tree_rank=10,
                        time_samples_per_chunk=256, beams_per_gpu=2, beams_per_batch=2, num_active_batches=1,
                        frequency_subband_counts=[0, 1, 1], primary_trees=[pt(0, 16, 32, 32), pt(0, 8, 32, 32)],
                        max_gpu_clag=1), 10),
}


@pytest.mark.parametrize('name', sorted(SPILL))
def test_spilled_ring_bit_exact(name):
    cfg, nchunks = SPILL[name]
    log(f'[gpu spill] case={name} nchunks={nchunks}')
    bad, dd, plan = run_and_compare(cfg, nchunks, seed=3, log=log)
    assert not bad, f'{len(bad)} mismatching (seq, tree) outputs, first: {bad[0]}'
    assert dd.host_bytes() > 0, 'nothing was spilled: the case does not test the host path'
    log(f'[gpu spill] {name}: {dd.host_bytes() / 2**20:.1f} MiB host, {dd.gpu_bytes() / 2**20:.1f} MiB device')


def test_spill_moves_ring_memory_to_host():
    """Same config with and without spilling: device bytes fall, host bytes appear."""
    import pirate_amd as pa
    cfg, _ = SPILL['clag0_fp32']
    sizes = {}
    for clag in (0, 10000):
        plan = pa.DedispersionPlan(pa.DedispersionConfig.from_dict(dict(cfg, max_gpu_clag=clag)))
        dd = pa.GpuDedisperser(plan)
        dd.allocate()
        sizes[clag] = (dd.gpu_bytes(), dd.host_bytes())
        del dd
    log(f'[gpu spill] (device, host) bytes: spill {sizes[0]}, no spill {sizes[10000]}')
    assert sizes[10000][1] == 0
    assert sizes[0][1] > 0 and sizes[0][0] < sizes[10000][0]
