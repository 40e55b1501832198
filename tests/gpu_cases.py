"""Shared configurations for GPU parity tests, smoke() and benchmarks.

Each case is a dedispersion config in the reference's YAML grammar (as a dict), small enough
for the CPU oracle to check in seconds, chosen to cover: one/several primary trees, early
triggers, subband searches with level-0, staggered and full-band multiplets, several
frequency zones, beam batching (several batches), and the 1-beam 1024-channel rank-10 shape.
"""

BASE = dict(time_sample_ms=1.0, dtype='float32', future_write_max_samples=0)


def pt(num_et, max_width, wt_dm, wt_t):
    return dict(num_early_triggers=num_et, max_width=max_width, dm_downsampling=0, time_downsampling=0,
                wt_dm_downsampling=wt_dm, wt_time_downsampling=wt_t)


CASES = {
    # Reference-toy-like: 3 primary trees, early triggers, subbands [2,2,1], 2 batches of 2 beams.
    'toy': dict(BASE, zone_nfreq=[640], zone_freq_edges=[400, 800], toplevel_tree_rank=8, time_samples_per_chunk=256,
                beams_per_gpu=4, beams_per_batch=2, num_active_batches=2, frequency_subband_counts=[2, 2, 1],
                primary_trees=[pt(0, 16, 16, 16), pt(1, 8, 16, 16), pt(1, 4, 16, 16)]),
    # One primary tree, no subbands (the structure of the rank-10 1024-channel case with a shorter chunk).
    'one_tree': dict(BASE, zone_nfreq=[1024], zone_freq_edges=[400, 800], toplevel_tree_rank=10,
                     time_samples_per_chunk=256, beams_per_gpu=1, beams_per_batch=1, num_active_batches=1,
                     frequency_subband_counts=[1], primary_trees=[pt(0, 16, 64, 64)]),
    # Four primary trees over three zones, subbands [3,3,1], early triggers.
    'zones4': dict(BASE, zone_nfreq=[200, 300, 500], zone_freq_edges=[300, 500, 800, 1500], toplevel_tree_rank=9,
                   time_samples_per_chunk=512, beams_per_gpu=2, beams_per_batch=1, num_active_batches=1,
                   frequency_subband_counts=[3, 3, 1],
                   primary_trees=[pt(0, 8, 32, 32), pt(1, 4, 32, 32), pt(0, 4, 32, 32), pt(2, 2, 16, 16)]),
    # Wide subband search (every level populated), Wmax 32, rank 12.
    'subbands': dict(BASE, zone_nfreq=[4096], zone_freq_edges=[300, 1500], toplevel_tree_rank=12,
                     time_samples_per_chunk=512, beams_per_gpu=2, beams_per_batch=2, num_active_batches=1,
                     frequency_subband_counts=[4, 5, 3, 1], primary_trees=[pt(0, 32, 64, 64), pt(1, 16, 64, 64)]),
    # Input channels much narrower than tree channels (2048 ch onto 256 tree channels, up to
    # ~10 per tree channel): the gridding kernel stores the gridded rows for stage 1.
    'wide_ch': dict(BASE, zone_nfreq=[2048], zone_freq_edges=[400, 800], toplevel_tree_rank=8,
                    time_samples_per_chunk=256, beams_per_gpu=2, beams_per_batch=2, num_active_batches=1,
                    frequency_subband_counts=[1], primary_trees=[pt(0, 8, 16, 16), pt(0, 4, 16, 16)]),
    # float32 chunks that are a multiple of 32 but not of 64 samples: primary tree 0 (192
    # samples) ends in a full tile, primary tree 1 (96 samples) in a half tile; subband
    # levels with staggered streams put stage 2 on the general kernel.
    'half_tile': dict(BASE, zone_nfreq=[300, 200], zone_freq_edges=[400, 600, 800], toplevel_tree_rank=8,
                      time_samples_per_chunk=192, beams_per_gpu=2, beams_per_batch=1, num_active_batches=1,
                      frequency_subband_counts=[2, 2, 1], primary_trees=[pt(1, 8, 16, 16), pt(0, 4, 16, 16)]),
    # Chunks shorter than the peak finder's history + final lags (64 samples, max_width 32:
    # 64 history samples), on a rank-9 tree whose stage 2 runs on the register kernel.
    'short_chunk': dict(BASE, zone_nfreq=[512], zone_freq_edges=[400, 800], toplevel_tree_rank=9,
                        time_samples_per_chunk=64, beams_per_gpu=2, beams_per_batch=2, num_active_batches=1,
                        frequency_subband_counts=[0, 3, 1], primary_trees=[pt(0, 32, 32, 64), pt(1, 16, 32, 32)]),
    # Rank-8 stage 2 (toplevel rank 15, as the reference's chime configs): 9 chunks of 4096
    # samples cover the largest ring lag (255 * 127 = 32385 samples).
    'rank15': dict(BASE, zone_nfreq=[1024], zone_freq_edges=[400, 800], toplevel_tree_rank=15,
                   time_samples_per_chunk=4096, beams_per_gpu=1, beams_per_batch=1, num_active_batches=1,
                   frequency_subband_counts=[0, 0, 0, 0, 1], primary_trees=[pt(0, 16, 64, 64)]),
    # Rank-8 stage 1 and stage 2 (toplevel rank 16, as the reference's chord configs), two
    # primary trees.
    'rank16': dict(BASE, zone_nfreq=[2048], zone_freq_edges=[400, 800], toplevel_tree_rank=16,
                   time_samples_per_chunk=1024, beams_per_gpu=1, beams_per_batch=1, num_active_batches=1,
                   frequency_subband_counts=[1], primary_trees=[pt(0, 8, 64, 64), pt(0, 4, 64, 64)]),
    # Rank-8 stage 2 with a subband search (chime_sb2-like counts [0,5,7,3,1]) and two early
    # triggers (ranks 7 and 6 on the LDS-tile kernel): the two-launch split kernel. (The
    # 'rank15' case covers every ring lag; this one covers the split kernel's passes,
    # histories and emissions over a few chunks.)
    'rank15_sb': dict(BASE, zone_nfreq=[1024], zone_freq_edges=[400, 800], toplevel_tree_rank=15,
                      time_samples_per_chunk=2048, beams_per_gpu=1, beams_per_batch=1, num_active_batches=1,
                      frequency_subband_counts=[0, 5, 7, 3, 1], primary_trees=[pt(2, 16, 64, 64)]),
    # Rank-8 stages 1 and 2 with chord_sb2-like counts [5,9,7,3,1] (level-0 copies and every
    # staggered level), two primary trees, float32.
    'rank16_sb': dict(BASE, zone_nfreq=[2048], zone_freq_edges=[400, 800], toplevel_tree_rank=16,
                      time_samples_per_chunk=1024, beams_per_gpu=1, beams_per_batch=1, num_active_batches=1,
                      frequency_subband_counts=[5, 9, 7, 3, 1], primary_trees=[pt(0, 8, 64, 64), pt(0, 4, 64, 64)]),
    # Five and six primary trees: the gridding kernel's downsampling groups of 16 and 32
    # samples leave 2 and 1 tiles per channel pair, so one workgroup takes 2 and 4 pairs.
    'pri5': dict(BASE, zone_nfreq=[300], zone_freq_edges=[400, 800], toplevel_tree_rank=8,
                 time_samples_per_chunk=2048, beams_per_gpu=2, beams_per_batch=2, num_active_batches=1,
                 frequency_subband_counts=[1], primary_trees=[pt(0, 16, 16, 16), pt(0, 8, 16, 16), pt(0, 4, 16, 16),
                                                              pt(0, 2, 16, 16), pt(0, 1, 16, 16)]),
    'pri6': dict(BASE, zone_nfreq=[200], zone_freq_edges=[400, 800], toplevel_tree_rank=7,
                 time_samples_per_chunk=1024, beams_per_gpu=1, beams_per_batch=1, num_active_batches=1,
                 frequency_subband_counts=[1], primary_trees=[pt(0, 8, 16, 16), pt(0, 8, 16, 16), pt(0, 4, 16, 16),
                                                              pt(0, 2, 16, 16), pt(0, 2, 16, 16), pt(0, 1, 16, 16)]),
    # 1 beam, 1024 channels 400-800 MHz, rank 10, one tree, chunks of 2048.
    'r10_1024ch': dict(BASE, zone_nfreq=[1024], zone_freq_edges=[400, 800], toplevel_tree_rank=10,
                       time_samples_per_chunk=2048, beams_per_gpu=1, beams_per_batch=1, num_active_batches=1,
                       frequency_subband_counts=[1], primary_trees=[pt(0, 16, 64, 64)]),
}

# Cases whose dense CPU checks (e.g. the analytic-variance oracle) would take minutes.
LARGE = {'rank15', 'rank16', 'rank15_sb', 'rank16_sb'}

# chord16k (bench workload): 16 beams x 16384 channels 300-1500 MHz, rank 14, four
# primary trees (1/2/4/8 ms), chunks of 2048 samples, float32. Not fixed by the workload and
# chosen here: no subband search ([1]), max_width 16/8/4/2 (16 ms at every downsampling),
# weight downsampling 64x64.
CHORD16K = dict(BASE, zone_nfreq=[16384], zone_freq_edges=[300, 1500], toplevel_tree_rank=14,
                time_samples_per_chunk=2048, beams_per_gpu=16, beams_per_batch=16, num_active_batches=1,
                frequency_subband_counts=[1],
                primary_trees=[pt(0, 16, 64, 64), pt(0, 8, 64, 64), pt(0, 4, 64, 64), pt(0, 2, 64, 64)])


def chord16k_with_beams(nbeams, batch=None):
    d = dict(CHORD16K)
    d['beams_per_gpu'] = nbeams
    d['beams_per_batch'] = batch or nbeams
    return d


# Batched heterogeneous beams: three beam groups of 32 beams at tree ranks 12/13/14
# (4096/8192/16384 channels over 300-1500 MHz) with 3/4/5 primary trees (downsampling levels),
# float16 storage, chunks of 2048 samples, run as one GpuDedisperserGroup (each group's chain on
# a branch stream of its own per chunk). Not fixed by the workload and chosen here: no subband search ([1]), max_width 16 ms
# at every downsampling level, weight downsampling 64x64.
def batched_group(rank, nbeams, dtype='float16'):
    nfreq = {12: 4096, 13: 8192, 14: 16384}[rank]
    npri = {12: 3, 13: 4, 14: 5}[rank]
    widths = [16, 8, 4, 2, 1][:npri]
    return dict(BASE, dtype=dtype, zone_nfreq=[nfreq], zone_freq_edges=[300, 1500], toplevel_tree_rank=rank,
                time_samples_per_chunk=2048, beams_per_gpu=nbeams, beams_per_batch=nbeams, num_active_batches=1,
                frequency_subband_counts=[1], primary_trees=[pt(0, w, 64, 64) for w in widths])


def batched_groups(nbeams=32, dtype='float16'):
    return [batched_group(r, nbeams, dtype) for r in (12, 13, 14)]
