"""bench.py's contract, checked on the CPU: more than one GPU is refused with a message, the
roofline's per-launch HBM traffic is read from the committed rocprofv3 summary, and the
algorithmic byte counts that the roofline and the chain rate divide by the measured times
equal the counts worked out from the sizes of the chord16k workload (16 beams x 16384
channels, tree rank 14, 4 primary trees at 1/2/4/8 ms, 2048-sample chunks, float32)."""

import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench                                   # noqa: E402
from gpu_cases import CHORD16K                 # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, 'bench.py'), '--gpus', '2'],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 2
    assert 'one GPU' in r.stderr
    assert r.stdout == ''


def test_traffic_comes_from_the_committed_profile():
    with open(os.path.join(ROOT, 'profiles', 'chain_chord16k_fp32_rocprof.json')) as f:
        rec = json.load(f)['kernels']['grid_lds']
    assert bench.measured_traffic('grid_lds') == int(rec['hbm_read_bytes'] + rec['hbm_write_bytes'])
    assert bench.measured_traffic('no_such_kernel') is None


def test_algorithmic_bytes_of_the_chord16k_step():
    import pirate_amd as pa
    from pirate_amd.resources import chain_alg_bytes, kernel_alg_bytes
    plan = pa.DedispersionPlan(pa.DedispersionConfig.from_dict(CHORD16K))
    beams, S = 16, 4
    # gridding: reads 16384 channels x 2048 samples, writes 2^14 gridded rows x 2048 and, for
    # primary trees 1-3, 2^13 downsampled rows of 1024, 512 and 256 samples
    per_beam = 16384 * 2048 + 16384 * 2048 + 8192 * (1024 + 512 + 256)
    assert kernel_alg_bytes(plan, 'grid_lds') == beams * per_beam * S == 5_234_491_392
    # stage 1 of tree 0 reads and writes 2^14 rows of 2048 samples; tree 1: 2^13 rows of 1024
    assert kernel_alg_bytes(plan, 'stage1_p0') == beams * 2 * 16384 * 2048 * S
    assert kernel_alg_bytes(plan, 'stage1_p1') == beams * 2 * 8192 * 1024 * S
    # the two-stage algorithm: input read, lag ring written and read once, and the outputs
    # (float32 max + uint32 argmax) of trees of 1024 x 128, 512 x 64, 512 x 32, 512 x 16 bins
    ring = 16384 * 2048 + 8192 * (1024 + 512 + 256)
    outs = (1024 * 128 + 512 * 64 + 512 * 32 + 512 * 16) * 8
    assert chain_alg_bytes(plan) == beams * (16384 * 2048 * S + 2 * ring * S + outs) == 8_345_616_384
