"""GPU dedisperser vs the CPU oracle (reference ReferenceDedisperser, sophistication 1).

Bar: bit-exact out_max and out_argmax for float32 storage, on several chunks (exercising every
kernel's persistent state), several beam batches, early triggers, subband searches and
several primary trees. Every mismatch record also reports whether the GPU tokens evaluate
(oracle eval_tokens) to the GPU maxima, to localize a failure to the max or the argmax path.
"""

import sys

import numpy as np
import pytest

from gpu_cases import CASES
from gpu_harness import run_and_compare

pytestmark = pytest.mark.gpu

NCHUNKS = {'toy': 4, 'one_tree': 3, 'zones4': 3, 'subbands': 3, 'wide_ch': 3, 'half_tile': 5, 'short_chunk': 12,
           'rank15': 9, 'rank16': 3, 'rank15_sb': 4, 'rank16_sb': 2, 'r10_1024ch': 4, 'pri5': 5, 'pri6': 7}


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize('name', ['toy', 'one_tree', 'zones4', 'subbands', 'wide_ch', 'half_tile', 'short_chunk', 'rank15',
                                  'rank16', 'rank15_sb', 'rank16_sb', 'r10_1024ch', 'pri5', 'pri6'])
def test_gpu_bit_exact_vs_oracle(name):
    log(f'[gpu parity] case={name} nchunks={NCHUNKS[name]}')
    bad, dd, plan = run_and_compare(CASES[name], NCHUNKS[name], seed=11, log=log)
    assert not bad, f'{len(bad)} mismatching (seq, tree) outputs, first: {bad[0]}'


def test_gpu_gaussian_input_r10_1024ch():
    log('[gpu parity] case=r10_1024ch gaussian input')
    bad, dd, plan = run_and_compare(CASES['r10_1024ch'], 2, seed=5, input_kind='normal', log=log)
    assert not bad, f'first mismatch: {bad[0]}'
