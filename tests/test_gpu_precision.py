"""Reduced-precision storage (float16 / bfloat16 storage, float32 arithmetic) against the
FP64 oracle: every (DM, time) output of every tree, on unit-variance noise with bare-kernel
weights, so out_max is in SNR units, without and with narrowband RFI (1% of the channels at a
constant +5 sigma, seeded). The inputs are rounded to the storage type first and the FP64
oracle sees the same rounded values, so the errors below are those of the chain. Errors are
reported per (DM quarter, peak-finding profile) trial class as well.

Per-level scaling: every tree pass scales by 1/sqrt(2), so every level's values keep unit
variance (white-noise input) and float16 stays far inside its range; a power-of-two scale
cannot change a floating-point type's relative rounding, which alone sets the error below.

Bound: 1% of SNR (relative to max(|SNR|, 1)) for the maximum, 2e-3 / 1e-2
for the rms. float16 meets it (measured max ~3e-3). bfloat16 has an 8-bit significand: each
of the ~4 storage roundings on the path is up to 2^-9 relative, and the measured maximum is
~2e-2, so its maximum-error case is marked xfail (its rms meets 1e-2). The GPU's argmax
candidate, re-evaluated in FP64, must be within the same bound of the FP64 maximum.
"""

import pytest

from gpu_cases import CASES
from gpu_harness import run_precision

pytestmark = pytest.mark.gpu

TOL = {'float16': dict(max=1e-2, rms=2e-3), 'bfloat16': dict(max=1e-2, rms=1e-2)}
BF16_MAX = pytest.mark.xfail(reason='bfloat16 storage: 8-bit significand, measured max SNR error ~2e-2 > 1e-2',
                             strict=False)


def _check(stats, dtype, check_max=True):
    tol = TOL[dtype]
    for it, s in stats.items():
        assert s['rms'] <= tol['rms'], (it, s)
    for it, s in stats.items():
        if check_max:
            assert s['max'] <= tol['max'], (it, s)
            assert s['tok_max'] <= tol['max'], (it, s)


@pytest.mark.parametrize('dtype', ['float16', pytest.param('bfloat16', marks=BF16_MAX)])
@pytest.mark.parametrize('name', ['r10_1024ch', 'toy', 'zones4'])
def test_reduced_precision_vs_fp64_oracle(name, dtype):
    nchunks = 2 if name == 'r10_1024ch' else 4
    _check(run_precision(CASES[name], dtype, nchunks, seed=7), dtype)


@pytest.mark.parametrize('dtype', ['float16', pytest.param('bfloat16', marks=BF16_MAX)])
def test_reduced_precision_with_rfi_vs_fp64_oracle(dtype):
    """The 1024-channel rank-10 configuration with narrowband RFI (1% of the channels at +5
    sigma), 3 chunks."""
    _check(run_precision(CASES['r10_1024ch'], dtype, 3, seed=8, rfi=True), dtype)


def test_bfloat16_rms_vs_fp64_oracle():
    """bfloat16: the rms error bound (1e-2) holds, with and without RFI."""
    _check(run_precision(CASES['r10_1024ch'], 'bfloat16', 2, seed=7), 'bfloat16', check_max=False)
    _check(run_precision(CASES['r10_1024ch'], 'bfloat16', 2, seed=8, rfi=True), 'bfloat16', check_max=False)


@pytest.mark.parametrize('dtype', ['float16', 'bfloat16'])
@pytest.mark.parametrize('name', ['toy', 'zones4', 'subbands', 'r10_1024ch', 'pri5'])
def test_16bit_storage_bit_exact_vs_rounding_oracle(name, dtype):
    """With 16-bit storage the GPU rounds to the storage type only where it stores (gridded
    rows, downsampled rows, lag ring, subband rows, out_max) and computes in float32 in
    between; the oracle rounding at the same points gives identical out_max and argmax."""
    from gpu_cases import CASES
    from gpu_harness import run_and_compare
    bad, dd, plan = run_and_compare(dict(CASES[name], dtype=dtype), 3, seed=23, log=print)
    assert not bad, f'{len(bad)} mismatching (seq, tree) outputs, first: {bad[0]}'


@pytest.mark.parametrize('name,dtype', [('rank16', 'float16'), ('rank16_sb', 'bfloat16'), ('rank15_sb', 'float16')])
def test_16bit_storage_bit_exact_rank8(name, dtype):
    """The rank-8 kernels (register tree, and the two-launch split for subband searches with
    its float32 scratch) with 16-bit storage, against the rounding oracle: 2 chunks."""
    from gpu_cases import CASES
    from gpu_harness import run_and_compare
    print(f'[gpu precision] case={name} dtype={dtype}', flush=True)
    bad, dd, plan = run_and_compare(dict(CASES[name], dtype=dtype), 2, seed=29, log=print)
    assert not bad, f'{len(bad)} mismatching (seq, tree) outputs, first: {bad[0]}'
