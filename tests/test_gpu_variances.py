"""Optional float64 out_var of the GPU peak finder (enable_variances): per chunk, the mean
square of each (multiplet, profile)'s weighted candidates, as the reference peak finder's
out_var (include/pirate/PeakFindingKernel.hpp:163-168, src_lib/PeakFindingKernel.cu:244-249,
390-447; ReferenceDedisperserBase::Params::enable_variances, include/pirate/Dedisperser.hpp:425).
The GPU sums the same float32 candidates in the same order in float64, so out_var must equal
the oracle's exactly; out_max and out_argmax stay bit-exact with the feature on."""

import sys

import numpy as np
import pytest

from gpu_cases import CASES
from gpu_harness import run_and_compare

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize('name', ['toy', 'zones4', 'r10_1024ch'])
def test_gpu_out_var_equals_oracle(name):
    log(f'[gpu variances] case={name}')
    bad, dd, plan = run_and_compare(CASES[name], 3, seed=17, log=log, variances=True)
    assert not bad, f'{len(bad)} mismatching records, first: {bad[0]}'


def test_out_var_requires_flag():
    import pirate_amd as pa
    plan = pa.DedispersionPlan(pa.DedispersionConfig.from_dict(CASES['one_tree']))
    dd = pa.GpuDedisperser(plan)
    dd.allocate()
    for it in range(plan.ntrees):
        dd.fill_all_weights(it, np.ones(dd.wt_shape(it), np.float32))
    with dd.get_input(0) as arr:
        arr.zero_()
    with dd.get_output(0) as out:
        assert out.out_var is None
    import ctypes
    from pirate_amd import _lib as L
    vp = (ctypes.c_void_p * plan.ntrees)()
    with pytest.raises(L.PddError, match='enable_variances'):
        L.check(L.lib().pdd_output_variances(dd._h, 0, vp, plan.ntrees))
