"""GpuDedisperser call-order semantics on the GPU: cursors, error latching, stop(),
several consumers, and back-pressure of release_input on unreleased outputs.

Reference: include/pirate/Dedisperser.hpp:170-219 (acquire/release contract, stop),
src_lib/GpuDedisperser.cpp:742-984 (cursor checks, consumer cursors, back-pressure
at :1080-1081), pirate_frb/GpuDedisperser.py:85-89.
"""

import threading
import time

import numpy as np
import pytest

from gpu_cases import CASES

pytestmark = pytest.mark.gpu


def make(name='toy', overrides=None, **kw):
    import pirate_amd as pa
    cfg = pa.DedispersionConfig.from_dict(dict(CASES[name], **(overrides or {})))
    plan = pa.DedispersionPlan(cfg)
    dd = pa.GpuDedisperser(plan, **kw)
    dd.allocate()
    for it in range(plan.ntrees):
        dd.fill_all_weights(it, np.ones(dd.wt_shape(it), np.float32))
    return dd


def push(dd, seq):
    with dd.get_input(seq) as arr:
        arr.zero_()


def test_input_out_of_order_latches_first_error():
    dd = make()
    with pytest.raises(RuntimeError, match='expected seq_id=0'):
        dd._acquire_input(1, 0)
    # every later call reports the first error
    with pytest.raises(RuntimeError, match='expected seq_id=0'):
        dd._acquire_input(0, 0)
    with pytest.raises(RuntimeError, match='expected seq_id=0'):
        dd._acquire_output(0, 0, 0)


def test_stop_latches_reason():
    dd = make()
    push(dd, 0)
    dd.stop('operator halt')
    with pytest.raises(RuntimeError, match='operator halt'):
        dd._acquire_input(1, 0)
    with pytest.raises(RuntimeError, match='operator halt'):
        dd._acquire_output(0, 0, 0)


def test_release_output_out_of_order():
    dd = make()
    push(dd, 0)
    out = dd._acquire_output(0, 0, 0, sync=True)
    assert out.ichunk_zero_based == 0
    with pytest.raises(RuntimeError, match='expected seq_id=0'):
        dd._release_output(0, 1, 0)


def test_negative_weights_rejected():
    dd = make()
    w = np.ones(dd.wt_shape(0), np.float32)
    w.flat[3] = -1.0
    with pytest.raises(RuntimeError, match='finite and >= 0'):
        dd.fill_all_weights(0, w)


def test_two_consumers_and_back_pressure():
    import torch
    # one active batch: nbatches_out may not be below num_active_batches (as in the reference)
    dd = make('toy', overrides=dict(num_active_batches=1), num_consumers=2, nbatches_out=1)
    push(dd, 0)
    dd._acquire_input(1, 0)
    res = {'done': False, 'err': None}

    def release():
        try:
            dd._release_input_and_launch_dd_kernels(1, 0)
            res['done'] = True
        except Exception as e:       # reported below
            res['err'] = e

    # Output slot 0 (nbatches_out = 1) is held until both consumers release seq 0.
    th = threading.Thread(target=release, daemon=True)
    th.start()
    try:
        time.sleep(0.5)
        assert not res['done'], 'release_input_and_launch did not wait for unreleased outputs'
        outs = [dd._acquire_output(c, 0, 0, sync=True) for c in range(2)]
        assert torch.equal(outs[0].out_max[0].cpu(), outs[1].out_max[0].cpu())
        dd._release_output(0, 0, 0)
        time.sleep(0.2)
        assert not res['done'], 'release_input_and_launch ran before every consumer released'
        dd._release_output(1, 0, 0)
        th.join(timeout=30)
        assert res['err'] is None and res['done']
    finally:
        if th.is_alive():
            dd.stop('test failed: waking the blocked release')   # stop() wakes every waiter
            th.join(timeout=10)
    for c in range(2):
        dd._acquire_output(c, 1, 0, sync=True)
        dd._release_output(c, 1, 0)


def test_fill_host_weights_bare_and_variances():
    from oracle import oracle as orc
    dd = make('toy')
    for it in range(dd.ntrees):
        nb, B, D, T, P, N = dd.wt_shape(it)
        w = dd.fill_host_weights(it)
        assert np.array_equal(w, np.broadcast_to(orc.bare_weights(B, D, T, P, N), w.shape))
        var = np.random.default_rng(it).uniform(0.5, 2.0, size=(N, D, P))
        w = dd.fill_host_weights(it, var)
        want = (np.float32(1) / np.sqrt(var.astype(np.float32))).transpose(1, 2, 0)      # (D, P, N)
        assert np.array_equal(w[0, 0, :, 0], want)
        w = dd.fill_host_weights(it, var, randomize=True, rng=np.random.default_rng(1))
        assert np.all(w >= 0) and np.all(w <= np.broadcast_to(want[None, None, :, None], w.shape))


def test_fill_analytic_weights_equals_explicit_variance_weights():
    """fill_analytic_weights(fv) uploads 1/sqrt(PfAvarApproximation(plan, fv).tree_variance)
    to every beam and weight slot (src_lib/GpuDedisperser.cpp:1162-1238): the chain's outputs
    equal those with the same weights built on the host and uploaded by fill_all_weights."""
    import torch
    import pirate_amd as pa
    a, b = make('toy'), make('toy')
    fv = np.random.default_rng(4).uniform(0.5, 2.0, a.nfreq)
    a.fill_analytic_weights(fv)
    av = pa.PfAvarApproximation(a.plan, fv)
    for it in range(b.ntrees):
        w = b.fill_host_weights(it, av.tree_variance[it])
        assert np.all(w > 0)
    gen = torch.Generator(device='cuda')
    for seq in range(3 * a.nbatches):
        gen.manual_seed(seq)
        x = torch.randn((a.beams_per_batch, a.nfreq, a.nt_in), generator=gen, device='cuda')
        for dd in (a, b):
            with dd.get_input(seq) as arr:
                arr.copy_(x)
        with a.get_output(seq) as oa, b.get_output(seq) as ob:
            for it in range(a.ntrees):
                assert torch.equal(oa.out_max[it], ob.out_max[it])
                assert torch.equal(oa.out_argmax[it], ob.out_argmax[it])


def test_fill_analytic_weights_bad_variances_stop():
    dd = make('toy')
    with pytest.raises(RuntimeError, match='positive'):
        dd.fill_analytic_weights(np.zeros(dd.nfreq))
    with pytest.raises(RuntimeError, match='stopped'):
        dd._acquire_input(0, 0)


def test_views_keep_the_dedisperser_alive():
    # Input and output tensors hold the dedisperser, so dropping the last Python reference to
    # it while a view lives does not free the memory under the view.
    import gc
    import weakref
    dd = make()
    ref = weakref.ref(dd)
    push(dd, 0)
    out = dd._acquire_output(0, 0, 0, sync=True)
    om = out.out_max[0]
    dd._release_output(0, 0, 0)
    del dd, out
    gc.collect()
    assert ref() is not None
    om.sum().item()       # the memory is still mapped and owned
    del om
    gc.collect()
    assert ref() is None
