"""Output ring shared with a consumer process over HIP IPC (the reference's
FrbServer -> FrbGrouper handshake, src_lib/FrbServer.cpp:1583-1612 and
src_lib/FrbGrouper.cpp:565-599): the consumer maps the exported region and reads every
tree's out_max and out_argmax slot; they must equal what the producer's own outputs hold.

The consumer is forked from a forkserver started when this module is imported (during
collection, before any test touches the GPU), so no process that has initialised the GPU
forks or execs another program.
"""

import multiprocessing as mp
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

_CTX = mp.get_context('forkserver')
_CTX.set_forkserver_preload([])
try:
    from multiprocessing import forkserver as _fs
    _fs.ensure_running()
except Exception:      # noqa: BLE001 -- the test below reports a missing server
    pass


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _consumer(handle, layout, q):
    try:
        import torch
        import pirate_amd as pa
        v = pa.OutputRingView(handle, layout)
        res = []
        for it in range(len(layout['trees'])):
            for slot in range(layout['nbatches_out']):
                res.append((v.out_max(it, slot).float().cpu().numpy(), v.out_argmax(it, slot).cpu().numpy()))
        torch.cuda.synchronize()
        v.close()
        q.put(('ok', res))
    except Exception as e:     # noqa: BLE001 -- reported to the parent
        q.put(('err', repr(e)))


def test_output_ring_ipc_to_consumer_process():
    import torch
    import pirate_amd as pa
    from gpu_cases import CASES
    from gpu_harness import random_weights

    cfg = pa.DedispersionConfig.from_dict(CASES['toy'])
    plan = pa.DedispersionPlan(cfg)
    dd = pa.GpuDedisperser(plan, nbatches_out=2)
    dd.allocate()
    rng = np.random.default_rng(4)
    for it in range(plan.ntrees):
        dd.fill_all_weights(it, random_weights(rng, dd, it))
    B, F, T = plan.beams_per_batch, plan.nfreq, plan.nt_in
    mine = {}
    for seq in range(2):
        with dd.get_input(seq) as arr:
            arr.copy_(torch.from_numpy(rng.uniform(-1, 1, size=(B, F, T)).astype(np.float32)).to(arr.device))
        with dd.get_output(seq) as out:
            for it in range(plan.ntrees):
                mine[(it, seq % 2)] = (out.out_max[it].float().cpu().numpy(), out.out_argmax[it].cpu().numpy())
    torch.cuda.synchronize()

    handle, layout = dd.output_ipc_handle()
    assert len(handle) == 64 and layout['nbytes'] > 0
    log(f'[gpu ipc] exporting {layout["nbytes"]} bytes, {plan.ntrees} trees x {layout["nbatches_out"]} slots')
    q = _CTX.Queue()
    p = _CTX.Process(target=_consumer, args=(handle, layout, q))
    p.start()
    try:
        status, res = q.get(timeout=150)
    finally:
        p.join(timeout=60)
        if p.is_alive():
            p.kill()
    assert status == 'ok', res
    k = 0
    for it in range(plan.ntrees):
        for slot in range(2):
            gm, ga = res[k]
            k += 1
            rm, ra = mine[(it, slot)]
            assert np.array_equal(gm, rm), (it, slot)
            assert np.array_equal(ga, ra), (it, slot)
    log('[gpu ipc] consumer saw the producer outputs bit for bit')
