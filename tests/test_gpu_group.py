"""Heterogeneous beam groups (GpuDedisperserGroup): every member's chain of a chunk runs on a
branch stream of its own, forked from and joined into one launch stream. (HIP graph replay of
the chunk is refused in pirate_amd; its test case is skipped with the reason.)

Checks, bit for bit:
  * a small group of three shapes (ranks 8/9/10, 2/3/4 primary trees, float32 and float16
    storage) against the oracle over enough chunks to cover every member's longest lag;
  * the batched heterogeneous workload (beam groups at ranks 12/13/14 with 4096/8192/16384
    channels over 300-1500 MHz and 3/4/5 primary trees, float16 storage, chunks of 2048;
    here 2 beams per group) against standalone GpuDedispersers of the same plans over 129
    chunks, which covers every ring row's lag of every group (the longest, rank-14 primary
    tree 4: 16129 samples of 128 per chunk), and against the oracle (rounding to float16
    where the chain stores) on the chunks that cover the rank-12 and rank-13 groups' longest
    lags (10 and 34 chunks) and the rank-14 group's primary trees 0 and 1 (18 chunks).
"""

import sys

import numpy as np
import pytest

from gpu_cases import BASE, batched_groups, pt
from gpu_harness import OracleWorker, random_weights

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def small_groups(dtype):
    mk = lambda R, F, T, trees, sub, B: dict(BASE, dtype=dtype, zone_nfreq=[F], zone_freq_edges=[400, 800],
                                             toplevel_tree_rank=R, time_samples_per_chunk=T, beams_per_gpu=B,
                                             beams_per_batch=B, num_active_batches=1, frequency_subband_counts=sub,
                                             primary_trees=trees)
    return [mk(8, 512, 512, [pt(0, 8, 16, 16), pt(1, 4, 16, 16)], [2, 2, 1], 2),
            mk(9, 1024, 512, [pt(0, 16, 16, 16), pt(0, 8, 16, 16), pt(0, 4, 16, 16)], [1], 3),
            mk(10, 1024, 512, [pt(0, 16, 16, 16), pt(0, 8, 16, 16), pt(0, 4, 16, 16), pt(0, 2, 16, 16)], [1], 1)]


def max_lag_chunks(plan, trees=None):
    """Chunks after which every ring row of the given trees (default all) has wrapped around
    its capacity (largest stage-2 lag + one chunk) and been read past the wrap: the largest
    ceil(lag / Tp) + 2, with lag = (2^dd - 1) (2^amb - 1 + [p > 0] 2^amb) (the largest rb_lag)."""
    nmax = 0
    for it, t in enumerate(plan.trees):
        if trees is not None and it not in trees:
            continue
        p = t.primary_tree_index
        lag = (2**t.dd_rank - 1) * (2**t.amb_rank - 1 + (2**t.amb_rank if p > 0 else 0))
        nmax = max(nmax, -(-lag // t.nt_ds) + 2)
    return nmax


def host_outputs(out, b):
    import torch
    return ([m[b:b + 1].to(torch.float32).cpu().numpy() for m in out.out_max],
            [a[b:b + 1].cpu().numpy().view(np.uint32) for a in out.out_argmax])


GRAPH_SKIP = pytest.mark.skip(reason='HIP graph replay of the group chain is refused in pirate_amd: its first run '
                                     'ended in a segmentation fault inside release_input_and_launch whose cause is '
                                     'not identified')


@pytest.mark.parametrize('dtype', ['float32', 'float16'])
@pytest.mark.parametrize('use_graph', [pytest.param(True, marks=GRAPH_SKIP), False])
def test_small_group_bit_exact_vs_oracle(dtype, use_graph):
    import torch
    import pirate_amd as pa
    from oracle import oracle as orc
    cfgs = small_groups(dtype)
    plans = [pa.DedispersionPlan(pa.DedispersionConfig.from_dict(d)) for d in cfgs]
    nchunks = max(max_lag_chunks(p) for p in plans) + 2
    log(f'[gpu group] small group {dtype} use_graph={use_graph}: ranks {[d["toplevel_tree_rank"] for d in cfgs]}, '
        f'{nchunks} chunks')
    g = pa.GpuDedisperserGroup(plans, use_graph=use_graph)
    g.allocate()
    rng = np.random.default_rng(7)
    workers = []
    for m, (d, mem) in enumerate(zip(cfgs, g.members)):
        W = [random_weights(rng, mem, it) for it in range(mem.ntrees)]
        for it in range(mem.ntrees):
            mem.fill_all_weights(it, W[it])
        for b in range(mem.beams_per_batch):
            w = OracleWorker(d, [W[it][0, b:b + 1] for it in range(mem.ntrees)], dtype, f'm{m}b{b}')
            w.start()
            workers.append((m, b, w))
    for ichunk in range(nchunks):
        xs = []
        with g.get_input(ichunk) as slots:
            for s in slots:
                x = rng.uniform(-1, 1, size=tuple(s.shape)).astype(np.float32)
                if dtype != 'float32':
                    x = orc.round_to_storage(x, dtype).reshape(x.shape)
                s.copy_(torch.from_numpy(x).to(s.device))
                xs.append(x)
        with g.get_output(ichunk) as outs:
            for m, b, w in workers:
                gm, ga = host_outputs(outs[m], b)
                w.q.put((ichunk, xs[m][b:b + 1], gm, ga))
    for _, _, w in workers:
        w.q.put(None)
    for _, _, w in workers:
        w.join()
        assert w.err is None, w.err
        assert w.nchecked == nchunks
    bad = [r for _, _, w in workers for r in w.bad]
    assert not bad, bad[:5]
    if use_graph:
        assert g.graph_nodes() > 0


def test_group_protocol_errors_latch():
    import pirate_amd as pa
    plans = [pa.DedispersionPlan(pa.DedispersionConfig.from_dict(d)) for d in small_groups('float32')]
    g = pa.GpuDedisperserGroup(plans)
    g.allocate()
    with pytest.raises(RuntimeError, match='expected ichunk=0'):
        g._acquire_input(1, 0)
    with pytest.raises(RuntimeError, match='expected ichunk=0'):
        g._acquire_input(0, 0)
    # a member driven by the group refuses its own protocol
    with pytest.raises(RuntimeError, match='GpuDedisperserGroup'):
        g.members[0]._acquire_input(0, 0)


def test_batched_groups_bit_exact():
    import torch
    import pirate_amd as pa
    cfgs = batched_groups(nbeams=2)
    plans = [pa.DedispersionPlan(pa.DedispersionConfig.from_dict(d)) for d in cfgs]
    nchunks = max(max_lag_chunks(p) for p in plans)
    ncheck = [max_lag_chunks(plans[0]), max_lag_chunks(plans[1]), max_lag_chunks(plans[2], trees=(0, 1))]
    log(f'[gpu group] batched shapes, 2 beams per group, float16: {nchunks} chunks vs standalone dedispersers, '
        f'oracle on the first {ncheck} chunks per group')
    g = pa.GpuDedisperserGroup(plans, use_graph=False)
    g.allocate()
    solo = []
    for p in plans:
        dd = pa.GpuDedisperser(p)
        dd.allocate()
        solo.append(dd)
    rng = np.random.default_rng(11)
    workers = []
    for m, (d, mem) in enumerate(zip(cfgs, g.members)):
        W = [random_weights(rng, mem, it) for it in range(mem.ntrees)]
        for it in range(mem.ntrees):
            mem.fill_all_weights(it, W[it])
            solo[m].fill_all_weights(it, W[it][:solo[m].nbatches_wt])
        for b in range(mem.beams_per_batch):
            w = OracleWorker(d, [W[it][0, b:b + 1] for it in range(mem.ntrees)], 'float16', f'm{m}b{b}')
            w.start()
            workers.append((m, b, w))
    gen = torch.Generator(device='cuda')
    gen.manual_seed(11)
    mism = []
    for ichunk in range(nchunks):
        xs = []
        with g.get_input(ichunk) as slots:
            for m, s in enumerate(slots):
                x = (torch.rand(tuple(s.shape), generator=gen, device=s.device) * 2 - 1).to(torch.float16)
                s.copy_(x)
                with solo[m].get_input(ichunk) as a:
                    a.copy_(x)
                xs.append(x)
        with g.get_output(ichunk) as outs:
            for m in range(len(plans)):
                with solo[m].get_output(ichunk) as so:
                    for it in range(len(so.out_max)):
                        nm = int((outs[m].out_max[it] != so.out_max[it]).sum().item())
                        na = int((outs[m].out_argmax[it] != so.out_argmax[it]).sum().item())
                        if nm or na:
                            mism.append(dict(ichunk=ichunk, member=m, itree=it, n_max=nm, n_argmax=na))
            for m, b, w in workers:
                if ichunk < ncheck[m]:
                    gm, ga = host_outputs(outs[m], b)
                    w.q.put((ichunk, xs[m][b:b + 1].to(torch.float32).cpu().numpy(), gm, ga))
        if ichunk % 16 == 0:
            log(f'[gpu group] batched chunk {ichunk}: {len(mism)} mismatches vs standalone so far')
    for _, _, w in workers:
        w.q.put(None)
    for m, b, w in workers:
        w.join()
        assert w.err is None, w.err
        assert w.nchecked == ncheck[m]
    assert not mism, mism[:5]
    bad = [r for _, _, w in workers for r in w.bad]
    assert not bad, bad[:5]
