"""Runs the GPU dedisperser and the CPU oracle on the same seeded inputs and weights."""

import numpy as np

from oracle import oracle as orc


def oracle_config(d):
    return orc.OracleConfig.from_yaml_dict(d, nelts_per_segment=32 if d['dtype'] == 'float32' else 64)


def random_weights(rng, dd, itree):
    """Sparse random non-negative weights for every weight slot (nbatches_wt, B, ndm_wt, nt_wt, P, N)."""
    shape = dd.wt_shape(itree)
    w = np.stack([orc.sparse_random_weights(rng, *shape[1:]) for _ in range(shape[0])])
    return w.astype(np.float32)


def run_and_compare(cfg_dict, nchunks, seed=0, input_kind='uniform', log=print):
    """Returns a list of per-(seq, tree) mismatch records (empty if bit-exact)."""
    import torch
    import pirate_amd as pa

    cfg = pa.DedispersionConfig.from_dict(cfg_dict)
    plan = pa.DedispersionPlan(cfg)
    dd = pa.GpuDedisperser(plan)
    dd.allocate()
    ref = orc.Dedisperser(oracle_config(cfg_dict), 1)
    rng = np.random.default_rng(seed)

    W = [random_weights(rng, dd, it) for it in range(plan.ntrees)]
    for it in range(plan.ntrees):
        dd.fill_all_weights(it, W[it])

    bad = []
    B, F, T = plan.beams_per_batch, plan.nfreq, plan.nt_in
    for ichunk in range(nchunks):
        for ib in range(dd.nbatches):
            seq = ichunk * dd.nbatches + ib
            if input_kind == 'uniform':
                x = rng.uniform(-1, 1, size=(B, F, T)).astype(np.float32)
            else:
                x = rng.standard_normal((B, F, T)).astype(np.float32)
            with dd.get_input(seq) as arr:
                arr.copy_(torch.from_numpy(x).to(arr.device))
            for it in range(plan.ntrees):
                ref.set_weights(it, W[it][seq % dd.nbatches_wt])
            ref.dedisperse(x, ichunk, ib)
            with dd.get_output(seq) as out:
                for it in range(plan.ntrees):
                    gm = out.out_max[it].cpu().numpy()
                    ga = out.out_argmax[it].cpu().numpy().view(np.uint32)
                    rm, ra = ref.output(it)
                    nm = int(np.sum(gm != rm))
                    na = int(np.sum(ga != ra))
                    if nm or na:
                        ev = ref.eval_tokens(it, ga)
                        rec = dict(seq=seq, itree=it, n_max_mismatch=nm, n_argmax_mismatch=na,
                                   max_abs=float(np.max(np.abs(gm.astype(np.float64) - rm))),
                                   eval_tokens_mismatch=int(np.sum(ev != gm)), total=int(gm.size))
                        bad.append(rec)
                        log(f'MISMATCH {rec}')
    return bad, dd, plan
