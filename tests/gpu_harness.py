"""Runs the GPU dedisperser and the CPU oracle on the same seeded inputs and weights."""

import queue
import threading

import numpy as np

from oracle import oracle as orc


def oracle_config(d):
    return orc.OracleConfig.from_yaml_dict(d, nelts_per_segment=32 if d['dtype'] == 'float32' else 64)


def random_weights(rng, dd, itree):
    """Sparse random non-negative weights for every weight slot (nbatches_wt, B, ndm_wt, nt_wt, P, N)."""
    shape = dd.wt_shape(itree)
    w = np.stack([orc.sparse_random_weights(rng, *shape[1:]) for _ in range(shape[0])])
    return w.astype(np.float32)


def run_and_compare(cfg_dict, nchunks, seed=0, input_kind='uniform', log=print, variances=False, nbatches_out=0,
                    nbatches_wt=0):
    """Returns a list of per-(seq, tree) mismatch records (empty if bit-exact). With
    variances=True the peak finder's float64 out_var is compared too (exactly). With a 16-bit
    dtype in cfg_dict, the oracle rounds to that type where the GPU chain stores it.
    nbatches_out / nbatches_wt: output and weight ring depths (0 = the dedisperser's default)."""
    import torch
    import pirate_amd as pa

    cfg = pa.DedispersionConfig.from_dict(cfg_dict)
    plan = pa.DedispersionPlan(cfg)
    dd = pa.GpuDedisperser(plan, enable_variances=variances, nbatches_out=nbatches_out, nbatches_wt=nbatches_wt)
    dd.allocate()
    storage = cfg_dict['dtype']
    ref = orc.Dedisperser(oracle_config(cfg_dict), 1, enable_variances=variances, storage=storage)
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
            if storage != 'float32':
                x = orc.round_to_storage(x, storage).reshape(x.shape)
            with dd.get_input(seq) as arr:
                arr.copy_(torch.from_numpy(x).to(arr.device))
            for it in range(plan.ntrees):
                ref.set_weights(it, W[it][seq % dd.nbatches_wt])
            ref.dedisperse(x, ichunk, ib)
            with dd.get_output(seq) as out:
                for it in range(plan.ntrees):
                    gm = out.out_max[it].float().cpu().numpy()
                    ga = out.out_argmax[it].cpu().numpy().view(np.uint32)
                    rm, ra = ref.output(it)
                    nm = int(np.sum(gm != rm))
                    na = int(np.sum(ga != ra))
                    if variances:
                        gv = out.out_var[it].cpu().numpy()
                        rv = ref.variances(it)
                        nv = int(np.sum(gv != rv))
                        if nv:
                            rec = dict(seq=seq, itree=it, n_var_mismatch=nv, total=int(gv.size),
                                       max_rel=float(np.max(np.abs(gv - rv) / np.maximum(np.abs(rv), 1e-300))))
                            bad.append(rec)
                            log(f'MISMATCH {rec}')
                    if nm or na:
                        ev = ref.eval_tokens(it, ga)
                        rec = dict(seq=seq, itree=it, n_max_mismatch=nm, n_argmax_mismatch=na,
                                   max_abs=float(np.max(np.abs(gm.astype(np.float64) - rm))),
                                   eval_tokens_mismatch=int(np.sum(ev != gm)), total=int(gm.size))
                        bad.append(rec)
                        log(f'MISMATCH {rec}')
    return bad, dd, plan


def quantize(x, dtype):
    """Rounds float32 x to the storage dtype (round-to-nearest-even) and returns it as float64."""
    import torch
    tdt = {'float16': torch.float16, 'bfloat16': torch.bfloat16, 'float32': torch.float32}[dtype]
    return torch.from_numpy(x).to(tdt).to(torch.float64).numpy()


def run_precision(cfg_dict, dtype, nchunks, seed=0, log=print, rfi=False):
    """Runs the GPU chain with 16-bit storage and the FP64 oracle on the same (quantized) inputs
    with bare-kernel weights (outputs in SNR units for unit-variance noise). Returns per-tree
    error statistics: max and rms of |gpu - fp64| / max(|fp64|, 1) over out_max, and the same
    for the FP64 value of the GPU's argmax token (how far the chosen candidate is from the max);
    'by_trial': the max error per (DM quarter of the tree, peak-finding profile of the GPU's
    argmax), i.e. per (DM, width) trial class. rfi=True adds narrowband RFI: 1% of the channels
    of every beam (drawn once from the seed) carry a constant +5 (sigma) offset."""
    import torch
    import pirate_amd as pa

    d16 = dict(cfg_dict, dtype=dtype)
    cfg = pa.DedispersionConfig.from_dict(d16)
    plan = pa.DedispersionPlan(cfg)
    plan32 = pa.DedispersionPlan(pa.DedispersionConfig.from_dict(dict(cfg_dict, dtype='float32')))
    for t16, t32 in zip(plan.trees, plan32.trees):
        assert vars(t16) == vars(t32), 'storage dtype changed the tree plan'
    dd = pa.GpuDedisperser(plan)
    dd.allocate()
    ref = orc.Dedisperser(oracle_config(dict(cfg_dict, dtype='float32')), 1, dbl=True)
    W = [orc.bare_weights(*dd.wt_shape(it)[1:]) for it in range(plan.ntrees)]
    for it in range(plan.ntrees):
        dd.fill_all_weights(it, np.broadcast_to(W[it], dd.wt_shape(it)).copy())
        ref.set_weights(it, W[it])
    tdt = {'float16': torch.float16, 'bfloat16': torch.bfloat16}[dtype]

    rng = np.random.default_rng(seed)
    B, F, T = plan.beams_per_batch, plan.nfreq, plan.nt_in
    errs = {it: [] for it in range(plan.ntrees)}
    tok_errs = {it: [] for it in range(plan.ntrees)}
    trial = {it: {} for it in range(plan.ntrees)}
    offs = np.zeros((B, F, 1), dtype=np.float32)
    if rfi:
        for b in range(B):
            offs[b, rng.choice(F, size=max(1, F // 100), replace=False), 0] = 5.0
    for ichunk in range(nchunks):
        for ib in range(dd.nbatches):
            seq = ichunk * dd.nbatches + ib
            x = rng.standard_normal((B, F, T)).astype(np.float32) + offs
            xq = quantize(x, dtype)
            with dd.get_input(seq) as arr:
                arr.copy_(torch.from_numpy(xq).to(tdt).to(arr.device))
            ref.dedisperse(xq, ichunk, ib)
            with dd.get_output(seq) as out:
                for it in range(plan.ntrees):
                    gm = out.out_max[it].to(torch.float64).cpu().numpy()
                    ga = out.out_argmax[it].cpu().numpy().view(np.uint32)
                    rm, _ = ref.output(it)
                    scale = np.maximum(np.abs(rm), 1.0)
                    e = np.abs(gm - rm) / scale
                    errs[it].append(e.ravel())
                    p = (ga >> 8) & 0xff
                    q = (np.arange(e.shape[1]) * 4 // e.shape[1])[None, :, None]
                    for qq in range(4):
                        for pp in np.unique(p):
                            m = (np.broadcast_to(q, e.shape) == qq) & (p == pp)
                            if m.any():
                                k = (qq, int(pp))
                                trial[it][k] = max(trial[it].get(k, 0.0), float(e[m].max()))
                    ev = ref.eval_tokens(it, ga)
                    tok_errs[it].append((np.abs(ev - rm) / scale).ravel())
    stats = {}
    for it in range(plan.ntrees):
        e = np.concatenate(errs[it])
        te = np.concatenate(tok_errs[it])
        stats[it] = dict(max=float(e.max()), rms=float(np.sqrt(np.mean(e**2))), tok_max=float(te.max()), n=int(e.size),
                         by_trial=dict(sorted(trial[it].items())))
        log(f'{dtype} tree {it}: max rel err {stats[it]["max"]:.2e}, rms {stats[it]["rms"]:.2e}, '
            f'argmax candidate max rel err {stats[it]["tok_max"]:.2e} over {e.size} (dm, t) outputs')
        log(f'{dtype} tree {it}: max rel err per (DM quarter, profile): '
            + ', '.join(f'{k}: {v:.1e}' for k, v in stats[it]['by_trial'].items()))
    return stats


class OracleWorker(threading.Thread):
    """One oracle instance per (member, beam), fed (ichunk, x, gpu_max, gpu_argmax) items; the
    oracle's C calls release the GIL, so workers run in parallel."""

    def __init__(self, cfg_dict, weights, storage, tag):
        super().__init__(daemon=True)
        from oracle import oracle as orc
        d1 = dict(cfg_dict, beams_per_gpu=1, beams_per_batch=1)
        self.ref = orc.Dedisperser(oracle_config(d1), 1, storage=storage)
        for it, w in enumerate(weights):
            self.ref.set_weights(it, w)
        self.q = queue.Queue(maxsize=4)
        self.bad = []
        self.nchecked = 0
        self.tag = tag
        self.err = None

    def run(self):
        try:
            while True:
                item = self.q.get()
                if item is None:
                    return
                ichunk, x, gmax, garg = item
                self.ref.dedisperse(x, ichunk, 0)
                for it in range(len(gmax)):
                    rm, ra = self.ref.output(it)
                    nm, na = int(np.sum(gmax[it] != rm)), int(np.sum(garg[it] != ra))
                    if nm or na:
                        self.bad.append(dict(tag=self.tag, ichunk=ichunk, itree=it, n_max=nm, n_argmax=na))
                self.nchecked += 1
        except Exception as e:       # noqa: BLE001 -- reported by the test
            self.err = e
