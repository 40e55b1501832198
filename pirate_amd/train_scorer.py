"""Trains the learned peak scorer on the GPU and compares it with the plain peak finder.

    python -m pirate_amd.train_scorer [--train-tiles 4000000] [--epochs 6] [--clean-frac 0.5] [--eval-chunks 400]
                                      [--out scorer_params.npz] [--json summary.json]

Data: pirate_amd.scorer_data (the 1024-channel rank-10 chain, 16 beams per chunk, simulated
pulses in half of the beams, noise + RFI everywhere in a fraction 1 - clean_frac of the
training chunks, pure noise in the rest). Training and evaluation use disjoint
pulse banks and noise streams. Prints progress lines and one JSON summary: training loss per
epoch, completeness of the scorer, of the scorer with its head (pirate_amd.ScoreHead: the
network's logit and the peak finder's statistic combined by a logistic regression fitted on
the training tiles) and of the peak finder at false-alarm rates 1e-2, 1e-3 and 1e-4 on
evaluation negatives with RFI and, separately, on pure-noise evaluation data (every beam
without RFI), overall and per SNR bin, the data-generation rate, and the
scorer kernels' throughput measured with HIP events (forward: tiles/s and bf16 MFMA TFLOP/s at
2 * 16 * 256 * (64 + 16 + 16) * 9 flop per tile; training step: tiles/s).
"""

import argparse
import json
import sys
import time

import numpy as np

FLOP_PER_TILE_FWD = 2 * 16 * 256 * (64 + 16 + 16) * 9
MFMA_BF16_PEAK_TFLOPS = 2500.0


def log(msg):
    print(f'[train_scorer] {msg}', file=sys.stderr, flush=True)


def kernel_rates(scorer, tiles, labels, reps=5):
    """Average forward and training-step (gradient + Adam) times over 'reps' launches on the
    current stream, from HIP events."""
    import torch
    from .scorer import TileSource
    src = TileSource.from_tiles(tiles)
    scorer.forward(src)
    scorer.grad(src, labels)
    torch.cuda.synchronize()
    e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
    e0.record()
    for _ in range(reps):
        scorer.forward(src)
    e1.record()
    for _ in range(reps):
        scorer.grad(src, labels)
        scorer.adam_step(lr=0.0)
    e2.record()
    torch.cuda.synchronize()
    n = tiles.shape[0]
    fwd_s = e0.elapsed_time(e1) / 1e3 / reps
    trn_s = e1.elapsed_time(e2) / 1e3 / reps
    return {'tiles': n, 'forward_ms': fwd_s * 1e3, 'forward_tiles_per_s': n / fwd_s,
            'forward_tflops': n * FLOP_PER_TILE_FWD / fwd_s / 1e12,
            'forward_mfma_frac': n * FLOP_PER_TILE_FWD / fwd_s / 1e12 / MFMA_BF16_PEAK_TFLOPS,
            'train_step_ms': trn_s * 1e3, 'train_tiles_per_s': n / trn_s}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split('\n')[0])
    ap.add_argument('--train-tiles', type=int, default=4000000)
    ap.add_argument('--epochs', type=int, default=6)
    ap.add_argument('--batch', type=int, default=512)
    ap.add_argument('--lr', type=float, default=2e-3)
    ap.add_argument('--bank', type=int, default=512, help='simulated pulses in the training bank')
    ap.add_argument('--eval-bank', type=int, default=128)
    ap.add_argument('--eval-chunks', type=int, default=400)
    ap.add_argument('--seed', type=int, default=5)
    ap.add_argument('--clean-frac', type=float, default=0.5,
                    help='fraction of the training chunks drawn without RFI (pure noise + pulses)')
    ap.add_argument('--head-clean-chunks', type=int, default=100,
                    help='chunks of pure-noise tiles (no RFI) added to the head fit')
    ap.add_argument('--nproc', type=int, default=None, help='worker processes for the pulse bank (1: none)')
    ap.add_argument('--out', default=None, help='save the trained parameters (.npz)')
    ap.add_argument('--json', default=None, help='also write the summary here')
    a = ap.parse_args(argv)

    import torch
    import pirate_amd as pa
    from . import scorer_data as sd
    t0 = time.perf_counter()
    bank = sd.make_template_bank(a.bank, seed=1000 * a.seed + 1, nproc=a.nproc)
    ebank = sd.make_template_bank(a.eval_bank, seed=1000 * a.seed + 2, nproc=a.nproc)
    t_bank = time.perf_counter() - t0
    log(f'{len(bank)} + {len(ebank)} simulated pulses in {t_bank:.1f} s')

    stream = sd.TileStream(bank, nbeams=16, seed=a.seed, neg_per_beam=8)
    per_chunk = stream.npulse_beams * (stream.T // sd.TILE_T) + (stream.B - stream.npulse_beams) * 8
    nchunks = -(-a.train_tiles // per_chunk) + 1
    t0 = time.perf_counter()
    nclean = int(round(a.clean_frac * nchunks))
    extra = [(sd.TileStream(bank, nbeams=16, seed=a.seed + 4, neg_per_beam=8, rfi=False), nclean)] if nclean else []
    train = sd.build_dataset(stream, nchunks - nclean, log=log, extra=extra)
    del extra
    t_data = time.perf_counter() - t0
    n = train['tiles'].shape[0]
    log(f'{n} training tiles from {nchunks} chunks in {t_data:.1f} s')
    del stream

    scr = pa.LearnedScorer(seed=a.seed)
    t0 = time.perf_counter()
    hist = sd.train(scr, train, epochs=a.epochs, batch=a.batch, lr=a.lr, seed=a.seed, log=log)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    rates = kernel_rates(scr, train['tiles'][:65536].contiguous(), train['labels'][:65536].contiguous())
    log(f'kernels: {rates}')
    # the head over (network logit, peak-finder statistic), fitted on the training tiles
    # (on the training tiles plus pure-noise tiles of the training pulse bank, so that it sees
    # negatives both with and without RFI)
    zs, ps, ys = [sd.score(scr, train['tiles'])], [train['pf']], [train['labels'].cpu().numpy()]
    if a.head_clean_chunks > 0:
        clean = sd.build_dataset(sd.TileStream(bank, nbeams=16, seed=a.seed + 3, neg_per_beam=8, rfi=False),
                                 a.head_clean_chunks)
        zs.append(sd.score(scr, clean['tiles']))
        ps.append(clean['pf'])
        ys.append(clean['labels'].cpu().numpy())
        del clean
    head = pa.ScoreHead()
    head_loss = head.fit(np.concatenate(zs), np.concatenate(ps), np.concatenate(ys))
    log(f'head: log-loss {head_loss[0]:.4f} -> {head_loss[-1]:.4f}, weights {np.round(head.w, 3).tolist()}')
    del train

    def evaluate(ev, what):
        s = sd.score(scr, ev['tiles'])
        y = ev['labels'].cpu().numpy() > 0.5
        sh = head(s, ev['pf'])
        res = {'pulses': int(y.sum()), 'negatives': int((~y).sum()), 'far': {}}
        for far in (1e-2, 1e-3, 1e-4):
            c_nn, thr_nn = sd.completeness_at_far(s[y], s[~y], far)
            c_hd, _ = sd.completeness_at_far(sh[y], sh[~y], far)
            c_pf, thr_pf = sd.completeness_at_far(ev['pf'][y], ev['pf'][~y], far)
            bins = {}
            for lo, hi in ((5, 10), (10, 15), (15, 20), (20, 30)):
                m = (ev['snr'][y] >= lo) & (ev['snr'][y] < hi)
                bins[f'snr_{lo}_{hi}'] = {'scorer': float(np.mean(s[y][m] > thr_nn)),
                                          'peak_finder': float(np.mean(ev['pf'][y][m] > thr_pf)), 'n': int(m.sum())}
            res['far'][f'{far:g}'] = {'scorer': c_nn, 'scorer_with_head': c_hd, 'peak_finder': c_pf,
                                      'scorer_threshold_logit': thr_nn, 'peak_finder_threshold_snr': thr_pf,
                                      'by_snr': bins}
            log(f'{what}: FAR {far:g}: completeness scorer {c_nn:.3f}, scorer + head {c_hd:.3f}, '
                f'peak finder {c_pf:.3f} (peak-finder threshold SNR {thr_pf:.1f})')
        return res

    ev = sd.build_dataset(sd.TileStream(ebank, nbeams=16, seed=a.seed + 1, neg_per_beam=128), a.eval_chunks, log=log)
    res_rfi = evaluate(ev, 'noise + RFI')
    del ev
    ev = sd.build_dataset(sd.TileStream(ebank, nbeams=16, seed=a.seed + 2, neg_per_beam=128, rfi=False),
                          a.eval_chunks, log=log)
    res_clean = evaluate(ev, 'noise only')
    del ev
    summary = {'train_tiles': n, 'clean_frac': a.clean_frac, 'epochs': a.epochs, 'batch': a.batch, 'lr': a.lr, 'loss_per_epoch': hist,
               'bank': len(bank), 'eval_bank': len(ebank), 'bank_s': t_bank, 'data_s': t_data,
               'data_tiles_per_s': n / t_data, 'train_s': t_train, 'kernels': rates,
               'head': head.state_dict(), 'head_log_loss': head_loss,
               'eval_noise_rfi': res_rfi, 'eval_noise_only': res_clean}
    if a.out:
        np.savez(a.out, params=scr.params.cpu().numpy(), head_w=head.w, head_mean=head.mean, head_std=head.std)
    print(json.dumps(summary), flush=True)
    if a.json:
        with open(a.json, 'w') as f:
            json.dump(summary, f, indent=1)
    return summary


if __name__ == '__main__':
    main()
