"""The learned scorer trained on the GPU by its own kernels, and measured against the plain peak finder.

Data (pirate_amd.scorer_data): 64 x 256 tiles of the 1024-channel rank-10 chain's dedispersed
output (400-800 MHz, 1 ms, chunks of 2048, 16 beams per chunk), with simulated pulses
(pirate_amd.simpulse: SNR 5-30, width 1-32 ms, scattering 0-10 ms) in half of the beams and
noise + RFI (persistent narrowband +5 sigma channels, narrowband bursts, broadband zero-DM
impulses) everywhere. Training and evaluation use disjoint pulse banks and disjoint noise
streams.

The peak finder's statistic for a tile is the maximum of its out_max over the tile's region
(bare-kernel weights: a boxcar SNR). At a false-alarm rate of 1e-3 on the evaluation
negatives, the trained scorer's completeness, alone and with its head (pirate_amd.ScoreHead,
fitted on the training tiles), must be at least the peak finder's, and the training loss
must fall. This is a reduced run (about 77k training tiles, 3 epochs); the full
4 x 10^6-tile run is python -m pirate_amd.train_scorer."""

import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def log(msg):
    print(f'[scorer training] {msg}', file=sys.stderr, flush=True)


@pytest.mark.timeout(900)
def test_trained_scorer_completeness_vs_peak_finder():
    import pirate_amd as pa
    from pirate_amd import scorer_data as sd
    log('template banks (simpulse)')
    bank_train = sd.make_template_bank(192, seed=5001)
    bank_eval = sd.make_template_bank(64, seed=5002)
    log('training tiles')
    train = sd.build_dataset(sd.TileStream(bank_train, nbeams=16, seed=5, neg_per_beam=8), 600, log=log)
    n = train['tiles'].shape[0]
    npos = int(train['labels'].sum())
    log(f'{n} training tiles, {npos} with pulses')
    assert 0.4 < npos / n < 0.6
    # the positive tiles hold their pulses: bright pulses stand out in the peak finder's statistic
    yb = train['labels'].cpu().numpy() > 0.5
    bright = yb & (train['snr'] > 20)
    log(f'peak-finder SNR of pulses with SNR > 20: median {np.median(train["pf"][bright]):.1f}; '
        f'of negatives: median {np.median(train["pf"][~yb]):.1f}')
    assert np.median(train['pf'][bright]) > np.median(train['pf'][~yb]) + 5
    scr = pa.LearnedScorer(seed=5)
    hist = sd.train(scr, train, epochs=3, batch=512, lr=2e-3, seed=5, log=log)
    assert hist[-1] < 0.8 * hist[0], hist
    head = pa.ScoreHead()
    hl = head.fit(sd.score(scr, train['tiles']), train['pf'], train['labels'].cpu().numpy())
    assert hl[-1] < hl[0]
    del train
    log('evaluation tiles')
    ev = sd.build_dataset(sd.TileStream(bank_eval, nbeams=16, seed=6, neg_per_beam=128), 120, log=log)
    s = sd.score(scr, ev['tiles'])
    y = ev['labels'].cpu().numpy() > 0.5
    assert np.all(np.isfinite(s))
    sh = head(s, ev['pf'])
    for far in (1e-2, 1e-3):
        c_nn, thr_nn = sd.completeness_at_far(s[y], s[~y], far)
        c_hd, _ = sd.completeness_at_far(sh[y], sh[~y], far)
        c_pf, thr_pf = sd.completeness_at_far(ev['pf'][y], ev['pf'][~y], far)
        log(f'FAR {far:g} over {int((~y).sum())} negatives: completeness scorer {c_nn:.3f} (threshold {thr_nn:.2f}), '
            f'scorer + head {c_hd:.3f}, peak finder {c_pf:.3f} (threshold {thr_pf:.2f} SNR) over {int(y.sum())} pulses')
        if far == 1e-3:
            assert c_nn >= c_pf
            assert c_hd >= c_pf
