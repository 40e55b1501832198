"""ORACLE — test infrastructure only.

Dense float64 restatement of the reference's analytic peak-finder variances
(PfAvarApproximation: src_lib/PfVariance.cpp:238-366, pirate_frb/slow_avar/PfVariance.py:
465-503). The reference propagates sparse tiles (src_lib/SparseTile.cpp); this module
propagates each input channel's impulse densely instead:

  1. grid a one-hot input channel i (time 0) with the oracle's tree gridding;
  2. run k passes of the tree recursion on the whole (2^R, nt) array
     (out[F, 2d+b, t] = (in[2F, d, t-d-b] + in[2F+1, d, t]) / sqrt(2),
     notes/tree_dedispersion.tex:113-121, src_lib/ReferenceTree.cpp:310-345), giving
     (2^(R-k) bands, 2^k rows, nt);
  3. filter every row with each peak-finding kernel, built here as the impulse response of
     the boxcar recursion b_(l+1)[t] = b_l[t] + b_l[t - 2^l] and the profiles
     y0 = x3, y1 = x2 + x3, y2 = x1/2 + x2 + x3/2, y3 = x0/2 + x1 + x2 + x3/2 on taps spaced
     2^l (notes/tree_dedispersion.tex:326-356), and sum the squares;
  4. combine as the reference does (PfVariance.cpp:337-362, 300-334): tree t stops at level
     k = r - L (+1 for primary trees p > 0, which keep the upper half of the rows), bands
     F < 2^L are summed with weight 2^-(L-R) into pf-rank band F >> (L-R), weighted by the
     channel's variance, and subband n averages its pf-rank bands [n_to_flo, n_to_fhi).

Only tests/ may use this module, and only as the checker.
"""

import numpy as np

from . import oracle as orc

C = 1.0 / np.sqrt(2.0)


def tree_levels(x, k):
    """x: (..., nchan, nt) float64 -> (..., nchan >> k, 2^k, nt): k passes of the tree
    recursion with zero history (rows in natural DM order)."""
    a = np.asarray(x, dtype=np.float64)[..., :, None, :]
    nt = a.shape[-1]
    for _ in range(k):
        lo, up = a[..., 0::2, :, :], a[..., 1::2, :, :]
        nr = lo.shape[-2]
        out = np.empty(lo.shape[:-2] + (2 * nr, nt))
        for d in range(nr):
            for b in (0, 1):
                s = d + b
                sh = np.zeros(lo.shape[:-2] + (nt,))
                if s < nt:
                    sh[..., s:] = lo[..., d, :nt - s]
                out[..., 2 * d + b, :] = C * (sh + up[..., d, :])
        a = out
    return a


def pf_kernels(P):
    """The first P peak-finding kernels (impulse responses), from the boxcar recursion."""
    nlev = (P - 1) // 3
    n = 4 << max(nlev - 1, 0)
    delta = np.zeros(n)
    delta[0] = 1.0
    b = delta.copy()                       # b_0: one sample
    ks = [delta[:1].copy()]
    for lev in range(nlev):
        w = 1 << lev
        tap = lambda j: np.concatenate([np.zeros(j * w), b])[:n]   # x_(3-j) = b_l[t - j w]
        x3, x2, x1, x0 = tap(0), tap(1), tap(2), tap(3)
        for y in (x2 + x3, 0.5 * x1 + x2 + 0.5 * x3, 0.5 * x0 + x1 + x2 + 0.5 * x3):
            nz = np.nonzero(y)[0]
            ks.append(y[:nz[-1] + 1].copy())
        b = b + np.concatenate([np.zeros(w), b])[:n]                 # b_(l+1)
    return ks[:P]


def _filtered_sumsq(rows, kernels):
    """rows: (..., nt) -> (..., P): sum over t of (h_p * row)^2 (full convolution), by
    Parseval's theorem on a zero-padded DFT."""
    nt = rows.shape[-1]
    n = nt + max(len(h) for h in kernels)
    G = np.fft.rfft(rows, n=n, axis=-1)
    g2 = G.real ** 2 + G.imag ** 2
    wgt = np.full(g2.shape[-1], 2.0)                  # rfft keeps half the spectrum
    wgt[0] = 1.0
    if n % 2 == 0:
        wgt[-1] = 1.0
    out = np.empty(rows.shape[:-1] + (len(kernels),))
    for p, h in enumerate(kernels):
        H = np.fft.rfft(h, n=n)
        out[..., p] = (g2 @ (wgt * (H.real ** 2 + H.imag ** 2))) / n
    return out


def subband_ranges(subband_counts):
    """(n_to_flo, n_to_fhi) in pf-rank bands (src_lib/FrequencySubbands.cpp:32-60)."""
    flo, fhi = [], []
    for lev, cnt in enumerate(subband_counts):
        sp = 1 << max(lev - 1, 0)
        for s in range(cnt):
            flo.append(s * sp)
            fhi.append(s * sp + (1 << lev))
    return flo, fhi


def pf_variances(cfg, freq_variances, chunk=512):
    """cfg: oracle.OracleConfig. Returns (tree_variance, per_tf): per tree, float64 arrays
    of shape (N, ndm_wt, P) and (2^pf_rank, ndm_wt, P)."""
    fv = np.asarray(freq_variances, dtype=np.float64)
    assert fv.shape == (cfg.nfreq,) and np.all(fv > 0)
    cm = cfg.channel_map()
    R = cfg.R
    nchan = 1 << R
    trees = cfg.trees()
    specs = []
    for t in trees:
        r = t['amb_rank'] + t['dd_rank']
        L = int(np.log2(t['wt_dm_ds']))
        Rp = len(t['subband_counts']) - 1
        up = t['ipri'] > 0
        k = r - L + (1 if up else 0)
        assert 0 <= Rp <= L <= r and k <= R
        specs.append(dict(r=r, L=L, Rp=Rp, k=k, up=up, P=t['nprofiles'], kern=pf_kernels(t['nprofiles']),
                          ptf=np.zeros((1 << Rp, 1 << (r - L), t['nprofiles']))))
    # Gridding columns of every input channel: (nfreq, nchan).
    gcol = np.zeros((cfg.nfreq, nchan))
    for i0 in range(0, cfg.nfreq, 256):
        i1 = min(cfg.nfreq, i0 + 256)
        onehot = np.zeros((i1 - i0, cfg.nfreq, 1))
        onehot[np.arange(i1 - i0), np.arange(i0, i1), 0] = 1.0
        gcol[i0:i1] = orc.tree_gridding(onehot, cm)[:, :, 0]

    for s in specs:
        k, L, sh = s['k'], s['L'], s['L'] - s['Rp']
        nrow = 1 << (s['r'] - L)
        # (input channel, level-k band) pairs with a nonzero gridding weight, bands F < 2^L
        bands = gcol[:, :(1 << L) << k].reshape(cfg.nfreq, 1 << L, 1 << k)
        ii, FF = np.nonzero(np.any(bands != 0.0, axis=2))
        for c0 in range(0, len(ii), chunk):
            i, F = ii[c0:c0 + chunk], FF[c0:c0 + chunk]
            x = np.zeros((len(i), 1 << k, (1 << k) + 1))       # delays reach 2^k - 1
            x[:, :, 0] = bands[i, F]
            a = tree_levels(x, k)[:, 0]                         # (npairs, 2^k, nt): one band
            rows = a[:, nrow:, :] if s['up'] else a[:, :nrow, :]
            e = _filtered_sumsq(rows, s['kern']) * fv[i][:, None, None]
            np.add.at(s['ptf'], F >> sh, (2.0 ** -sh) * e)

    tv_all, ptf_all = [], []
    for t, s in zip(trees, specs):
        flo, fhi = subband_ranges(t['subband_counts'])
        tv = np.stack([s['ptf'][a:b].mean(axis=0) for a, b in zip(flo, fhi)])
        tv_all.append(tv)
        ptf_all.append(s['ptf'])
    return tv_all, ptf_all
