"""Training and evaluation data for the learned peak scorer, made by the library's own chain.

Tiles are 64 DM rows x 256 samples of the dedispersed output (the peak-finder input rows,
GpuDedisperser.tree_rows) of the 1024-channel rank-10 tree (400-800 MHz, 1 ms samples, chunks of
2048, one primary tree, max_width 16), run on B beams per chunk:

  * every beam: unit-variance white noise plus radio-frequency interference (RFI), drawn per
    beam and chunk: a persistent narrowband set (1% of the channels at a constant +5 sigma, on
    half of the beams), narrowband bursts (1-8 adjacent channels, 8-256 samples, 2-8 sigma)
    and broadband zero-DM impulses (all channels, 1-3 samples, zero-DM SNR 8-60);
  * "pulse" beams (a fraction pulse_frac of them): one simulated pulse per 256-sample window,
    from a bank of pirate_amd.simpulse pulses (DM 2-50, intrinsic width 1-32 ms log-uniform,
    scattering time 0-10 ms at 600 MHz, spectral index -2..2), scaled to an SNR drawn from
    U(5, 30) and placed at a random DM trial and time; the tile around each pulse is a
    positive (label 1);
  * tiles of the other beams are negatives (label 0).

Every tile also carries the plain peak finder's statistic over the same region: the maximum
of out_max over its 8 coarse DMs x 32 coarse times (bare-kernel weights, so out_max is a
boxcar SNR for unit-variance noise). completeness_at_far compares the scorer with it.

Pulses span up to three chunks (dispersion sweep up to ~1 s plus scattering tails): a pulse
whose dedispersed peak lies in chunk c+1 is added to a three-chunk accumulator while chunk c is
formed. Dedispersion is linear, so the chain sees exactly noise + RFI + the simulated pulses.
"""

import concurrent.futures
import multiprocessing
import os

import numpy as np

from .simpulse import SinglePulse, dispersion_delay

CHAIN_CONFIG = dict(
    zone_nfreq=[1024], zone_freq_edges=[400.0, 800.0], time_sample_ms=1.0, toplevel_tree_rank=10,
    time_samples_per_chunk=2048, dtype='float32', beams_per_gpu=16, beams_per_batch=16, num_active_batches=1,
    frequency_subband_counts=[1], future_write_max_samples=0,
    primary_trees=[dict(num_early_triggers=0, max_width=16, dm_downsampling=0, time_downsampling=0,
                        wt_dm_downsampling=64, wt_time_downsampling=64)])

TILE_R, TILE_T = 64, 256


class Template:
    """One simulated pulse at SNR 1 on a dense (nfreq, L) window; the dedispersed peak (pulse
    centre at the bottom of the band) is sample peak_rel of the window."""

    def __init__(self, dm, width_ms, tau600_ms, alpha, data, peak_rel):
        self.dm, self.width_ms, self.tau600_ms, self.alpha = dm, width_ms, tau600_ms, alpha
        self.data, self.peak_rel = data, peak_rel


def _make_template(args):
    dm, width_ms, tau600_ms, alpha, nfreq, flo, fhi = args
    sm = tau600_ms * (600.0 / 1000.0) ** 4          # scattering time at 1 GHz, in ms
    t_inf = 0.05
    p = SinglePulse(dm=dm, sm=sm, intrinsic_width=1.0e-3 * width_ms, spectral_index=alpha,
                    undispersed_arrival_time_sec=t_inf, time_sample_ms=1.0, snr=1.0,
                    freq_edges_MHz=np.linspace(flo, fhi, nfreq + 1))
    L = p.it_end - p.it_start
    dense = np.zeros((nfreq, L), dtype=np.float32)
    p.add_to_timestream(dense, p.it_start)
    peak = int(round((t_inf + dispersion_delay(dm, flo)) / 1.0e-3))
    return dm, width_ms, tau600_ms, alpha, dense, peak - p.it_start


def make_template_bank(n, seed, nproc=None, dm_range=(2.0, 50.0), width_ms=(1.0, 32.0), tau600_ms=(0.0, 10.0),
                       alpha=(-2.0, 2.0), nfreq=1024, band=(400.0, 800.0)):
    """n simulated pulses with parameters drawn from numpy's PCG64(seed), built in nproc worker
    processes (fresh interpreters: no GPU state is inherited)."""
    rng = np.random.default_rng(seed)
    args = [(float(rng.uniform(*dm_range)), float(np.exp(rng.uniform(np.log(width_ms[0]), np.log(width_ms[1])))),
             float(rng.uniform(*tau600_ms)), float(rng.uniform(*alpha)), nfreq, band[0], band[1]) for _ in range(n)]
    nproc = nproc or min(16, len(os.sched_getaffinity(0)) if hasattr(os, 'sched_getaffinity') else 4)
    if nproc <= 1:
        res = [_make_template(a) for a in args]
    else:
        ctx = multiprocessing.get_context('spawn')
        with concurrent.futures.ProcessPoolExecutor(nproc, mp_context=ctx) as ex:
            res = list(ex.map(_make_template, args, chunksize=max(1, n // (4 * nproc))))
    return [Template(*r) for r in res]


class TileStream:
    """Runs the chain on synthetic chunks and yields labelled tile origins (see module docstring).

    step() forms and dedisperses one chunk of all B beams and returns a dict:
        'source':   scorer.TileSource over the chunk's tree rows (valid until the next step)
        'pos', 'neg': (n, 3) int32 origins (beam, row0, t0) of positive / negative tiles
        'pf_pos', 'pf_neg': peak-finder statistic of each tile
        'snr', 'dm', 'width_ms', 'tau600_ms': parameters of each positive's pulse
    Negatives are neg_per_beam random tiles of every non-pulse beam. The first chunk returns no
    tiles (its rows lack their history)."""

    def __init__(self, bank, nbeams=16, pulse_frac=0.5, neg_per_beam=8, seed=0, snr_range=(5.0, 30.0),
                 rfi=True, device='cuda'):
        import torch
        from . import DedispersionConfig, DedispersionPlan, GpuDedisperser
        d = dict(CHAIN_CONFIG, beams_per_gpu=nbeams, beams_per_batch=nbeams)
        self.cfg = DedispersionConfig.from_dict(d)
        self.plan = DedispersionPlan(self.cfg)
        self.dd = GpuDedisperser(self.plan)
        self.dd.allocate()
        self.dd.fill_host_weights(0)
        self.B, self.F, self.T = nbeams, self.plan.nfreq, self.plan.nt_in
        self.nrows = self.dd.tree_rows(0).struct.nrows      # DM trials (one full-band subband)
        self.bank = bank
        self.gpu_bank = [torch.from_numpy(t.data).to(device) for t in bank]
        self.npulse_beams = int(round(pulse_frac * nbeams))
        self.neg_per_beam = neg_per_beam
        self.snr_range = snr_range
        self.rfi = rfi
        self.rng = np.random.default_rng(seed)
        self.gen = torch.Generator(device=device)
        self.gen.manual_seed(int(seed))
        self.accum = torch.zeros((nbeams, self.F, 3 * self.T), dtype=torch.float32, device=device)
        self.dm_unit = self.cfg.dm_per_unit_delay()
        # persistent narrowband RFI: 1% of the channels at +5 sigma on half of the beams
        self.rfi_chan = []
        for b in range(nbeams):
            if rfi and self.rng.random() < 0.5:
                self.rfi_chan.append(self.rng.choice(self.F, size=max(1, self.F // 100), replace=False))
            else:
                self.rfi_chan.append(None)
        self.pending = None
        self.ichunk = 0

    def _plan_pulses(self):
        """Pulses whose dedispersed peak lies in the next chunk: added to the accumulator now."""
        rng = self.rng
        T = self.T
        pos, meta = [], []
        for b in range(self.npulse_beams):
            for j in range(T // TILE_T):
                k = int(rng.integers(len(self.bank)))
                tpl = self.bank[k]
                snr = float(rng.uniform(*self.snr_range))
                d = int(round(tpl.dm / self.dm_unit))
                prel = TILE_T * j + int(rng.integers(96, 160))
                start = T + prel - tpl.peak_rel
                L = tpl.data.shape[1]
                if start < 0 or start + L > 3 * T:
                    continue
                self.accum[b, :, start:start + L] += snr * self.gpu_bank[k]
                row0 = int(np.clip(8 * ((d - int(rng.integers(8, 56))) // 8), 0, self.nrows - TILE_R))
                t0 = int(np.clip(8 * ((prel - int(rng.integers(64, 192))) // 8), 0, T - TILE_T))
                pos.append((b, row0, t0))
                meta.append((snr, tpl.dm, tpl.width_ms, tpl.tau600_ms))
        return np.array(pos, dtype=np.int32).reshape(-1, 3), np.array(meta, dtype=np.float64).reshape(-1, 4)

    def _add_rfi(self, x):
        import torch
        rng = self.rng
        F, T = self.F, self.T
        for b in range(self.B):
            if self.rfi_chan[b] is not None:
                idx = torch.from_numpy(self.rfi_chan[b]).to(x.device)
                x[b, idx, :] += 5.0
            if not self.rfi:
                continue
            for _ in range(rng.poisson(2.0)):          # narrowband bursts
                nch = int(rng.integers(1, 9))
                f0 = int(rng.integers(0, F - nch))
                dur = int(rng.integers(8, 257))
                t0 = int(rng.integers(0, T - dur))
                x[b, f0:f0 + nch, t0:t0 + dur] += float(rng.uniform(2.0, 8.0))
            for _ in range(rng.poisson(2.0)):          # broadband zero-DM impulses
                ln = int(rng.integers(1, 4))
                t0 = int(rng.integers(0, T - ln))
                snr = float(rng.uniform(8.0, 60.0))
                x[b, :, t0:t0 + ln] += snr / np.sqrt(F * ln)

    def _pf_stat(self, out_max, origins):
        import torch
        if len(origins) == 0:
            return np.zeros(0)
        o = torch.from_numpy(origins.astype(np.int64)).to(out_max.device)
        r = o[:, 1:2, None] // 8 + torch.arange(8, device=out_max.device)[None, :, None]
        t = o[:, 2:3, None] // 8 + torch.arange(TILE_T // 8, device=out_max.device)[None, None, :]
        v = out_max[o[:, 0:1, None], r, t]
        return v.amax(dim=(1, 2)).float().cpu().numpy()

    def step(self):
        import torch
        c = self.ichunk
        T = self.T
        pos_next, meta_next = self._plan_pulses()
        with self.dd.get_input(c) as arr:
            arr.normal_(generator=self.gen)
            arr += self.accum[:, :, :T]
            self._add_rfi(arr)
        self.accum[:, :, :2 * T] = self.accum[:, :, T:].clone()
        self.accum[:, :, 2 * T:] = 0.0
        with self.dd.get_output(c) as out:
            out_max = out.out_max[0].float().clone()
        res = {'source': self.dd.tree_rows(0)}
        if c == 0 or self.pending is None:
            pos = np.zeros((0, 3), dtype=np.int32)
            meta = np.zeros((0, 4))
            neg = np.zeros((0, 3), dtype=np.int32)
        else:
            pos, meta = self.pending
            nb = self.B - self.npulse_beams
            nn = nb * self.neg_per_beam
            neg = np.stack([np.repeat(np.arange(self.npulse_beams, self.B), self.neg_per_beam),
                            8 * self.rng.integers(0, (self.nrows - TILE_R) // 8 + 1, nn),
                            8 * self.rng.integers(0, (T - TILE_T) // 8 + 1, nn)], axis=1).astype(np.int32)
        self.pending = (pos_next, meta_next)
        res.update(pos=pos, neg=neg, pf_pos=self._pf_stat(out_max, pos), pf_neg=self._pf_stat(out_max, neg),
                   snr=meta[:, 0], dm=meta[:, 1], width_ms=meta[:, 2], tau600_ms=meta[:, 3])
        self.ichunk += 1
        return res


def gather_tiles(source, origins):
    """Copies the tiles at 'origins' ((n, 3) int32: beam, row0, t0) of a tree-rows TileSource
    into a contiguous (n, 64, 256) bfloat16 device tensor (round to nearest even)."""
    import torch
    from .dedisperser import device_view
    s = source.struct
    n = len(origins)
    if n == 0:
        return torch.zeros((0, TILE_R, TILE_T), dtype=torch.bfloat16, device='cuda')
    esize = {0: 4, 1: 2, 2: 2}[s.dtype]
    dt = {0: 'float32', 1: 'float16', 2: 'bfloat16'}[s.dtype]
    lo = -s.t_lo
    rows = device_view(s.base - lo * esize, (s.nbeams, s.nrows, s.row_stride), dt, owner=source)
    o = torch.from_numpy(np.asarray(origins, dtype=np.int64)).to(rows.device)
    r = o[:, 1:2, None] + torch.arange(TILE_R, device=rows.device)[None, :, None]
    t = lo + o[:, 2:3, None] + torch.arange(TILE_T, device=rows.device)[None, None, :]
    return rows[o[:, 0:1, None], r, t].to(torch.bfloat16)


def _tiles_per_chunk(stream):
    """Upper bound of the tiles one chunk of a TileStream yields (pulse tiles + negatives)."""
    return stream.npulse_beams * (stream.T // TILE_T) + (stream.B - stream.npulse_beams) * stream.neg_per_beam


def build_dataset(stream, nchunks, log=None, extra=()):
    """Runs nchunks chunks of a TileStream (then nchunks' chunks of each (stream', nchunks') in
    extra); returns a dict of device tiles (n, 64, 256) bf16, labels (n,) float32 and numpy
    arrays pf (n,), snr, dm, width_ms, tau600_ms (NaN for negatives). The tiles go straight
    into one preallocated tensor (no concatenation, whose copy would double the peak memory of
    a dataset of 10^6 tiles or more)."""
    import torch
    jobs = [(stream, nchunks)] + list(extra)
    cap = sum(n * _tiles_per_chunk(s) for s, n in jobs)
    tiles = torch.empty((cap, TILE_R, TILE_T), dtype=torch.bfloat16, device='cuda')
    k = 0
    labels, pf, meta = [], [], []
    ntot = sum(n for _, n in jobs)
    done = 0
    for st, n in jobs:
        for i in range(n):
            r = st.step()
            for kind in ('pos', 'neg'):
                o = r[kind]
                if len(o) == 0:
                    continue
                if k + len(o) > cap:
                    raise RuntimeError('build_dataset: more tiles than the per-chunk bound')
                tiles[k:k + len(o)] = gather_tiles(r['source'], o)
                k += len(o)
                labels.append(np.full(len(o), 1.0 if kind == 'pos' else 0.0, dtype=np.float32))
                pf.append(r['pf_' + kind])
                if kind == 'pos':
                    meta.append(np.stack([r['snr'], r['dm'], r['width_ms'], r['tau600_ms']], axis=1))
                else:
                    meta.append(np.full((len(o), 4), np.nan))
            done += 1
            if log and done % max(1, ntot // 10) == 0:
                log(f'dataset: {done}/{ntot} chunks, {k} tiles')
    torch.cuda.synchronize()
    meta = np.concatenate(meta) if meta else np.zeros((0, 4))
    return dict(tiles=tiles[:k], labels=torch.from_numpy(np.concatenate(labels) if labels else np.zeros(0, np.float32)).cuda(),
                pf=np.concatenate(pf) if pf else np.zeros(0), snr=meta[:, 0], dm=meta[:, 1], width_ms=meta[:, 2],
                tau600_ms=meta[:, 3])


def train(scorer, data, epochs=1, batch=512, lr=2e-3, seed=0, log=None):
    """Minibatch Adam on data['tiles'] / data['labels']; returns the mean loss per epoch."""
    import torch
    from .scorer import TileSource
    n = data['tiles'].shape[0]
    g = torch.Generator(device='cuda')
    g.manual_seed(int(seed))
    hist = []
    for ep in range(epochs):
        perm = torch.randperm(n, generator=g, device='cuda')
        tot = torch.zeros((), dtype=torch.float64, device='cuda')
        nb = 0
        for i in range(0, n - batch + 1, batch):
            idx = perm[i:i + batch]
            x = data['tiles'][idx].contiguous()
            y = data['labels'][idx].contiguous()
            tot += scorer.train_step(TileSource.from_tiles(x), y, lr=lr).double()
            nb += 1
        hist.append(float(tot) / max(nb, 1))
        if log:
            log(f'train: epoch {ep + 1}/{epochs}, mean loss {hist[-1]:.4f}')
    return hist


def score(scorer, tiles, batch=8192):
    """Logits of (n, 64, 256) device tiles, as a numpy array."""
    import torch
    out = []
    for i in range(0, tiles.shape[0], batch):
        out.append(scorer.forward(tiles[i:i + batch].contiguous()))
    return torch.cat(out).cpu().numpy() if out else np.zeros(0)


def completeness_at_far(score_pos, score_neg, far):
    """Threshold at which a fraction far of the negatives score above it (the empirical
    (1 - far) quantile of score_neg) and the fraction of positives above it."""
    neg = np.sort(np.asarray(score_neg, dtype=np.float64))
    k = int(np.floor(far * len(neg)))                # negatives allowed above the threshold
    thr = neg[len(neg) - 1 - k] if k < len(neg) else -np.inf
    return float(np.mean(np.asarray(score_pos) > thr)), float(thr)
