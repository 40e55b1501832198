"""Random dedispersion configs, drawn the way the reference's randomized tests draw them.

Follows DedispersionConfig::make_random() (src_lib/DedispersionConfig.cpp:785-1007) with
gpu_valid=false (no kernel registry here): a base stage-2 rank, tree rank, 1-5 frequency zones,
1-4 primary trees, random peak-finder widths, output and weight downsampling factors, early
triggers within a total budget, and a random (time chunk, beams per batch, batch count)
split of a fixed sample budget. The reference's subband-count draw and restriction are
FrequencySubbands::make_random_subband_counts / restrict_subband_counts
(src_lib/FrequencySubbands.cpp:174-231).

Two draws are narrowed to what the reference's GPU kernels accept (its gpu_valid=true mode
draws from a kernel registry instead): the peak finder's output time downsampling is at most
32 (makefile_helper.py:194-330 compiles no larger Dout), and a rank-2 toplevel tree has one
primary tree (a second one would need a stage-1 kernel of rank 0, which
GpuDedispersionKernel rejects: src_lib/DedispersionKernel.cu:496).

Every draw comes from a numpy Generator, so a seed names a config. gpu_supported() reports
whether pirate_amd's GPU chain accepts a config, and the reason if not.
"""

import math

import numpy as np

MAX_PF_WIDTH = 32


def make_random_subband_counts(rng, pf_rank):
    """FrequencySubbands::make_random_subband_counts(pf_rank) (src_lib/FrequencySubbands.cpp:218-231)."""
    counts = []
    for level in range(pf_rank):
        cap = (2 ** (pf_rank + 1 - level) - 1) if level > 0 else 2 ** pf_rank
        counts.append(int(rng.integers(0, cap + 1)))
    counts.append(1)
    return counts


def restrict_subband_counts(counts, et_level, new_pf_rank):
    """FrequencySubbands::restrict_subband_counts (src_lib/FrequencySubbands.cpp:174-215):
    first drop et_level top levels (an early trigger sees a shorter band), then re-index to
    new_pf_rank levels; level 0 holds disjoint bands, so converting between level 0 and an
    overlapping level maps n bands to 2n-1 (or n to n/2+1)."""
    src_rank = len(counts) - 1
    early_rank = max(src_rank - et_level, 0)
    early = [0] * (early_rank + 1)
    early[early_rank] = 1
    for lv in range(early_rank):
        cap = (2 ** (early_rank + 1 - lv) - 1) if lv > 0 else 2 ** early_rank
        early[lv] = min(counts[lv], cap)
    dst = [0] * (new_pf_rank + 1)
    dst[new_pf_rank] = 1
    for lv in range(new_pf_rank):
        s = lv - new_pf_rank + early_rank
        if s < 0 or early[s] == 0:
            dst[lv] = 0
        elif s == 0 and lv > 0:
            dst[lv] = 2 * early[s] - 1
        elif s > 0 and lv == 0:
            dst[lv] = early[s] // 2 + 1
        else:
            dst[lv] = early[s]
    return dst


def integers_with_bounded_product(rng, n, bound):
    """n integers >= 1 whose product is <= bound, each factor log-uniform over what the
    remaining budget allows (the role of ksgpu::random_integers_with_bounded_product in the
    reference's make_random; this draw is this file's own)."""
    out = []
    rem = bound
    order = rng.permutation(n)
    vals = [1] * n
    for k, i in enumerate(order):
        if k == n - 1:
            hi = rem
        else:
            hi = max(1, int(math.floor(rem ** rng.uniform(0.0, 1.0))))
        v = int(math.floor(math.exp(rng.uniform(0.0, math.log(hi + 1)))))
        v = min(max(v, 1), hi)
        vals[i] = v
        rem //= v
    out = vals
    return out


def _pt(num_et, max_width, dm_ds, t_ds, wt_dm, wt_t):
    return dict(num_early_triggers=num_et, max_width=max_width, dm_downsampling=dm_ds, time_downsampling=t_ds,
                wt_dm_downsampling=wt_dm, wt_time_downsampling=wt_t)


def make_random_config(rng, max_toplevel_rank=10, max_early_triggers=5, dtypes=('float32', 'float16', 'bfloat16'),
                       sample_budget=8192):
    """A reference-valid config dict (the YAML grammar's keys) drawn from 'rng'."""
    max_stage2_rank = (max_toplevel_rank + 1) // 2
    dtype = str(rng.choice(list(dtypes)))
    nbits = 32 if dtype == 'float32' else 16
    nelts = 1024 // nbits                      # 128-byte segments

    def random_key(dd_rank):
        wmax = 2 ** int(rng.integers(0, int(math.log2(MAX_PF_WIDTH)) + 1))
        i = int(rng.integers(0, 6 if nbits == 32 else 5))
        j = int(rng.integers(0, 6 - i))
        return dict(dd_rank=dd_rank, Wmax=wmax, Tinner=2 ** j, Dout=2 ** i * (32 // nbits),
                    sc=make_random_subband_counts(rng, (dd_rank + 1) // 2))

    keys = [random_key(int(rng.integers(1, max_stage2_rank + 1)))]
    subband_counts = keys[0]['sc']
    dd0 = keys[0]['dd_rank']
    rank = int(rng.integers(max(2 * dd0 - 1, 2), 2 * dd0 + 1))

    nzones = int(rng.integers(1, 6))
    zlo, zhi = max(2 ** rank // 4, 1), 2 ** rank
    zone_nfreq = [int(x) for x in rng.integers(zlo, zhi, size=nzones)] if zhi > zlo else [zlo] * nzones
    edges = sorted(float(x) for x in rng.uniform(200.0, 2000.0, size=nzones + 1))
    time_sample_ms = float(rng.uniform(1.0, 10.0))

    ds_dd = rank // 2
    npri = 1
    if ds_dd >= 1 and rank >= 3:
        ds_sc = restrict_subband_counts(subband_counts, 0, (ds_dd + 1) // 2)
        npri = int(rng.integers(1, 5))
        for _ in range(1, npri):
            k = random_key(ds_dd)
            k['sc'] = ds_sc
            keys.append(k)

    nt_div = nelts * 2 ** (npri - 1)
    v = integers_with_bounded_product(rng, 3, max(sample_budget // nt_div, 1))
    nt_chunk = v[0] * nt_div
    beams_per_batch = v[1]
    beams_per_gpu = v[1] * v[2]
    num_active_batches = int(rng.integers(1, v[2] + 1))
    max_delay = 2 ** (rank + npri - 1)
    max_gpu_clag = int(rng.integers(0, max_delay // nt_chunk + 2))
    fw = 0 if rng.uniform() < 0.25 else int(rng.integers(0, 4 * nt_chunk + 1))
    while (nt_chunk // nt_div) % 2 == 0:
        nt_div *= 2

    trees, max_et = [], []
    for ipri, k in enumerate(keys):
        prank = rank - 1 if ipri else rank
        st1 = prank // 2
        st2 = prank - st1
        nt_ds = nt_div // 2 ** ipri
        min_wtds = 1024 // (k['Tinner'] * nbits)
        lo_t = int(math.log2(min_wtds))
        hi_t = int(math.log2(nt_ds)) if k['Tinner'] == 1 else lo_t
        lo_d, hi_d = (st2 + 1) // 2, prank
        pt = _pt(0, k['Wmax'], 0, k['Dout'], 2 ** int(rng.integers(lo_d, hi_d + 1)), 2 ** int(rng.integers(lo_t, hi_t + 1)))
        trees.append(pt)
        min_et_rank = max(st1 + 1, int(math.log2(pt['wt_dm_downsampling'])))
        m = 0
        while prank - (m + 1) >= min_et_rank:
            m += 1
        max_et.append(m)

    budget = max_early_triggers
    for ipri in rng.permutation(npri):
        cap = min(max_et[ipri], budget)
        n = int(rng.integers(0, cap + 1))
        trees[ipri]['num_early_triggers'] = n
        budget -= n

    return dict(zone_nfreq=zone_nfreq, zone_freq_edges=edges, time_sample_ms=time_sample_ms, toplevel_tree_rank=rank,
                time_samples_per_chunk=nt_chunk, dtype=dtype, beams_per_gpu=beams_per_gpu,
                beams_per_batch=beams_per_batch, num_active_batches=num_active_batches,
                future_write_max_samples=fw, max_gpu_clag=max_gpu_clag, frequency_subband_counts=subband_counts,
                primary_trees=trees)


def gpu_supported(cfg_dict):
    """(True, '') if pirate_amd's GPU chain accepts the config, else (False, reason). Builds
    the plan and the dedisperser object (no GPU memory is touched before allocate())."""
    import pirate_amd as pa
    try:
        plan = pa.DedispersionPlan(pa.DedispersionConfig.from_dict(cfg_dict))
        pa.GpuDedisperser(plan)
    except Exception as e:      # noqa: BLE001 -- the reason is the point
        return False, str(e)
    return True, ''


def draw_supported(seed, max_tries=200, **kw):
    """The first GPU-supported config drawn from default_rng(seed), with the number of
    rejected draws before it."""
    rng = np.random.default_rng(seed)
    for ntry in range(max_tries):
        d = make_random_config(rng, **kw)
        ok, _ = gpu_supported(d)
        if ok:
            return d, ntry
    raise RuntimeError(f'no supported config in {max_tries} draws (seed {seed})')
