"""Autotuner of the chain's launch shapes.

    python -m pirate_amd.autotune [--workloads chord16k_fp32,chord16k_fp16,batched_r12,batched_r13,batched_r14]
                                  [--out pirate_amd/tuning_table.json] [--steps 10]

The launch shapes change no output, only speed (GpuDedisperser.set_tuning):
  * pf_target_blocks: the peak finder splits each chunk's time axis into the fewest segments
    (one work group each, warmed up over the history samples) that give at least this many
    work groups;
  * grid_xcd_run: consecutive channel pairs that the fused gridding kernel gives each of the 8
    XCDs (0: plain launch order), which sets how the input rows of neighbouring tree channels
    share an XCD's L2;
  * fork: where primary trees p >= 1 leave for their own streams (0 after the gridding, 1 after
    tree 0's stage 2, -1 automatic).
For each workload (one dedisperser, inputs resident in its input slot) the tuner runs a
coordinate sweep (each parameter in turn, the others at their best so far, up to two rounds):
each candidate is timed as the median of three runs of 'steps' chunk steps against the current
best re-timed beside it, and replaces it only when more than 1.5% faster, so that clock noise
does not pick the shape; it records the best shape, its time and the defaults' time. The
table is keyed by a workload signature (tree rank, channels, primary trees, storage type,
beams per batch); GpuDedisperser.apply_tuning_table() looks a dedisperser's signature up in it.
"""

import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_TABLE = os.path.join(HERE, 'tuning_table.json')

SWEEP = {
    'pf_target_blocks': [512, 1024, 2048, 4096, 8192],
    'grid_xcd_run': [0, 2, 4, 8, 16],
    'fork': [-1, 0, 1],
}


def signature(plan):
    c = plan.config
    return (f'R{c.toplevel_tree_rank}_F{plan.nfreq}_P{plan.num_primary_trees}_{plan.dtype}'
            f'_B{plan.beams_per_batch}_T{plan.nt_in}')


def load_table(path=DEFAULT_TABLE):
    try:
        with open(path) as f:
            return json.load(f)
    except OSError:
        return {}


def _workload_configs():
    """Named workloads: the CHORD-like 16-beam rank-14 chain (float32 and float16) and the three
    beam groups of the batched heterogeneous workload (32 beams each, float16)."""
    base = dict(time_sample_ms=1.0, future_write_max_samples=0, zone_freq_edges=[300.0, 1500.0],
                time_samples_per_chunk=2048, num_active_batches=1, frequency_subband_counts=[1])

    def pt(w):
        return dict(num_early_triggers=0, max_width=w, dm_downsampling=0, time_downsampling=0,
                    wt_dm_downsampling=64, wt_time_downsampling=64)

    chord = dict(base, dtype='float32', zone_nfreq=[16384], toplevel_tree_rank=14, beams_per_gpu=16,
                 beams_per_batch=16, primary_trees=[pt(16), pt(8), pt(4), pt(2)])
    out = {'chord16k_fp32': chord, 'chord16k_fp16': dict(chord, dtype='float16')}
    for R, F, npri in ((12, 4096, 3), (13, 8192, 4), (14, 16384, 5)):
        out[f'batched_r{R}'] = dict(base, dtype='float16', zone_nfreq=[F], toplevel_tree_rank=R, beams_per_gpu=32,
                                    beams_per_batch=32, primary_trees=[pt(max(16 >> p, 1)) for p in range(npri)])
    return out


def time_steps(dd, steps, warmup=2):
    """Seconds per chunk step (resident input, one output acquired per step)."""
    import torch
    s = torch.cuda.current_stream().cuda_stream
    seq = dd._tune_seq

    def step(seq):
        dd._acquire_input(seq, s)
        dd._release_input_and_launch_dd_kernels(seq, s)
        dd._acquire_output(0, seq, s)
        dd._release_output(0, seq, s)

    for _ in range(warmup):
        step(seq)
        seq += 1
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step(seq)
        seq += 1
    torch.cuda.synchronize()
    dd._tune_seq = seq
    return (time.perf_counter() - t0) / steps


def _median_time(dd, steps, reps):
    return sorted(time_steps(dd, steps) for _ in range(reps))[reps // 2]


def tune(dd, steps=10, rounds=2, reps=3, margin=0.015, log=print):
    """Coordinate descent over SWEEP. Each candidate is timed as the median of 'reps' runs of
    'steps' chunk steps, and the current best is re-timed right after it the same way; the
    candidate replaces it only when faster by more than 'margin' (run-to-run drift of the
    clocks is about 1%), so noise does not move the shape away from the defaults."""
    defaults = {k: dd.get_tuning(k) for k in SWEEP}
    best = dict(defaults)
    for k, v in best.items():
        dd.set_tuning(k, v)
    t_default = _median_time(dd, steps, reps)
    points = []
    for r in range(rounds):
        changed = False
        for name, values in SWEEP.items():
            for v in values:
                if v == best[name]:
                    continue
                dd.set_tuning(name, v)
                t = _median_time(dd, steps, reps)
                dd.set_tuning(name, best[name])
                t_ref = _median_time(dd, steps, reps)
                points.append({'round': r, name: v, 'others': dict(best), 'ms': round(t * 1e3, 4),
                               'best_ms': round(t_ref * 1e3, 4)})
                if t < (1.0 - margin) * t_ref:
                    best[name] = v
                    dd.set_tuning(name, v)
                    changed = True
        log(f'round {r}: best {best}')
        if not changed:
            break
    t_best = _median_time(dd, steps, reps)
    log(f'best {best}: {t_best * 1e3:.3f} ms/step; defaults {defaults}: {t_default * 1e3:.3f} ms/step')
    return best, t_best, t_default, points


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split('\n')[0])
    ap.add_argument('--workloads', default=','.join(_workload_configs()))
    ap.add_argument('--out', default=DEFAULT_TABLE)
    ap.add_argument('--steps', type=int, default=10)
    a = ap.parse_args(argv)

    import torch
    import pirate_amd as pa

    def log(msg):
        print(f'[autotune] {msg}', file=sys.stderr, flush=True)

    table = load_table(a.out)
    cfgs = _workload_configs()
    for name in a.workloads.split(','):
        plan = pa.DedispersionPlan(pa.DedispersionConfig.from_dict(cfgs[name]))
        dd = pa.GpuDedisperser(plan)
        dd.allocate()
        for it in range(plan.ntrees):
            dd.fill_host_weights(it)
        with dd.get_input(0) as arr:
            x = torch.empty(tuple(arr.shape), dtype=torch.float32, device=arr.device)
            x.normal_()
            arr.copy_(x)
            del x
        with dd.get_output(0):
            pass
        dd._tune_seq = 1
        log(f'{name} ({signature(plan)})')
        best, t_best, t_default, points = tune(dd, steps=a.steps, log=log)
        table[signature(plan)] = {'workload': name, 'tuning': best, 'ms_per_step': round(t_best * 1e3, 4),
                                  'default_ms_per_step': round(t_default * 1e3, 4), 'sweep': points}
        del dd
        torch.cuda.synchronize()
        with open(a.out, 'w') as f:
            json.dump(table, f, indent=1, sort_keys=True)
    print(json.dumps({k: {kk: v[kk] for kk in ('workload', 'tuning', 'ms_per_step', 'default_ms_per_step')}
                      for k, v in table.items()}))


if __name__ == '__main__':
    main()
