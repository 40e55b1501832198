"""Benchmark: real-time beams of the MI355X dedispersion chain at CHORD-like scale.

Workload (one "step" = one chunk of one beam batch through the whole chain: tree gridding
fused with lagged downsampling, stage-1 trees, lag ring, stage-2 trees + subbands, peak finding):
16 beams x 16384 channels (300-1500 MHz), 1 ms samples, toplevel tree rank 14, 4 primary
trees (1/2/4/8 ms), chunks of 2048 samples, float32 storage and arithmetic.

    real-time beams = beams_per_batch * chunk_duration / (wall time per step)

(the reference's formula, src_lib/GpuDedisperser.cpp:1676). The input chunk is resident
in the dedisperser's input slot (HBM) before the timed region. Per-kernel times come from
HIP events recorded on the dedisperser's compute stream (a separate, shorter pass), and
the roofline object reports the dominant kernel's algorithmic bytes / average duration
against MI355X's 8 TB/s HBM peak; its "traffic" (HBM bytes per launch) comes from the
rocprofv3 FETCH_SIZE / WRITE_SIZE summary committed in profiles/ for this workload, when
present. The CPU baseline runs the oracle (a single-threaded C++ restatement of the
reference's CPU dedisperser) on one beam of the same shapes.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
"""

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, 'tests'))

HBM_PEAK_GBPS = 8000.0
TRAFFIC_JSON = os.path.join(ROOT, 'profiles', 'chain_chord16k_fp32_rocprof.json')
K_DM = 4.148808e6   # ms MHz^2 / (pc cm^-3)


def kernel_alg_bytes(plan, dd, name):
    """Algorithmic (compulsory) HBM bytes of one launch of kernel 'name' (one beam batch)."""
    S = 4 if plan.dtype == 'float32' else 2
    B, F, T = plan.beams_per_batch, plan.nfreq, plan.nt_in
    R = plan.config.toplevel_tree_rank
    npri = plan.num_primary_trees
    if name == 'grid_lds':
        return B * (F * T + 2**R * T + sum(2**(R - 1) * (T >> p) for p in range(1, npri))) * S
    if name.startswith('stage1_p'):
        p = int(name[len('stage1_p'):])
        rp = R - (1 if p else 0)
        return B * 2 * 2**rp * (T >> p) * S
    if name.startswith('stage2_t') or name.startswith('pf_t'):
        it = int(name.split('_t')[1])
        t = plan.trees[it]
        streams = t.ndm_out * t.nmultiplets_M
        if name.startswith('stage2_t'):
            return B * (2**(t.amb_rank + t.dd_rank) + streams) * t.nt_ds * S
        return B * (streams * t.nt_ds * S + t.ndm_out * t.nt_out * (S + 4))
    if name.startswith('spill_p'):
        return 0    # host-spill scatter: its traffic is PCIe copies, not part of the HBM byte model
    raise ValueError(name)


def chain_alg_bytes(plan):
    """B_alg of the whole chain per batch: input read + lag-ring write and read + outputs."""
    S = 4 if plan.dtype == 'float32' else 2
    B, F, T = plan.beams_per_batch, plan.nfreq, plan.nt_in
    R = plan.config.toplevel_tree_rank
    ring = sum(2**(R - (1 if p else 0)) * (T >> p) for p in range(plan.num_primary_trees))
    outs = sum(t.ndm_out * t.nt_out * (S + 4) for t in plan.trees)
    return B * (F * T * S + 2 * ring * S + outs)


def inject_pulses(x, cfg, rng, npulses=8):
    """Adds npulses dispersed boxcar pulses per beam to x (B, F, T) torch tensor (in place).
    DM ~ U(0, 3000), width ~ logU(1, 16) ms, spectral index ~ U(-2, 2), arrival times spread
    over the chunk (delays wrap within the chunk: the chunk is reused at every step)."""
    import torch
    B, F, T = x.shape
    edges = np.asarray(cfg.zone_freq_edges, dtype=np.float64)
    nz = np.asarray(cfg.zone_nfreq)
    freqs = np.concatenate([np.linspace(edges[i], edges[i + 1], nz[i], endpoint=False) + 0.5 * (edges[i + 1] - edges[i]) / nz[i]
                            for i in range(len(nz))])
    fhi = edges[-1]
    for b in range(B):
        for _ in range(npulses):
            dm = rng.uniform(0, 3000)
            width = int(round(np.exp(rng.uniform(0, np.log(16)))))
            alpha = rng.uniform(-2, 2)
            t0 = rng.integers(0, T)
            delay = np.round(K_DM * dm * (freqs**-2 - fhi**-2) / cfg.time_sample_ms).astype(np.int64)
            amp = 20.0 / np.sqrt(F * width) * (freqs / freqs.mean())**alpha
            for w in range(width):
                tt = torch.from_numpy((t0 + delay + w) % T).to(x.device)
                ff = torch.arange(F, device=x.device)
                x[b].index_put_((ff, tt), torch.from_numpy(amp.astype(np.float32)).to(x.device, x.dtype), accumulate=True)


def cpu_baseline(cfg_dict, nchunks=24, seed=0):
    """Oracle (single-threaded C++ restatement of the reference CPU dedisperser,
    sophistication 1) on one beam of the bench shapes; returns real-time beams."""
    from oracle import oracle as orc
    from gpu_cases import cfg2_with_beams
    d = cfg2_with_beams(1)
    ocfg = orc.OracleConfig.from_yaml_dict(d)
    dd = orc.Dedisperser(ocfg, 1)
    for it in range(len(dd.trees)):
        dd.set_weights(it, orc.bare_weights(*dd.wt_shape(it)))
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((1, ocfg.nfreq, ocfg.nt_in)).astype(np.float32)
    dd.dedisperse(x, 0, 0)   # untimed warm-up chunk
    t0 = time.perf_counter()
    for c in range(1, nchunks + 1):
        dd.dedisperse(x, c, 0)
    dt = (time.perf_counter() - t0) / nchunks
    chunk_s = ocfg.nt_in * cfg_dict['time_sample_ms'] * 1e-3
    return dict(value=round(1 * chunk_s / dt, 4), unit='beams', cores=1, kind='port',
                sample=f'1 beam x {nchunks} chunks of the same shapes (16384 ch, rank 14, 4 trees, T=2048, fp32), '
                       f'{dt:.3f} s per beam-chunk on one core of {platform.processor() or platform.machine()}')


def measured_traffic(kernel):
    """HBM bytes per launch (read + write) of 'kernel' from the committed rocprofv3 summary."""
    try:
        with open(TRAFFIC_JSON) as f:
            rec = json.load(f)['kernels'][kernel]
        return int(rec['hbm_read_bytes'] + rec['hbm_write_bytes'])
    except (OSError, KeyError, ValueError):
        return None


def measure(pa, torch, cfg_dict, steps, warmup, timing_steps):
    """Runs the chain on one resident chunk; returns (plan, seconds per step, per-kernel dict, GiB)."""
    cfg = pa.DedispersionConfig.from_dict(cfg_dict)
    plan = pa.DedispersionPlan(cfg)
    dd = pa.GpuDedisperser(plan)
    dd.allocate()
    gib = dd.gpu_bytes() / 2**30
    print(f'[bench] {plan.dtype}: allocated {gib:.2f} GiB of device memory, {dd.host_bytes() / 2**30:.2f} GiB of '
          f'pinned host memory', file=sys.stderr, flush=True)

    from oracle import oracle as orc   # weights helper only (bare-kernel weights)
    for it in range(plan.ntrees):
        w = orc.bare_weights(*dd.wt_shape(it)[1:])
        dd.fill_all_weights(it, np.broadcast_to(w, dd.wt_shape(it)))

    rng = np.random.default_rng(2)
    seq = 0
    with dd.get_input(seq) as arr:
        gen = torch.Generator(device=arr.device)
        gen.manual_seed(2)
        arr.normal_(generator=gen)
        inject_pulses(arr, cfg, rng)
    with dd.get_output(seq):
        pass
    seq += 1

    def step():
        nonlocal seq
        # The input slot keeps the resident chunk (stale-slot reuse): no host or HBM copy.
        s = torch.cuda.current_stream().cuda_stream
        dd._acquire_input(seq, s)
        dd._release_input_and_launch_dd_kernels(seq, s)
        dd._acquire_output(0, seq, s)
        dd._release_output(0, seq, s)
        seq += 1

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    print(f'[bench] {plan.dtype}: {steps} steps, {dt * 1e3:.3f} ms/step', file=sys.stderr, flush=True)

    # Per-kernel times (HIP events on the compute stream), separate pass.
    dd.set_kernel_timing(True)
    for _ in range(timing_steps):
        step()
    torch.cuda.synchronize()
    kt = dd.kernel_timing()
    dd.set_kernel_timing(False)
    per_kernel = {}
    for name, (n, ms) in kt.items():
        avg_ms = ms / n
        nbytes = kernel_alg_bytes(plan, dd, name)
        per_kernel[name] = dict(avg_ms=round(avg_ms, 4), alg_bytes=int(nbytes), GBps=round(nbytes / (avg_ms * 1e-3) / 1e9, 1))
    del dd
    torch.cuda.synchronize()
    return plan, dt, per_kernel, gib


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=20)
    ap.add_argument('--warmup', type=int, default=5)
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--timing-steps', type=int, default=4)
    ap.add_argument('--no-fp16', action='store_true', help='skip the float16-storage measurement')
    ap.add_argument('--calib-copy', action='store_true',
                    help='run one 1 GiB device-to-device copy first (reference dispatch for HBM counters)')
    args = ap.parse_args()

    if args.gpus != 1:
        print(f'bench.py: this build measures one GPU only (--gpus {args.gpus} requested). The chain shards '
              'by beam with no communication; run one process per GPU.', file=sys.stderr)
        sys.exit(2)

    import torch
    import pirate_amd as pa
    from gpu_cases import CFG2

    if args.calib_copy:
        src = torch.ones(2**28, dtype=torch.float32, device='cuda')
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        print('[bench] calibration copy: 1 GiB read + 1 GiB write', file=sys.stderr, flush=True)

    plan, dt, per_kernel, gib = measure(pa, torch, CFG2, args.steps, args.warmup, args.timing_steps)
    B, F, T = plan.beams_per_batch, plan.nfreq, plan.nt_in
    cfg = plan.config
    dom = max(per_kernel, key=lambda k: per_kernel[k]['avg_ms'])
    for k, v in sorted(per_kernel.items(), key=lambda kv: -kv[1]['avg_ms']):
        print(f'[bench] kernel {k:12s} {v["avg_ms"]:8.3f} ms  {v["GBps"]:8.1f} GB/s (alg bytes {v["alg_bytes"] / 1e6:.1f} MB)',
              file=sys.stderr, flush=True)

    chunk_s = T * cfg.time_sample_ms * 1e-3
    value = B * chunk_s / dt
    chain_gbps = chain_alg_bytes(plan) / dt / 1e9
    result = {
        'metric': 'real-time beams/GPU (16k-ch tree dd+peaks) FP16-batched vs FP32 plain; HBM GB/s',
        'value': round(value, 2),
        'unit': 'beams',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': round(dt * 1e3, 4),
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f32',
        'data': 'synthetic: N(0,1) noise + 8 dispersed boxcar pulses per beam (DM U(0,3000), width logU(1,16) ms, '
                'spectral index U(-2,2)), one resident chunk reused every step',
        'config': {'workload': 'chord16k plain FP32 chain: 16 beams x 16384 ch 300-1500 MHz, 1 ms, '
                               'tree rank 14, 4 primary trees (1/2/4/8 ms), chunk 2048, no subbands, max_width 16/8/4/2',
                   'beams_per_batch': B, 'nfreq': F, 'nt_chunk': T, 'toplevel_tree_rank': cfg.toplevel_tree_rank,
                   'num_primary_trees': plan.num_primary_trees, 'chain_alg_GBps': round(chain_gbps, 1),
                   'chain_alg_bytes_per_step': int(chain_alg_bytes(plan)), 'per_kernel': per_kernel},
        'roofline': {'bound': 'hbm', 'kernel': dom, 'achieved': per_kernel[dom]['GBps'], 'peak': HBM_PEAK_GBPS,
                     'unit': 'GB/s', 'frac': round(per_kernel[dom]['GBps'] / HBM_PEAK_GBPS, 4),
                     'traffic': measured_traffic(dom)},
    }
    if not args.no_fp16:
        # Same workload with float16 storage (float32 arithmetic): the reduced-precision tier.
        p16, dt16, pk16, gib16 = measure(pa, torch, dict(CFG2, dtype='float16'), max(args.steps // 2, 1),
                                         args.warmup, args.timing_steps)
        dom16 = max(pk16, key=lambda k: pk16[k]['avg_ms'])
        result['fp16'] = {'value': round(B * chunk_s / dt16, 2), 'unit': 'beams', 'ms_per_step': round(dt16 * 1e3, 4),
                          'ratio_vs_fp32': round(dt / dt16, 3), 'dtype': 'f16 storage, f32 arithmetic',
                          'gpu_GiB': round(gib16, 2), 'dominant_kernel': dom16,
                          'dominant_alg_GBps': pk16[dom16]['GBps'], 'per_kernel': pk16}
    if not args.no_cpu_baseline:
        result['cpu_baseline'] = cpu_baseline(CFG2)
    print(json.dumps(result), flush=True)


if __name__ == '__main__':
    main()
