"""Benchmark: real-time beams of the MI355X dedispersion chain, FP32 plain vs FP16 batched.

Headline (value): the plain FP32 chain at CHORD-like scale ("chord16k"): one step = one chunk
of 16 beams x 16384 channels (300-1500 MHz), 1 ms samples, toplevel tree rank 14, 4 primary
trees (1/2/4/8 ms), chunks of 2048 samples, float32 storage and arithmetic, through tree
gridding fused with lagged downsampling, stage-1 trees, lag ring, stage-2 trees and peak
finding.

fp16_batched: the batched heterogeneous workload in one GpuDedisperserGroup: 96 beams split
32/32/32 over tree ranks 12/13/14 (4096/8192/16384 channels, 300-1500 MHz) with 3/4/5 primary
trees, float16 storage (float32 arithmetic), chunks of 2048 samples; one step = one chunk of
all 96 beams, each group's chain on a branch stream of its own.

    real-time beams = beams * chunk_duration / (wall time per step)

(the reference's formula, src_lib/GpuDedisperser.cpp:1676). Inputs are resident in the input
slots (HBM) before the timed region. Per-kernel times come from HIP events around each kernel
in a separate pass with every kernel run alone on one stream (launch shape fork=2), and the
roofline object reports the dominant kernel's algorithmic bytes / average duration against
MI355X's 8 TB/s HBM peak, and beside it against a plain 1 GiB device copy timed in the same
run (device_copy_GBps); its "traffic" (HBM bytes per launch) comes from the rocprofv3
FETCH_SIZE / WRITE_SIZE summary committed in profiles/ for this workload. The CPU baseline runs
the oracle (a single-threaded C++ restatement of the reference's CPU dedisperser) on bounded
samples of the same shapes: one beam per thread on all host cores, one core, and the 1-beam
1024-channel rank-10 shape.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline] [--no-batched]
"""

import argparse
import json
import os
import platform
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, 'tests'))

HBM_PEAK_GBPS = 8000.0
HBM_GIB = 288.0 * 1e9 / 2**30       # MI355X HBM3E: 288 GB
TRAFFIC_JSON = os.path.join(ROOT, 'profiles', 'chain_chord16k_fp32_rocprof.json')
K_DM = 4.148808e6   # ms MHz^2 / (pc cm^-3)


def log(msg):
    print(f'[bench] {msg}', file=sys.stderr, flush=True)


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
            ff = torch.arange(F, device=x.device)
            vals = torch.from_numpy(amp.astype(np.float32)).to(x.device, x.dtype)
            for w in range(width):
                tt = torch.from_numpy((t0 + delay + w) % T).to(x.device)
                x[b].index_put_((ff, tt), vals, accumulate=True)


def fill_resident(arr, cfg, seed):
    """Resident synthetic chunk: N(0,1) noise + 8 pulses per beam."""
    import torch
    gen = torch.Generator(device=arr.device)
    gen.manual_seed(seed)
    if arr.dtype == torch.float32:
        arr.normal_(generator=gen)
    else:
        x = torch.empty(tuple(arr.shape), dtype=torch.float32, device=arr.device)
        x.normal_(generator=gen)
        arr.copy_(x)
        del x
    inject_pulses(arr, cfg, np.random.default_rng(seed))


# ---------------------------------------------------------------------------------------------
# CPU baseline (oracle)

def host_cpu():
    model = platform.processor() or platform.machine()
    try:
        with open('/proc/cpuinfo') as f:
            for line in f:
                if line.startswith('model name'):
                    model = line.split(':', 1)[1].strip()
                    break
    except OSError:
        pass
    ncpu = len(os.sched_getaffinity(0)) if hasattr(os, 'sched_getaffinity') else (os.cpu_count() or 1)
    return model, ncpu


def _oracle_chunks(cfg_dict, nchunks, seed, out, idx):
    from oracle import oracle as orc
    ocfg = orc.OracleConfig.from_yaml_dict(cfg_dict)
    dd = orc.Dedisperser(ocfg, 1)
    for it in range(len(dd.trees)):
        dd.set_weights(it, orc.bare_weights(*dd.wt_shape(it)))
    x = np.random.default_rng(seed).standard_normal((1, ocfg.nfreq, ocfg.nt_in)).astype(np.float32)
    dd.dedisperse(x, 0, 0)   # untimed warm-up chunk
    t0 = time.perf_counter()
    for c in range(1, nchunks + 1):
        dd.dedisperse(x, c, 0)
    out[idx] = (time.perf_counter() - t0) / nchunks


def cpu_timing(cfg_dict, nthreads, nchunks):
    """nthreads oracle instances (one beam each, threads: the oracle releases the GIL) for
    nchunks chunks after a warm-up chunk; returns (seconds per beam-chunk per thread, wall)."""
    res = [0.0] * nthreads
    th = [threading.Thread(target=_oracle_chunks, args=(cfg_dict, nchunks, 100 + i, res, i)) for i in range(nthreads)]
    t0 = time.perf_counter()
    for t in th:
        t.start()
    for t in th:
        t.join()
    return float(np.mean(res)), time.perf_counter() - t0


def cpu_baseline():
    from gpu_cases import CASES, chord16k_with_beams
    model, ncpu = host_cpu()
    nthr = max(1, min(ncpu, 16))
    d = chord16k_with_beams(1)
    chunk_s = d['time_samples_per_chunk'] * d['time_sample_ms'] * 1e-3
    t1, _ = cpu_timing(d, 1, 12)
    log(f'cpu baseline: chord16k shapes, 1 thread: {t1:.3f} s per beam-chunk')
    tn, wall = cpu_timing(d, nthr, 4)
    log(f'cpu baseline: chord16k shapes, {nthr} threads: {tn:.3f} s per beam-chunk per thread ({wall:.1f} s wall)')
    d1 = dict(CASES['r10_1024ch'])
    c1, _ = cpu_timing(d1, 1, 40)
    chunk1 = d1['time_samples_per_chunk'] * d1['time_sample_ms'] * 1e-3
    log(f'cpu baseline: 1024 ch, rank 10, 1 tree, 1 thread: {c1 * 1e3:.1f} ms per beam-chunk')
    return dict(value=round(nthr * chunk_s / tn, 4), unit='beams', cores=nthr, kind='port',
                sample=f'chord16k shapes (16384 ch, rank 14, 4 trees, T=2048, fp32): {nthr} threads x 1 beam x 4 chunks '
                       f'(after a warm-up chunk), {tn:.3f} s per beam-chunk per thread, on {nthr} of {ncpu} usable '
                       f'cores of "{model}"',
                single_core=dict(value=round(chunk_s / t1, 4), cores=1, s_per_beam_chunk=round(t1, 4),
                                 sample='1 beam x 12 chunks of the chord16k shapes'),
                r10_1024ch=dict(value=round(chunk1 / c1, 2), cores=1, ms_per_beam_chunk=round(c1 * 1e3, 3),
                                sample='1 beam, 1024 ch 400-800 MHz, rank 10, 1 tree, T=2048, fp32, 40 chunks'),
                host_cpu=model, host_cpus_usable=ncpu)


def measured_traffic(kernel):
    """HBM bytes per launch (read + write) of 'kernel' from the committed rocprofv3 summary."""
    try:
        with open(TRAFFIC_JSON) as f:
            rec = json.load(f)['kernels'][kernel]
        return int(rec['hbm_read_bytes'] + rec['hbm_write_bytes'])
    except (OSError, KeyError, ValueError):
        return None


def device_copy_gbps(torch, nbytes=2**30, reps=10):
    """HBM bytes (read + write) per second of a plain nbytes device-to-device copy (torch's copy
    kernel, on torch's current stream, timed with events on that stream): the practical ceiling
    that an HBM-bound kernel is compared with beside the published peak."""
    src = torch.ones(nbytes // 4, dtype=torch.float32, device='cuda')
    dst = torch.empty_like(src)
    for _ in range(2):
        dst.copy_(src)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    e1.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del src, dst
    return 2 * nbytes / (ms * 1e-3) / 1e9


# ---------------------------------------------------------------------------------------------
# GPU legs

def per_kernel_table(plan, kt):
    from pirate_amd.resources import kernel_alg_bytes
    out = {}
    for name, (n, ms) in kt.items():
        avg_ms = ms / n
        nbytes = kernel_alg_bytes(plan, name)
        out[name] = dict(avg_ms=round(avg_ms, 4), alg_bytes=int(nbytes), GBps=round(nbytes / (avg_ms * 1e-3) / 1e9, 1))
    return out


def measure(pa, torch, cfg_dict, steps, warmup, timing_steps):
    """Plain chain on one resident chunk; returns (plan, s/step, per-kernel dict, GiB, dd)."""
    cfg = pa.DedispersionConfig.from_dict(cfg_dict)
    plan = pa.DedispersionPlan(cfg)
    dd = pa.GpuDedisperser(plan)
    dd.allocate()
    gib = dd.gpu_bytes() / 2**30
    log(f'{plan.dtype}: allocated {gib:.2f} GiB of device memory, {dd.host_bytes() / 2**30:.2f} GiB of pinned host memory')
    for it in range(plan.ntrees):
        dd.fill_host_weights(it)          # bare-kernel weights: out_max in SNR units for unit-variance noise
    tuned = dd.apply_tuning_table()      # launch shapes recorded by python -m pirate_amd.autotune
    log(f'{plan.dtype}: launch shapes {tuned if tuned else "default (no tuning-table entry)"}')

    seq = 0
    with dd.get_input(seq) as arr:
        fill_resident(arr, cfg, 2)
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
    log(f'{plan.dtype}: {steps} steps, {dt * 1e3:.3f} ms/step')

    # Per-kernel times: HIP events on the stream each kernel runs on, every kernel alone.
    fork0 = dd.get_tuning('fork')
    dd.set_tuning('fork', 2)
    dd.set_kernel_timing(True)
    for _ in range(timing_steps):
        step()
    torch.cuda.synchronize()
    kt = dd.kernel_timing()
    dd.set_kernel_timing(False)
    dd.set_tuning('fork', fork0)
    return plan, dt, per_kernel_table(plan, kt), gib, dd


def measure_batched(pa, torch, steps, warmup, timing_steps):
    """The 96-beam heterogeneous group, every group's chain on its own branch stream, resident inputs."""
    from gpu_cases import batched_groups
    from pirate_amd.resources import chain_alg_bytes
    cfgs = batched_groups(32)
    plans = [pa.DedispersionPlan(pa.DedispersionConfig.from_dict(d)) for d in cfgs]
    g = pa.GpuDedisperserGroup(plans)
    g.allocate()
    gib = g.gpu_bytes() / 2**30
    nbeams = sum(p.beams_per_batch for p in plans)
    log(f'fp16 batched: {nbeams} beams in {len(plans)} groups, allocated {gib:.2f} GiB of device memory')
    for m in g.members:
        for it in range(m.ntrees):
            m.fill_host_weights(it)
        tuned = m.apply_tuning_table()
        log(f'fp16 batched member rank {m.plan.config.toplevel_tree_rank}: launch shapes '
            f'{tuned if tuned else "default (no tuning-table entry)"}')
    ichunk = 0
    for _ in range(2):
        with g.get_input(ichunk) as slots:      # both input slots hold a resident chunk
            for m, (s, p) in enumerate(zip(slots, plans)):
                fill_resident(s, p.config, 10 + m)
        with g.get_output(ichunk):
            pass
        ichunk += 1

    def step():
        nonlocal ichunk
        s = torch.cuda.current_stream().cuda_stream
        g._acquire_input(ichunk, s)
        g._release_input_and_launch(ichunk, s)
        g._acquire_output(ichunk, s)
        g._release_output(ichunk, s)
        ichunk += 1

    def timed(n):
        for _ in range(warmup):
            step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(n):
            step()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / n

    dt = timed(steps)
    log(f'fp16 batched: {dt * 1e3:.3f} ms/step')

    # Per-kernel: each member alone (the other members' launches paused), its kernels one at a
    # time on one stream, timed with HIP events. The group issues every member per chunk, so
    # the timing runs on standalone dedispersers of the same plans.
    per_member = []
    del g
    torch.cuda.synchronize()
    for d in cfgs:
        _, _, pk, _, dd = measure(pa, torch, d, 2, 1, timing_steps)
        per_member.append(pk)
        del dd
        torch.cuda.synchronize()
    serial_ms = sum(v['avg_ms'] for pk in per_member for v in pk.values())
    alg = sum(chain_alg_bytes(p) for p in plans)
    kern_bytes = sum(v['alg_bytes'] for pk in per_member for v in pk.values())
    return dict(plans=plans, dt=dt, per_member=per_member, gib=gib, nbeams=nbeams,
                chain_alg_bytes=alg, kernel_alg_bytes=kern_bytes, serial_ms=serial_ms)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=20)
    ap.add_argument('--warmup', type=int, default=5)
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--timing-steps', type=int, default=4)
    ap.add_argument('--no-fp16', action='store_true', help='skip the float16-storage chord16k measurement')
    ap.add_argument('--no-batched', action='store_true', help='skip the 96-beam heterogeneous float16 group')
    ap.add_argument('--no-input-delivery', action='store_true', help='skip the int4 upload + dequantization variant')
    ap.add_argument('--calib-copy', action='store_true',
                    help='run one 1 GiB device-to-device copy first (reference dispatch for HBM counters)')
    args = ap.parse_args()

    if args.gpus != 1:
        print(f'bench.py measures one GPU only (--gpus {args.gpus} requested). The chain shards '
              'by beam with no communication; run one process per GPU.', file=sys.stderr)
        sys.exit(2)

    import torch
    import pirate_amd as pa
    from gpu_cases import CHORD16K
    from pirate_amd.resources import chain_alg_bytes

    if args.calib_copy:
        src = torch.ones(2**28, dtype=torch.float32, device='cuda')
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        log('calibration copy: 1 GiB read + 1 GiB write')

    plan, dt, per_kernel, gib, dd = measure(pa, torch, CHORD16K, args.steps, args.warmup, args.timing_steps)
    delivery = None
    if not args.no_input_delivery:
        # the reference's time() loop: int4 upload + dequantization inside every step
        res = dd.time(2 * dd.nstreams + max(args.steps // 2, 4), log=log)
        delivery = {k: round(v, 3) if isinstance(v, float) else v for k, v in res.items()}
    del dd
    torch.cuda.synchronize()
    B, F, T = plan.beams_per_batch, plan.nfreq, plan.nt_in
    cfg = plan.config
    dom = max(per_kernel, key=lambda k: per_kernel[k]['avg_ms'])
    for k, v in sorted(per_kernel.items(), key=lambda kv: -kv[1]['avg_ms']):
        log(f'kernel {k:12s} {v["avg_ms"]:8.3f} ms  {v["GBps"]:8.1f} GB/s (alg bytes {v["alg_bytes"] / 1e6:.1f} MB)')

    copy_gbps = device_copy_gbps(torch)
    log(f'plain 1 GiB device copy: {copy_gbps:.1f} GB/s (read + write)')
    chunk_s = T * cfg.time_sample_ms * 1e-3
    value = B * chunk_s / dt
    chain_gbps = chain_alg_bytes(plan) / dt / 1e9
    gib_per_beam = gib / B
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
                   'chain_alg_bytes_per_step': int(chain_alg_bytes(plan)),
                   'per_kernel_serialized': per_kernel,
                   'per_kernel_sum_ms': round(sum(v['avg_ms'] for v in per_kernel.values()), 4),
                   'gpu_GiB': round(gib, 3), 'gpu_GiB_per_beam': round(gib_per_beam, 4),
                   'beams_fit_in_288GB': int(HBM_GIB // gib_per_beam),
                   'real_time_beams_capacity_capped': round(min(value, HBM_GIB // gib_per_beam), 1)},
        'roofline': {'bound': 'hbm', 'kernel': dom, 'achieved': per_kernel[dom]['GBps'], 'peak': HBM_PEAK_GBPS,
                     'unit': 'GB/s', 'frac': round(per_kernel[dom]['GBps'] / HBM_PEAK_GBPS, 4),
                     'traffic': measured_traffic(dom),
                     'device_copy_GBps': round(copy_gbps, 1),
                     'frac_of_device_copy': round(per_kernel[dom]['GBps'] / copy_gbps, 4)},
    }
    if delivery:
        result['config']['with_input_delivery'] = dict(
            delivery, note='int4 upload (pinned host -> HBM, own stream) + dequantization inside every step, '
                           'as the reference time() loop; PCIe-inclusive, not the headline value')
    if not args.no_batched:
        r = measure_batched(pa, torch, args.steps, args.warmup, args.timing_steps)
        v16 = r['nbeams'] * chunk_s / r['dt']
        alg_gbps = r['chain_alg_bytes'] / r['dt'] / 1e9
        kern_gbps = r['kernel_alg_bytes'] / r['dt'] / 1e9
        result['fp16_batched'] = {
            'workload': '96 beams: 32 x (rank 12, 4096 ch, 3 primary trees) + 32 x (rank 13, 8192 ch, 4) + '
                        '32 x (rank 14, 16384 ch, 5), 300-1500 MHz, 1 ms, chunk 2048, float16 storage, '
                        'each group\'s chain on a branch stream of its own per chunk',
            'value': round(v16, 2), 'unit': 'beams', 'ms_per_step': round(r['dt'] * 1e3, 4),
            'ratio_vs_fp32_plain': round(v16 / value, 3), 'dtype': 'f16 storage, f32 arithmetic',
            'chain_alg_GBps': round(alg_gbps, 1), 'chain_alg_frac_of_peak': round(alg_gbps / HBM_PEAK_GBPS, 4),
            'kernel_alg_GBps': round(kern_gbps, 1), 'kernel_alg_frac_of_peak': round(kern_gbps / HBM_PEAK_GBPS, 4),
            'per_kernel_serialized_sum_ms': round(r['serial_ms'], 4),
            'gpu_GiB': round(r['gib'], 2), 'gpu_GiB_per_beam': round(r['gib'] / r['nbeams'], 4),
            'beams_fit_in_288GB': int(HBM_GIB // (r['gib'] / r['nbeams'])),
            'per_kernel_serialized': {f'r{p.config.toplevel_tree_rank}': pk for p, pk in zip(r['plans'], r['per_member'])},
        }
        log(f'fp16 batched: {v16:.1f} real-time beams ({v16 / value:.2f}x fp32 plain), chain {alg_gbps:.0f} GB/s alg')
    if not args.no_fp16:
        # chord16k shapes with float16 storage (float32 arithmetic)
        p16, dt16, pk16, gib16, dd16 = measure(pa, torch, dict(CHORD16K, dtype='float16'), max(args.steps // 2, 1),
                                               args.warmup, args.timing_steps)
        del dd16
        dom16 = max(pk16, key=lambda k: pk16[k]['avg_ms'])
        result['fp16'] = {'value': round(B * chunk_s / dt16, 2), 'unit': 'beams', 'ms_per_step': round(dt16 * 1e3, 4),
                          'ratio_vs_fp32': round(dt / dt16, 3), 'dtype': 'f16 storage, f32 arithmetic',
                          'workload': 'chord16k shapes with float16 storage',
                          'gpu_GiB': round(gib16, 2), 'dominant_kernel': dom16,
                          'dominant_alg_GBps': pk16[dom16]['GBps'], 'per_kernel_serialized': pk16}
    if not args.no_cpu_baseline:
        result['cpu_baseline'] = cpu_baseline()
    print(json.dumps(result), flush=True)


if __name__ == '__main__':
    main()
