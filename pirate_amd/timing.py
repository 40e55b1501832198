"""Timing harness of GpuDedisperser (the reference's GpuDedisperser::time,
src_lib/GpuDedisperser.cpp:1573-1690).

Per iteration (seq_id = i, input slot s = i % num_active_batches):
  copy stream: wait until the dequantization of iteration i - S has read slot s's raw
               buffers, then upload the int4 chunk (B, F, T/2 bytes) and its float16
               (scale, offset) pairs (B, F, T/256, 2) from pinned host memory;
  slot stream s: wait for the upload, acquire_input(i), dequantize into the input slot,
               release_input_and_launch(i), acquire_output(0, i), release_output(0, i).
Real-time beams = B * chunk duration / (wall time per iteration), over the iterations after
the first 2 S (warm-up), with the host waiting for completion only at the ends.
"""

import ctypes
import time

import numpy as np

from . import _lib as L


def time_dedisperser(dd, niterations, log=print, seed=0):
    import torch
    S = dd.nstreams
    if niterations <= 2 * S:
        raise ValueError(f'time(): niterations must be > 2 * num_active_batches = {2 * S}')
    B, F, T = dd.beams_per_batch, dd.nfreq, dd.nt_in
    if T % 256:
        raise ValueError('time(): time_samples_per_chunk must be a multiple of 256')
    dev = torch.device('cuda', dd.cuda_device_id)
    rng = np.random.default_rng(seed)
    raw_cpu = torch.from_numpy(rng.integers(0, 256, size=(S, B, F, T // 2), dtype=np.uint8)).pin_memory()
    sc = np.empty((S, B, F, T // 256, 2), dtype=np.float16)
    sc[..., 0] = 0.25
    sc[..., 1] = rng.uniform(-0.1, 0.1, size=sc.shape[:-1])
    scoff_cpu = torch.from_numpy(sc).pin_memory()
    raw_gpu = torch.empty(raw_cpu.shape, dtype=torch.uint8, device=dev)
    scoff_gpu = torch.empty(scoff_cpu.shape, dtype=torch.float16, device=dev)
    h2g = torch.cuda.Stream(device=dev)
    cstreams = [torch.cuda.Stream(device=dev) for _ in range(S)]
    ev_raw = [torch.cuda.Event() for _ in range(S)]
    ev_dq = [torch.cuda.Event() for _ in range(S)]
    raw_nbytes = raw_cpu[0].numel()
    scoff_nbytes = scoff_cpu[0].numel() * 2
    rt = dd.resource_tracker()
    gmem = rt.get_gmem_bw() + B * F * T * (0.5 + (4 if dd.dtype == 'float32' else 2)) + scoff_nbytes
    chunk_s = T * dd.config.time_sample_ms * 1e-3
    lib = L.lib()
    dt_name = dd.dtype.encode()

    seq0 = getattr(dd, '_next_seq', 0)      # the dedisperser may have run chunks before
    nwarm = 2 * S
    t0 = None
    for i in range(niterations):
        seq = seq0 + i
        s = i % S
        cs = cstreams[s]
        if i == nwarm:
            torch.cuda.synchronize(dev)
            t0 = time.perf_counter()
        with torch.cuda.stream(h2g):
            if i >= S:
                h2g.wait_event(ev_dq[s])
            raw_gpu[s].copy_(raw_cpu[s], non_blocking=True)
            scoff_gpu[s].copy_(scoff_cpu[s], non_blocking=True)
            ev_raw[s].record(h2g)
        cs.wait_event(ev_raw[s])
        slot = dd._acquire_input(seq, cs.cuda_stream)
        L.check(lib.pdd_dequantize(dt_name, B, F, T, ctypes.c_void_p(raw_gpu[s].data_ptr()),
                                   ctypes.c_void_p(scoff_gpu[s].data_ptr()), ctypes.c_void_p(slot.data_ptr()),
                                   ctypes.c_void_p(cs.cuda_stream)))
        ev_dq[s].record(cs)
        dd._release_input_and_launch_dd_kernels(seq, cs.cuda_stream)
        dd._acquire_output(0, seq, cs.cuda_stream)
        dd._release_output(0, seq, cs.cuda_stream)
    torch.cuda.synchronize(dev)
    dt = (time.perf_counter() - t0) / (niterations - nwarm)
    res = dict(ms_per_step=dt * 1e3, real_time_beams=B * chunk_s / dt, gmem_GBps=gmem / dt / 1e9,
               h2g_GBps=(raw_nbytes + scoff_nbytes + rt.get_h2g_bw()) / dt / 1e9, g2h_GBps=rt.get_g2h_bw() / dt / 1e9,
               iterations=niterations - nwarm)
    log(f'GpuDedisperser.time(): {res["ms_per_step"]:.3f} ms per batch, real-time beams = {res["real_time_beams"]:.1f}, '
        f'gmem_bw = {res["gmem_GBps"]:.1f}, h2g_bw = {res["h2g_GBps"]:.2f}, g2h_bw = {res["g2h_GBps"]:.2f} GB/s')
    return res
