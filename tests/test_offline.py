"""OfflineDedisperser (reference pirate_frb/OfflineDedisperser.py:13-257): int4 frames ->
dequantization into the input slot -> GpuDedisperser with analytic weights.

CPU: per-zone channel variances; the failure latch (a bad frame, then every later call
raises). GPU: the frame metadata overrides the config's zones and sampling; outputs equal,
bit for bit, a GpuDedisperser fed with the oracle's dequantization of the same frames and
the same analytic weights; an injected pulse comes out as an SNR close to the injected one;
a gap in time_chunk_index raises and latches.
"""

import numpy as np
import pytest

from gpu_cases import CASES


def _pack(vals):
    u = (np.asarray(vals, dtype=np.int64) & 0xf).astype(np.uint8)
    return (u[..., 0::2] | (u[..., 1::2] << 4)).astype(np.uint8)


def make_frames(x, zone_nfreq, zone_edges, nt, beam_id=3, tci0=100, sigma=1.0):
    """Quantizes float data x (nfreq, nchunks*nt) to int4 with one (scale, offset) per
    256-sample block: v = clip(round(x / s), -7, 7), s = sigma / 2.5, offset 0."""
    import pirate_amd as pa
    F = x.shape[0]
    md = pa.XEngineMetadata(zone_nfreq=list(zone_nfreq), zone_freq_edges=list(zone_edges), dt_ns_per_seq=2500,
                            seq_per_frb_time_sample=400, noise_variance=[sigma ** 2] * len(zone_nfreq),
                            unix_ns_at_seq_0=12345, beamset=1, beam_ids=[beam_id])
    s = np.float16(sigma / 2.5)
    frames = []
    for c in range(x.shape[1] // nt):
        blk = x[:, c * nt:(c + 1) * nt]
        v = np.clip(np.rint(blk / np.float32(s)), -7, 7).astype(np.int64)
        so = np.zeros((F, nt // 256, 2), dtype=np.float16)
        so[..., 0] = s
        frames.append(pa.AssembledFrame(beam_id=beam_id, time_chunk_index=tci0 + c, data=_pack(v),
                                        scales_offsets=so, metadata=md))
    return md, frames


def test_channel_variances():
    import pirate_amd as pa
    md = pa.XEngineMetadata(zone_nfreq=[3, 2], zone_freq_edges=[400, 500, 800], dt_ns_per_seq=1,
                            seq_per_frb_time_sample=1, noise_variance=[2.0, 0.5])
    assert md.get_channel_variances().tolist() == [2.0, 2.0, 2.0, 0.5, 0.5]


def test_failure_latches():
    import pirate_amd as pa
    od = pa.OfflineDedisperser(pa.DedispersionConfig.from_dict(CASES['one_tree']))
    md, frames = make_frames(np.zeros((1024, 256), np.float32), [1024], [400, 800], 256)
    bad = pa.AssembledFrame(beam_id=7, time_chunk_index=0, data=frames[0].data,
                            scales_offsets=frames[0].scales_offsets, metadata=md)   # beam_ids = [3]
    with pytest.raises(AssertionError, match='beam_ids'):
        with od.dedisperse(bad):
            pass
    with pytest.raises(RuntimeError, match='previous dedisperse'):
        with od.dedisperse(frames[0]):
            pass


@pytest.mark.gpu
def test_offline_equals_direct_and_detects():
    import torch
    import pirate_amd as pa
    from pirate_amd import simpulse
    from oracle import oracle as orc

    base = dict(CASES['one_tree'])                  # 1024 ch 400-800 MHz, rank 10, T = 256
    zones, edges = [512, 512], [400.0, 560.0, 800.0]     # the frames' zones override the config's
    T, nchunks = base['time_samples_per_chunk'], 24
    rng = np.random.default_rng(5)
    x = rng.standard_normal((1024, nchunks * T)).astype(np.float32)
    fedges = np.concatenate([np.linspace(400, 560, 513), np.linspace(560, 800, 513)[1:]])
    pulse = simpulse.SinglePulse(dm=40.0, sm=0.0, intrinsic_width=2e-3, spectral_index=0.0,
                                 undispersed_arrival_time_sec=1.9, time_sample_ms=1.0, snr=40.0,
                                 freq_edges_MHz=fedges)
    pulse.add_to_timestream(x, 0)
    md, frames = make_frames(x, zones, edges, T)

    od = pa.OfflineDedisperser(pa.DedispersionConfig.from_dict(base))
    ref_cfg = pa.DedispersionConfig.from_dict(dict(base, zone_nfreq=zones, zone_freq_edges=edges, time_sample_ms=1.0))
    ref = pa.GpuDedisperser(pa.DedispersionPlan(ref_cfg))
    ref.allocate()
    ref.fill_analytic_weights(md.get_channel_variances())

    best = 0.0
    for c, fr in enumerate(frames):
        with od.dedisperse(fr) as out:
            got = [(m.clone(), a.clone()) for m, a in zip(out.out_max, out.out_argmax)]
        xin = orc.dequantize(fr.data[None], fr.scales_offsets[None], 1, 1024, T)
        with ref.get_input(c) as arr:
            arr.copy_(torch.from_numpy(xin).to(arr.device))
        with ref.get_output(c) as out:
            for it in range(ref.ntrees):
                assert torch.equal(got[it][0], out.out_max[it]) and torch.equal(got[it][1], out.out_argmax[it])
        best = max(best, max(float(m.max()) for m, _ in got))
    assert od.config.zone_nfreq == zones and od.config.zone_freq_edges == edges
    assert abs(od.time_sample_ms - 1.0) < 1e-12
    print(f'offline: max SNR {best:.1f} (injected 40)')
    assert 25.0 < best < 46.0


@pytest.mark.gpu
def test_offline_gap_raises_and_latches():
    import pirate_amd as pa
    od = pa.OfflineDedisperser(pa.DedispersionConfig.from_dict(CASES['one_tree']))
    md, frames = make_frames(np.zeros((1024, 3 * 256), np.float32), [1024], [400, 800], 256)
    with od.dedisperse(frames[0]):
        pass
    with pytest.raises(AssertionError, match='non-consecutive'):
        with od.dedisperse(frames[2]):
            pass
    with pytest.raises(RuntimeError, match='previous dedisperse'):
        with od.dedisperse(frames[1]):
            pass
