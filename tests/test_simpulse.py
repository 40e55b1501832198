"""Pulse simulator (reference: src_lib/simpulse.cpp, pirate_frb/simpulse/test_pulse_upsampling.py).

Pinned by: the SNR normalization (sum sample^2/variance == snr^2), per-channel fluence
(the samples of a channel telescope to freq_wt/dt times the global scale, so the spectral
index is checked channel by channel), the dispersion delay of each channel's centroid
(k_dm DM nu^-2 plus the mean scattering delay), the invariance under time/frequency
upsampling that the reference's own test checks, and add_to_timestream's bounds rule."""

import numpy as np
import pytest

from pirate_amd import simpulse as sp


def make(**kw):
    p = dict(dm=100.0, sm=0.5, intrinsic_width=2e-3, spectral_index=-1.5, undispersed_arrival_time_sec=0.05,
             time_sample_ms=1.0, snr=20.0, freq_edges_MHz=np.linspace(400, 800, 65))
    p.update(kw)
    return sp.SinglePulse(**p)


def dense(p):
    out = np.zeros((p.nfreq, p.it_end - p.it_start), dtype=np.float32)
    p.add_to_timestream(out, p.it_start)
    return out


def test_snr_normalization():
    var = np.linspace(0.5, 2.0, 64)
    p = make(freq_variances=var)
    d = dense(p).astype(np.float64)
    assert np.isclose(np.sum(d**2 / var[:, None]), 20.0**2, rtol=1e-5)


def test_channel_fluence_follows_spectral_index():
    p = make()
    d = dense(p).astype(np.float64)
    edges = p.freq_edges_MHz
    nu_c = 0.5 * (edges[:-1] + edges[1:])
    fw = (nu_c / (0.5 * (edges[0] + edges[-1]))) ** p.spectral_index
    flu = d.sum(axis=1)
    assert np.allclose(flu / flu.max(), fw / fw.max(), rtol=1e-5)


def test_channel_centroid_delay():
    p = make(sm=0.0, intrinsic_width=1e-3, time_sample_ms=0.25, dm=300.0)
    d = dense(p).astype(np.float64)
    dt = 1e-3 * p.time_sample_ms
    t = (np.arange(p.it_start, p.it_end) + 0.5) * dt
    edges = p.freq_edges_MHz
    for f in range(p.nfreq):
        centroid = np.sum(d[f] * t) / np.sum(d[f])
        want = p.undispersed_arrival_time_sec + 0.5 * (sp.dispersion_delay(p.dm, edges[f]) + sp.dispersion_delay(p.dm, edges[f + 1]))
        assert abs(centroid - want) < 0.05 * dt + 1e-6, (f, centroid, want)


def test_upsampling_invariance():
    # Same pulse at time_sample_ms and time_sample_ms/4, and on 64 vs 256 channels: the fine
    # pulse summed down to the coarse grid must match (reference test_pulse_upsampling).
    e = np.linspace(500, 900, 257)
    kw = dict(dm=80.0, sm=0.2, intrinsic_width=3e-3, spectral_index=0.7, undispersed_arrival_time_sec=0.02, snr=30.0)
    coarse = sp.SinglePulse(time_sample_ms=1.0, freq_edges_MHz=e[::4], **kw)
    fine = sp.SinglePulse(time_sample_ms=0.25, freq_edges_MHz=e, **kw)
    nt = 4 * (max(coarse.it_end, fine.it_end // 4 + 1) + 2)
    c = np.zeros((64, nt // 4), np.float32)
    coarse.add_to_timestream(c, 0)
    fz = np.zeros((256, nt), np.float32)
    fine.add_to_timestream(fz, 0)
    # samples are intensities per unit time: summing 4 time samples and 4 channels of the fine
    # pulse gives 16x the coarse sample up to the SNR normalization, so compare shapes
    fd = fz.reshape(64, 4, nt // 4, 4).sum(axis=(1, 3))
    r = np.corrcoef(c.ravel(), fd.ravel())[0, 1]
    assert r > 0.999, r


def test_add_to_timestream_requires_full_span():
    p = make()
    out = np.zeros((p.nfreq, p.it_end - p.it_start - 1), np.float32)
    with pytest.raises(ValueError, match='does not cover'):
        p.add_to_timestream(out, p.it_start)


def test_dispersion_delay_constant():
    # k_dm = 4.148808e6 ms MHz^2 (include/pirate/constants.hpp:49)
    assert sp.dispersion_delay(1.0, 1000.0) == pytest.approx(4.148808e-3, rel=1e-12)
    assert sp.scattering_time(2.0, 500.0) == pytest.approx(2e-3 * 16, rel=1e-12)
