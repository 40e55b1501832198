"""Simulated FRB pulses in the (frequency, time) plane, in float64.

Restates the reference's pulse model (src_lib/simpulse.cpp:27-408, include/pirate/simpulse.hpp):
a Gaussian of width ``intrinsic_width``, broadened within each channel by the dispersion sweep
across the channel (a boxcar, sinc in Fourier space) and by a one-sided exponential scattering
tail (time ``scattering_time(sm, nu)``), synthesized per channel with an inverse real FFT,
integrated over each time sample through a normalized cumulative sum, weighted by the spectral
index, and normalized so that sum(sample^2 / variance) == snr^2 over all channels.

Used for synthetic test data (injected pulses, detection tests). numpy's unnormalized inverse
FFT (norm='forward') has the same sign convention as the reference's c2r transform; the
normalization is irrelevant because each channel's cumulative sum is normalized to 1.
"""

import numpy as np

K_DM = 4.148808e6              # ms MHz^2 / (pc cm^-3)   (include/pirate/constants.hpp:49)
FRB_SCATTERING_INDEX = 4.0     # include/pirate/constants.hpp:50


def dispersion_delay(dm, freq_MHz):
    """Dispersion delay in seconds (src_lib/simpulse.cpp:32-38)."""
    if dm < 0 or freq_MHz <= 0:
        raise ValueError('dispersion_delay: need dm >= 0 and freq > 0')
    return 1.0e-3 * K_DM * dm / (freq_MHz * freq_MHz)


def scattering_time(sm, freq_MHz):
    """Scattering time in seconds; sm is the scattering time in ms at 1 GHz (simpulse.cpp:41-47)."""
    if sm < 0 or freq_MHz <= 0:
        raise ValueError('scattering_time: need sm >= 0 and freq > 0')
    return 1.0e-3 * sm / (freq_MHz / 1000.0) ** FRB_SCATTERING_INDEX


def _interp_cumsum(arr, s):
    """Linear interpolation of the cumulative sum (length n+1) at sample coordinate s
    (simpulse.cpp:81-94), vectorized."""
    n = len(arr) - 1
    s = np.asarray(s, dtype=np.float64)
    i = np.clip(np.floor(s).astype(np.int64), 0, n - 1)
    ds = s - i
    out = (1.0 - ds) * arr[i] + ds * arr[i + 1]
    out = np.where(s < 1.0e-10, 0.0, out)
    return np.where(s > n - 1.0e-10, arr[n], out)


class SinglePulse:
    """One dispersed, scattered pulse on a fixed channelization and zero-based time grid
    (sample 'it' spans [it*dt, (it+1)*dt] seconds). Frequencies are ordered low to high.

    Attributes mirror the reference: freq_it0, freq_nt, freq_sd_off (per channel), sparse_data
    (packed float32 samples), it_start, it_end."""

    def __init__(self, dm, sm, intrinsic_width, spectral_index, undispersed_arrival_time_sec,
                 time_sample_ms, snr, freq_edges_MHz, freq_variances=None,
                 subband_freq_lo_MHz=0.0, subband_freq_hi_MHz=1.0e9, internal_nt=1024):
        edges = np.asarray(freq_edges_MHz, dtype=np.float64)
        nfreq = len(edges) - 1
        var = np.ones(nfreq) if freq_variances is None else np.asarray(freq_variances, dtype=np.float64)
        if nfreq < 1 or np.any(edges <= 0) or np.any(np.diff(edges) <= 0):
            raise ValueError('SinglePulse: freq_edges_MHz must be positive and strictly increasing')
        if var.shape != (nfreq,) or np.any(var <= 0):
            raise ValueError('SinglePulse: freq_variances must be positive, one per channel')
        if internal_nt < 64 or (internal_nt & (internal_nt - 1)):
            raise ValueError('SinglePulse: internal_nt must be a power of two >= 64')
        if dm < 0 or sm < 0 or intrinsic_width < 0 or snr < 0 or time_sample_ms <= 0:
            raise ValueError('SinglePulse: dm, sm, intrinsic_width, snr must be >= 0 and time_sample_ms > 0')
        if not subband_freq_lo_MHz < subband_freq_hi_MHz:
            raise ValueError('SinglePulse: empty subband')
        if subband_freq_hi_MHz <= edges[0] or subband_freq_lo_MHz >= edges[-1]:
            raise ValueError('SinglePulse: subband does not overlap the frequency band')
        if dm == 0 and sm == 0 and intrinsic_width == 0:
            raise ValueError('SinglePulse: delta-function pulse (dm=sm=width=0) is not allowed')

        self.dm, self.sm, self.intrinsic_width = dm, sm, intrinsic_width
        self.spectral_index, self.snr = spectral_index, snr
        self.undispersed_arrival_time_sec = undispersed_arrival_time_sec
        self.time_sample_ms, self.internal_nt = time_sample_ms, internal_nt
        self.freq_edges_MHz, self.freq_variances = edges, var
        self.nfreq = nfreq

        dt = 1.0e-3 * time_sample_ms
        nu0 = 0.5 * (edges[0] + edges[-1])
        nfft = 2 * internal_nt
        j = np.arange(nfft // 2 + 1)

        it0 = np.zeros(nfreq, dtype=np.int64)
        nt = np.zeros(nfreq, dtype=np.int64)
        sd_off = np.zeros(nfreq, dtype=np.int64)
        chunks = []
        snr_sq = 0.0
        total = 0
        it_start, it_end = None, None
        for f in range(nfreq):
            nu_lo, nu_hi = edges[f], edges[f + 1]
            sd_off[f] = total
            if nu_hi <= subband_freq_lo_MHz or nu_lo >= subband_freq_hi_MHz:
                continue
            nu_c = 0.5 * (nu_lo + nu_hi)
            d0 = dispersion_delay(dm, nu_hi)
            d1 = dispersion_delay(dm, nu_lo)
            dmw = d1 - d0
            ts = scattering_time(sm, nu_c)
            t0 = d0 - 0.1 * dmw - 4.0 * intrinsic_width - ts
            t1 = d1 + 0.1 * dmw + 4.0 * intrinsic_width + 10.0 * ts
            tc = 0.5 * (d0 + d1)
            dtc = tc - (t0 + (t1 - t0) / (2.0 * internal_nt))
            fw = (nu_c / nu0) ** spectral_index

            omega = (2.0 * np.pi * internal_nt / nfft / (t1 - t0)) * j
            x = 0.5 * dmw * omega
            sinc = np.where(x * x > 1.0e-100, np.sin(x) / np.where(x == 0, 1.0, x), 1.0)
            z = np.exp(-(intrinsic_width * omega) ** 2 / 2.0) * sinc
            z = z / (1.0 + 1j * ts * omega)
            z = z * (np.cos(dtc * omega) - 1j * np.sin(dtc * omega))
            prof = np.fft.irfft(z, n=nfft, norm='forward')[:internal_nt]
            cs = np.concatenate([[0.0], np.cumsum(np.maximum(prof, 0.0))])
            if cs[-1] <= 0:
                raise RuntimeError(f'SinglePulse: channel {f} produced a degenerate pulse profile')
            cs = cs / cs[-1]

            pt0 = undispersed_arrival_time_sec + t0
            pt1 = undispersed_arrival_time_sec + t1
            i0 = int(np.floor(pt0 / dt))
            i1 = int(np.ceil(pt1 / dt))
            its = np.arange(i0, i1, dtype=np.float64)
            a = _interp_cumsum(cs, internal_nt * (its * dt - pt0) / (t1 - t0))
            b = _interp_cumsum(cs, internal_nt * ((its + 1) * dt - pt0) / (t1 - t0))
            sk = (fw / dt) * (b - a)
            chunks.append(sk.astype(np.float32).astype(np.float64))
            snr_sq += float(np.sum(sk * sk / var[f]))
            it0[f], nt[f] = i0, i1 - i0
            total += i1 - i0
            it_start = i0 if it_start is None else min(it_start, i0)
            it_end = i1 if it_end is None else max(it_end, i1)

        it0[nt == 0] = it_start
        scale = snr / np.sqrt(snr_sq)
        self.freq_it0, self.freq_nt, self.freq_sd_off = it0, nt, sd_off
        self.sparse_data = (np.concatenate(chunks) * scale).astype(np.float32)
        self.it_start, self.it_end = it_start, it_end

    def add_to_timestream(self, out, out_it0, weight=1.0):
        """Adds weight * pulse to out (nfreq, out_nt) float32, whose column i is grid sample
        out_it0 + i; out must cover [it_start, it_end) (simpulse.cpp:331-369)."""
        if out.ndim != 2 or out.shape[0] != self.nfreq:
            raise ValueError(f'add_to_timestream: out must have shape ({self.nfreq}, nt)')
        if out_it0 > self.it_start or out_it0 + out.shape[1] < self.it_end:
            raise ValueError(f'add_to_timestream: output spans [{out_it0}, {out_it0 + out.shape[1]}), which does '
                             f'not cover the pulse range [{self.it_start}, {self.it_end})')
        wf = np.float32(weight)
        for f in range(self.nfreq):
            n = self.freq_nt[f]
            if n:
                base = self.freq_it0[f] - out_it0
                src = self.sparse_data[self.freq_sd_off[f]:self.freq_sd_off[f] + n]
                out[f, base:base + n] += wf * src

    def shift_samples(self, delta_it):
        """Shifts the pulse by delta_it samples (simpulse.cpp:372-382)."""
        self.freq_it0 = self.freq_it0 + delta_it
        self.it_start += delta_it
        self.it_end += delta_it
        self.undispersed_arrival_time_sec += 1.0e-3 * delta_it * self.time_sample_ms
