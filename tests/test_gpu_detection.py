"""End-to-end detection on the GPU: a simulated dispersed pulse (pirate_amd.simpulse) injected
into unit-variance noise is found by the chain at the injected DM and arrival time, decoded
from the argmax token with the plan's decode_argmax / decode_argmax2 (reference:
src_lib/DedispersionPlan.cpp:481-615; the decoded timestamp is the pulse centre at the
bottom of the band, where the tree's output time is referenced)."""

import numpy as np
import pytest

from gpu_cases import CASES

pytestmark = pytest.mark.gpu


def test_injected_pulse_is_recovered():
    import torch
    import pirate_amd as pa
    from pirate_amd import simpulse
    from oracle import oracle as orc

    d = CASES['r10_1024ch']                   # 1 beam, 1024 ch 400-800 MHz, rank 10, T = 2048
    cfg = pa.DedispersionConfig.from_dict(d)
    plan = pa.DedispersionPlan(cfg)
    dd = pa.GpuDedisperser(plan)
    dd.allocate()
    dd.fill_all_weights(0, np.broadcast_to(orc.bare_weights(*dd.wt_shape(0)[1:]), dd.wt_shape(0)).copy())

    F, T, nchunks = plan.nfreq, plan.nt_in, 3
    dm, t_inf, snr = 30.0, 2.6, 40.0
    pulse = simpulse.SinglePulse(dm=dm, sm=0.0, intrinsic_width=2e-3, spectral_index=0.0,
                                 undispersed_arrival_time_sec=t_inf, time_sample_ms=1.0, snr=snr,
                                 freq_edges_MHz=np.linspace(400.0, 800.0, F + 1))
    rng = np.random.default_rng(11)
    x = rng.standard_normal((F, nchunks * T)).astype(np.float32)
    pulse.add_to_timestream(x, 0)

    best = None
    for c in range(nchunks):
        with dd.get_input(c) as arr:
            arr.copy_(torch.from_numpy(x[None, :, c * T:(c + 1) * T].copy()).to(arr.device))
        with dd.get_output(c) as out:
            mx = out.out_max[0][0].cpu().numpy()
            am = out.out_argmax[0][0].cpu().numpy().view(np.uint32)
        i = np.unravel_index(np.argmax(mx), mx.shape)
        if best is None or mx[i] > best[0]:
            best = (float(mx[i]), c, int(i[0]), int(i[1]), int(am[i]))

    val, c, idm, itime, tok = best
    assert c == 1 and val > 0.5 * snr, best
    fmin, fmax, tlo, thi, p = plan.decode_argmax(tok, 0, idm, itime)
    flo, fhi, dm_hat, ts, width = plan.decode_argmax2(0, fmin, fmax, tlo, thi, p)
    t_expect = (t_inf + simpulse.dispersion_delay(dm, 400.0)) / 1e-3 - c * T
    print(f'peak {val:.1f} at chunk {c}, dm {dm_hat:.2f} (injected {dm}), t {ts:.1f} (expected {t_expect:.1f}), width {width}')
    assert abs(dm_hat - dm) < 1.0
    assert abs(ts - t_expect) < 4.0
