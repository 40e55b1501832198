"""CPU checks of the scorer's data helpers (pirate_amd.scorer_data): simulated pulse templates
(unit SNR, dedispersed peak where the template says) and the completeness-at-false-alarm-rate
rule used to compare the scorer with the peak finder."""

import numpy as np

from pirate_amd import scorer_data as sd
from pirate_amd.simpulse import dispersion_delay


def test_templates_unit_snr_and_peak_position():
    bank = sd.make_template_bank(3, seed=7, nproc=1, width_ms=(2.0, 4.0), tau600_ms=(0.0, 0.5))
    for t in bank:
        # unit-variance channels: sum of squares over all channels = SNR^2 = 1 (float32 samples)
        assert abs(float(np.sum(t.data.astype(np.float64) ** 2)) - 1.0) < 1e-4
        # bottom channel (400 MHz) peaks at the template's peak sample, within the pulse width
        assert abs(int(np.argmax(t.data[0])) - t.peak_rel) <= 3
        # top channel peaks earlier by the dispersion delay across the band
        sweep = (dispersion_delay(t.dm, 400.0) - dispersion_delay(t.dm, 800.0)) / 1e-3
        assert abs((t.peak_rel - int(np.argmax(t.data[-1]))) - sweep) <= 4


def test_completeness_at_far():
    neg = np.arange(1000, dtype=np.float64)          # 0 .. 999
    pos = np.array([990.5, 998.5, 999.5, 10.0])
    c, thr = sd.completeness_at_far(pos, neg, 1e-2)  # 10 negatives (990..999) may exceed
    assert thr == 989.0 and c == 0.75
    c, thr = sd.completeness_at_far(pos, neg, 1e-3)  # one negative (999) may exceed
    assert thr == 998.0 and c == 0.5
    c, thr = sd.completeness_at_far(pos, neg, 0.0)
    assert thr == 999.0 and c == 0.25
