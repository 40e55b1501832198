"""The random config generator (the reference's DedispersionConfig::make_random()) on the
CPU: every draw is a valid config and plan, its subband-count restriction agrees with the
plan's per-tree subband counts, and pirate_amd's GPU chain accepts every draw (the
dedisperser object is built without touching GPU memory)."""

import numpy as np
import pytest

import pirate_amd as pa
from random_configs import gpu_supported, make_random_config, restrict_subband_counts

NDRAWS = 300


def draws():
    rng = np.random.default_rng(2024)
    return [make_random_config(rng) for _ in range(NDRAWS)]


def test_draws_are_valid_plans_with_matching_subband_counts():
    for d in draws():
        cfg = pa.DedispersionConfig.from_dict(d)
        cfg.validate()
        plan = pa.DedispersionPlan(cfg)
        for t in plan.trees:
            pf_rank = (t.dd_rank + 1) // 2
            want = restrict_subband_counts(d['frequency_subband_counts'], t.early_trigger_level, pf_rank)
            assert t.frequency_subband_counts == want, (d, t)


def test_gpu_chain_accepts_every_draw():
    rejected = []
    for d in draws():
        ok, why = gpu_supported(d)
        if not ok:
            rejected.append((d, why))
    assert not rejected, f'{len(rejected)} of {NDRAWS} rejected, first: {rejected[0]}'


def test_draw_covers_the_interesting_corners():
    ds = draws()
    assert {d['dtype'] for d in ds} == {'float32', 'float16', 'bfloat16'}
    assert any(d['time_samples_per_chunk'] % 64 for d in ds)                # half tiles
    assert any(len(d['primary_trees']) == 4 for d in ds)
    assert any(p['num_early_triggers'] > 0 for d in ds for p in d['primary_trees'])
    assert any(len(d['frequency_subband_counts']) >= 4 for d in ds)
    assert any(len(d['zone_nfreq']) >= 4 for d in ds)
