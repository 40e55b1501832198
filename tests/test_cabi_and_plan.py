"""CPU tests of the C-ABI library and the host model (config, plan), no GPU calls.

  * libpirate_amd.so loads and exports every function declared in include/pirate_amd.h.
  * DedispersionConfig parses the reference's YAML grammar, rejects unknown/old keys, and
    round-trips through to_yaml_string().
  * DedispersionPlan derives the same trees as the reference's own example plan
    (golden: tests/golden/reference_plans.json) and as the independent oracle restatement.
  * decode_argmax / decode_argmax2 / compute_steady_state_it0 closed forms.
"""

import json
import os
import re

import numpy as np
import pytest

import pirate_amd as pa
from pirate_amd import _lib

from gpu_cases import CASES, CHORD16K

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def test_library_exports_every_declared_symbol():
    with open(os.path.join(ROOT, 'include', 'pirate_amd.h')) as f:
        text = f.read()
    declared = set(re.findall(r'^\s*(?:int64_t|int|void|const char)\s*\*?\s*(pdd_\w+)\s*\(', text, re.M))
    assert len(declared) >= 25
    L = _lib.lib()
    for name in sorted(declared):
        assert hasattr(L, name), f'{name} declared in pirate_amd.h but not exported'
    assert declared == set(_lib.EXPORTED_SYMBOLS)
    assert L.pdd_abi_version() == 4


def test_plain_c_caller(tmp_path):
    """The header compiles as C (gcc -std=c99) and a C program linked against the library
    drives the host-side entry points (tests/c/capi_host.c)."""
    import shutil
    import subprocess
    if shutil.which('gcc') is None:
        pytest.skip('gcc not available')
    _lib.lib()                                  # the library is built
    exe = str(tmp_path / 'capi_host')
    libdir = os.path.dirname(_lib.LIBPATH)
    subprocess.run(['gcc', '-std=c99', '-Wall', '-Werror', '-I', os.path.join(ROOT, 'include'),
                    os.path.join(HERE, 'c', 'capi_host.c'), '-o', exe, '-L', libdir, '-lpirate_amd',
                    '-Wl,-rpath,' + libdir], check=True, capture_output=True, text=True)
    env = dict(os.environ)
    torch_lib = os.path.join(os.path.dirname(__import__('torch').__file__), 'lib')
    env['LD_LIBRARY_PATH'] = ':'.join(x for x in (torch_lib, '/opt/rocm/lib', env.get('LD_LIBRARY_PATH', '')) if x)
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=60)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count('ok ') == 9


def golden():
    with open(os.path.join(HERE, 'golden', 'reference_plans.json')) as f:
        return json.load(f)


YAML_TEXT = """
# comment line
zone_nfreq: [ 8192, 8192, 6144, 2048, 3584 ]   # trailing comment
zone_freq_edges: [300, 350, 450, 600, 800, 1500]
time_sample_ms: 1.0
beams_per_gpu: 60
dtype: float16
time_samples_per_chunk: 2048
frequency_subband_counts: [5,9,7,3,1]
max_gpu_clag: 5
beams_per_batch: 2
num_active_batches: 2
toplevel_tree_rank: 16
primary_trees:
  - {num_early_triggers: 0, max_width: 32, dm_downsampling: 0, time_downsampling: 0, wt_dm_downsampling: 64, wt_time_downsampling: 64}
  - {num_early_triggers: 1, max_width: 16, dm_downsampling: 0, time_downsampling: 0,
     wt_dm_downsampling: 64, wt_time_downsampling: 64}
  - num_early_triggers: 2
    max_width: 8
    wt_dm_downsampling: 64
    wt_time_downsampling: 64
  - {num_early_triggers: 3, max_width:  4, wt_dm_downsampling: 64, wt_time_downsampling: 64}
future_write_max_samples: 464000
"""


def test_config_yaml_grammar_and_round_trip():
    c = pa.DedispersionConfig.from_yaml_string(YAML_TEXT)
    g = golden()['configs']['chord_sb2_et']
    assert c.zone_nfreq == g['zone_nfreq']
    assert c.zone_freq_edges == [float(x) for x in g['zone_freq_edges']]
    assert c.dtype == 'float16' and c.toplevel_tree_rank == 16 and c.max_gpu_clag == 5
    assert [pt.max_width for pt in c.primary_trees] == [32, 16, 8, 4]
    c2 = pa.DedispersionConfig.from_yaml_string(c.to_yaml_string())
    assert c2.to_dict() == c.to_dict()
    c.test()


@pytest.mark.parametrize('bad,msg', [
    ('tree_rank: 3\n', 'old config syntax'),
    ('early_triggers: []\n', 'old config syntax'),
    ('bogus_key: 1\n', "invalid key 'bogus_key'"),
])
def test_config_rejects_bad_keys(bad, msg):
    with pytest.raises(_lib.PddError, match=msg):
        pa.DedispersionConfig.from_yaml_string(YAML_TEXT + bad)


def test_config_validation_errors():
    d = dict(CASES['toy'])
    d['time_samples_per_chunk'] = 100
    with pytest.raises(_lib.PddError, match='must be a multiple of'):
        pa.DedispersionConfig.from_dict(d)
    d = dict(CASES['toy'])
    d['frequency_subband_counts'] = [2, 2, 2]
    with pytest.raises(_lib.PddError, match='full band'):
        pa.DedispersionConfig.from_dict(d)


def test_plan_matches_reference_example_plan():
    G = golden()
    cfg = pa.DedispersionConfig.from_dict(G['configs']['chord_sb2_et'])
    plan = pa.DedispersionPlan(cfg)
    gp = G['plans']['chord_sb2_et']
    assert plan.ntrees == gp['ntrees']
    assert plan.stage1_dd_rank == gp['stage1_dd_rank'] and plan.stage1_amb_rank == gp['stage1_amb_rank']
    assert plan.nfreq == gp['nfreq'] and plan.nt_in == gp['nt_in']
    for t, g in zip(plan.trees, gp['trees']):
        for k in ['ndm_out', 'nt_out', 'amb_rank', 'dd_rank', 'nt_ds', 'max_width', 'dm_downsampling',
                  'time_downsampling', 'Dcore', 'nprofiles', 'ndm_wt', 'nt_wt', 'primary_tree_index',
                  'early_trigger_level', 'wt_dm_downsampling', 'wt_time_downsampling']:
            assert getattr(t, k) == g[k], (k, getattr(t, k), g[k])
        assert t.frequency_subband_counts == g['frequency_subband_counts']
        assert t.dm_min == pytest.approx(g['dm_min'], rel=1e-12, abs=1e-9)
        assert t.dm_max == pytest.approx(g['dm_max'], rel=1e-12)
        assert t.trigger_frequency == pytest.approx(g['trigger_frequency'], rel=1e-12)


@pytest.mark.parametrize('name', sorted(CASES) + ['chord16k'])
def test_plan_trees_match_oracle(orc, name):
    d = CHORD16K if name == 'chord16k' else CASES[name]
    plan = pa.DedispersionPlan(pa.DedispersionConfig.from_dict(d))
    ot = orc.OracleConfig.from_yaml_dict(d).trees()
    assert len(ot) == plan.ntrees
    for t, o in zip(plan.trees, ot):
        assert (t.primary_tree_index, t.early_trigger_level, t.amb_rank, t.dd_rank, t.nt_ds) == \
               (o['ipri'], o['et'], o['amb_rank'], o['dd_rank'], o['nt_ds'])
        assert (t.ndm_out, t.nt_out, t.ndm_wt, t.nt_wt, t.nprofiles, t.Dcore, t.nmultiplets_M, t.nsubbands_N) == \
               (o['ndm_out'], o['nt_out'], o['ndm_wt'], o['nt_wt'], o['nprofiles'], o['Dcore'], o['M'], o['N'])
        assert t.frequency_subband_counts == o['subband_counts']


def test_channel_map_matches_oracle(orc):
    d = CASES['zones4']
    cm = pa.DedispersionConfig.from_dict(d).make_channel_map()
    om = orc.OracleConfig.from_yaml_dict(d).channel_map()
    np.testing.assert_array_equal(cm, om)


def test_decode_argmax_closed_forms():
    plan = pa.DedispersionPlan(pa.DedispersionConfig.from_dict(CASES['toy']))
    for itree, t in enumerate(plan.trees):
        Dout = t.nt_ds // t.nt_out
        R = plan.config.toplevel_tree_rank
        rr = R - t.early_trigger_level
        M = t.nmultiplets_M
        # Full-band multiplet (last) with p=0 at the first fine sample: the subband is the
        # whole tree band, time range collapses to one sample when p=0 and the tree is not
        # downsampled (reference DedispersionPlan.hpp: "length 1 iff p == 0 and ipri == 0").
        m = M - 1
        for idm in [0, t.ndm_out - 1]:
            fmin, fmax, tlo, thi, p = plan.decode_argmax((m << 16) | 0, itree, idm, 1)
            assert p == 0 and fmin == 0 and fmax == 2**rr - 1
            assert tlo <= thi <= plan.nt_in
            if t.primary_tree_index == 0:
                # thi = trailing edge at the band top: sample (Dout + 0) + 1 (exclusive)
                assert thi == 1 * Dout + 1
        with pytest.raises(_lib.PddError, match='bad token'):
            plan.decode_argmax((M << 16), itree, 0, 0)
        with pytest.raises(_lib.PddError, match='bad token'):
            plan.decode_argmax((t.nprofiles << 8), itree, 0, 0)
        fl, fh, dm, ts, w = plan.decode_argmax2(itree, 0, 2**rr - 1, 0, 2**rr - 1, 0)
        assert fh == pytest.approx(plan.config.zone_freq_edges[-1])
        assert dm == pytest.approx(2**R * plan.config.dm_per_unit_delay(), rel=1e-12)


def test_decode_argmax_batch_equals_scalar():
    plan = pa.DedispersionPlan(pa.DedispersionConfig.from_dict(CASES['toy']))
    rng = np.random.default_rng(0)
    for itree, t in enumerate(plan.trees):
        Dout = t.nt_ds // t.nt_out
        shape = (2, t.ndm_out, t.nt_out)
        m = rng.integers(0, t.nmultiplets_M, size=shape)
        p = rng.integers(0, t.nprofiles, size=shape)
        lev = np.where(p == 0, 0, (p - 1) // 3)
        step = np.minimum(t.Dcore, 2 ** lev)                  # token times are multiples of this
        tt = rng.integers(0, Dout, size=shape) // step * step
        tok = (tt | (p << 8) | (m << 16)).astype(np.uint32)
        got = plan.decode_argmax_batch(itree, tok)
        assert got.shape == shape + (5,)
        for idx in [(0, 0, 0), (1, t.ndm_out - 1, t.nt_out - 1), (1, t.ndm_out // 2, t.nt_out // 3)]:
            assert tuple(got[idx]) == plan.decode_argmax(int(tok[idx]), itree, idx[1], idx[2])
        flat = got.reshape(-1, t.ndm_out, t.nt_out, 5)
        for b in range(2):
            for d in range(0, t.ndm_out, max(1, t.ndm_out // 5)):
                for i in range(0, t.nt_out, max(1, t.nt_out // 5)):
                    assert tuple(flat[b, d, i]) == plan.decode_argmax(int(tok[b, d, i]), itree, d, i)
    with pytest.raises(ValueError, match='last two axes'):
        plan.decode_argmax_batch(0, np.zeros((3, 4), np.uint32))


def test_steady_state_it0_closed_form():
    plan = pa.DedispersionPlan(pa.DedispersionConfig.from_dict(CASES['zones4']))
    R = plan.config.toplevel_tree_rank
    for itree, t in enumerate(plan.trees):
        p, e = t.primary_tree_index, t.early_trigger_level
        d0 = (2**(R + p - 1) if p > 0 else 0) // 2**(e + p)
        it0 = plan.compute_steady_state_it0(itree)
        for idm in range(t.ndm_out):
            dmax = d0 + (idm + 1) * t.dm_downsampling - 1
            assert it0[idm] == -(-(dmax + 4 * t.max_width) // t.time_downsampling)


def test_reference_configs_accepted_and_limits_reported_cleanly():
    # Every config of the reference's configs/dedispersion/ (ranks up to 16, subband
    # searches, early triggers) builds a GPU dedisperser (no device memory is touched before
    # allocate()); a shape beyond pirate_amd's kernels (peak-finder time downsampling 64)
    # fails at creation with a clear message, not at launch time.
    for name, d in golden()['configs'].items():
        pa.GpuDedisperser(pa.DedispersionPlan(pa.DedispersionConfig.from_dict(d)))
    d = dict(golden()['configs']['chord_sb2_et'])
    d['primary_trees'] = [dict(t, time_downsampling=64, wt_time_downsampling=64) for t in d['primary_trees']]
    plan = pa.DedispersionPlan(pa.DedispersionConfig.from_dict(d))
    with pytest.raises(_lib.PddError, match='not supported'):
        pa.GpuDedisperser(plan)


def test_bare_profile_variances_match_oracle(orc):
    from pirate_amd.dedisperser import bare_profile_variances
    for P in (1, 4, 7, 10, 13, 16):
        assert np.array_equal(bare_profile_variances(P), orc.bare_profile_variances(P))


def test_nbatches_defaults_and_lower_bounds():
    # Reference defaults: nbatches_out and nbatches_wt both default to num_active_batches and
    # may not be below it (src_lib/GpuDedisperser.cpp:116-123); construction touches no device.
    d = dict(golden()['configs']['toy'])
    d.update(beams_per_gpu=4, beams_per_batch=2, num_active_batches=1)
    plan = pa.DedispersionPlan(pa.DedispersionConfig.from_dict(d))
    dd = pa.GpuDedisperser(plan, nbatches_out=2)
    assert dd.nbatches_out == 2 and dd.nbatches_wt == 1
    assert dd.wt_shape(0)[0] == 1
    d.update(num_active_batches=2)
    plan2 = pa.DedispersionPlan(pa.DedispersionConfig.from_dict(d))
    with pytest.raises(ValueError, match='num_active_batches'):
        pa.GpuDedisperser(plan2, nbatches_out=1)
    with pytest.raises(ValueError, match='num_active_batches'):
        pa.GpuDedisperser(plan2, nbatches_wt=1)


def test_device_view_keeps_its_owner_alive():
    # A tensor viewing dedisperser memory holds the _DeviceView (torch keeps the object that
    # provides __cuda_array_interface__), and the view holds its owner, so the owner (whose
    # destructor frees the memory) outlives every view.
    import gc
    import weakref
    from pirate_amd.dedisperser import _DeviceView

    class Owner:
        pass

    o = Owner()
    ref = weakref.ref(o)
    v = _DeviceView(0x1000, (2, 3), '<f4', owner=o)
    del o
    gc.collect()
    assert ref() is not None
    assert v.__cuda_array_interface__['shape'] == (2, 3)
    del v
    gc.collect()
    assert ref() is None


def test_group_refuses_graph_replay():
    """HIP graph replay of a group's chunk is refused (its capture path is switched off): the
    group is created with use_graph=False by default, and use_graph=True fails at creation."""
    import pirate_amd as pa
    plan = pa.DedispersionPlan(pa.DedispersionConfig.from_dict(CASES['one_tree']))
    with pytest.raises(RuntimeError, match='graph replay is disabled'):
        pa.GpuDedisperserGroup([plan], use_graph=True)
