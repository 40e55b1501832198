"""DedispersionConfig and DedispersionPlan (Python surface of the reference's pybind11
classes, src_pybind11/pirate_pybind11.cpp:128-415), backed by libpirate_amd.so."""

import ctypes

import numpy as np
import yaml

from . import _lib as L


class PrimaryTree:
    """One primary tree (DM range) of a config. Reference: DedispersionConfig::PrimaryTree
    (include/pirate/DedispersionConfig.hpp:62-70)."""

    _keys = ('num_early_triggers', 'max_width', 'dm_downsampling', 'time_downsampling',
             'wt_dm_downsampling', 'wt_time_downsampling')

    def __init__(self, num_early_triggers=0, max_width=0, dm_downsampling=0, time_downsampling=0,
                 wt_dm_downsampling=0, wt_time_downsampling=0):
        self.num_early_triggers = int(num_early_triggers)
        self.max_width = int(max_width)
        self.dm_downsampling = int(dm_downsampling)
        self.time_downsampling = int(time_downsampling)
        self.wt_dm_downsampling = int(wt_dm_downsampling)
        self.wt_time_downsampling = int(wt_time_downsampling)

    def to_dict(self):
        return {k: getattr(self, k) for k in self._keys}

    def __repr__(self):
        return 'PrimaryTree(' + ', '.join(f'{k}={getattr(self, k)}' for k in self._keys) + ')'


class DedispersionConfig:
    """Dedispersion configuration (YAML-backed). Same fields, defaults and validation as the
    reference (include/pirate/DedispersionConfig.hpp:21-95); validation runs in C++."""

    _scalar_keys = ('time_sample_ms', 'toplevel_tree_rank', 'time_samples_per_chunk', 'dtype',
                    'beams_per_gpu', 'beams_per_batch', 'num_active_batches', 'future_write_max_samples',
                    'max_gpu_clag')

    def __init__(self):
        self.zone_nfreq = []
        self.zone_freq_edges = []
        self.time_sample_ms = 0.0
        self.toplevel_tree_rank = -1
        self.time_samples_per_chunk = 0
        self.dtype = 'float32'
        self.frequency_subband_counts = [1]
        self.primary_trees = []
        self.beams_per_gpu = 0
        self.beams_per_batch = 0
        self.num_active_batches = 0
        self.future_write_max_samples = 0
        self.max_gpu_clag = 10000

    # ---- construction ----

    @staticmethod
    def from_yaml(filename):
        h = ctypes.c_void_p()
        L.check(L.lib().pdd_config_from_yaml_file(filename.encode(), ctypes.byref(h)))
        try:
            return DedispersionConfig._from_handle(h)
        finally:
            L.lib().pdd_config_destroy(h)

    @staticmethod
    def from_yaml_string(text):
        h = ctypes.c_void_p()
        L.check(L.lib().pdd_config_from_yaml_string(text.encode(), ctypes.byref(h)))
        try:
            return DedispersionConfig._from_handle(h)
        finally:
            L.lib().pdd_config_destroy(h)

    @staticmethod
    def from_dict(d):
        return DedispersionConfig.from_yaml_string(yaml.safe_dump(d, sort_keys=False, width=1 << 30))

    @staticmethod
    def _from_handle(h):
        d = yaml.safe_load(L.get_string(L.lib().pdd_config_to_yaml, h, 0))
        c = DedispersionConfig()
        c.zone_nfreq = [int(x) for x in d['zone_nfreq']]
        c.zone_freq_edges = [float(x) for x in d['zone_freq_edges']]
        for k in DedispersionConfig._scalar_keys:
            if k in d:
                setattr(c, k, d[k])
        c.frequency_subband_counts = [int(x) for x in d['frequency_subband_counts']]
        c.primary_trees = [PrimaryTree(**pt) for pt in d['primary_trees']]
        return c

    def to_dict(self):
        d = {'zone_nfreq': [int(x) for x in self.zone_nfreq],
             'zone_freq_edges': [float(x) for x in self.zone_freq_edges]}
        for k in self._scalar_keys:
            d[k] = getattr(self, k)
        d['frequency_subband_counts'] = [int(x) for x in self.frequency_subband_counts]
        d['primary_trees'] = [pt.to_dict() for pt in self.primary_trees]
        return d

    def _handle(self):
        text = yaml.safe_dump(self.to_dict(), sort_keys=False, default_flow_style=None, width=1 << 30)
        h = ctypes.c_void_p()
        L.check(L.lib().pdd_config_from_yaml_string(text.encode(), ctypes.byref(h)))
        return h

    def _with_handle(self, fn):
        h = self._handle()
        try:
            return fn(h)
        finally:
            L.lib().pdd_config_destroy(h)

    # ---- reference methods ----

    def validate(self):
        self._with_handle(lambda h: None)

    def clone(self):
        return DedispersionConfig.from_dict(self.to_dict())

    def to_yaml_string(self, verbose=False):
        return self._with_handle(lambda h: L.get_string(L.lib().pdd_config_to_yaml, h, int(verbose)))

    def get_nelts_per_segment(self):
        self.validate()
        return 128 // (4 if self.dtype in ('float32', 'fp32', 'f32') else 2)

    def get_total_nfreq(self):
        return int(sum(self.zone_nfreq))

    @property
    def num_primary_trees(self):
        return len(self.primary_trees)

    def _eval(self, what, x=0.0):
        def f(h):
            out = ctypes.c_double()
            L.check(L.lib().pdd_config_eval(h, what.encode(), float(x), ctypes.byref(out)))
            return out.value
        return self._with_handle(f)

    def frequency_to_index(self, f):
        return self._eval('frequency_to_index', f)

    def index_to_frequency(self, index):
        return self._eval('index_to_frequency', index)

    def delay_to_frequency(self, delay):
        return self._eval('delay_to_frequency', delay)

    def frequency_to_delay(self, f):
        return self._eval('frequency_to_delay', f)

    def dm_per_unit_delay(self):
        return self._eval('dm_per_unit_delay')

    def max_dm_of_all_trees(self):
        return self._eval('max_dm_of_all_trees')

    def max_width_of_base_tree(self):
        return self.primary_trees[0].max_width

    def make_channel_map(self):
        def f(h):
            out = np.zeros(2**int(self.toplevel_tree_rank) + 1, dtype=np.float64)
            L.check(L.lib().pdd_config_make_channel_map(h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), out.size))
            return out
        return self._with_handle(f)

    def test(self):
        self._with_handle(lambda h: L.check(L.lib().pdd_config_test(h)))


class DedispersionTree:
    """Read-only description of one output tree (include/pirate/DedispersionTree.hpp:19-70)."""

    def __init__(self, info, subband_counts):
        for name, _ in L.TreeInfo._fields_:
            setattr(self, name, getattr(info, name))
        self.frequency_subband_counts = list(subband_counts)

    def total_rank(self):
        return self.amb_rank + self.dd_rank

    def __repr__(self):
        return (f'DedispersionTree(p={self.primary_tree_index}, et={self.early_trigger_level}, '
                f'amb_rank={self.amb_rank}, dd_rank={self.dd_rank}, ndm_out={self.ndm_out}, nt_out={self.nt_out})')


class DedispersionPlan:
    """Dedispersion plan: stage-1/stage-2 trees derived from a config
    (reference: src_lib/DedispersionPlan.cpp:29-365)."""

    def __init__(self, config, gpu_runnable=True):
        self.config = config.clone()
        self._h = ctypes.c_void_p()
        ch = config._handle()
        try:
            L.check(L.lib().pdd_plan_create(ch, int(gpu_runnable), ctypes.byref(self._h)))
        finally:
            L.lib().pdd_config_destroy(ch)
        self.gpu_runnable = bool(gpu_runnable)
        y = yaml.safe_load(self.to_yaml_string())
        self.dtype = y['dtype']
        self.nfreq = int(y['nfreq'])
        self.nt_in = int(y['nt_in'])
        self.num_primary_trees = int(y['num_primary_trees'])
        self.beams_per_gpu = int(y['beams_per_gpu'])
        self.beams_per_batch = int(y['beams_per_batch'])
        self.num_active_batches = int(y['num_active_batches'])
        self.stage1_dd_rank = list(y['stage1_dd_rank'])
        self.stage1_amb_rank = list(y['stage1_amb_rank'])
        self.ntrees = int(y['ntrees'])
        self.nbits = 32 if self.dtype == 'float32' else 16
        self.nelts_per_segment = 128 * 8 // self.nbits
        self.nbytes_per_segment = 128
        self.trees = []
        for i in range(self.ntrees):
            info = L.TreeInfo()
            L.check(L.lib().pdd_plan_tree_info(self._h, i, ctypes.byref(info)))
            self.trees.append(DedispersionTree(info, y['trees'][i]['frequency_subband_counts']))

    def __del__(self):
        h = getattr(self, '_h', None)
        if h and L._lib is not None:
            L._lib.pdd_plan_destroy(h)
            self._h = None

    def to_yaml_string(self, verbose=False):
        return L.get_string(L.lib().pdd_plan_to_yaml, self._h, int(verbose))

    def decode_argmax(self, token, itree, idm_coarse, itime_coarse):
        """Returns (fmin, fmax, tlo, thi, p) (reference DedispersionPlan::decode_argmax)."""
        out = (ctypes.c_int64 * 5)()
        L.check(L.lib().pdd_plan_decode_argmax(self._h, int(token), itree, idm_coarse, itime_coarse, out))
        return tuple(int(x) for x in out)

    def decode_argmax2(self, itree, fmin, fmax, tlo, thi, p):
        """Returns (freq_lo_MHz, freq_hi_MHz, dm, timestamp_samp, width_samp)."""
        out = (ctypes.c_double * 5)()
        L.check(L.lib().pdd_plan_decode_argmax2(self._h, itree, fmin, fmax, tlo, thi, p, out))
        return tuple(float(x) for x in out)

    def decode_argmax_batch(self, itree, out_argmax):
        """Vector version of decode_argmax over an (..., ndm_out, nt_out) token array (decoded in
        the library). Returns an int64 array of shape (..., ndm_out, nt_out, 5)."""
        if hasattr(out_argmax, 'detach'):
            out_argmax = out_argmax.detach().cpu().numpy()
        a = np.ascontiguousarray(np.asarray(out_argmax).view(np.uint32) if np.asarray(out_argmax).dtype == np.int32
                                 else np.asarray(out_argmax, dtype=np.uint32))
        t = self.trees[itree]
        if a.ndim < 2 or a.shape[-2:] != (t.ndm_out, t.nt_out):
            raise ValueError(f'decode_argmax_batch: last two axes must be (ndm_out, nt_out) = {(t.ndm_out, t.nt_out)}, got {a.shape}')
        res = np.zeros(a.shape + (5,), dtype=np.int64)
        L.check(L.lib().pdd_plan_decode_argmax_batch(self._h, itree, a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), a.size,
                                                     res.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
        return res

    def compute_steady_state_it0(self, itree):
        t = self.trees[itree]
        out = np.zeros(t.ndm_out, dtype=np.int64)
        L.check(L.lib().pdd_plan_steady_state_it0(self._h, itree, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), out.size))
        return out
