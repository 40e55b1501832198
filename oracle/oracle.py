"""ORACLE — test infrastructure only.

ctypes wrapper over oracle/build/liboracle.so, the CPU restatement of the reference's
dedispersion chain (see oracle/oracle.hpp). Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg may use this module, and only as the checker.
"""

import ctypes
import os
import subprocess

import numpy as np

_here = os.path.dirname(os.path.abspath(__file__))
_libpath = os.path.join(_here, 'build', 'liboracle.so')
_lib = None

c_long_p = ctypes.POINTER(ctypes.c_long)
c_double_p = ctypes.POINTER(ctypes.c_double)


def build():
    """Compiles the oracle (g++, a few seconds)."""
    subprocess.check_call(['make', '-s', '-C', _here])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_libpath):
            build()
        L = ctypes.CDLL(_libpath)
        L.orc_last_error.restype = ctypes.c_char_p
        for name in ['orc_tree_create', 'orc_lagbuf_create', 'orc_lds_create', 'orc_pf_create', 'orc_dd_create']:
            getattr(L, name).restype = ctypes.c_void_p
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        L.orc_trees.restype = ctypes.c_long
        L.orc_tree_subband_counts.restype = ctypes.c_long
        L.orc_dedispersion_delay.restype = ctypes.c_long
        L.orc_dedispersion_delay.argtypes = [ctypes.c_int, ctypes.c_long, ctypes.c_long]
        L.orc_rb_lag.restype = ctypes.c_long
        L.orc_rb_lag.argtypes = [ctypes.c_long, ctypes.c_long, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise RuntimeError(lib().orc_last_error().decode())


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class _OrcCfg(ctypes.Structure):
    _fields_ = [('nzones', ctypes.c_long), ('zone_nfreq', c_long_p), ('zone_freq_edges', c_double_p),
                ('R', ctypes.c_long), ('nt_in', ctypes.c_long), ('beams_per_gpu', ctypes.c_long),
                ('beams_per_batch', ctypes.c_long), ('nelts_per_segment', ctypes.c_long),
                ('n_sbc', ctypes.c_long), ('subband_counts', c_long_p), ('npri', ctypes.c_long),
                ('pt', c_long_p), ('reference_dcore', ctypes.c_long)]


class OracleConfig:
    """Config fields the oracle derives its trees from (restating the reference plan).

    'primary_trees' is a list of dicts with the reference's keys (num_early_triggers,
    max_width, dm_downsampling, time_downsampling, wt_dm_downsampling, wt_time_downsampling).
    """

    def __init__(self, zone_nfreq, zone_freq_edges, toplevel_tree_rank, time_samples_per_chunk,
                 beams_per_gpu, beams_per_batch, frequency_subband_counts, primary_trees,
                 nelts_per_segment=32, reference_dcore=False):
        self.zone_nfreq = np.asarray(zone_nfreq, dtype=np.int64)
        self.zone_freq_edges = np.asarray(zone_freq_edges, dtype=np.float64)
        self.R = int(toplevel_tree_rank)
        self.nt_in = int(time_samples_per_chunk)
        self.beams_per_gpu = int(beams_per_gpu)
        self.beams_per_batch = int(beams_per_batch)
        self.nelts_per_segment = int(nelts_per_segment)
        self.sbc = np.asarray(frequency_subband_counts, dtype=np.int64)
        pts = []
        for pt in primary_trees:
            pts.append([pt['num_early_triggers'], pt['max_width'], pt.get('dm_downsampling', 0),
                        pt.get('time_downsampling', 0), pt['wt_dm_downsampling'], pt['wt_time_downsampling']])
        self.pt = np.asarray(pts, dtype=np.int64).reshape(-1, 6)
        self.reference_dcore = bool(reference_dcore)
        self.nfreq = int(self.zone_nfreq.sum())

    @staticmethod
    def from_yaml_dict(d, **kw):
        return OracleConfig(d['zone_nfreq'], d['zone_freq_edges'], d['toplevel_tree_rank'],
                            d['time_samples_per_chunk'], d['beams_per_gpu'], d['beams_per_batch'],
                            d['frequency_subband_counts'], d['primary_trees'], **kw)

    def _c(self):
        c = _OrcCfg()
        c.nzones = len(self.zone_nfreq)
        c.zone_nfreq = self.zone_nfreq.ctypes.data_as(c_long_p)
        c.zone_freq_edges = self.zone_freq_edges.ctypes.data_as(c_double_p)
        c.R, c.nt_in = self.R, self.nt_in
        c.beams_per_gpu, c.beams_per_batch = self.beams_per_gpu, self.beams_per_batch
        c.nelts_per_segment = self.nelts_per_segment
        c.n_sbc = len(self.sbc)
        c.subband_counts = self.sbc.ctypes.data_as(c_long_p)
        c.npri = len(self.pt)
        c.pt = self.pt.ctypes.data_as(c_long_p)
        c.reference_dcore = int(self.reference_dcore)
        return c

    def channel_map(self):
        out = np.zeros(2**self.R + 1, dtype=np.float64)
        c = self._c()
        _check(lib().orc_channel_map(ctypes.byref(c), _ptr(out)))
        return out

    _tree_keys = ['ipri', 'et', 'amb_rank', 'dd_rank', 'nt_ds', 'M', 'N', 'Wmax', 'dm_ds', 'time_ds',
                  'wt_dm_ds', 'wt_time_ds', 'Dcore', 'nprofiles', 'ndm_out', 'ndm_wt', 'nt_out', 'nt_wt']

    def trees(self):
        buf = np.zeros((64, 18), dtype=np.int64)
        c = self._c()
        n = lib().orc_trees(ctypes.byref(c), _ptr(buf), 64)
        if n < 0:
            raise RuntimeError(lib().orc_last_error().decode())
        ret = []
        for i in range(n):
            t = dict(zip(self._tree_keys, [int(x) for x in buf[i]]))
            sc = np.zeros(8, dtype=np.int64)
            k = lib().orc_tree_subband_counts(ctypes.byref(c), i, _ptr(sc))
            t['subband_counts'] = [int(x) for x in sc[:k]]
            ret.append(t)
        return ret


_STORAGE = {'float32': 0, 'float16': 1, 'bfloat16': 2}


def round_to_storage(x, storage):
    """float32 array rounded to 'float16' / 'bfloat16' (round-to-nearest-even), as float32."""
    a = np.ascontiguousarray(x, dtype=np.float32).copy()
    _check(lib().orc_round_storage(_ptr(a), ctypes.c_long(a.size), _STORAGE[storage]))
    return a


def _dt(dbl):
    return np.float64 if dbl else np.float32


def dedispersion_delay(rank, freq, dm_brev):
    return lib().orc_dedispersion_delay(rank, freq, dm_brev)


def rb_lag(freq_coarse, dm_brev, stage1_rank, stage2_rank, uflag):
    return lib().orc_rb_lag(freq_coarse, dm_brev, stage1_rank, stage2_rank, int(uflag))


def bare_profile_variances(P):
    out = np.zeros(P, dtype=np.float64)
    _check(lib().orc_bare_profile_variances(ctypes.c_long(P), _ptr(out)))
    return out


def dedisperse_non_incremental(arr, nspec=1):
    """In-place whole-array tree on a (nfreq, ninner) contiguous array."""
    assert arr.flags.c_contiguous and arr.ndim == 2
    dbl = arr.dtype == np.float64
    _check(lib().orc_dedisperse_non_incremental(_ptr(arr), ctypes.c_long(arr.shape[0]),
                                                 ctypes.c_long(arr.shape[1]), ctypes.c_long(nspec), int(dbl)))
    return arr


def tree_gridding(inp, channel_map):
    """inp: (B, nfreq, ntime) -> (B, nchan, ntime), nchan = len(channel_map) - 1."""
    inp = np.ascontiguousarray(inp)
    B, F, T = inp.shape
    N = len(channel_map) - 1
    out = np.zeros((B, N, T), dtype=inp.dtype)
    cm = np.ascontiguousarray(channel_map, dtype=np.float64)
    _check(lib().orc_tree_gridding(_ptr(out), _ptr(inp), _ptr(cm), ctypes.c_long(B), ctypes.c_long(N),
                                   ctypes.c_long(F), ctypes.c_long(T), int(inp.dtype == np.float64)))
    return out


class _Obj:
    def __del__(self):
        h = getattr(self, '_h', None)
        if h and _lib is not None:
            _lib.orc_destroy(ctypes.c_void_p(h))
            self._h = None


class Tree(_Obj):
    """Incremental tree with subband outputs (reference ReferenceTree)."""

    def __init__(self, num_beams, amb_rank, dd_rank, ntime, nspec=1, subband_counts=(1,), dbl=False):
        sc = np.asarray(subband_counts, dtype=np.int64)
        self.B, self.A, self.D = num_beams, 2**amb_rank, 2**dd_rank
        self.amb_rank, self.dd_rank = amb_rank, dd_rank
        self.ntime, self.nspec, self.dbl = ntime, nspec, dbl
        self.pf_rank = len(sc) - 1
        self.M = _subbands_M(sc)
        self.Dpf = 2**(amb_rank + dd_rank - self.pf_rank)
        self._h = lib().orc_tree_create(ctypes.c_long(num_beams), ctypes.c_long(amb_rank), ctypes.c_long(dd_rank),
                                        ctypes.c_long(ntime), ctypes.c_long(nspec), _ptr(sc), ctypes.c_long(len(sc)), int(dbl))
        if not self._h:
            raise RuntimeError(lib().orc_last_error().decode())

    def dedisperse(self, buf, out=None):
        """buf: (B, A, D, T*S) contiguous, dedispersed in place; out: (B, Dpf, M, T*S) or None."""
        assert buf.flags.c_contiguous and buf.dtype == _dt(self.dbl)
        TS = self.ntime * self.nspec
        assert buf.shape == (self.B, self.A, self.D, TS)
        optr, osb, osd, osm = None, 0, 0, 0
        if out is not None:
            assert out.flags.c_contiguous and out.shape == (self.B, self.Dpf, self.M, TS)
            optr, osb, osd, osm = _ptr(out), self.Dpf * self.M * TS, self.M * TS, TS
        _check(lib().orc_tree_dedisperse(ctypes.c_void_p(self._h), _ptr(buf), ctypes.c_long(self.A * self.D * TS),
                                         ctypes.c_long(self.D * TS), ctypes.c_long(TS), optr,
                                         ctypes.c_long(osb), ctypes.c_long(osd), ctypes.c_long(osm)))


class Lagbuf(_Obj):
    def __init__(self, lags, ntime, dbl=False):
        lags = np.ascontiguousarray(lags, dtype=np.int64).reshape(-1)
        self.n, self.ntime, self.dbl = len(lags), ntime, dbl
        self._h = lib().orc_lagbuf_create(_ptr(lags), ctypes.c_long(len(lags)), ctypes.c_long(ntime), int(dbl))
        if not self._h:
            raise RuntimeError(lib().orc_last_error().decode())

    def apply(self, arr):
        assert arr.flags.c_contiguous and arr.size == self.n * self.ntime and arr.dtype == _dt(self.dbl)
        _check(lib().orc_lagbuf_apply(ctypes.c_void_p(self._h), _ptr(arr)))


class LaggedDownsampler(_Obj):
    """Lagged downsampling for primary trees p >= 1 (reference ReferenceLaggedDownsamplingKernel)."""

    def __init__(self, B, R, output_dd_rank, npri, ntime, dbl=False):
        self.B, self.R, self.npri, self.ntime, self.dbl = B, R, npri, ntime, dbl
        self._h = lib().orc_lds_create(ctypes.c_long(B), ctypes.c_long(R), ctypes.c_long(output_dd_rank),
                                       ctypes.c_long(npri), ctypes.c_long(ntime), int(dbl))
        if not self._h:
            raise RuntimeError(lib().orc_last_error().decode())

    def apply(self, inp):
        inp = np.ascontiguousarray(inp, dtype=_dt(self.dbl))
        assert inp.shape == (self.B, 2**self.R, self.ntime)
        outs = [np.zeros((self.B, 2**(self.R - 1), self.ntime // 2**p), dtype=_dt(self.dbl)) for p in range(1, self.npri)]
        arr = (ctypes.c_void_p * max(1, len(outs)))(*[o.ctypes.data for o in outs])
        _check(lib().orc_lds_apply(ctypes.c_void_p(self._h), _ptr(inp), arr))
        return outs


class PeakFinder(_Obj):
    """Reference peak finder (ReferencePeakFindingKernel)."""

    def __init__(self, subband_counts, max_width, beams_per_batch, total_beams, ndm_out, ndm_wt, nt_out, nt_in,
                 nt_wt, Dcore, dbl=False):
        sc = np.asarray(subband_counts, dtype=np.int64)
        self.B, self.ndm_out, self.ndm_wt, self.nt_out, self.nt_in, self.nt_wt = beams_per_batch, ndm_out, ndm_wt, nt_out, nt_in, nt_wt
        self.M, self.N = _subbands_M(sc), _subbands_N(sc)
        self.P = 1 + 3 * int(round(np.log2(max_width)))
        self.dbl = dbl
        self._h = lib().orc_pf_create(_ptr(sc), ctypes.c_long(len(sc)), ctypes.c_long(max_width), ctypes.c_long(beams_per_batch),
                                      ctypes.c_long(total_beams), ctypes.c_long(ndm_out), ctypes.c_long(ndm_wt),
                                      ctypes.c_long(nt_out), ctypes.c_long(nt_in), ctypes.c_long(nt_wt), ctypes.c_long(Dcore), int(dbl))
        if not self._h:
            raise RuntimeError(lib().orc_last_error().decode())

    def apply(self, inp, wt, ibatch, want_var=False):
        inp = np.ascontiguousarray(inp, dtype=_dt(self.dbl))
        wt = np.ascontiguousarray(wt, dtype=np.float32)
        assert inp.shape == (self.B, self.ndm_out, self.M, self.nt_in)
        assert wt.shape == (self.B, self.ndm_wt, self.nt_wt, self.P, self.N)
        mx = np.zeros((self.B, self.ndm_out, self.nt_out), dtype=_dt(self.dbl))
        am = np.zeros((self.B, self.ndm_out, self.nt_out), dtype=np.uint32)
        var = np.zeros((self.B, self.ndm_out, self.M, self.P), dtype=np.float64) if want_var else None
        _check(lib().orc_pf_apply(ctypes.c_void_p(self._h), _ptr(mx), _ptr(am), _ptr(var) if want_var else None,
                                  _ptr(inp), _ptr(wt), ctypes.c_long(ibatch)))
        return (mx, am, var) if want_var else (mx, am)

    def eval_tokens(self, tokens, wt):
        tokens = np.ascontiguousarray(tokens, dtype=np.uint32)
        wt = np.ascontiguousarray(wt, dtype=np.float32)
        out = np.zeros((self.B, self.ndm_out, self.nt_out), dtype=_dt(self.dbl))
        _check(lib().orc_pf_eval_tokens(ctypes.c_void_p(self._h), _ptr(out), _ptr(tokens), _ptr(wt)))
        return out


class Dedisperser(_Obj):
    """Reference dedisperser (ReferenceDedisperser), sophistication 0 or 1."""

    def __init__(self, cfg, sophistication, dbl=False, enable_variances=False, tree_domain_input=False,
                 storage='float32'):
        """storage 'float16' / 'bfloat16' (float oracle, sophistication 1): values are rounded to
        the storage type where the GPU chain stores them (oracle.hpp, RefDedisperser::dedisperse)."""
        self.cfg, self.soph, self.dbl = cfg, sophistication, dbl
        self.trees = cfg.trees()
        self.B = cfg.beams_per_batch
        self.nbatches = cfg.beams_per_gpu // cfg.beams_per_batch
        self.nchan_in = 2**cfg.R if tree_domain_input else cfg.nfreq
        self.enable_variances = enable_variances
        c = cfg._c()
        self._h = lib().orc_dd_create(ctypes.byref(c), int(sophistication), int(dbl), int(enable_variances),
                                      int(tree_domain_input), _STORAGE[storage])
        if not self._h:
            raise RuntimeError(lib().orc_last_error().decode())

    def wt_shape(self, itree):
        t = self.trees[itree]
        return (self.B, t['ndm_wt'], t['nt_wt'], t['nprofiles'], t['N'])

    def out_shape(self, itree):
        t = self.trees[itree]
        return (self.B, t['ndm_out'], t['nt_out'])

    def set_weights(self, itree, wt):
        wt = np.ascontiguousarray(wt, dtype=np.float32)
        assert wt.shape == self.wt_shape(itree)
        _check(lib().orc_dd_set_weights(ctypes.c_void_p(self._h), ctypes.c_long(itree), _ptr(wt)))

    def dedisperse(self, inp, ichunk, ibatch):
        inp = np.ascontiguousarray(inp, dtype=_dt(self.dbl))
        assert inp.shape == (self.B, self.nchan_in, self.cfg.nt_in)
        _check(lib().orc_dd_set_input(ctypes.c_void_p(self._h), _ptr(inp)))
        _check(lib().orc_dd_dedisperse(ctypes.c_void_p(self._h), ctypes.c_long(ichunk), ctypes.c_long(ibatch)))

    def output(self, itree):
        mx = np.zeros(self.out_shape(itree), dtype=_dt(self.dbl))
        am = np.zeros(self.out_shape(itree), dtype=np.uint32)
        _check(lib().orc_dd_get_output(ctypes.c_void_p(self._h), ctypes.c_long(itree), _ptr(mx), _ptr(am), None))
        return mx, am

    def variances(self, itree):
        t = self.trees[itree]
        var = np.zeros((self.B, t['ndm_out'], t['M'], t['nprofiles']), dtype=np.float64)
        _check(lib().orc_dd_get_output(ctypes.c_void_p(self._h), ctypes.c_long(itree), None, None, _ptr(var)))
        return var

    def eval_tokens(self, itree, tokens):
        tokens = np.ascontiguousarray(tokens, dtype=np.uint32)
        out = np.zeros(self.out_shape(itree), dtype=_dt(self.dbl))
        _check(lib().orc_dd_eval_tokens(ctypes.c_void_p(self._h), ctypes.c_long(itree), _ptr(tokens), _ptr(out)))
        return out


def _subbands_M(sc):
    return int(sum(int(c) * 2**lev for lev, c in enumerate(sc)))


def _subbands_N(sc):
    return int(sum(int(c) for c in sc))


def bare_weights(B, ndm_wt, nt_wt, P, N):
    """Peak-finding weights for unit-variance input: 1/sqrt(zero-lag autocorrelation) per profile
    (reference PeakFindingKernelParams::fill_host_weights with empty variances, randomize=false)."""
    v = bare_profile_variances(P).astype(np.float32)
    w = (1.0 / np.sqrt(v)).astype(np.float32)
    return np.broadcast_to(w[None, None, None, :, None], (B, ndm_wt, nt_wt, P, N)).copy()


def sparse_random_weights(rng, B, ndm_wt, nt_wt, P, N):
    """Sparse random weights in the style of the reference's randomized fill_host_weights:
    per (b, d, t) an occupancy p0 ~ U(0.01, 1.1), then each (p, n) weight is zero with
    probability ~(1 - p0), else U(0,1) times the bare-kernel weight."""
    base = bare_weights(1, 1, 1, P, N)[0, 0, 0]
    p0 = rng.uniform(0.01, 1.1, size=(B, ndm_wt, nt_wt, 1, 1)).astype(np.float32)
    r = rng.uniform(0.0, 1.0, size=(B, ndm_wt, nt_wt, P, N)).astype(np.float32)
    x = rng.uniform(0.0, 1.0, size=(B, ndm_wt, nt_wt, P, N)).astype(np.float32)
    return np.where(r < p0, x, np.float32(0.0)).astype(np.float32) * base[None, None, None]


def dequantize(data_u8, scoff_f16, nbeams, nfreq, ntime):
    """Reference dequantization (ReferenceDequantizationKernel::apply,
    src_lib/GpuDequantizationKernel.cu:326-380). data_u8: uint8 (nbeams, nfreq, ntime/2), two int4
    per byte (low nibble = even sample); scoff_f16: float16 (nbeams, nfreq, ntime/256, 2).
    Returns float32 (nbeams, nfreq, ntime)."""
    d = np.ascontiguousarray(data_u8, dtype=np.uint8)
    s = np.ascontiguousarray(scoff_f16, dtype=np.float16).view(np.uint16)
    assert d.size * 2 == nbeams * nfreq * ntime and s.size == nbeams * nfreq * (ntime // 256) * 2
    out = np.zeros((nbeams, nfreq, ntime), dtype=np.float32)
    _check(lib().orc_dequantize(_ptr(d), _ptr(s), ctypes.c_long(nbeams), ctypes.c_long(nfreq),
                                ctypes.c_long(ntime), _ptr(out)))
    return out


def ringbuf_copy(ringbuf_u8, octuples, seg_bytes, nbeams, iframe):
    """Reference ring-buffer copies (CpuRingbufCopyKernel::apply, src_lib/RingbufCopyKernel.cu:
    60-123), in place on a uint8 array. octuples: uint32 (N, 8)."""
    assert ringbuf_u8.dtype == np.uint8 and ringbuf_u8.flags.c_contiguous
    o = np.ascontiguousarray(octuples, dtype=np.uint32).reshape(-1, 8)
    _check(lib().orc_ringbuf_copy(_ptr(ringbuf_u8), ctypes.c_long(ringbuf_u8.size), _ptr(o), ctypes.c_long(len(o)),
                                  ctypes.c_long(seg_bytes), ctypes.c_long(nbeams), ctypes.c_ulonglong(iframe)))
    return ringbuf_u8
