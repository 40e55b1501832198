"""GpuDedisperser: Python surface of the reference's GpuDedisperser
(src_pybind11/pirate_pybind11.cpp:453-594, pirate_frb/GpuDedisperser.py:105-197),
backed by the C-ABI of libpirate_amd.so. Arrays are exposed as torch tensors that view
device memory owned by the dedisperser (torch on ROCm is used for streams and views).
"""

import ctypes
from contextlib import contextmanager

import numpy as np

from . import _lib as L

_TYPESTR = {'float32': '<f4', 'float16': '<f2', 'bfloat16': '<V2', 'float64': '<f8'}


class _DeviceView:
    """Minimal __cuda_array_interface__ provider for torch.as_tensor(). torch keeps this
    object alive as long as any tensor viewing its memory, and this object keeps 'owner'
    (the object whose device allocation the view points into) alive, so the allocation is
    not freed under a live tensor."""

    def __init__(self, ptr, shape, typestr, owner=None):
        self.owner = owner
        self.__cuda_array_interface__ = {'shape': tuple(int(s) for s in shape), 'typestr': typestr,
                                         'data': (int(ptr), False), 'version': 2, 'strides': None}


def _torch():
    import torch
    return torch


def device_view(ptr, shape, dtype, owner=None):
    """torch tensor viewing 'ptr' (no copy). dtype in {'float32','float16','bfloat16','uint32'}.
    The tensor holds a reference to 'owner' (the allocation's owner) while it lives."""
    torch = _torch()
    if dtype == 'uint32':
        t = torch.as_tensor(_DeviceView(ptr, shape, '<i4', owner), device='cuda')
        return t   # int32 view of the uint32 tokens (same bits)
    if dtype == 'bfloat16':
        t = torch.as_tensor(_DeviceView(ptr, shape, '<i2', owner), device='cuda')
        return t.view(torch.bfloat16)
    return torch.as_tensor(_DeviceView(ptr, shape, _TYPESTR[dtype], owner), device='cuda')


def _stream_ptr(stream):
    if stream is None:
        return _torch().cuda.current_stream().cuda_stream
    if isinstance(stream, int):
        return stream
    return stream.cuda_stream


def bare_profile_variances(P):
    """Zero-lag autocorrelation of each peak-finding profile for unit-variance white input
    (what the reference's PfVarianceConvolver returns for a single unit sample,
    src_lib/PfVariance.cpp): profile 0 is one sample; profiles 3l+1, 3l+2, 3l+3 are the kernels
    [1,1], [1/2,1,1/2], [1/2,1,1,1/2] on the level-l grid, each tap a boxcar of 2^l samples."""
    v = np.zeros(P, dtype=np.float64)
    for p in range(P):
        if p == 0:
            v[p] = 1.0
        else:
            l, q = (p - 1) // 3, (p - 1) % 3 + 1
            v[p] = 2.0**l * {1: 2.0, 2: 1.5, 3: 2.5}[q]
    return v


class GpuDedisperserOutputs:
    """Outputs of one beam batch: out_max[itree], out_argmax[itree] (tensor views valid until
    the matching release). Reference: pirate_frb/core/GpuDedisperserOutputs.py."""

    def __init__(self, out_max, out_argmax, ichunk_zero_based, ichunk_fpga_based, ibeam, out_var=None):
        self.out_max = out_max
        self.out_argmax = out_argmax
        self.out_var = out_var      # [itree] float64 (B, ndm_out, M, nprofiles) if enable_variances
        self.out_score = None       # float32 (B, rows / 64, samples / 256) learned-scorer logits (set_scorer)
        self.ichunk_zero_based = ichunk_zero_based
        self.ichunk_fpga_based = ichunk_fpga_based
        self.ibeam = ibeam


class GpuDedisperser:
    """MI355X dedisperser with the reference's acquire/release interface.

    Example::

        dd = GpuDedisperser(plan, cuda_device_id=0)
        dd.allocate()
        dd.fill_all_weights(itree, w)           # per tree
        for seq_id in range(nchunks * dd.nbatches):
            with dd.get_input(seq_id) as arr:    # (beams_per_batch, nfreq, nt_in) tensor
                arr.copy_(data)
            with dd.get_output(seq_id) as out:
                out.out_max[itree], out.out_argmax[itree]
    """

    def __init__(self, plan, stream_pool=None, cuda_device_id=0, num_consumers=1, nbatches_out=0, nbatches_wt=0,
                 initial_chunk=0, synchronous=False, enable_variances=False):
        self._setup(plan, cuda_device_id, num_consumers, nbatches_out, nbatches_wt, initial_chunk, enable_variances)
        p = L.DedisperserParams(nbatches_out, nbatches_wt, self.num_consumers, cuda_device_id, int(synchronous),
                                initial_chunk, int(self.enable_variances))
        self._h = ctypes.c_void_p()
        L.check(L.lib().pdd_dedisperser_create(plan._h, ctypes.byref(p), ctypes.byref(self._h)))

    @classmethod
    def _wrap(cls, plan, handle, cuda_device_id, nbatches_out):
        """A GpuDedisperser over an existing C handle (a member of a GpuDedisperserGroup)."""
        self = cls.__new__(cls)
        self._setup(plan, cuda_device_id, 1, nbatches_out, 0, 0, False)
        self._h = handle
        return self

    def _setup(self, plan, cuda_device_id, num_consumers, nbatches_out, nbatches_wt, initial_chunk, enable_variances):
        self.plan = plan
        self.config = plan.config
        self.dtype = plan.dtype
        self.nfreq = plan.nfreq
        self.nt_in = plan.nt_in
        self.total_beams = plan.beams_per_gpu
        self.beams_per_batch = plan.beams_per_batch
        self.nstreams = plan.num_active_batches
        self.nbatches = self.total_beams // self.beams_per_batch
        self.ntrees = plan.ntrees
        self.trees = plan.trees
        self.num_consumers = num_consumers if num_consumers > 0 else 1
        # defaults and lower bounds as the reference's (src_lib/GpuDedisperser.cpp:116-123)
        self.nbatches_out = nbatches_out if nbatches_out > 0 else self.nstreams
        self.nbatches_wt = nbatches_wt if nbatches_wt > 0 else self.nstreams
        if self.nbatches_out < self.nstreams or self.nbatches_wt < self.nstreams:
            raise ValueError(f'GpuDedisperser: nbatches_out ({self.nbatches_out}) and nbatches_wt ({self.nbatches_wt}) '
                             f'must be >= num_active_batches ({self.nstreams})')
        self.initial_chunk = initial_chunk
        self.cuda_device_id = cuda_device_id
        self.enable_variances = bool(enable_variances)

    def __del__(self):
        h = getattr(self, '_h', None)
        if h and L._lib is not None:
            L._lib.pdd_dedisperser_destroy(h)
            self._h = None

    @property
    def Dcore(self):
        return [t.Dcore for t in self.trees]

    def allocate(self, gpu_allocator=None, host_allocator=None):
        L.check(L.lib().pdd_dedisperser_allocate(self._h))

    def output_ipc_handle(self):
        """(handle, layout) exporting the output ring to a consumer process of this node:
        handle is 64 bytes (hipIpcMemHandle_t) of the device allocation holding every tree's
        out_max and out_argmax ring; layout is a picklable dict (region size, dtype, slot
        count, per-tree byte offsets and shapes) for OutputRingView. Output seq_id lives in
        slot seq_id % nbatches_out; which seq_ids are ready is agreed out of band."""
        h = ctypes.create_string_buffer(64)
        n = ctypes.c_int64()
        L.check(L.lib().pdd_output_ipc_handle(self._h, h, ctypes.byref(n)))
        trees = []
        for it, t in enumerate(self.trees):
            offs = []
            for which in (0, 1):
                o = ctypes.c_int64()
                L.check(L.lib().pdd_output_array_offset(self._h, it, which, ctypes.byref(o)))
                offs.append(o.value)
            trees.append(dict(max_offset=offs[0], argmax_offset=offs[1],
                              shape=(self.nbatches_out * self.beams_per_batch, t.ndm_out, t.nt_out)))
        layout = dict(nbytes=n.value, dtype=self.dtype, nbatches_out=self.nbatches_out,
                      beams_per_batch=self.beams_per_batch, device=self.cuda_device_id, trees=trees)
        return h.raw, layout

    def host_bytes(self):
        """Pinned host memory of lag-ring rows spilled to the host (config max_gpu_clag)."""
        n = ctypes.c_int64()
        L.check(L.lib().pdd_dedisperser_host_bytes(self._h, ctypes.byref(n)))
        return n.value

    def gpu_bytes(self):
        n = ctypes.c_int64()
        L.check(L.lib().pdd_dedisperser_gpu_bytes(self._h, ctypes.byref(n)))
        return n.value

    def resource_tracker(self):
        """Per-kernel HBM bytes per beam batch, footprints and host-spill PCIe bytes (the
        reference's GpuDedisperser::resource_tracker, include/pirate/ResourceTracker.hpp:14-75).
        Needs allocate()."""
        from .resources import ResourceTracker
        n = ctypes.c_int64()
        L.check(L.lib().pdd_dedisperser_spill_bytes(self._h, ctypes.byref(n)))
        return ResourceTracker(self.plan, gmem_footprint=self.gpu_bytes(), hmem_footprint=self.host_bytes(),
                               memcpy_h2g=n.value, memcpy_g2h=n.value)

    def tree_rows(self, itree):
        """scorer.TileSource over the peak-finder input rows of tree itree (its stage-2 output):
        row idm_out * M + m of every global beam, samples t_lo <= t < t_hi (t_hi = the tree's
        chunk samples; t_lo = 0 without subband final lags) of the latest chunk of the beam's
        batch. With one full-band subband, row r is DM trial r. Valid after get_output of that
        chunk and before the batch's next launch. Add tile origins with .with_origins()."""
        from .scorer import TileSource
        s = L.TileSource()
        L.check(L.lib().pdd_dedisperser_tree_rows(self._h, itree, ctypes.byref(s)))
        return TileSource(s, [self])

    def set_scorer(self, scorer, itree=0):
        """Runs the learned scorer (a pirate_amd.LearnedScorer or its parameter vector) inside
        the chain from the next launch on: every chunk's rows of tree itree are cut into tiles
        of 64 DM trials x 256 samples on a fixed grid and scored right after the tree's stage 2;
        the acquired outputs then carry out_score, a (B, rows / 64, samples / 256) float32
        tensor of logits."""
        p = scorer.params.detach().cpu().numpy() if hasattr(scorer, 'params') else np.asarray(scorer)
        p = np.ascontiguousarray(p, dtype=np.float32)
        L.check(L.lib().pdd_dedisperser_set_scorer(self._h, itree, p.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                                   p.size))
        self._scorer_tree = itree

    def time(self, niterations, log=print):
        """Timing loop of the reference's GpuDedisperser::time (src_lib/GpuDedisperser.cpp:1573-1690):
        every iteration uploads an int4 chunk and its float16 scales/offsets from pinned host
        memory on a copy stream, dequantizes it into the acquired input slot on the slot's
        stream and runs the chain. Returns the steady-state figures (see timing.time_dedisperser)."""
        from .timing import time_dedisperser
        return time_dedisperser(self, niterations, log=log)

    # ---- low-level interface (reference: _acquire_input etc.) ----

    def _acquire_input(self, seq_id, stream_ptr):
        slot = ctypes.c_void_p()
        L.check(L.lib().pdd_acquire_input(self._h, seq_id, ctypes.c_void_p(stream_ptr), ctypes.byref(slot)))
        self._next_seq = seq_id + 1
        return device_view(slot.value, (self.beams_per_batch, self.nfreq, self.nt_in), self.dtype, owner=self)

    def _release_input_and_launch_dd_kernels(self, seq_id, stream_ptr):
        L.check(L.lib().pdd_release_input_and_launch(self._h, seq_id, ctypes.c_void_p(stream_ptr)))

    def _acquire_output(self, consumer_id, seq_id, stream_ptr, sync=False):
        n = self.ntrees
        om = (ctypes.c_void_p * n)()
        oa = (ctypes.c_void_p * n)()
        L.check(L.lib().pdd_acquire_output(self._h, consumer_id, seq_id, ctypes.c_void_p(stream_ptr), int(sync),
                                           om, oa, n))
        mx, am = [], []
        for i, t in enumerate(self.trees):
            shape = (self.beams_per_batch, t.ndm_out, t.nt_out)
            mx.append(device_view(om[i], shape, self.dtype, owner=self))
            am.append(device_view(oa[i], shape, 'uint32', owner=self))
        ov = None
        if self.enable_variances:
            vp = (ctypes.c_void_p * n)()
            L.check(L.lib().pdd_output_variances(self._h, seq_id, vp, n))
            ov = [device_view(vp[i], (self.beams_per_batch, t.ndm_out, t.nmultiplets_M, t.nprofiles), 'float64', owner=self)
                  for i, t in enumerate(self.trees)]
        ichunk = seq_id // self.nbatches
        outs = GpuDedisperserOutputs(mx, am, ichunk, ichunk + self.initial_chunk,
                                     (seq_id % self.nbatches) * self.beams_per_batch, ov)
        if getattr(self, '_scorer_tree', None) is not None:
            sp = ctypes.c_void_p()
            ntr = ctypes.c_int64()
            ntt = ctypes.c_int64()
            L.check(L.lib().pdd_output_scores(self._h, seq_id, ctypes.byref(sp), ctypes.byref(ntr), ctypes.byref(ntt)))
            outs.out_score = device_view(sp.value, (self.beams_per_batch, ntr.value, ntt.value), 'float32', owner=self)
        return outs

    def _release_output(self, consumer_id, seq_id, stream_ptr):
        L.check(L.lib().pdd_release_output(self._h, consumer_id, seq_id, ctypes.c_void_p(stream_ptr)))

    # ---- context managers (reference: pirate_frb/GpuDedisperser.py:105-197) ----

    @contextmanager
    def get_input(self, seq_id, stream=None):
        sp = _stream_ptr(stream)
        arr = self._acquire_input(seq_id, sp)
        try:
            yield arr
        finally:
            self._release_input_and_launch_dd_kernels(seq_id, sp)

    @contextmanager
    def get_output(self, seq_id, stream=None, consumer_id=0):
        sp = _stream_ptr(stream)
        out = self._acquire_output(consumer_id, seq_id, sp)
        try:
            yield out
        finally:
            self._release_output(consumer_id, seq_id, sp)

    # ---- weights ----

    def wt_shape(self, itree):
        t = self.trees[itree]
        return (self.nbatches_wt, self.beams_per_batch, t.ndm_wt, t.nt_wt, t.nprofiles, t.nsubbands_N)

    def fill_all_weights(self, itree, weights):
        """weights: (nbatches_wt, beams_per_batch, ndm_wt, nt_wt, nprofiles, N) float, >= 0."""
        w = weights
        if hasattr(w, 'detach'):
            w = w.detach().cpu().numpy()
        w = np.ascontiguousarray(w, dtype=np.float32)
        if w.shape != self.wt_shape(itree):
            raise ValueError(f'fill_all_weights: expected shape {self.wt_shape(itree)}, got {w.shape}')
        L.check(L.lib().pdd_fill_all_weights(self._h, itree, w.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), w.size))

    def fill_analytic_weights(self, freq_variances):
        """Weights 1/sqrt(analytic variance) for every tree, beam and weight slot, so that
        out_max is an SNR for white-noise input with these per-channel variances
        (float64, (nfreq,), positive). Reference: GpuDedisperser::fill_analytic_weights
        (include/pirate/Dedisperser.hpp:244, src_lib/GpuDedisperser.cpp:1162-1238)."""
        fv = np.ascontiguousarray(freq_variances, dtype=np.float64)
        if fv.shape != (self.nfreq,):
            raise ValueError(f'fill_analytic_weights: freq_variances must have shape ({self.nfreq},), got {fv.shape}')
        L.check(L.lib().pdd_fill_analytic_weights(self._h, fv.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), fv.size))

    def fill_host_weights(self, itree, variances=None, randomize=False, rng=None):
        """Builds and uploads tree itree's weights the way the reference's
        PeakFindingKernelParams::fill_host_weights does (src_lib/PeakFindingKernel.cu:588-691):
        base[d, n, p] = 1/sqrt(variances[n, d, p]) for variances of shape (N, ndm_wt, nprofiles),
        or, with variances=None, the bare-kernel weights 1/sqrt(zero-lag autocorrelation of
        profile p) that give SNR units for unit-variance input. randomize=True multiplies each
        weight by a sparse random factor (per (beam, d, t) an occupancy p0 ~ U(0.01, 1.1); each
        (p, n) weight is 0 with probability ~(1 - p0), else U(0, 1)), drawn from 'rng' (numpy
        Generator). Returns the (nbatches_wt, B, ndm_wt, nt_wt, P, N) float32 array uploaded."""
        t = self.trees[itree]
        nb, B, D, T, P, N = self.wt_shape(itree)
        if variances is None:
            # rsqrtf(float(var)) as in the reference: the square root is taken in float32
            bw = np.float32(1.0) / np.sqrt(bare_profile_variances(P).astype(np.float32))
            base = np.broadcast_to(bw[None, None, :], (D, N, P))
        else:
            v = np.asarray(variances, dtype=np.float64)
            if v.shape != (N, D, P):
                raise ValueError(f'fill_host_weights: variances must have shape (N, ndm_wt, P) = {(N, D, P)}, got {v.shape}')
            if not np.all(v > 0):
                raise ValueError('fill_host_weights: variances must be positive')
            base = (np.float32(1.0) / np.sqrt(v.astype(np.float32))).transpose(1, 0, 2)      # (D, N, P)
        w = np.broadcast_to(base.transpose(0, 2, 1)[None, None, :, None, :, :], (nb, B, D, T, P, N)).astype(np.float32)
        if randomize:
            rng = rng if rng is not None else np.random.default_rng()
            p0 = rng.uniform(0.01, 1.1, size=(nb, B, D, T, 1, 1))
            r = rng.uniform(0.0, 1.0, size=w.shape)
            x = rng.uniform(0.0, 1.0, size=w.shape)
            w = (np.where(r < p0, x, 0.0).astype(np.float32) * w).astype(np.float32)
        self.fill_all_weights(itree, w)
        return w

    def stop(self, reason=''):
        L.check(L.lib().pdd_stop(self._h, reason.encode()))

    def set_tuning(self, name, value):
        """Sets a launch shape that changes no output, only speed: 'pf_target_blocks',
        'grid_xcd_run' or 'fork' (see include/pirate_amd.h); applies from the next launch."""
        L.check(L.lib().pdd_dedisperser_set_tuning(self._h, name.encode(), int(value)))

    def apply_tuning_table(self, table=None):
        """Sets the launch shapes that pirate_amd.autotune recorded for this dedisperser's
        workload signature (tree rank, channels, primary trees, dtype, beams, chunk); returns
        them, or None (defaults kept) when the table has no entry for it."""
        from . import autotune
        t = autotune.load_table() if table is None else table
        ent = t.get(autotune.signature(self.plan))
        if ent is None:
            return None
        for k, v in ent['tuning'].items():
            self.set_tuning(k, v)
        return dict(ent['tuning'])

    def get_tuning(self, name):
        v = ctypes.c_int64()
        L.check(L.lib().pdd_dedisperser_get_tuning(self._h, name.encode(), ctypes.byref(v)))
        return v.value

    def set_kernel_timing(self, enable):
        L.check(L.lib().pdd_kernel_timing(self._h, int(enable)))

    def kernel_timing(self):
        """{kernel name: (launches, total_ms)} accumulated since set_kernel_timing(True)."""
        rep = L.get_string(L.lib().pdd_kernel_timing_report, self._h)
        out = {}
        for line in rep.splitlines():
            name, n, ms = line.split()
            out[name] = (int(n), float(ms))
        return out

    def debug_copy(self, name):
        """Host copy of an internal buffer (tests): 'grid', 'lds<p>', 'sb<itree>', 'ring<p>'."""
        n = ctypes.c_size_t()
        L.check(L.lib().pdd_debug_copy(self._h, name.encode(), None, 0, ctypes.byref(n)))
        buf = np.zeros(n.value, dtype=np.uint8)
        L.check(L.lib().pdd_debug_copy(self._h, name.encode(), buf.ctypes.data_as(ctypes.c_void_p), n.value, ctypes.byref(n)))
        return buf


class OutputRingView:
    """Consumer-side view of a dedisperser's output ring exported by
    GpuDedisperser.output_ipc_handle() in another process of the node (the reference's
    grouper side, src_lib/FrbGrouper.cpp:565-599). out_max(itree, slot) and
    out_argmax(itree, slot) are (beams_per_batch, ndm_out, nt_out) tensors on the GPU
    (argmax tokens as int32 with the uint32 bits). Each returned tensor keeps this view
    alive, so the mapping outlives every tensor unless close() is called explicitly; close()
    waits for the device's queued work and then unmaps, after which the tensors must not be
    used."""

    def __init__(self, handle, layout):
        self.layout = layout
        self.device = int(layout['device'])
        self._base = ctypes.c_void_p()
        L.check(L.lib().pdd_ipc_open(bytes(handle), self.device, ctypes.byref(self._base)))

    def _view(self, itree, slot, which):
        if self._base is None or not self._base.value:
            raise RuntimeError('OutputRingView: closed')
        t = self.layout['trees'][itree]
        if not 0 <= slot < self.layout['nbatches_out']:
            raise IndexError(f'OutputRingView: slot {slot} out of range')
        B = self.layout['beams_per_batch']
        nslot, ndm, nt = t['shape']
        dtype = self.layout['dtype'] if which == 0 else 'uint32'
        esize = 2 if dtype in ('float16', 'bfloat16') else 4
        off = t['max_offset' if which == 0 else 'argmax_offset'] + slot * B * ndm * nt * esize
        return device_view(self._base.value + off, (B, ndm, nt), dtype, owner=self)

    def out_max(self, itree, slot):
        return self._view(itree, slot, 0)

    def out_argmax(self, itree, slot):
        return self._view(itree, slot, 1)

    def close(self):
        if self._base is not None and self._base.value:
            L.check(L.lib().pdd_ipc_close(self._base, self.device))
        self._base = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        if getattr(self, '_base', None) is not None and L._lib is not None:
            try:
                self.close()
            except Exception:     # noqa: BLE001 -- interpreter shutdown
                pass
