"""GpuDedisperserGroup: beam groups of different shapes (channel counts, tree ranks, primary
trees, downsampling levels), each with its own DedispersionPlan, run as one chunk pipeline on
one GPU: per chunk, every group's chain is a branch of one HIP graph launch.

The reference batches homogeneous beams only (one plan, beams_per_batch beams per launch
sequence: include/pirate/DedispersionConfig.hpp:81-84, src_lib/GpuDedisperser.cpp:988-1105);
this class has no counterpart there. Its members are GpuDedispersers (weights, launch shapes
and timing are set per member); the chunk protocol mirrors GpuDedisperser's get_input /
get_output with one chunk index for all members.

Example::

    g = GpuDedisperserGroup([plan_r12, plan_r13, plan_r14])
    g.allocate()
    for m in g.members:
        for it in range(m.ntrees):
            m.fill_host_weights(it)
    for ichunk in range(nchunks):
        with g.get_input(ichunk) as slots:       # slots[m]: (B_m, F_m, T) tensor
            ...
        with g.get_output(ichunk) as outs:       # outs[m].out_max[itree], outs[m].out_argmax[itree]
            ...
"""

import ctypes
from contextlib import contextmanager

from . import _lib as L
from .dedisperser import GpuDedisperser, GpuDedisperserOutputs, device_view, _stream_ptr


class GpuDedisperserGroup:
    def __init__(self, plans, nslots=2, nbatches_out=2, cuda_device_id=0, use_graph=False):
        plans = list(plans)
        if not plans:
            raise ValueError('GpuDedisperserGroup: no plans')
        self.plans = plans
        self.cuda_device_id = cuda_device_id
        arr = (ctypes.c_void_p * len(plans))(*[p._h.value for p in plans])
        prm = L.GroupParams(nslots, nbatches_out, cuda_device_id, int(bool(use_graph)))
        self._h = ctypes.c_void_p()
        L.check(L.lib().pdd_group_create(arr, len(plans), ctypes.byref(prm), ctypes.byref(self._h)))
        self.members = []
        for m, plan in enumerate(plans):
            h = ctypes.c_void_p()
            L.check(L.lib().pdd_group_member(self._h, m, ctypes.byref(h)))
            self.members.append(GpuDedisperser._wrap(plan, h, cuda_device_id, nbatches_out))
        self.nbatches_out = self.members[0].nbatches_out
        self.nt_in = plans[0].nt_in

    def __del__(self):
        h = getattr(self, '_h', None)
        if h and L._lib is not None:
            L._lib.pdd_group_destroy(h)
            self._h = None

    def allocate(self):
        L.check(L.lib().pdd_group_allocate(self._h))

    def gpu_bytes(self):
        n = ctypes.c_int64()
        L.check(L.lib().pdd_group_gpu_bytes(self._h, ctypes.byref(n)))
        return n.value

    def set_use_graph(self, on):
        L.check(L.lib().pdd_group_set_use_graph(self._h, int(bool(on))))

    def set_tuning(self, imember, name, value):
        L.check(L.lib().pdd_group_set_tuning(self._h, imember, name.encode(), int(value)))

    def graph_nodes(self):
        n = ctypes.c_int64()
        L.check(L.lib().pdd_group_graph_nodes(self._h, ctypes.byref(n)))
        return n.value

    def stop(self, reason=''):
        L.check(L.lib().pdd_group_stop(self._h, reason.encode()))

    # ---- chunk protocol ----

    def _acquire_input(self, ichunk, stream_ptr):
        n = len(self.members)
        slots = (ctypes.c_void_p * n)()
        L.check(L.lib().pdd_group_acquire_input(self._h, ichunk, ctypes.c_void_p(stream_ptr), slots, n))
        return [device_view(slots[i], (m.beams_per_batch, m.nfreq, m.nt_in), m.dtype, owner=m)
                for i, m in enumerate(self.members)]

    def _release_input_and_launch(self, ichunk, stream_ptr):
        L.check(L.lib().pdd_group_release_input_and_launch(self._h, ichunk, ctypes.c_void_p(stream_ptr)))

    def _acquire_output(self, ichunk, stream_ptr, sync=False):
        n = sum(m.ntrees for m in self.members)
        om = (ctypes.c_void_p * n)()
        oa = (ctypes.c_void_p * n)()
        L.check(L.lib().pdd_group_acquire_output(self._h, ichunk, ctypes.c_void_p(stream_ptr), int(sync), om, oa, n))
        outs, k = [], 0
        for m in self.members:
            mx, am = [], []
            for t in m.trees:
                shape = (m.beams_per_batch, t.ndm_out, t.nt_out)
                mx.append(device_view(om[k], shape, m.dtype, owner=m))
                am.append(device_view(oa[k], shape, 'uint32', owner=m))
                k += 1
            outs.append(GpuDedisperserOutputs(mx, am, ichunk, ichunk, 0))
        return outs

    def _release_output(self, ichunk, stream_ptr):
        L.check(L.lib().pdd_group_release_output(self._h, ichunk, ctypes.c_void_p(stream_ptr)))

    @contextmanager
    def get_input(self, ichunk, stream=None):
        sp = _stream_ptr(stream)
        slots = self._acquire_input(ichunk, sp)
        try:
            yield slots
        finally:
            self._release_input_and_launch(ichunk, sp)

    @contextmanager
    def get_output(self, ichunk, stream=None, sync=False):
        sp = _stream_ptr(stream)
        outs = self._acquire_output(ichunk, sp, sync)
        try:
            yield outs
        finally:
            self._release_output(ichunk, sp)
