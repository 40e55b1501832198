"""Single-beam offline dedisperser driven one int4 frame at a time.

Mirrors pirate_frb/OfflineDedisperser.py:13-257: the first frame's metadata fixes the
frequency zones and time sampling (overriding the config's), builds a one-beam plan and
GpuDedisperser, and fills analytic weights from the per-zone noise variances, so out_max is
an SNR; each frame is uploaded, dequantized into the dedisperser's input slot (the int4
dequantization kernel) and dedispersed, and its outputs are yielded inside a context manager.
Later frames must be the same beam, consecutive in time_chunk_index and carry the same
metadata. A failure latches: every later call raises, pointing at the first error.

The frame and metadata types carry only the fields this driver reads, with the reference's
names (include/pirate/AssembledFrame.hpp:40-95, include/pirate/XEngineMetadata.hpp:40-160);
any objects with these attributes can be passed instead.
"""

from contextlib import contextmanager
from dataclasses import dataclass, field

import numpy as np

from .config import DedispersionPlan
from .dedisperser import GpuDedisperser
from .kernels import GpuDequantizationKernel


@dataclass
class XEngineMetadata:
    zone_nfreq: list
    zone_freq_edges: list
    dt_ns_per_seq: int
    seq_per_frb_time_sample: int
    noise_variance: list                  # per zone
    unix_ns_at_seq_0: int = 0
    beamset: int = 0
    beam_ids: list = field(default_factory=lambda: [0])

    def get_channel_variances(self):
        """Per-channel noise variance: each zone's value repeated over its channels
        (src_lib/XEngineMetadata.cpp:116-133)."""
        if len(self.noise_variance) != len(self.zone_nfreq):
            raise ValueError('XEngineMetadata: noise_variance must have one entry per zone')
        return np.repeat(np.asarray(self.noise_variance, dtype=np.float64), np.asarray(self.zone_nfreq, dtype=np.int64))


@dataclass
class AssembledFrame:
    """One beam, one time chunk of int4 data: data (nfreq, ntime/2) uint8, two samples per
    byte, low nibble first, -8 = missing; scales_offsets (nfreq, ntime/256, 2) float16."""
    beam_id: int
    time_chunk_index: int
    data: np.ndarray
    scales_offsets: np.ndarray
    metadata: XEngineMetadata

    @property
    def nfreq(self):
        return int(self.data.shape[0])

    @property
    def ntime(self):
        return int(self.data.shape[1]) * 2


class OfflineDedisperser:
    """OfflineDedisperser(config, cuda_device_id=0); `with od.dedisperse(frame) as outputs:`
    yields GpuDedisperserOutputs (out_max[itree], out_argmax[itree], valid inside the block)."""

    def __init__(self, config, cuda_device_id=0):
        self.original_config = config
        self.config = None
        self.cuda_device_id = cuda_device_id
        self.dtype = config.dtype
        self.plan = self.dd = self.dqk = None
        self.nfreq = self.nt_in = self.ntrees = self.trees = None
        self.zone_nfreq = self.zone_freq_edges = None
        self.dt_ns_per_seq = self.seq_per_frb_time_sample = None
        self.time_sample_ms = self.noise_variance = None
        self.unix_ns_at_seq_0 = self.beamset = None
        self._beam_id = None
        self._initial_tci = None
        self._seq_id = 0
        self._failed = None

    def _initialize(self, md):
        cfg = self.original_config.clone()
        cfg.beams_per_gpu = 1
        cfg.beams_per_batch = 1
        cfg.num_active_batches = 1
        cfg.zone_nfreq = [int(x) for x in md.zone_nfreq]
        cfg.zone_freq_edges = [float(x) for x in md.zone_freq_edges]
        cfg.time_sample_ms = md.dt_ns_per_seq * md.seq_per_frb_time_sample / 1.0e6
        cfg.validate()
        self.config = cfg
        self.plan = DedispersionPlan(cfg)
        self.nfreq, self.nt_in = self.plan.nfreq, self.plan.nt_in
        self.ntrees, self.trees = self.plan.ntrees, self.plan.trees
        self.dd = GpuDedisperser(self.plan, cuda_device_id=self.cuda_device_id, num_consumers=1,
                                 nbatches_out=1, nbatches_wt=1, initial_chunk=0)
        self.dd.allocate()
        self.dqk = GpuDequantizationKernel(self.dtype, 1, self.nfreq, self.nt_in)
        fv = np.ascontiguousarray(md.get_channel_variances(), dtype=np.float64)
        if fv.shape != (self.nfreq,):
            raise ValueError(f'OfflineDedisperser: metadata gives {fv.shape[0]} channel variances, plan has {self.nfreq} channels')
        self.dd.fill_analytic_weights(fv)
        self.zone_nfreq = list(md.zone_nfreq)
        self.zone_freq_edges = list(md.zone_freq_edges)
        self.dt_ns_per_seq = md.dt_ns_per_seq
        self.seq_per_frb_time_sample = md.seq_per_frb_time_sample
        self.time_sample_ms = cfg.time_sample_ms
        self.noise_variance = list(md.noise_variance)
        self.unix_ns_at_seq_0 = md.unix_ns_at_seq_0
        self.beamset = md.beamset

    def _check_metadata(self, md):
        for name in ('zone_nfreq', 'zone_freq_edges', 'noise_variance'):
            if list(getattr(md, name)) != getattr(self, name):
                raise AssertionError(f'OfflineDedisperser: metadata {name} changed mid-stream')
        for name in ('dt_ns_per_seq', 'seq_per_frb_time_sample', 'unix_ns_at_seq_0', 'beamset'):
            if getattr(md, name) != getattr(self, name):
                raise AssertionError(f'OfflineDedisperser: metadata {name} changed mid-stream')

    @contextmanager
    def dedisperse(self, frame, stream=None):
        """Dedisperses one frame; yields its outputs. All GPU work is issued on 'stream'
        (default: torch's current stream) and the outputs are tied to it."""
        if self._failed is not None:
            raise RuntimeError(f'OfflineDedisperser: a previous dedisperse() call raised ({self._failed!r}); '
                               'internal state may be out of sync -- construct a fresh OfflineDedisperser') from self._failed
        try:
            import torch
            md = frame.metadata
            if md is None:
                raise AssertionError('OfflineDedisperser: frame.metadata is None')
            if list(md.beam_ids) != [frame.beam_id]:
                raise AssertionError(f'OfflineDedisperser: frame.metadata.beam_ids {list(md.beam_ids)} != [{frame.beam_id}]')
            if self.dd is None:
                self._beam_id = frame.beam_id
                self._initial_tci = frame.time_chunk_index
                self._initialize(md)
            else:
                if frame.beam_id != self._beam_id:
                    raise AssertionError(f'OfflineDedisperser: beam_id changed mid-stream ({frame.beam_id} != {self._beam_id})')
                expect = self._initial_tci + self._seq_id
                if frame.time_chunk_index != expect:
                    raise AssertionError(f'OfflineDedisperser: non-consecutive time_chunk_index (got {frame.time_chunk_index}, expected {expect})')
                self._check_metadata(md)
            if frame.nfreq != self.nfreq or frame.ntime != self.nt_in:
                raise AssertionError(f'OfflineDedisperser: frame shape ({frame.nfreq}, {frame.ntime}) != ({self.nfreq}, {self.nt_in})')

            seq = self._seq_id
            if stream is None:
                stream = torch.cuda.current_stream()
            with torch.cuda.stream(stream):
                data = torch.from_numpy(np.ascontiguousarray(frame.data, dtype=np.uint8)[None]).to('cuda', non_blocking=False)
                so = np.ascontiguousarray(frame.scales_offsets, dtype=np.float16)[None]
                scoff = torch.from_numpy(so).to('cuda', non_blocking=False)
                with self.dd.get_input(seq, stream=stream) as arr:
                    self.dqk.launch(arr, scoff, data, stream=stream)
            self._seq_id += 1
            with self.dd.get_output(seq, stream=stream) as outputs:
                yield outputs
        except BaseException as e:
            self._failed = e
            raise
