"""Standalone kernels with the reference's Python surface (pirate_frb/kernels).

GpuDequantizationKernel mirrors pirate_frb/kernels/GpuDequantizationKernel.py and the
binding in src_pybind11/pirate_pybind11_kernels.cpp:46-95: constructed with
(dtype, nbeams, nfreq, ntime), launched with (out, scales_offsets, data_uint8, stream).
Arrays are torch tensors on the GPU instead of ksgpu arrays; the stream defaults to
torch's current stream.

GpuRingbufCopyKernel mirrors pirate's GpuRingbufCopyKernel (include/pirate/RingbufCopyKernel.hpp:
15-60): constructed from RingbufCopyKernelParams' fields (total_beams, beams_per_batch,
nelts_per_segment, octuples) plus the ring's element dtype, launched with
(ringbuf, ichunk, ibatch, stream).
"""

import ctypes

import numpy as np

from . import _lib as L


class GpuDequantizationKernel:
    """int4 -> float dequantization: out = scale * value + offset per (beam, freq, 256-sample
    minichunk), value -8 (missing sample) -> 0 (include/pirate/GpuDequantizationKernel.hpp:17-43)."""

    def __init__(self, dtype, nbeams, nfreq, ntime):
        dtype = str(dtype).replace('torch.', '')
        if dtype not in ('float32', 'float16', 'bfloat16'):
            raise ValueError(f'GpuDequantizationKernel: output dtype must be float32, float16 or bfloat16, got {dtype}')
        if ntime % 256:
            raise ValueError(f'GpuDequantizationKernel: ntime must be a multiple of 256, got {ntime}')
        self.dtype, self.nbeams, self.nfreq, self.ntime = dtype, int(nbeams), int(nfreq), int(ntime)

    def launch(self, out, scales_offsets, data_uint8, stream=None):
        """out: (nbeams, nfreq, ntime) of self.dtype; scales_offsets: (nbeams, nfreq, ntime//256, 2)
        float16; data_uint8: (nbeams, nfreq, ntime//2) uint8 (low nibble = even sample).
        All contiguous torch tensors on the GPU. Asynchronous on 'stream'."""
        import torch
        B, F, T = self.nbeams, self.nfreq, self.ntime
        want = {'out': ((B, F, T), getattr(torch, self.dtype)), 'scales_offsets': ((B, F, T // 256, 2), torch.float16),
                'data_uint8': ((B, F, T // 2), torch.uint8)}
        for name, arr in (('out', out), ('scales_offsets', scales_offsets), ('data_uint8', data_uint8)):
            shape, dt = want[name]
            if tuple(arr.shape) != shape or arr.dtype != dt or not arr.is_cuda or not arr.is_contiguous():
                raise ValueError(f'GpuDequantizationKernel.launch: {name} must be a contiguous GPU tensor of shape '
                                 f'{shape} and dtype {dt}, got {tuple(arr.shape)} {arr.dtype}')
        if stream is None:
            stream = torch.cuda.current_stream()
        sp = stream if isinstance(stream, int) else stream.cuda_stream
        L.check(L.lib().pdd_dequantize(self.dtype.encode(), B, F, T, ctypes.c_void_p(data_uint8.data_ptr()),
                                       ctypes.c_void_p(scales_offsets.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                                       ctypes.c_void_p(sp)))


_ITEMSIZE = {'float32': 4, 'float16': 2, 'bfloat16': 2, 'uint32': 4, 'uint16': 2, 'int32': 4, 'int16': 2}


class GpuRingbufCopyKernel:
    """Segment copies inside a ring buffer, one octuple per copy
    {src_offset, src_phase, src_len, src_nseg, dst_offset, dst_phase, dst_len, dst_nseg}
    (offsets in segments, len = frames in the zone, nseg = segments per frame); a launch for
    (ichunk, ibatch) copies beams_per_batch consecutive frames of each octuple
    (src_lib/RingbufCopyKernel.cu:60-123, 161-220). Segments are nelts_per_segment elements of
    'dtype' (128 or 256 bytes)."""

    def __init__(self, total_beams, beams_per_batch, nelts_per_segment, octuples, dtype='float32'):
        dtype = str(dtype).replace('torch.', '')
        if dtype not in _ITEMSIZE:
            raise ValueError(f'GpuRingbufCopyKernel: unsupported ring dtype {dtype}')
        o = np.ascontiguousarray(octuples, dtype=np.uint32)
        if o.size % 8 or (o.ndim == 2 and o.shape[1] not in (4, 8)):
            raise ValueError(f'GpuRingbufCopyKernel: octuples must have shape (2N, 4) or (N, 8), got {o.shape}')
        self.total_beams, self.beams_per_batch = int(total_beams), int(beams_per_batch)
        self.nelts_per_segment, self.dtype = int(nelts_per_segment), dtype
        self.nbytes_per_segment = self.nelts_per_segment * _ITEMSIZE[dtype]
        self.octuples = o.reshape(-1, 8)
        self._h = ctypes.c_void_p()
        L.check(L.lib().pdd_ringbuf_copy_create(self.total_beams, self.beams_per_batch, self.nbytes_per_segment,
                                                o.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), o.size,
                                                ctypes.byref(self._h)))

    def __del__(self):
        h = getattr(self, '_h', None)
        if h and L._lib is not None:
            L._lib.pdd_ringbuf_copy_destroy(h)
            self._h = None

    def launch(self, ringbuf, ichunk, ibatch, stream=None):
        """ringbuf: 1-D contiguous torch tensor on the GPU. Asynchronous on 'stream'."""
        import torch
        if ringbuf.dim() != 1 or not ringbuf.is_cuda or not ringbuf.is_contiguous():
            raise ValueError('GpuRingbufCopyKernel.launch: ringbuf must be a 1-D contiguous GPU tensor')
        if stream is None:
            stream = torch.cuda.current_stream()
        sp = stream if isinstance(stream, int) else stream.cuda_stream
        nbytes = ringbuf.numel() * ringbuf.element_size()
        L.check(L.lib().pdd_ringbuf_copy_launch(self._h, ctypes.c_void_p(ringbuf.data_ptr()), nbytes, int(ichunk),
                                                int(ibatch), ctypes.c_void_p(sp)))
