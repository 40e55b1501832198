"""Standalone kernels with the reference's Python surface (pirate_frb/kernels).

GpuDequantizationKernel mirrors pirate_frb/kernels/GpuDequantizationKernel.py and the
binding in src_pybind11/pirate_pybind11_kernels.cpp:46-95: constructed with
(dtype, nbeams, nfreq, ntime), launched with (out, scales_offsets, data_uint8, stream).
Arrays are torch tensors on the GPU instead of ksgpu arrays; the stream defaults to
torch's current stream.
"""

import ctypes

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
