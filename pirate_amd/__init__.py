"""pirate_amd: the pirate FRB-search GPU dedispersion chain, rebuilt for AMD MI355X (gfx950).

Python surface (mirrors pirate_frb's DedispersionConfig / DedispersionPlan / GpuDedisperser)
over the C-ABI of libpirate_amd.so (include/pirate_amd.h). Build the library with
``python -m pirate_amd.build``.
"""

from .config import DedispersionConfig, DedispersionPlan, DedispersionTree, PrimaryTree
from .dedisperser import GpuDedisperser, GpuDedisperserOutputs, OutputRingView
from .group import GpuDedisperserGroup
from .kernels import GpuDequantizationKernel, GpuRingbufCopyKernel
from .pf_variance import PfAvarApproximation
from .offline import AssembledFrame, OfflineDedisperser, XEngineMetadata
from .scorer import LearnedScorer, ScoreHead, TileSource

__all__ = ['DedispersionConfig', 'DedispersionPlan', 'DedispersionTree', 'PrimaryTree',
           'GpuDedisperser', 'GpuDedisperserOutputs', 'GpuDedisperserGroup', 'OutputRingView', 'GpuDequantizationKernel', 'GpuRingbufCopyKernel',
           'PfAvarApproximation', 'OfflineDedisperser', 'AssembledFrame', 'XEngineMetadata', 'LearnedScorer',
           'ScoreHead', 'TileSource']
