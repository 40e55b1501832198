"""Per-kernel byte accounting of a dedisperser (the role of the reference's ResourceTracker,
include/pirate/ResourceTracker.hpp:14-75, with per-kernel entries as in
src_lib/TreeGriddingKernel.cu:117, GpuLaggedDownsamplingKernel.cu:768-775,
DedispersionKernel.cu:515-516, CoalescedDdKernel2.cu:131-140).

Counts are per launch of one beam batch (one chunk of beams_per_batch beams). Two models:

* algorithmic ("alg"): the compulsory HBM reads and writes of each kernel of this library's
  chain -- the fused gridding + lagged downsampling reads the input chunk and writes the
  gridded rows and every downsampled level; stage 1 reads its input rows and writes the lag
  ring; stage 2 reads the lag ring and writes the peak-finder rows; the peak finder reads
  those rows and writes out_max / out_argmax. Weights, tree states and descriptor tables are
  not counted.
* chain: the compulsory traffic of the two-stage algorithm as a whole (input read, lag-ring
  write and read, outputs), independent of how the chain is split into kernels.
"""


def _S(dtype):
    return 4 if dtype == 'float32' else 2


def kernel_alg_bytes(plan, name):
    """Algorithmic HBM bytes of one launch of kernel 'name' (one beam batch). Names are those
    of GpuDedisperser.kernel_timing(): grid_lds, stage1_p<p>, stage2_t<i>, pf_t<i>, spill_p<p>."""
    S = _S(plan.dtype)
    B, F, T = plan.beams_per_batch, plan.nfreq, plan.nt_in
    R = plan.config.toplevel_tree_rank
    npri = plan.num_primary_trees
    if name == 'grid_lds':
        return B * (F * T + 2**R * T + sum(2**(R - 1) * (T >> p) for p in range(1, npri))) * S
    if name.startswith('stage1_p'):
        p = int(name[len('stage1_p'):])
        rp = R - (1 if p else 0)
        return B * 2 * 2**rp * (T >> p) * S
    if name.startswith('stage2_t') or name.startswith('pf_t'):
        it = int(name.split('_t')[1])
        t = plan.trees[it]
        streams = t.ndm_out * t.nmultiplets_M
        if name.startswith('stage2_t'):
            return B * (2**(t.amb_rank + t.dd_rank) + streams) * t.nt_ds * S
        return B * (streams * t.nt_ds * S + t.ndm_out * t.nt_out * (S + 4))
    if name.startswith('spill_p'):
        return 0    # host spill: PCIe copies, not HBM streaming
    raise ValueError(f'unknown kernel name {name!r}')


def kernel_names(plan):
    names = ['grid_lds'] + [f'stage1_p{p}' for p in range(plan.num_primary_trees)]
    names += [f'stage2_t{i}' for i in range(plan.ntrees)] + [f'pf_t{i}' for i in range(plan.ntrees)]
    return names


def chain_alg_bytes(plan):
    """Compulsory traffic of one beam batch through the two-stage algorithm: input read +
    lag-ring write and read + outputs (the B_alg byte model)."""
    S = _S(plan.dtype)
    B, F, T = plan.beams_per_batch, plan.nfreq, plan.nt_in
    R = plan.config.toplevel_tree_rank
    ring = sum(2**(R - (1 if p else 0)) * (T >> p) for p in range(plan.num_primary_trees))
    outs = sum(t.ndm_out * t.nt_out * (S + 4) for t in plan.trees)
    return B * (F * T * S + 2 * ring * S + outs)


class ResourceTracker:
    """Bytes per beam batch: kernels[name] (HBM, algorithmic), memcpy_h2g / memcpy_g2h (PCIe,
    e.g. the spilled lag-ring rows), and footprints (device / pinned host bytes allocated).
    Rates follow the reference's get_*_bw(): bytes per batch, to be divided by the time per
    batch."""

    def __init__(self, plan, gmem_footprint=0, hmem_footprint=0, memcpy_h2g=0, memcpy_g2h=0):
        self.plan = plan
        self.kernels = {n: kernel_alg_bytes(plan, n) for n in kernel_names(plan)}
        self.gmem_footprint = int(gmem_footprint)
        self.hmem_footprint = int(hmem_footprint)
        self.memcpy_h2g = int(memcpy_h2g)
        self.memcpy_g2h = int(memcpy_g2h)

    def get_gmem_bw(self):
        """HBM bytes per batch summed over the chain's kernels (algorithmic)."""
        return sum(self.kernels.values())

    def get_chain_bytes(self):
        return chain_alg_bytes(self.plan)

    def get_h2g_bw(self):
        return self.memcpy_h2g

    def get_g2h_bw(self):
        return self.memcpy_g2h

    def to_dict(self):
        return dict(kernels=dict(self.kernels), gmem_bytes=self.get_gmem_bw(), chain_bytes=self.get_chain_bytes(),
                    gmem_footprint=self.gmem_footprint, hmem_footprint=self.hmem_footprint,
                    memcpy_h2g=self.memcpy_h2g, memcpy_g2h=self.memcpy_g2h)
