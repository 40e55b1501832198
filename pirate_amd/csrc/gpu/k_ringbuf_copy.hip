// Segment copies inside a ring buffer, driven by "octuples" (gfx950).
//
// Reference semantics: CpuRingbufCopyKernel / GpuRingbufCopyKernel
// (src_lib/RingbufCopyKernel.cu:23-123, 161-220, include/pirate/RingbufCopyKernel.hpp:15-60).
// Each octuple names a source and a destination as (segment offset, frame phase, frames in
// zone, segments per frame); a launch for frame index iframe = ichunk * total_beams +
// ibatch * beams_per_batch copies, for beams b = 0..nbeams-1, the segment
//     offset_src + ((phase_src + iframe + b) mod len_src) * nseg_src
// to
//     offset_dst + ((phase_dst + iframe + b) mod len_dst) * nseg_dst.
// A segment is 128 or 256 bytes. Beams of one octuple are copied in order by the same lanes
// (a destination that is a later beam's source sees the copied data, as in the reference's
// sequential loop); copies of different octuples must not overlap, as the reference requires.
//
// Mapping: a group of L = segment_bytes / 16 lanes per octuple (a wave64 covers 64/L
// octuples), one 16-byte load and store per lane per beam. Segment indices are computed in
// 64 bits; the host checks every addressed segment against the buffer size before launch.

#include "device_common.hpp"
#include "kernels.hpp"

namespace pirate_amd {

template<int L>
__global__ void __launch_bounds__(256)
k_ringbuf_copy(uint4 *rb, const uint32_t *oct, long noct, int nbeams, unsigned long long iframe)
{
    const long gid = long(blockIdx.x) * 256 + threadIdx.x;
    const long io = gid / L;
    const int k = int(gid % L);
    if (io >= noct)
        return;
    const uint32_t *o = oct + 8 * io;
    const unsigned long long soff = o[0], slen = o[2], snseg = o[3];
    const unsigned long long doff = o[4], dlen = o[6], dnseg = o[7];
    unsigned long long sph = (o[1] + iframe) % slen;
    unsigned long long dph = (o[5] + iframe) % dlen;
    for (int b = 0; b < nbeams; b++) {
        const uint4 v = rb[(soff + sph * snseg) * L + k];
        rb[(doff + dph * dnseg) * L + k] = v;
        sph = (sph + 1 == slen) ? 0 : (sph + 1);
        dph = (dph + 1 == dlen) ? 0 : (dph + 1);
    }
}

// Host spill, device side: outbox row i of beam b (Tp elements at outbox_off + i * Tp of the
// beam's ring region, just filled from host memory) is appended to device ring row rows[i]
// at ring time chunk * Tp, i.e. at positions (chunk * Tp + t) mod cap, t < Tp. The row's
// capacity is at least Tp, so a copy wraps at most once. One workgroup per (row, beam).
template<typename E>
__global__ void __launch_bounds__(256)
k_spill_scatter(E *ring, long bstride, long outbox_off, const int *rows, const long *row_off, const long *row_cap,
                int Tp, long chunk, int beam0)
{
    const int i = blockIdx.x, b = blockIdx.y;
    const long r = rows[i];
    const long cap = row_cap[r];
    const long pos0 = (chunk * long(Tp)) % cap;
    E *base = ring + long(beam0 + b) * bstride;
    const E *src = base + outbox_off + long(i) * Tp;
    E *dst = base + row_off[r];
    for (int t = threadIdx.x; t < Tp; t += 256) {
        long pos = pos0 + t;
        if (pos >= cap)
            pos -= cap;
        dst[pos] = src[t];
    }
}

void launch_spill_scatter(void *ring, long bstride, int elt_bytes, long outbox_off, const int *rows, long nrows,
                          const long *row_off, const long *row_cap, long Tp, long chunk, int beam0, int B, hipStream_t st)
{
    if (nrows == 0)
        return;
    const dim3 grid{unsigned(nrows), unsigned(B), 1u}, block{256u, 1u, 1u};
    if (elt_bytes == 4)
        hipLaunchKernelGGL(k_spill_scatter<uint32_t>, grid, block, 0, st, reinterpret_cast<uint32_t *>(ring), bstride,
                           outbox_off, rows, row_off, row_cap, int(Tp), chunk, beam0);
    else
        hipLaunchKernelGGL(k_spill_scatter<uint16_t>, grid, block, 0, st, reinterpret_cast<uint16_t *>(ring), bstride,
                           outbox_off, rows, row_off, row_cap, int(Tp), chunk, beam0);
}

void launch_ringbuf_copy(void *ringbuf, const uint32_t *octuples, long noct, long seg_bytes, int nbeams,
                         unsigned long long iframe, hipStream_t st)
{
    if (noct == 0)
        return;
    const int L = int(seg_bytes / 16);
    const long nthreads = noct * L;
    dim3 grid((nthreads + 255) / 256), block(256);
    uint4 *rb = reinterpret_cast<uint4 *>(ringbuf);
    if (L == 8)
        hipLaunchKernelGGL(k_ringbuf_copy<8>, grid, block, 0, st, rb, octuples, noct, nbeams, iframe);
    else
        hipLaunchKernelGGL(k_ringbuf_copy<16>, grid, block, 0, st, rb, octuples, noct, nbeams, iframe);
}

}  // namespace pirate_amd
