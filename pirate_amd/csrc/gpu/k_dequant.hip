// Dequantization of packed int4 input on gfx950.
// Reference semantics: ReferenceDequantizationKernel::apply (src_lib/GpuDequantizationKernel.cu:326-380),
// format in include/pirate/GpuDequantizationKernel.hpp:17-43:
//   data (nbeams, nfreq, ntime) int4, two per byte, low nibble = even sample;
//   scales_offsets (nbeams, nfreq, ntime/256, 2) float16 {scale, offset}, converted to float
//   before any arithmetic;  out = (value == -8) ? 0 : scale * value + offset.
// The product of a float16 scale and a 4-bit integer is exact in float32, so the single
// rounding of fmaf equals the reference's multiply-then-add. 16-bit outputs round that float32
// value to the storage type.
//
// HBM-bound (0.5 B read, 2-4 B written per sample). One thread per 32-bit word (8 samples);
// 32 consecutive threads share one (scale, offset) pair; stores are 16-byte vectors.

#include "device_common.hpp"
#include "kernels.hpp"

namespace pirate_amd {

template<typename Tout>
__global__ void __launch_bounds__(256)
k_dequant(const uint32_t *data, const uint32_t *scoff, Tout *out, long nwords)
{
    const long w = long(blockIdx.x) * 256 + threadIdx.x;
    if (w >= nwords)
        return;
    const uint32_t word = data[w];
    const uint32_t so = scoff[w >> 5];       // 256 samples = 32 words per pair
    const float scale = __half2float(__ushort_as_half((unsigned short) (so & 0xffffu)));
    const float offset = __half2float(__ushort_as_half((unsigned short) (so >> 16)));
    float x[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const int nib = int((word >> (4 * k)) & 0xfu);
        const float v = float((nib ^ 8) - 8);
        x[k] = (nib == 8) ? 0.0f : fmaf(scale, v, offset);
    }
    Tout *o = out + 8 * w;
    st4_st(o, make_float4(x[0], x[1], x[2], x[3]));
    st4_st(o + 4, make_float4(x[4], x[5], x[6], x[7]));
}

void launch_dequant(Dtype dt, const void *data, const void *scoff, void *out, long nsamples, hipStream_t st)
{
    const long nwords = nsamples / 8;
    dim3 grid((nwords + 255) / 256), block(256);
    const uint32_t *d = (const uint32_t *) data;
    const uint32_t *s = (const uint32_t *) scoff;
    switch (dt) {
        case Dtype::float32: hipLaunchKernelGGL(k_dequant<float>, grid, block, 0, st, d, s, (float *) out, nwords); break;
        case Dtype::float16: hipLaunchKernelGGL(k_dequant<__half>, grid, block, 0, st, d, s, (__half *) out, nwords); break;
        case Dtype::bfloat16:
            hipLaunchKernelGGL(k_dequant<__hip_bfloat16>, grid, block, 0, st, d, s, (__hip_bfloat16 *) out, nwords);
            break;
    }
}

}  // namespace pirate_amd
