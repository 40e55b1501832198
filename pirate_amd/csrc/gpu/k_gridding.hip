// Tree gridding on gfx950: rebin input frequency channels into 2^R tree channels that
// are uniform in freq^-2, out[b,n,t] = sum_{f ascending} w[n,f] * in[b,f,t].
// Reference semantics: src_lib/TreeGriddingKernel.cu:62-102 (ReferenceTreeGriddingKernel).
//
// HBM-bound. Each thread owns 4 consecutive time samples of NPB consecutive tree channels;
// a block covers 1024 samples. The input channels of adjacent tree channels overlap at
// the boundaries, so the NPB-channel loop re-reads them from L1/L2, not HBM. The sum
// starts from 0.0f and adds w*x in ascending f with separate roundings (no FMA), exactly
// as the CPU reference does.

#include "device_common.hpp"
#include "kernels.hpp"

namespace pirate_amd {

static constexpr int kGridThreads = 256;
static constexpr int kGridNPB = 8;

template<typename Tst>
struct Vec4 {
    __device__ static void load(const Tst *p, float *x)
    {
        const float4 v = ld4_st(p);
        x[0] = v.x; x[1] = v.y; x[2] = v.z; x[3] = v.w;
    }
    __device__ static void store(Tst *p, const float *x) { st4_st(p, make_float4(x[0], x[1], x[2], x[3])); }
};

template<typename Tst>
__global__ void __launch_bounds__(kGridThreads)
k_gridding(GriddingArgs a)
{
    const int t = (blockIdx.x * kGridThreads + threadIdx.x) * 4;
    const int b = blockIdx.z;
    if (t >= a.ntime)
        return;

    const Tst *in = reinterpret_cast<const Tst *>(a.in) + b * a.in_bstride + t;
    Tst *out = reinterpret_cast<Tst *>(a.out) + b * a.out_bstride + t;

    const int n0 = blockIdx.y * kGridNPB;
    const int n1 = min(n0 + kGridNPB, a.nchan);

    for (int n = n0; n < n1; n++) {
        const int f0 = a.row_f0[n];
        const int cnt = a.row_cnt[n];
        const float *w = a.w + a.row_woff[n];
        float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        for (int k = 0; k < cnt; k++) {
            float x[4];
            Vec4<Tst>::load(in + long(f0 + k) * a.ntime, x);
            const float wk = w[k];
            for (int i = 0; i < 4; i++)
                acc[i] = acc[i] + wk * x[i];
        }
        Vec4<Tst>::store(out + long(n) * a.ntime, acc);
    }
}

void launch_gridding(Dtype dt, const GriddingArgs &a, hipStream_t s)
{
    dim3 grid((a.ntime / 4 + kGridThreads - 1) / kGridThreads, (a.nchan + kGridNPB - 1) / kGridNPB, a.B);
    dim3 block(kGridThreads);
    switch (dt) {
        case Dtype::float32: hipLaunchKernelGGL(k_gridding<float>, grid, block, 0, s, a); break;
        case Dtype::float16: hipLaunchKernelGGL(k_gridding<__half>, grid, block, 0, s, a); break;
        case Dtype::bfloat16: hipLaunchKernelGGL(k_gridding<__hip_bfloat16>, grid, block, 0, s, a); break;
    }
}

}  // namespace pirate_amd
