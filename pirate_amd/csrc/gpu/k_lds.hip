// Lagged downsampling on gfx950: builds the half-rank inputs of primary trees p >= 1.
// Reference semantics: src_lib/ReferenceLaggedDownsamplingKernel.cpp:27-139, i.e. for each
// level p = 1..npri-1, starting from the level-(p-1) time-downsampled tree-gridded rows:
//   ds_p[r][t]   = c*(ds_{p-1}[r][2t] + ds_{p-1}[r][2t+1])                  (c = 1/sqrt 2)
//   out_p[j][t'] = c*(ds_p[2j][t-1] + ds_p[2j+1][t]),  t' = t + L_j          (one-sample lag on
//                                                                            even rows, then a
//                                                                            "large" lag L_j)
//   L_j = 2^(s-1) - 1 - (j mod 2^(s-1)),   s = stage-1 dd_rank(p>=1) + 1.
//
// Mapping: one wave64 per output row j (row pair 2j, 2j+1). Lane l owns G = 2^(npri-1)
// consecutive level-0 samples, so every level's samples of a lane stay in registers; the
// one-sample lag crosses lanes with a shuffle (lane 0 takes the previous tile's lane 63,
// or the per-beam state at the chunk start). The large lag is applied as a write offset:
// outputs that land past the chunk end go to the per-beam state and are copied into the
// next chunk's first L_j samples. No LDS is needed.

#include "device_common.hpp"
#include "kernels.hpp"

namespace pirate_amd {

long lds_state_nelts(int R, int npri, int s)
{
    if (npri <= 1)
        return 0;
    long nrows = 1L << (R - 1);
    long lcap = 1L << (s - 1);
    return (npri - 1) * nrows * (1 + lcap);
}

template<typename Tst, int NPRI>
__global__ void __launch_bounds__(256)
k_lds(LdsArgs a)
{
    constexpr int G = 1 << (NPRI - 1);
    const int lane = threadIdx.x & 63;
    const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int b = blockIdx.y;
    const int nrows = 1 << (a.R - 1);
    if (j >= nrows)
        return;

    const int lmask = (1 << (a.s - 1)) - 1;
    const int Lj = lmask - (j & lmask);
    const int lcap = lmask + 1;

    float *st = a.state + long(a.beam0 + b) * a.state_bstride;
    float *st_small = st;                                     // [(p-1)*nrows + j]
    float *st_large = st + long(NPRI - 1) * nrows;            // [((p-1)*nrows + j)*lcap + k]

    const Tst *in_e = reinterpret_cast<const Tst *>(a.in) + long(b) * a.in_bstride + long(2 * j) * a.ntime;
    const Tst *in_o = in_e + a.ntime;

    Tst *outp[NPRI];
    float carry[NPRI];
#pragma unroll
    for (int p = 1; p < NPRI; p++) {
        const int Tp = a.ntime >> p;
        outp[p] = reinterpret_cast<Tst *>(a.out[p]) + long(b) * a.out_bstride[p] + long(j) * Tp;
        const float *sl = st_large + (long(p - 1) * nrows + j) * lcap;
        for (int k = lane; k < Lj; k += 64)
            st_st(outp[p] + k, sl[k]);
        carry[p] = st_small[long(p - 1) * nrows + j];
    }

    const int ngroups = a.ntime / G;               // level-0 sample groups per row
    for (int g0 = 0; g0 < ngroups; g0 += 64) {
        const int g = g0 + lane;
        const int nact = min(64, ngroups - g0);
        const bool active = (lane < nact);
        const long t0 = long(g) * G;

        float xe[G], xo[G];
#pragma unroll
        for (int k = 0; k < G; k++) {
            xe[k] = active ? ld_st(in_e + t0 + k) : 0.0f;
            xo[k] = active ? ld_st(in_o + t0 + k) : 0.0f;
        }

#pragma unroll
        for (int p = 1; p < NPRI; p++) {
            const int n = G >> p;
#pragma unroll
            for (int k = 0; k < (G >> 1); k++) {
                if (k < n) {
                    xe[k] = kRsqrt2 * (xe[2 * k] + xe[2 * k + 1]);
                    xo[k] = kRsqrt2 * (xo[2 * k] + xo[2 * k + 1]);
                }
            }
            // One-sample lag of the even row: previous level-p sample is the previous
            // lane's last one (lane 0: previous tile, or the state at the chunk start).
            float prev = __shfl_up(xe[n - 1], 1);
            if (lane == 0)
                prev = carry[p];
            carry[p] = __shfl(xe[n - 1], nact - 1);

            const int Tp = a.ntime >> p;
            const long tp = t0 >> p;
            float *sl = st_large + (long(p - 1) * nrows + j) * lcap;
            if (active) {
#pragma unroll
                for (int k = 0; k < (G >> 1); k++) {
                    if (k < n) {
                        const float se = (k == 0) ? prev : xe[k - 1];
                        const float o = kRsqrt2 * (se + xo[k]);
                        const long pos = tp + k + Lj;
                        if (pos < Tp)
                            st_st(outp[p] + pos, o);
                        else
                            sl[pos - Tp] = o;
                    }
                }
            }
        }
    }

    if (lane == 0) {
#pragma unroll
        for (int p = 1; p < NPRI; p++)
            st_small[long(p - 1) * nrows + j] = carry[p];
    }
}

template<typename Tst>
static void launch_lds_t(const LdsArgs &a, hipStream_t st)
{
    const int nrows = 1 << (a.R - 1);
    dim3 grid((nrows + 3) / 4, a.B);
    dim3 block(256);
    switch (a.npri) {
        case 2: hipLaunchKernelGGL((k_lds<Tst, 2>), grid, block, 0, st, a); break;
        case 3: hipLaunchKernelGGL((k_lds<Tst, 3>), grid, block, 0, st, a); break;
        case 4: hipLaunchKernelGGL((k_lds<Tst, 4>), grid, block, 0, st, a); break;
        case 5: hipLaunchKernelGGL((k_lds<Tst, 5>), grid, block, 0, st, a); break;
        case 6: hipLaunchKernelGGL((k_lds<Tst, 6>), grid, block, 0, st, a); break;
        case 7: hipLaunchKernelGGL((k_lds<Tst, 7>), grid, block, 0, st, a); break;
        default: break;
    }
}

void launch_lds(Dtype dt, const LdsArgs &a, hipStream_t st)
{
    if (a.npri <= 1)
        return;
    switch (dt) {
        case Dtype::float32: launch_lds_t<float>(a, st); break;
        case Dtype::float16: launch_lds_t<__half>(a, st); break;
        case Dtype::bfloat16: launch_lds_t<__hip_bfloat16>(a, st); break;
    }
}

}  // namespace pirate_amd
