// Tree gridding fused with lagged downsampling, on gfx950.
//
// Reference semantics:
//   gridding  out[n,t] = sum_{f ascending} w[n,f] * in[f,t]   (src_lib/TreeGriddingKernel.cu:62-102;
//             the sum starts from 0 and adds w*x with separate roundings)
//   lagged downsampling, p = 1..npri-1 (src_lib/ReferenceLaggedDownsamplingKernel.cpp:27-139),
//   from the level-(p-1) time-downsampled gridded rows:
//     ds_p[r][t]   = c*(ds_{p-1}[r][2t] + ds_{p-1}[r][2t+1])             (c = 1/sqrt 2)
//     out_p[j][t'] = c*(ds_p[2j][t-1] + ds_p[2j+1][t]),  t' = t + L_j     (one-sample lag on
//                    the even row, then the "large" lag L_j = 2^(s-1) - 1 - (j mod 2^(s-1)))
//
// Mapping: one workgroup of 4 waves per tree-channel pair (2j, 2j+1); wave w takes time
// tiles w, w+4, ... of 64 lanes x G samples (G = max(4, 2^(npri-1))), so a workgroup streams
// whole rows. Lane l owns G consecutive samples of both rows: it grids them from the input
// channels (16-byte loads), stores the gridded rows (the input of primary tree 0), and keeps
// them in registers for the downsampling levels, so the gridded chunk is never read back
// from HBM. The one-sample lag crosses lanes with a shuffle; lane 0 of each wave takes the
// previous tile's last value through LDS (the previous chunk's, from per-beam state, for the
// first tile). The large lag is a write offset: outputs past the chunk end go to per-beam
// state and are copied to the next chunk's first L_j samples.
// With 16-bit storage the gridded values are rounded to the storage type before the
// downsampling levels use them (as if stored and re-read); the levels themselves stay in
// float32 and round only at their output stores.
//
// Workgroups are mapped to channel ranges XCD by XCD (workgroup i runs on XCD i mod 8), so
// neighbouring channel pairs, which share boundary input channels, run on the same XCD and
// share its L2.

#include <type_traits>

#include "device_common.hpp"
#include "kernels.hpp"

namespace pirate_amd {

static constexpr int kGlWaves = 4;

long lds_state_nelts(int R, int npri, int s)
{
    if (npri <= 1)
        return 0;
    long nrows = 1L << (R - 1);
    long lcap = 1L << (s - 1);
    return (npri - 1) * nrows * (1 + lcap);
}

template<typename Tst>
__device__ __forceinline__ float quant(float x)
{
    if constexpr (std::is_same<Tst, float>::value)
        return x;
    else {
        Tst q;
        st_st(&q, x);
        return ld_st(&q);
    }
}

// Grids tree channels 2j ("e") and 2j+1 ("o") for G samples at 'in'. The two channels'
// input ranges [fe, fe+ce) and [fo, fo+co) are adjacent or overlapping: each channel of
// their union is loaded once, in batches of 4 whose loads are all issued before use, and
// added to each row whose range contains it, in ascending channel order.
template<typename Tst, int G>
__device__ __forceinline__ void grid_pair(const Tst *in, long ntime, int fe, int ce, const float *we,
                                          int fo, int co, const float *wo, float (&xe)[G], float (&xo)[G])
{
#pragma unroll
    for (int i = 0; i < G; i++)
        xe[i] = xo[i] = 0.0f;
    const int fa = min(fe, fo), fb = max(fe + ce, fo + co);
    for (int f1 = fa; f1 < fb; f1 += 4) {
        float4 v[4][G / 4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const Tst *p = in + long(min(f1 + u, fb - 1)) * ntime;
#pragma unroll
            for (int q = 0; q < G / 4; q++)
                v[u][q] = ld4_st(p + 4 * q);
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int f = f1 + u;
            if (f >= fb)
                break;
            if (f >= fe && f < fe + ce) {
                const float wk = we[f - fe];
#pragma unroll
                for (int q = 0; q < G / 4; q++) {
                    xe[4 * q] = xe[4 * q] + wk * v[u][q].x;
                    xe[4 * q + 1] = xe[4 * q + 1] + wk * v[u][q].y;
                    xe[4 * q + 2] = xe[4 * q + 2] + wk * v[u][q].z;
                    xe[4 * q + 3] = xe[4 * q + 3] + wk * v[u][q].w;
                }
            }
            if (f >= fo && f < fo + co) {
                const float wk = wo[f - fo];
#pragma unroll
                for (int q = 0; q < G / 4; q++) {
                    xo[4 * q] = xo[4 * q] + wk * v[u][q].x;
                    xo[4 * q + 1] = xo[4 * q + 1] + wk * v[u][q].y;
                    xo[4 * q + 2] = xo[4 * q + 2] + wk * v[u][q].z;
                    xo[4 * q + 3] = xo[4 * q + 3] + wk * v[u][q].w;
                }
            }
        }
    }
}

template<typename Tst, int NPRI>
__global__ void __launch_bounds__(kGlWaves * 64)
k_grid_lds(GridLdsArgs a)
{
    constexpr int G = ((1 << (NPRI - 1)) > 4) ? (1 << (NPRI - 1)) : 4;
    constexpr int NL = (NPRI > 1) ? NPRI : 2;          // array size for per-level values
    __shared__ float lastv[kGlWaves][NL];              // last even-row value of each wave's tile, per level
    __shared__ float carry[NL];                        // value before the current round's first tile

    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int nblk = gridDim.x;
    const int xb = blockIdx.x;
    const int j = (nblk % 8 == 0) ? ((xb % 8) * (nblk / 8) + xb / 8) : xb;    // channel pair
    const int b = blockIdx.y;
    const int nrows = a.nchan >> 1;

    const int fe = a.row_f0[2 * j], ce = a.row_cnt[2 * j];
    const int fo = a.row_f0[2 * j + 1], co = a.row_cnt[2 * j + 1];
    const float *we = a.w + a.row_woff[2 * j];
    const float *wo = a.w + a.row_woff[2 * j + 1];
    const Tst *in = reinterpret_cast<const Tst *>(a.in) + long(b) * a.in_bstride;
    Tst *gout_e = reinterpret_cast<Tst *>(a.grid) + long(b) * a.grid_bstride + long(2 * j) * a.ntime;
    Tst *gout_o = gout_e + a.ntime;

    // lagged-downsampling state (NPRI > 1): outputs that overflowed the previous chunk, and
    // the last even-row value of the previous chunk at every level
    const int lmask = (1 << (a.s - 1)) - 1;
    const int Lj = lmask - (j & lmask);
    const int lcap = lmask + 1;
    float *st = a.state + long(a.beam0 + b) * a.state_bstride;
    float *st_small = st;                                     // [(p-1)*nrows + j]
    float *st_large = st + long(NPRI - 1) * nrows;            // [((p-1)*nrows + j)*lcap + k]
    Tst *outp[NL];
#pragma unroll
    for (int p = 1; p < NPRI; p++) {
        const int Tp = a.ntime >> p;
        outp[p] = reinterpret_cast<Tst *>(a.out[p]) + long(b) * a.out_bstride[p] + long(j) * Tp;
        if (w == p % kGlWaves) {
            const float *sl = st_large + (long(p - 1) * nrows + j) * lcap;
            for (int k = lane; k < Lj; k += 64)
                st_st(outp[p] + k, sl[k]);
        }
        if (threadIdx.x == 0)
            carry[p] = st_small[long(p - 1) * nrows + j];
    }
    if constexpr (NPRI > 1)
        __syncthreads();     // state reads above precede the overflow writes of the last tile

    const int ngroups = a.ntime / G;
    const int ntiles = (ngroups + 63) / 64;
    for (int k0 = 0; k0 < ntiles; k0 += kGlWaves) {
        const int tile = k0 + w;
        const bool tile_on = tile < ntiles;
        const int g0 = tile * 64;
        const int g = g0 + lane;
        const int nact = tile_on ? min(64, ngroups - g0) : 0;
        const bool active = (lane < nact);
        const int t0 = g * G;

        float xe[G], xo[G];
        if (active) {
            grid_pair<Tst, G>(in + t0, a.ntime, fe, ce, we, fo, co, wo, xe, xo);
#pragma unroll
            for (int q = 0; q < G; q += 4) {
                st4_st(gout_e + t0 + q, make_float4(xe[q], xe[q + 1], xe[q + 2], xe[q + 3]));
                st4_st(gout_o + t0 + q, make_float4(xo[q], xo[q + 1], xo[q + 2], xo[q + 3]));
            }
        } else {
#pragma unroll
            for (int i = 0; i < G; i++)
                xe[i] = xo[i] = 0.0f;
        }

        if constexpr (NPRI > 1) {
#pragma unroll
            for (int i = 0; i < G; i++) {
                xe[i] = quant<Tst>(xe[i]);
                xo[i] = quant<Tst>(xo[i]);
            }
            float xo0[NL];     // lane 0: odd-row value of its first output, per level (needs the carry)
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
                if (tile_on && lane == nact - 1)
                    lastv[w][p] = xe[n - 1];
                const float prev = __shfl_up(xe[n - 1], 1);
                xo0[p] = xo[0];

                const int Tp = a.ntime >> p;
                const int tp = g * n;
                float *sl = st_large + (long(p - 1) * nrows + j) * lcap;
                const bool all_in = (g0 + nact) * n + Lj <= Tp;    // wave-uniform
                if (active) {
#pragma unroll
                    for (int k = 0; k < (G >> 1); k++) {
                        if (k < n && (k > 0 || lane > 0)) {
                            const float se = (k == 0) ? prev : xe[k - 1];
                            const float o = kRsqrt2 * (se + xo[k]);
                            const int pos = tp + k + Lj;
                            if (all_in || pos < Tp)
                                st_st(outp[p] + pos, o);
                            else
                                sl[pos - Tp] = o;
                        }
                    }
                }
            }
            __syncthreads();     // lastv of every wave of this round
            if (tile_on && lane == 0) {
#pragma unroll
                for (int p = 1; p < NPRI; p++) {
                    const int n = G >> p;
                    const int Tp = a.ntime >> p;
                    const float prev = (w == 0) ? carry[p] : lastv[w - 1][p];
                    const float o = kRsqrt2 * (prev + xo0[p]);
                    const int pos = g0 * n + Lj;
                    if (pos < Tp)
                        st_st(outp[p] + pos, o);
                    else
                        st_large[(long(p - 1) * nrows + j) * lcap + pos - Tp] = o;
                }
            }
            __syncthreads();     // every lane 0 has read carry / lastv
            if (threadIdx.x == 0) {
                const int wl = min(kGlWaves, ntiles - k0) - 1;     // last wave with a tile this round
#pragma unroll
                for (int p = 1; p < NPRI; p++)
                    carry[p] = lastv[wl][p];
            }
            __syncthreads();
        }
    }

    if constexpr (NPRI > 1) {
        if (threadIdx.x == 0) {
#pragma unroll
            for (int p = 1; p < NPRI; p++)
                st_small[long(p - 1) * nrows + j] = carry[p];
        }
    }
}

template<typename Tst>
static void launch_grid_lds_t(const GridLdsArgs &a, hipStream_t st)
{
    dim3 grid(a.nchan >> 1, a.B);      // one workgroup per tree-channel pair
    dim3 block(kGlWaves * 64);
    switch (a.npri) {
        case 1: hipLaunchKernelGGL((k_grid_lds<Tst, 1>), grid, block, 0, st, a); break;
        case 2: hipLaunchKernelGGL((k_grid_lds<Tst, 2>), grid, block, 0, st, a); break;
        case 3: hipLaunchKernelGGL((k_grid_lds<Tst, 3>), grid, block, 0, st, a); break;
        case 4: hipLaunchKernelGGL((k_grid_lds<Tst, 4>), grid, block, 0, st, a); break;
        case 5: hipLaunchKernelGGL((k_grid_lds<Tst, 5>), grid, block, 0, st, a); break;
        case 6: hipLaunchKernelGGL((k_grid_lds<Tst, 6>), grid, block, 0, st, a); break;
        case 7: hipLaunchKernelGGL((k_grid_lds<Tst, 7>), grid, block, 0, st, a); break;
        default: break;
    }
}

void launch_grid_lds(Dtype dt, const GridLdsArgs &a, hipStream_t st)
{
    switch (dt) {
        case Dtype::float32: launch_grid_lds_t<float>(a, st); break;
        case Dtype::float16: launch_grid_lds_t<__half>(a, st); break;
        case Dtype::bfloat16: launch_grid_lds_t<__hip_bfloat16>(a, st); break;
    }
}

}  // namespace pirate_amd
