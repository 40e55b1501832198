// Register-resident sub-trees shared by the tree kernels (gfx950).
//
// A rank-K sub-tree over 2^K rows holds one row per register, lane = time sample of a
// 64-sample tile. Pass r pairs rows e = fout*2^(r+1) + dmb and o = e + 2^r (dmb < 2^r):
//   e <- c*(x_e[t - L] + x_o[t]),  o <- c*(x_e[t - L - 1] + x_o[t]),  L = bitrev(dmb, r)
// (ReferenceTree.cpp:310-345), with lanes t < L reading the previous tile's value from a
// history register. Results are the direct tree's, operation for operation.

#pragma once

#include "device_common.hpp"

namespace pirate_amd {
namespace treg {

__host__ __device__ constexpr int brev_c(int x, int n)
{
    int r = 0;
    for (int i = 0; i < n; i++)
        if ((x >> i) & 1)
            r |= 1 << (n - 1 - i);
    return r;
}

// Rank-K sub-tree held in registers.
template<int K>
struct Sub {
    static constexpr int NR = 1 << K;                       // rows
    static constexpr int NP = (K > 0) ? NR / 2 : 0;         // pairs per pass
    static constexpr int NH = (K > 0) ? K * NP : 1;         // history registers (>= 1)
    static constexpr int NST = NP * (NR - 1);               // saved history floats
    // saved history of (pass r, pair q): the last 2^r samples
    static constexpr int hoff(int r, int q) { return NP * ((1 << r) - 1) + q * (1 << r); }
};

// Stream value at time t - L (L = 0..64): cur at lane t - L, or prev (the previous tile)
// at lane 64 + t - L when t < L. L is a compile-time constant after unrolling; L = 1 (about
// half of all shifts) is a whole-wave rotate by one lane in the VALU (DPP wave_ror:1, lane t
// reads lane t - 1 mod 64), the other lags a ds_bpermute through the LDS crossbar.
__device__ __forceinline__ float lane_shift(float cur, float prev, int L, int lane)
{
    const float w = (lane >= 64 - L) ? prev : cur;
    if (L == 1)
        return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(w), 0x13C, 0xf, 0xf, false));
    return __int_as_float(__builtin_amdgcn_ds_bpermute(((lane - L) & 63) << 2, __float_as_int(w)));
}

template<int K>
__device__ __forceinline__ void sub_passes(float (&x)[Sub<K>::NR], float (&h)[Sub<K>::NH], int lane)
{
    constexpr int NP = Sub<K>::NP;
#pragma unroll
    for (int r = 0; r < K; r++) {
#pragma unroll
        for (int q = 0; q < NP; q++) {
            const int fout = q >> r, dmb = q & ((1 << r) - 1);
            const int e = (fout << (r + 1)) + dmb, o = e + (1 << r);
            const int L = brev_c(dmb, r);
            const float xe = x[e], xo = x[o], hp = h[r * NP + q];
            const float a0 = (L == 0) ? xe : lane_shift(xe, hp, L, lane);
            const float a1 = lane_shift(xe, hp, L + 1, lane);
            h[r * NP + q] = xe;
            x[e] = kRsqrt2 * (a0 + xo);
            x[o] = kRsqrt2 * (a1 + xo);
        }
    }
}

template<int K>
__device__ __forceinline__ void load_hist(float (&h)[Sub<K>::NH], const float *src, int lane)
{
    constexpr int NP = Sub<K>::NP;
#pragma unroll
    for (int r = 0; r < K; r++) {
        const int n = 1 << r;
#pragma unroll
        for (int q = 0; q < NP; q++)
            h[r * NP + q] = (lane >= 64 - n) ? src[Sub<K>::hoff(r, q) + lane - (64 - n)] : 0.0f;
    }
    if constexpr (K == 0)
        h[0] = 0.0f;
}

// Saves the last 2^r samples of each pass history; 'hi' = valid lanes of the last tile
// (64, or 32 for a half tile).
template<int K>
__device__ __forceinline__ void save_hist(const float (&h)[Sub<K>::NH], float *dst, int lane, int hi)
{
    constexpr int NP = Sub<K>::NP;
#pragma unroll
    for (int r = 0; r < K; r++) {
        const int n = 1 << r;
#pragma unroll
        for (int q = 0; q < NP; q++)
            if (lane >= hi - n && lane < hi)
                dst[Sub<K>::hoff(r, q) + lane - (hi - n)] = h[r * NP + q];
    }
}

}  // namespace treg
}  // namespace pirate_amd
