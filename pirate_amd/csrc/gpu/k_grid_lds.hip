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
// Mapping: one wave per (tree-channel pair (2j, 2j+1), time tile of 64 x G samples),
// G = max(16, 2^(npri-1)); a workgroup holds 4 consecutive
// (pair, tile) waves in pair-major order (4 tiles of one pair, or 2 whole pairs when a pair
// has 2 tiles). The wave grids its tile in a coalesced layout (lane l: 16 bytes of
// consecutive samples in each block of 64 x 16 bytes, so every load and store instruction
// covers 1 KB of a row), stores the gridded rows (the input of primary tree 0), and
// transposes them through LDS (a per-wave buffer) to G consecutive samples per lane for the
// downsampling levels, so the gridded chunk is never read back from HBM. The one-sample lag
// crosses lanes with a shuffle; lane 0 takes the even row's value just before the tile from
// the previous wave of the workgroup when that wave holds the previous tile (one barrier per
// workgroup), else recomputes it from the input (for the first tile: the previous chunk's
// last value, from per-beam state). The large lag is a write offset: outputs past the chunk end go to per-beam state
// and are copied to the next chunk's first L_j samples. The state is double-buffered by
// chunk parity, so the first tile's reads and the last tile's writes of one launch never
// touch the same memory.
// With 16-bit storage the gridded values are rounded to the storage type before the
// downsampling levels use them (as if stored and re-read); the levels themselves stay in
// float32 and round only at their output stores.
//
// Workgroup-to-pair mapping: workgroups are dispatched to the 8 XCDs in turn; each XCD takes
// runs of a.xcd_run (default 8) consecutive workgroup slots (8 channel pairs when one
// workgroup covers a pair's whole chunk) and the 8 XCDs' runs are adjacent, so the device
// streams one contiguous region of 64 pairs at a time while neighbouring pairs, which share
// boundary input channels, share an XCD's L2. Measured on MI355X (chord16k): launch order
// (pair = workgroup index) 1.27 ms, runs of 8 pairs 1.18 ms, and each XCD a contiguous
// eighth of the channels 1.83 ms (the device then streams eight distant regions at once).

#include <algorithm>
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
    return 2 * (npri - 1) * nrows * (1 + lcap);     // two parities
}

// Samples per lane in the downsampling layout: at least 16, so that each wave's fixed work
// (the pair's set-up, the transposes, the hand-over and the downsampling levels) is spread
// over 64 or 32 bytes of each row per lane (16 beams x 16384 channels with 4 primary trees:
// float16 0.69 -> 0.63 ms and float32 1.07 -> 1.03 ms against 8 samples per lane). The chunk
// length is a multiple of 32 samples, so 16 divides it.
static constexpr int kGlMinGroup = 16;

// Grids tree channels 2j ("e") and 2j+1 ("o") for G samples, G/V groups of V consecutive
// samples at in + QS*q (q < nq; groups q >= nq are outside the chunk and left at zero). The
// two channels' input ranges [fe, fe+ce) and [fo, fo+co) are adjacent or overlapping: each
// channel of their union is loaded once, in batches of NB whose loads are all issued before
// use, and added to each row whose range contains it, in ascending channel order. NB is 2
// for G >= 16 (88 instead of 110 VGPRs for float16 with 5 primary trees, so 5 waves per SIMD
// instead of 4: 1.99 -> 1.64 ms for 32 beams x 16384 channels), else 4.
template<typename Tst, int G, int V, int QS>
__device__ __forceinline__ void grid_pair(const Tst *in, long ntime, int nq, int fe, int ce, const float *we,
                                          int fo, int co, const float *wo, float (&xe)[G], float (&xo)[G])
{
    constexpr int NB = (G >= 16) ? 2 : 4;
#pragma unroll
    for (int i = 0; i < G; i++)
        xe[i] = xo[i] = 0.0f;
    const int fa = min(fe, fo), fb = max(fe + ce, fo + co);
    for (int f1 = fa; f1 < fb; f1 += NB) {
        // as loaded, converted where used (16-bit types: each load keeps its own registers,
        // so a batch's loads are all in flight at once)
        rawv_st<Tst, V> v[NB][G / V];
#pragma unroll
        for (int u = 0; u < NB; u++) {
            const Tst *p = in + long(min(f1 + u, fb - 1)) * ntime;
#pragma unroll
            for (int q = 0; q < G / V; q++)
                v[u][q] = (q < nq) ? ldv_raw<Tst, V>(p + QS * q) : rawv_st<Tst, V>{};
        }
#pragma unroll
        for (int u = 0; u < NB; u++) {
            const int f = f1 + u;
            if (f >= fb)
                break;
            float x[G];
#pragma unroll
            for (int q = 0; q < G / V; q++)
                from_rawv<Tst, V>(v[u][q], &x[V * q]);
            if (f >= fe && f < fe + ce) {
                const float wk = we[f - fe];
#pragma unroll
                for (int i = 0; i < G; i++)
                    xe[i] = xe[i] + wk * x[i];
            }
            if (f >= fo && f < fo + co) {
                const float wk = wo[f - fo];
#pragma unroll
                for (int i = 0; i < G; i++)
                    xo[i] = xo[i] + wk * x[i];
            }
        }
    }
}

template<typename Tst, int NPRI>
__global__ void __launch_bounds__(kGlWaves * 64)
k_grid_lds(GridLdsArgs a)
{
    constexpr int G = std::max(kGlMinGroup, 1 << (NPRI - 1));
    constexpr int NL = (NPRI > 1) ? NPRI : 2;          // array size for per-level values

    const int lane = threadIdx.x & 63;
    const int ngroups = a.ntime / G;
    const int ntiles = (ngroups + 63) / 64;
    int blk = blockIdx.x;
    const int run = a.xcd_run;
    if (run > 0 && gridDim.x % (8 * run) == 0) {
        const int x = blk & 7, slot = blk >> 3;                 // workgroup blk runs on XCD blk mod 8
        blk = (slot / run) * (8 * run) + x * run + (slot % run);
    }
    // Waves are numbered pair-major over (pair, tile), so a workgroup holds 4 consecutive
    // tiles of one pair, or several whole pairs when a pair has fewer than 4 tiles. The wave
    // index is made wave-uniform (scalar), so that the pair's channel ranges, weights and
    // loop bounds stay in scalar registers.
    const int wv = blk * kGlWaves + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int j = wv / ntiles;                                     // channel pair
    const int tile = wv - j * ntiles;
    const int b = blockIdx.y;
    const int nrows = a.nchan >> 1;
    if (j >= nrows)
        return;

    const int fe = a.row_f0[2 * j], ce = a.row_cnt[2 * j];
    const int fo = a.row_f0[2 * j + 1], co = a.row_cnt[2 * j + 1];
    const float *we = a.w + a.row_woff[2 * j];
    const float *wo = a.w + a.row_woff[2 * j + 1];
    const Tst *in = reinterpret_cast<const Tst *>(a.in) + long(b) * a.in_bstride;
    Tst *gout_e = reinterpret_cast<Tst *>(a.grid) + long(b) * a.grid_bstride + long(2 * j) * a.ntime;
    Tst *gout_o = gout_e + a.ntime;

    const int g0 = tile * 64;
    const int g = g0 + lane;
    const int nact = min(64, ngroups - g0);
    const int t0 = g * G;

    // Gridding in the coalesced layout: lane l holds samples tb0 + 64 V q + V l + (0..V-1),
    // q < G/V, V = 16 bytes of samples (4 float32 or 8 16-bit values; V divides G), so that
    // each load and store instruction of the wave covers 64 V consecutive samples of a row.
    // Groups past the chunk end (a partial last tile) are skipped.
    constexpr int V = 16 / int(sizeof(Tst));
    constexpr int BLK = 64 * V;
    const int tb0 = g0 * G;
    const int nq = min(G / V, max(0, (a.ntime - tb0 - V * lane + BLK - 1) / BLK));
    float xe[G], xo[G];
    grid_pair<Tst, G, V, BLK>(in + tb0 + V * lane, a.ntime, nq, fe, ce, we, fo, co, wo, xe, xo);
#pragma unroll
    for (int q = 0; q < G / V; q++) {
        if (q < nq) {
            stv_st<Tst, V>(gout_e + tb0 + BLK * q + V * lane, &xe[V * q]);
            stv_st<Tst, V>(gout_o + tb0 + BLK * q + V * lane, &xo[V * q]);
        }
    }

    if constexpr (NPRI > 1) {
        // Per-beam state, double-buffered by chunk parity so that no wave reads what another
        // wave of the same launch writes: [parity][small (p-1, j) | large ((p-1, j), lcap)].
        const int lmask = (1 << (a.s - 1)) - 1;
        const int Lj = lmask - (j & lmask);
        const int lcap = lmask + 1;
        const long half = long(NPRI - 1) * nrows * (1 + lcap);
        float *st = a.state + long(a.beam0 + b) * a.state_bstride;
        const long chunk = launch_chunk(a);
        const float *rd = st + (chunk & 1) * half;              // written by the previous chunk
        float *wr = st + ((chunk + 1) & 1) * half;              // read by the next chunk
        const float *rd_small = rd, *rd_large = rd + long(NPRI - 1) * nrows;
        float *wr_small = wr, *wr_large = wr + long(NPRI - 1) * nrows;

        Tst *outp[NL];
#pragma unroll
        for (int p = 1; p < NPRI; p++)
            outp[p] = reinterpret_cast<Tst *>(a.out[p]) + long(b) * a.out_bstride[p] + long(j) * (a.ntime >> p);

        // Lane 0: the even row's level-p value just before this tile. For the first tile it
        // is the previous chunk's last value (per-beam state). When the previous tile is the
        // previous wave of this workgroup (every wave but the first, in a workgroup whose 4
        // waves all have work), that wave's last lane computes it from its own samples and
        // hands it over through LDS, after one workgroup barrier (below). Otherwise lane 0
        // recomputes it from the previous group of input samples.
        const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
        const bool full_wg = (long(blk) + 1) * kGlWaves <= long(nrows) * ntiles;
        const bool from_prev_wave = full_wg && tile > 0 && w > 0;
        float prevs[NL];
        if (lane == 0) {
            if (tile == 0) {
#pragma unroll
                for (int p = 1; p < NPRI; p++)
                    prevs[p] = rd_small[long(p - 1) * nrows + j];
            } else if (!from_prev_wave) {
                float xp[G], xq[G];
                grid_pair<Tst, G, 4, 4>(in + t0 - G, a.ntime, G / 4, fe, ce, we, fe, 0, we, xp, xq);
#pragma unroll
                for (int i = 0; i < G; i++)
                    xp[i] = round_st<Tst>(xp[i]);
#pragma unroll
                for (int p = 1; p < NPRI; p++) {
                    const int n = G >> p;
#pragma unroll
                    for (int k = 0; k < (G >> 1); k++)
                        if (k < n)
                            xp[k] = kRsqrt2 * (xp[2 * k] + xp[2 * k + 1]);
                    prevs[p] = xp[n - 1];
                }
            }
        }
        // The first tile also lays down the outputs that overflowed the previous chunk.
        if (tile == 0) {
#pragma unroll
            for (int p = 1; p < NPRI; p++) {
                const float *sl = rd_large + (long(p - 1) * nrows + j) * lcap;
                for (int k = lane; k < Lj; k += 64)
                    st_st_cached(outp[p] + k, sl[k]);
            }
        }

        // Per-wave LDS transpose buffer. First the gridded rows go from the coalesced layout
        // to G consecutive samples per lane (lane l: samples tb0 + G l + i), the layout of the
        // downsampling levels; then each level's outputs go back so that each global store
        // instruction writes 64 consecutive samples.
        __shared__ float tbuf[kGlWaves][64 * G];
        float *tb = tbuf[threadIdx.x >> 6];
#pragma unroll
        for (int r = 0; r < 2 && V != G; r++) {
            float (&x)[G] = r ? xo : xe;
#pragma unroll
            for (int q = 0; q < G / V; q++)
#pragma unroll
                for (int c = 0; c < V; c += 4)
                    *reinterpret_cast<float4 *>(tb + BLK * q + V * lane + c) =
                        make_float4(x[V * q + c], x[V * q + c + 1], x[V * q + c + 2], x[V * q + c + 3]);
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int i = 0; i < G; i += 4) {
                const float4 y = *reinterpret_cast<const float4 *>(tb + G * lane + i);
                x[i] = y.x;
                x[i + 1] = y.y;
                x[i + 2] = y.z;
                x[i + 3] = y.w;
            }
            __builtin_amdgcn_wave_barrier();
        }
#pragma unroll
        for (int i = 0; i < G; i++) {
            xe[i] = round_st<Tst>(xe[i]);
            xo[i] = round_st<Tst>(xo[i]);
        }
        // Hand-over of the even row's last value per level to the next wave (the tile's last
        // lane holds its last G samples, full tiles only: a tile followed by another is full).
        // The full_wg condition is uniform over the workgroup, so every wave reaches the barrier.
        __shared__ float tail[kGlWaves][NL];
        if (full_wg) {
            if (lane == 63) {
                float xp[G];
#pragma unroll
                for (int i = 0; i < G; i++)
                    xp[i] = xe[i];
#pragma unroll
                for (int p = 1; p < NPRI; p++) {
                    const int n = G >> p;
#pragma unroll
                    for (int k = 0; k < (G >> 1); k++)
                        if (k < n)
                            xp[k] = kRsqrt2 * (xp[2 * k] + xp[2 * k + 1]);
                    tail[w][p] = xp[n - 1];
                }
            }
            __syncthreads();
            if (from_prev_wave && lane == 0) {
#pragma unroll
                for (int p = 1; p < NPRI; p++)
                    prevs[p] = tail[w - 1][p];
            }
        }
        const bool last_tile = (tile == ntiles - 1);
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
            float prev = __shfl_up(xe[n - 1], 1);
            if (lane == 0)
                prev = prevs[p];
            if (last_tile && lane == nact - 1)
                wr_small[long(p - 1) * nrows + j] = xe[n - 1];

#pragma unroll
            for (int k = 0; k < (G >> 1); k++)
                if (k < n)
                    tb[lane * n + k] = kRsqrt2 * (((k == 0) ? prev : xe[k - 1]) + xo[k]);
            __builtin_amdgcn_wave_barrier();

            const int Tp = a.ntime >> p;
            const int base = g0 * n + Lj;                // position of this tile's first output
            const int nout = nact * n;
            float *sl = wr_large + (long(p - 1) * nrows + j) * lcap;
#pragma unroll
            for (int q = 0; q < (G >> 1); q++) {
                if (q < n) {
                    const int i = 64 * q + lane;
                    const float o = tb[i];
                    const int pos = base + i;
                    if (i < nout) {
                        if (pos < Tp)
                            st_st_cached(outp[p] + pos, o);
                        else
                            sl[pos - Tp] = o;
                    }
                }
            }
            __builtin_amdgcn_wave_barrier();     // the buffer is reused by the next level
        }
    }
}

template<typename Tst>
static void launch_grid_lds_t(const GridLdsArgs &a, hipStream_t st)
{
    const int G = std::max(kGlMinGroup, 1 << (a.npri - 1));
    const long nwaves = long(a.nchan >> 1) * ((a.ntime / G + 63) / 64);
    dim3 grid(unsigned((nwaves + kGlWaves - 1) / kGlWaves), a.B);   // 4 independent (pair, tile) waves each
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
