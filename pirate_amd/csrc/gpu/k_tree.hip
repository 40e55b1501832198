// Tree dedispersion on gfx950 (stage 1 and stage 2 of the two-stage tree).
//
// Reference semantics: ReferenceTree (src_lib/ReferenceTree.cpp:120-345) applied as in
// ReferenceDedisperser sophistication 1 (src_lib/ReferenceDedisperser.cpp:388-450):
//   pass r pairs rows e = fout*2^(r+1) + dmb and o = e + 2^r (in-place, bit-reversed DM):
//     e <- c*(x_e[t - L] + x_o[t]),  o <- c*(x_e[t - L - 1] + x_o[t]),  L = bitrev(dmb, r),
//   with x_e at negative chunk time taken from the per-pass history of the previous chunk.
//
// Layout: one workgroup per tree instance (beam, ambient index) sweeps the chunk in time
// tiles of TT = 64 samples (one wave64 row segment). The tile [D][TT] and the per-pass
// histories (D/2 pairs x 2^r samples at pass r) live in LDS; passes alternate a read
// phase (all lags resolved from tile or history into registers) and a write phase.
// Histories persist across chunks in an HBM state buffer (loaded/saved once per chunk).
// A chunk whose length is a multiple of 32 but not of 64 ends in a half tile (nv = 32
// valid samples): its emits stop at nv, and each pass history is refilled with the last
// 2^r valid samples, from the tile and, where 2^r > nv, from the older history.
//
// Stage 1 writes its output rows to the lag ring (per-row circular buffers whose
// capacity covers the row's stage-2 lag). Stage 2 reads ring rows at their full lag
// (rb_lag, no segment/residual split), runs the stage-2 passes, and emits the subband
// streams (ReferenceTree subband outputs, incl. the final lags) to the peak-finder input.

#include "device_common.hpp"
#include "kernels.hpp"
#include "tree_regs.hpp"

#include <type_traits>

namespace pirate_amd {

using namespace treg;

static constexpr int kTT = 64;
static constexpr int kTreeThreads = 512;   // 8 waves: 35% faster than 4 on the rank-7 subband kernel

template<int DD>
struct TreeDims {
    static constexpr int D = 1 << DD;
    static constexpr int NP = D / 2;
    static constexpr int HIST = NP * (D - 1);            // sum_r NP * 2^r
    static constexpr int WORK = NP * kTT;                 // pair-samples per pass
    static constexpr int ITEMS = (WORK + kTreeThreads - 1) / kTreeThreads;
};

__device__ __forceinline__ int hist_off(int np, int r) { return np * ((1 << r) - 1); }


// One in-place pass. If emit != nullptr, the outputs of main pairs with emit[p] >= 0
// are also sent to streams emit[p] (out0) and emit[p]+1 (out1) through 'sink'.
template<int DD, typename Sink>
__device__ __forceinline__ void tree_pass(int r, float *tile, float *hist, const int *emit, Sink &sink, int nv)
{
    using TD = TreeDims<DD>;
    const int tid = threadIdx.x;
    const int half = 1 << r;
    float *h = hist + hist_off(TD::NP, r);

    float o0[TD::ITEMS], o1[TD::ITEMS], sv[TD::ITEMS];
#pragma unroll
    for (int i = 0; i < TD::ITEMS; i++) {
        const int idx = i * kTreeThreads + tid;
        if (idx < TD::WORK) {
            const int p = idx / kTT, t = idx % kTT;
            const int dmb = p & (half - 1);
            const int e = ((p >> r) << (r + 1)) + dmb;
            const int o = e + half;
            const int L = bitrev_dev(dmb, r);
            const float xo = tile[o * kTT + t];
            const float xa = (t >= L) ? tile[e * kTT + t - L] : h[p * half + half + t - L];
            const float xb = (t >= L + 1) ? tile[e * kTT + t - L - 1] : h[p * half + half + t - L - 1];
            o0[i] = kRsqrt2 * (xa + xo);
            o1[i] = kRsqrt2 * (xb + xo);
            // new history slot t - (kTT - half) holds time nv - half + (that slot)
            const int ts = t - (kTT - nv);
            sv[i] = (t < kTT - half) ? 0.0f : (ts >= 0) ? tile[e * kTT + ts] : h[p * half + half + ts];
        }
    }
    sink.read_phase_extra(r);
    __syncthreads();
#pragma unroll
    for (int i = 0; i < TD::ITEMS; i++) {
        const int idx = i * kTreeThreads + tid;
        if (idx < TD::WORK) {
            // p is uniform over a wave (kTT = 64 samples per pair): scalar, so that the
            // emission-table and final-lag reads are scalar loads
            const int p = __builtin_amdgcn_readfirstlane(idx / kTT), t = idx % kTT;
            const int dmb = p & (half - 1);
            const int e = ((p >> r) << (r + 1)) + dmb;
            const int o = e + half;
            tile[e * kTT + t] = o0[i];
            tile[o * kTT + t] = o1[i];
            if (t >= kTT - half)
                h[p * half + (t - (kTT - half))] = sv[i];
            if (emit) {
                const int s0 = ld_table(emit, r * TD::NP + p);
                if (s0 >= 0) {
                    sink.emit(s0, t, o0[i]);
                    sink.emit(s0 + 1, t, o1[i]);
                }
            }
        }
    }
    sink.write_phase_extra(r);
    __syncthreads();
}


// ---------------------------------------------------------------------------------------
// Stage 2


long stage2_state_nelts_per_tree(int dd, int stag_hist_nelts)
{
    return (1L << (dd - 1)) * ((1L << dd) - 1) + stag_hist_nelts;
}

int stage2_lds_bytes(int dd, int stag_hist_nelts)
{
    long d = 1L << dd;
    return int(sizeof(float) * (d * kTT + (d / 2) * (d - 1) + stag_hist_nelts + 3) + 16 * d);   // + row descriptors
}

// Register pre-passes of k_stage2: passes 0..KP-1 (KP <= 4, all before the first subband
// emission) pair rows inside groups of 2^KP consecutive rows, so each group is a rank-KP
// sub-tree that one wave runs in registers (tree_regs.hpp) straight from the ring loads,
// before the tile goes to LDS for the remaining passes. Groups per wave:
template<int DD, int KP>
struct PrePass {
    static constexpr int NG = (1 << DD) >> KP;                          // groups
    static constexpr int NWV = kTreeThreads / 64;
    static constexpr int GPW = (NG + NWV - 1) / NWV;                    // groups per wave
};

// Emits subband stream samples (and runs the staggered butterflies and level-0 copies)
// into the peak-finder input rows. Holds copies of the fields it needs (not a reference
// to the kernel argument, which would force the argument block into scratch memory).
template<typename Tst, int DD>
struct Stage2Sink {
    Tst *sbb;            // sb + beam offset
    const int *stag, *fl;
    int M, rowlen, tpad;
    int dpf_base, t0, nv;     // nv: valid samples of the current tile
    const int *stag_begin;
    float *tile, *shist;

    // Staggered butterflies of pass r, run before the pass (they read its input rows, which
    // the pass overwrites only in its write phase) in rounds of SMAX entries per thread:
    // each round reads, waits for every thread, then emits and updates the histories.
    static constexpr int SMAX = 4;

    __device__ __forceinline__ void emit(int s, int t, float v)
    {
        if (t >= nv)
            return;
        const int pf_dm = s / M, m = s - pf_dm * M;
        const long row = long(dpf_base + pf_dm) * M + m;
        st_st(sbb + row * rowlen + tpad + t0 + t + ld_table(fl, s), v);
    }

    __device__ void read_phase_extra(int) { }
    __device__ void write_phase_extra(int) { }

    __device__ void stag_rounds(int r)
    {
        const int half = 1 << r;
        const int k0 = ld_table(stag_begin, r), n = (ld_table(stag_begin, r + 1) - k0) * kTT;
        for (int base = 0; base < n; base += SMAX * kTreeThreads) {
            float so0[SMAX], so1[SMAX], ssv[SMAX];
#pragma unroll
            for (int q = 0; q < SMAX; q++) {
                const int idx = base + q * kTreeThreads + threadIdx.x;
                if (idx < n) {
                    const int k = k0 + __builtin_amdgcn_readfirstlane(idx / kTT), t = idx % kTT;   // k: wave-uniform
                    const int row0 = ld_table(stag, 6 * k + 1), row1 = ld_table(stag, 6 * k + 2);
                    const int L = ld_table(stag, 6 * k + 3);
                    const float *h = shist + ld_table(stag, 6 * k + 5);
                    const float xo = tile[row1 * kTT + t];
                    const float xa = (t >= L) ? tile[row0 * kTT + t - L] : h[half + t - L];
                    const float xb = (t >= L + 1) ? tile[row0 * kTT + t - L - 1] : h[half + t - L - 1];
                    so0[q] = kRsqrt2 * (xa + xo);
                    so1[q] = kRsqrt2 * (xb + xo);
                    const int ts = t - (kTT - nv);
                    ssv[q] = (t < kTT - half) ? 0.0f : (ts >= 0) ? tile[row0 * kTT + ts] : h[half + ts];
                }
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < SMAX; q++) {
                const int idx = base + q * kTreeThreads + threadIdx.x;
                if (idx < n) {
                    const int k = k0 + __builtin_amdgcn_readfirstlane(idx / kTT), t = idx % kTT;
                    const int s0 = ld_table(stag, 6 * k + 4);
                    emit(s0, t, so0[q]);
                    emit(s0 + 1, t, so1[q]);
                    if (t >= kTT - half)
                        shist[ld_table(stag, 6 * k + 5) + (t - (kTT - half))] = ssv[q];
                }
            }
            __syncthreads();
        }
    }
};

// 128 VGPRs (4 waves per SIMD; the rank-7 instances spill a few registers at this limit,
// against 3 waves at 129-134): chime_sb1-shaped chain, 8 beams, float16, 15.15 -> 14.86 ms.
template<typename Tst, int DD>
__global__ void __launch_bounds__(kTreeThreads) __attribute__((amdgpu_waves_per_eu(4)))
k_stage2(Stage2Args a)
{
    using TD = TreeDims<DD>;
    extern __shared__ float smem[];
    float *tile = smem;
    float *hist = smem + TD::D * kTT;
    float *shist = hist + TD::HIST;
    const long nst = TD::HIST + a.stag_hist_nelts;
    int4 *desc = reinterpret_cast<int4 *>(smem + ((TD::D * kTT + nst + 3) & ~3L));   // 16-byte aligned

    const int j = blockIdx.x, b = blockIdx.y;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wid = __builtin_amdgcn_readfirstlane(tid >> 6);
    float *st = a.state + long(a.beam0 + b) * a.state_bstride + long(j) * nst;

    for (int i = tid; i < nst; i += kTreeThreads)
        hist[i] = st[i];     // hist and shist are contiguous
    // ring row of tree row fc at this chunk's time 0: (offset + position, samples to the
    // wrap, capacity)
    const long chunk = launch_chunk(a);
    for (int fc = tid; fc < TD::D; fc += kTreeThreads) {
        const long rr = (long(fc) << a.amb) + j;
        const long cap = a.ring.row_cap[rr];
        const long pos = ring_pos(chunk * a.Tp - a.row_lag[rr], cap);
        desc[fc] = make_int4(int(a.ring.row_off[rr] + pos), int(cap - pos), int(cap), 0);
    }

    const int pf_ndm = 1 << (DD - a.pf_rank);
    const int dpf_base = bitrev_dev(j, a.amb) * pf_ndm;
    const Tst *ring = reinterpret_cast<const Tst *>(a.ring.base) + long(a.beam0 + b) * a.ring.bstride;
    Tst *sbb = reinterpret_cast<Tst *>(a.sb) + long(a.beam0 + b) * a.sb_bstride;
    Stage2Sink<Tst, DD> sink;
    sink.sbb = sbb;
    sink.stag = a.stag;
    sink.fl = a.fl;
    sink.M = a.M;
    sink.rowlen = a.rowlen;
    sink.tpad = a.tpad;
    sink.dpf_base = dpf_base;
    sink.t0 = 0;
    sink.nv = kTT;
    sink.tile = tile;
    sink.shist = shist;
    sink.stag_begin = a.stag_begin;
    const int r_lvl0 = DD - a.pf_rank;
    const int nlvl0 = a.nlvl0;
    const int *lvl0_tab = a.lvl0;
    const int *main_emit = a.main_emit;
    const int Tp = a.Tp;
    __syncthreads();

    auto run = [&](auto kp) {
        constexpr int KP = decltype(kp)::value;
        using PP = PrePass<DD, (KP > 0 ? KP : 1)>;
        using SK = Sub<(KP > 0 ? KP : 1)>;
        // register histories of the pre-passes, in the LDS history layout when saved:
        // pass r, global pair p = group * 2^(KP-1) + q, slot k at hist_off(NP, r) + p*2^r + k
        float hreg[PP::GPW][SK::NH];
        if constexpr (KP > 0) {
#pragma unroll
            for (int gg = 0; gg < PP::GPW; gg++) {
                const int grp = wid + gg * PP::NWV;
#pragma unroll
                for (int r = 0; r < KP; r++) {
                    const int n = 1 << r;
#pragma unroll
                    for (int q = 0; q < SK::NP; q++) {
                        const int p = grp * SK::NP + q;
                        hreg[gg][r * SK::NP + q] = (grp < PP::NG && lane >= 64 - n)
                                                   ? hist[hist_off(TD::NP, r) + p * n + lane - (64 - n)] : 0.0f;
                    }
                }
            }
        }

        for (int t0 = 0; t0 < Tp; t0 += kTT) {
            const int nv = min(kTT, Tp - t0);
            sink.t0 = t0;
            sink.nv = nv;
            const int tl = t0 + lane;
            if constexpr (KP > 0) {
                const bool live = tl < Tp;
#pragma unroll
                for (int gg = 0; gg < PP::GPW; gg++) {
                    const int grp = wid + gg * PP::NWV;
                    if (grp < PP::NG) {
                        // rows as loaded, converted after all of them are issued (raw_st);
                        // lanes past the chunk end load the last sample (in bounds) and take
                        // zero, so that no load sits under a per-lane branch
                        const int tlc = live ? tl : Tp - 1;
                        raw_st<Tst> vr[SK::NR];
#pragma unroll
                        for (int l = 0; l < SK::NR; l++) {
                            const int4 d = desc[grp * SK::NR + l];
                            vr[l] = ld_raw(ring + (d.x + tlc - ((tlc >= d.y) ? d.z : 0)));
                        }
                        float v[SK::NR];
#pragma unroll
                        for (int l = 0; l < SK::NR; l++)
                            v[l] = live ? from_raw<Tst>(vr[l]) : 0.0f;
                        sub_passes<KP>(v, hreg[gg], lane);
#pragma unroll
                        for (int l = 0; l < SK::NR; l++)
                            tile[(grp * SK::NR + l) * kTT + lane] = v[l];
                    }
                }
            } else {
                // each wave's rows: all loads issued, then converted and written to the tile
                constexpr int NWV = kTreeThreads / 64, RPW = (TD::D + NWV - 1) / NWV;
                raw_st<Tst> vr[RPW];
#pragma unroll
                for (int k = 0; k < RPW; k++) {
                    const int fc = wid + k * NWV;
                    if (fc < TD::D) {
                        const int4 d = desc[fc];
                        vr[k] = ld_raw(ring + (d.x + tl - ((tl >= d.y) ? d.z : 0)));
                    }
                }
#pragma unroll
                for (int k = 0; k < RPW; k++) {
                    const int fc = wid + k * NWV;
                    if (fc < TD::D)
                        tile[fc * kTT + lane] = from_raw<Tst>(vr[k]);
                }
            }
            __syncthreads();

#pragma unroll 1
            for (int r = KP; r < DD; r++) {
                if (r == r_lvl0) {
                    for (int idx = tid; idx < nlvl0 * kTT; idx += kTreeThreads) {
                        const int k = __builtin_amdgcn_readfirstlane(idx / kTT), t = idx % kTT;
                        sink.emit(ld_table(lvl0_tab, 2 * k + 1), t, tile[ld_table(lvl0_tab, 2 * k) * kTT + t]);
                    }
                }
                sink.stag_rounds(r);
                tree_pass<DD>(r, tile, hist, main_emit, sink, nv);
            }
            if (r_lvl0 == DD) {
                for (int idx = tid; idx < nlvl0 * kTT; idx += kTreeThreads) {
                    const int k = __builtin_amdgcn_readfirstlane(idx / kTT), t = idx % kTT;
                    sink.emit(ld_table(lvl0_tab, 2 * k + 1), t, tile[ld_table(lvl0_tab, 2 * k) * kTT + t]);
                }
                __syncthreads();
            }
        }

        if constexpr (KP > 0) {
            const int hi = (Tp & 63) ? (Tp & 63) : 64;
#pragma unroll
            for (int gg = 0; gg < PP::GPW; gg++) {
                const int grp = wid + gg * PP::NWV;
#pragma unroll
                for (int r = 0; r < KP; r++) {
                    const int n = 1 << r;
#pragma unroll
                    for (int q = 0; q < SK::NP; q++) {
                        const int p = grp * SK::NP + q;
                        if (grp < PP::NG && lane >= hi - n && lane < hi)
                            hist[hist_off(TD::NP, r) + p * n + lane - (hi - n)] = hreg[gg][r * SK::NP + q];
                    }
                }
            }
            __syncthreads();
        }
    };

    // Pre-passes: up to 4 passes, all before the first subband emission (pass DD - pf_rank).
    const int kpre = min(min(DD - a.pf_rank, 4), DD);
    switch (kpre) {
        case 1: run(std::integral_constant<int, 1>{}); break;
        case 2: run(std::integral_constant<int, (DD >= 2 ? 2 : 1)>{}); break;
        case 3: run(std::integral_constant<int, (DD >= 3 ? 3 : 1)>{}); break;
        case 4: run(std::integral_constant<int, (DD >= 4 ? 4 : 1)>{}); break;
        default: run(std::integral_constant<int, 0>{}); break;
    }

    for (int i = tid; i < nst; i += kTreeThreads)
        st[i] = hist[i];
}

// ---------------------------------------------------------------------------------------
// Rank-8 stage 2 with subband searches, in two launches
//
// At rank 8 the LDS tile of k_stage2 would need 256 x 64 floats plus 128 x 255 history floats
// (191 KB). The tree is split by rows instead: passes 0..3 pair rows inside groups of 16
// consecutive rows (h = row >> 4), passes 4..7 pair rows with equal low 4 bits (j = row & 15),
// and every subband stream comes from passes >= dd - pf_rank >= 4. Part A (register-tree
// code, k_tree_fast.hip) runs passes 0..3 per (tree instance, h) on rows read from the lag
// ring and writes the tree after pass 3 to a float32 scratch X (B, 2^amb, 256, Tp) of the
// current chunk; part B runs passes 4..7 per
// (tree instance, j) on X with the full tree's lags, pair histories, staggered butterflies and
// subband emissions. Every operation is the direct tree's, in its order: bit-identical.
// A history longer than a tile (2^7 = 128 samples at pass 7) is refilled from the tile and
// from its own older part.

static constexpr int kSplitNPL = 8;                       // pairs per pass in a 16-row group
static constexpr int kSplitHistA = kSplitNPL * 15;        // part A pair history (passes 0..3, k_tree_fast.hip)
static constexpr int kSplitHistB = kSplitNPL * (256 - 16);   // part B pair history (passes 4..7)
static constexpr int kSplitGPB = 4;                       // part B row groups per workgroup

__host__ __device__ static inline long split_nst(int stag_hist_nelts)
{
    return 16L * kSplitHistA + 16L * (kSplitHistB + stag_hist_nelts);
}

long split_state_nelts_per_tree(int stag_hist_nelts) { return split_nst(stag_hist_nelts); }



int split_b_lds_bytes(int stag_hist_g)
{
    return int(sizeof(float) * (kSplitGPB * (16 * kTT + kSplitHistB + stag_hist_g)));
}

// New content of history slot k (time nv - half + k relative to this tile) of a row whose
// tile row is 'row' and whose old history is 'old' (time -half + k' at slot k').
__device__ __forceinline__ float hist_refill(const float *tile, int row, const float *old, int half, int nv, int k)
{
    const int tm = nv - half + k;
    return (tm >= 0) ? tile[row * kTT + tm] : old[nv + k];
}

// Part B handles kSplitGPB row groups (equal low 4 bits g0..g0+GPB-1) per workgroup: tile
// rows [GPB][16][64]; per group in LDS, in the state buffer's layout: pair histories
// (kSplitHistB floats) then the group's own staggered histories (stag_hist_g floats, offsets
// stag_goff[k]).

// Part-B emissions: subband streams into the peak-finder rows, with staggered butterflies
// and level-0 copies restricted to the workgroup's groups.
template<typename Tst>
struct SplitSink {
    Tst *sbb;
    const int *stag, *fl, *stag_begin, *stag_goff;
    int M, rowlen, tpad, dpf_base, t0, nv, g0, nstg;
    float *tile, *stl;
    static constexpr int SMAX = 4, HB = 2;     // entries per thread per round; 64-slot history blocks

    __device__ __forceinline__ void emit(int s, int t, float v)
    {
        if (t >= nv)
            return;
        const int pf_dm = s / M, m = s - pf_dm * M;
        const long row = long(dpf_base + pf_dm) * M + m;
        st_st(sbb + row * rowlen + tpad + t0 + t + ld_table(fl, s), v);
    }

    // Staggered butterflies of pass r whose rows are in this workgroup's groups, before the
    // pass (as in Stage2Sink::stag_rounds).
    __device__ void stag_rounds(int r)
    {
        const int half = 1 << r;
        const int k0 = ld_table(stag_begin, r), n = (ld_table(stag_begin, r + 1) - k0) * kTT;
        for (int base = 0; base < n; base += SMAX * kTreeThreads) {
            float so0[SMAX], so1[SMAX], ssv[SMAX][HB];
#pragma unroll
            for (int q = 0; q < SMAX; q++) {
                const int idx = base + q * kTreeThreads + threadIdx.x;
                if (idx < n) {
                    const int k = k0 + __builtin_amdgcn_readfirstlane(idx / kTT), t = idx % kTT;   // k: wave-uniform
                    const int e1 = ld_table(stag, 6 * k + 1);
                    const unsigned gi = unsigned((e1 & 15) - g0);
                    if (gi < unsigned(kSplitGPB)) {
                        const float *tg = tile + gi * (16 * kTT);
                        const int row0 = e1 >> 4, row1 = ld_table(stag, 6 * k + 2) >> 4, L = ld_table(stag, 6 * k + 3);
                        const float *h = stl + gi * nstg + kSplitHistB + ld_table(stag_goff, k);
                        const float xo = tg[row1 * kTT + t];
                        const float xa = (t >= L) ? tg[row0 * kTT + t - L] : h[half + t - L];
                        const float xb = (t >= L + 1) ? tg[row0 * kTT + t - L - 1] : h[half + t - L - 1];
                        so0[q] = kRsqrt2 * (xa + xo);
                        so1[q] = kRsqrt2 * (xb + xo);
                        if (half <= kTT) {
                            if (t >= kTT - half)
                                ssv[q][0] = hist_refill(tg, row0, h, half, nv, t - (kTT - half));
                        } else {
#pragma unroll
                            for (int m = 0; m < HB; m++)
                                if (m * kTT + t < half)
                                    ssv[q][m] = hist_refill(tg, row0, h, half, nv, m * kTT + t);
                        }
                    }
                }
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < SMAX; q++) {
                const int idx = base + q * kTreeThreads + threadIdx.x;
                if (idx < n) {
                    const int k = k0 + __builtin_amdgcn_readfirstlane(idx / kTT), t = idx % kTT;
                    const unsigned gi = unsigned((ld_table(stag, 6 * k + 1) & 15) - g0);
                    if (gi < unsigned(kSplitGPB)) {
                        const int s0 = ld_table(stag, 6 * k + 4);
                        emit(s0, t, so0[q]);
                        emit(s0 + 1, t, so1[q]);
                        float *h = stl + gi * nstg + kSplitHistB + ld_table(stag_goff, k);
                        if (half <= kTT) {
                            if (t >= kTT - half)
                                h[t - (kTT - half)] = ssv[q][0];
                        } else {
#pragma unroll
                            for (int m = 0; m < HB; m++)
                                if (m * kTT + t < half)
                                    h[m * kTT + t] = ssv[q][m];
                        }
                    }
                }
            }
            __syncthreads();
        }
    }
};

// Pass r (4..7) of the rank-8 tree on the workgroup's groups g0 + gi (tile row l of group gi
// = global row (l << 4) + g0 + gi).
template<typename Sink>
__device__ __forceinline__ void split_pass(int r, int g0, float *tile, float *stl, int nstg, const int *emit, Sink &sink,
                                           int nv)
{
    constexpr int NPL = kSplitNPL, PER_G = NPL * kTT, ITEMS = kSplitGPB * PER_G / kTreeThreads, HB = 2;
    const int tid = threadIdx.x;
    const int rl = r - 4, hl = 1 << rl, half = 1 << r;

    float o0[ITEMS], o1[ITEMS], sv[ITEMS][HB];
#pragma unroll
    for (int i = 0; i < ITEMS; i++) {
        const int idx = i * kTreeThreads + tid;
        const int gi = idx / PER_G, rem = idx - gi * PER_G;
        const int q = rem / kTT, t = rem % kTT;
        const float *tg = tile + gi * (16 * kTT);
        const int el = ((q >> rl) << (rl + 1)) + (q & (hl - 1)), ol = el + hl;
        const int e = (el << 4) + g0 + gi;
        const int L = bitrev_dev(e & (half - 1), r);
        const float *hq = stl + gi * nstg + NPL * (half - 16) + q * half;
        const float xo = tg[ol * kTT + t];
        const float xa = (t >= L) ? tg[el * kTT + t - L] : hq[half + t - L];
        const float xb = (t >= L + 1) ? tg[el * kTT + t - L - 1] : hq[half + t - L - 1];
        o0[i] = kRsqrt2 * (xa + xo);
        o1[i] = kRsqrt2 * (xb + xo);
        if (half <= kTT) {
            sv[i][0] = (t >= kTT - half) ? hist_refill(tg, el, hq, half, nv, t - (kTT - half)) : 0.0f;
        } else {
#pragma unroll
            for (int m = 0; m < HB; m++)
                sv[i][m] = (m * kTT + t < half) ? hist_refill(tg, el, hq, half, nv, m * kTT + t) : 0.0f;
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < ITEMS; i++) {
        const int idx = i * kTreeThreads + tid;
        const int gi = idx / PER_G, rem = idx - gi * PER_G;
        const int q = rem / kTT, t = rem % kTT;
        float *tg = tile + gi * (16 * kTT);
        const int el = ((q >> rl) << (rl + 1)) + (q & (hl - 1)), ol = el + hl;
        const int e = (el << 4) + g0 + gi;
        float *hq = stl + gi * nstg + NPL * (half - 16) + q * half;
        tg[el * kTT + t] = o0[i];
        tg[ol * kTT + t] = o1[i];
        if (half <= kTT) {
            if (t >= kTT - half)
                hq[t - (kTT - half)] = sv[i][0];
        } else {
#pragma unroll
            for (int m = 0; m < HB; m++)
                if (m * kTT + t < half)
                    hq[m * kTT + t] = sv[i][m];
        }
        const int dmb = e & (half - 1);
        const int p = __builtin_amdgcn_readfirstlane(((e >> (r + 1)) << r) + dmb);   // wave-uniform (q, gi are)
        const int s0 = ld_table(emit, r * 128 + p);
        if (s0 >= 0) {
            sink.emit(s0, t, o0[i]);
            sink.emit(s0 + 1, t, o1[i]);
        }
    }
    __syncthreads();
}

template<typename Tst>
__global__ void __launch_bounds__(kTreeThreads)
k_stage2_split_b(Stage2Args a, const float *X)
{
    constexpr int GPB = kSplitGPB;
    extern __shared__ float smem[];
    float *tile = smem;                      // [GPB][16][kTT]
    float *stl = tile + GPB * 16 * kTT;      // [GPB][nstg]
    const int tr = blockIdx.x / (16 / GPB), g0 = (blockIdx.x % (16 / GPB)) * GPB, b = blockIdx.y;
    const int tid = threadIdx.x, lane = tid & 63;
    const long gbeam = long(a.beam0) + b;
    const int nstg = kSplitHistB + a.stag_hist_nelts;     // stag_hist_nelts = per-group size here
    float *st = a.state + gbeam * a.state_bstride + long(tr) * split_nst(a.stag_hist_nelts)
                + 16L * kSplitHistA + long(g0) * nstg;
    for (int i = tid; i < GPB * nstg; i += kTreeThreads)
        stl[i] = st[i];

    const int pf_ndm = 1 << (8 - a.pf_rank);
    SplitSink<Tst> sink;
    sink.sbb = reinterpret_cast<Tst *>(a.sb) + gbeam * a.sb_bstride;
    sink.stag = a.stag;
    sink.fl = a.fl;
    sink.stag_begin = a.stag_begin;
    sink.stag_goff = a.stag_goff;
    sink.M = a.M;
    sink.rowlen = a.rowlen;
    sink.tpad = a.tpad;
    sink.dpf_base = bitrev_dev(tr, a.amb) * pf_ndm;
    sink.g0 = g0;
    sink.nstg = nstg;
    sink.tile = tile;
    sink.stl = stl;
    const float *x = X + ((long(b) << a.amb) + tr) * (256L * a.Tp);
    const int r_lvl0 = 8 - a.pf_rank;
    __syncthreads();

    const int nlvl0 = a.nlvl0;
    const int *lvl0_tab = a.lvl0;
    auto lvl0 = [&sink, tile, tid, g0, nlvl0, lvl0_tab]() {
        for (int idx = tid; idx < nlvl0 * kTT; idx += kTreeThreads) {
            const int k = __builtin_amdgcn_readfirstlane(idx / kTT), t = idx % kTT;
            const int row = ld_table(lvl0_tab, 2 * k);
            const unsigned gi = unsigned((row & 15) - g0);
            if (gi < unsigned(GPB))
                sink.emit(ld_table(lvl0_tab, 2 * k + 1), t, tile[gi * (16 * kTT) + (row >> 4) * kTT + t]);
        }
    };

    for (int t0 = 0; t0 < a.Tp; t0 += kTT) {
        const int nv = min(kTT, a.Tp - t0);
        sink.t0 = t0;
        sink.nv = nv;
        {
            // each wave loads its RPW rows, then writes them to LDS (all loads in flight at once)
            constexpr int NWV = kTreeThreads / 64, RPW = GPB * 16 / NWV;
            static_assert(GPB * 16 % NWV == 0, "rows per wave");
            const int w0 = __builtin_amdgcn_readfirstlane(tid >> 6);
            float xv[RPW];
#pragma unroll
            for (int k = 0; k < RPW; k++) {
                const int rw = w0 + k * NWV, gi = rw >> 4, l = rw & 15;
                xv[k] = (lane < nv) ? x[long((l << 4) + g0 + gi) * a.Tp + t0 + lane] : 0.0f;
            }
#pragma unroll
            for (int k = 0; k < RPW; k++)
                tile[(w0 + k * NWV) * kTT + lane] = xv[k];
        }
        __syncthreads();
#pragma unroll 1
        for (int r = 4; r < 8; r++) {
            if (r == r_lvl0)
                lvl0();
            sink.stag_rounds(r);
            split_pass(r, g0, tile, stl, nstg, a.main_emit, sink, nv);
        }
        if (r_lvl0 == 8) {
            lvl0();
            __syncthreads();
        }
    }

    for (int i = tid; i < GPB * nstg; i += kTreeThreads)
        st[i] = stl[i];
}

template<typename Tst>
static void launch_stage2_split_t(const Stage2Args &a, float *X, hipStream_t st)
{
    const dim3 grid{unsigned((16 / kSplitGPB) << a.amb), unsigned(a.B), 1u}, block{unsigned(kTreeThreads), 1u, 1u};
    const int shm = split_b_lds_bytes(a.stag_hist_nelts);
    if (shm > 65536)
        (void) hipFuncSetAttribute((const void *) k_stage2_split_b<Tst>, hipFuncAttributeMaxDynamicSharedMemorySize, shm);
    hipLaunchKernelGGL((k_stage2_split_b<Tst>), grid, block, size_t(shm), st, a, (const float *) X);
}

void launch_stage2_split(Dtype dt, const Stage2Args &a, float *X, hipStream_t st)
{
    launch_stage2_split_a_reg(dt, a, X, split_nst(a.stag_hist_nelts), st);
    switch (dt) {
        case Dtype::float32: launch_stage2_split_t<float>(a, X, st); break;
        case Dtype::float16: launch_stage2_split_t<__half>(a, X, st); break;
        case Dtype::bfloat16: launch_stage2_split_t<__hip_bfloat16>(a, X, st); break;
    }
}

template<typename Tst>
static void launch_stage2_t(const Stage2Args &a, hipStream_t st)
{
    dim3 grid(1 << a.amb, a.B);
    dim3 block(kTreeThreads);
    size_t shm = stage2_lds_bytes(a.dd, a.stag_hist_nelts);
#define PDD_S2(DDV)                                                                           \
    case DDV: {                                                                               \
        (void) hipFuncSetAttribute((const void *) k_stage2<Tst, DDV>, hipFuncAttributeMaxDynamicSharedMemorySize, int(shm)); \
        hipLaunchKernelGGL((k_stage2<Tst, DDV>), grid, block, shm, st, a);                    \
        break;                                                                                \
    }
    switch (a.dd) {
        PDD_S2(1) PDD_S2(2) PDD_S2(3) PDD_S2(4) PDD_S2(5) PDD_S2(6) PDD_S2(7)
        default: break;
    }
#undef PDD_S2
}

void launch_stage2(Dtype dt, const Stage2Args &a, hipStream_t st)
{
    switch (dt) {
        case Dtype::float32: launch_stage2_t<float>(a, st); break;
        case Dtype::float16: launch_stage2_t<__half>(a, st); break;
        case Dtype::bfloat16: launch_stage2_t<__hip_bfloat16>(a, st); break;
    }
}

}  // namespace pirate_amd
