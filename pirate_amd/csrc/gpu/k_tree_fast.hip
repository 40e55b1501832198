// Register-resident tree dedispersion on gfx950 (stage 1, and stage 2 when every subband
// stream of the tree is a final-pass output).
//
// Reference semantics: ReferenceTree (src_lib/ReferenceTree.cpp:120-345) as applied by
// ReferenceDedisperser sophistication 1 (src_lib/ReferenceDedisperser.cpp:388-450):
//   pass r pairs rows e = fout*2^(r+1) + dmb and o = e + 2^r (in place, bit-reversed DM):
//     e <- c*(x_e[t - L] + x_o[t]),  o <- c*(x_e[t - L - 1] + x_o[t]),  L = bitrev(dmb, r).
//
// Factorization (the same identity that splits the full tree into stage 1 and stage 2):
// a rank-DD tree over rows i = h*2^K1 + l (K1 + K2 = DD) equals
//   phase A: 2^K2 independent rank-K1 trees over l (one per h),
//   a delay of (2^K2 - 1 - h) * bitrev(j, K1) samples on row (h, j), j = phase-A output row,
//   phase B: 2^K1 independent rank-K2 trees over h (one per j); final row = h'*2^K1 + j.
// The floating-point operations are exactly those of the direct tree, so results are
// bit-identical.
//
// Mapping: one workgroup per tree instance sweeps the chunk in tiles of 64 samples; lane = time
// sample. A wave holds the rows of one sub-tree in registers; a pass's lag L is a rotation
// of the e-row across lanes (ds_bpermute), with lanes t < L taking the previous tile's values
// (kept in registers: the per-pass history). Between the phases, rows go through an LDS delay
// line per row (length lag + 64, shifted down by 64 after each read). Histories and the
// delay-line contents persist across chunks in an HBM state buffer.
//
// Chunks whose length is a multiple of 32 but not of 64 end in a half tile: its loads past
// the chunk read zeros, its lanes 32..63 store nothing, the delay lines shift down by 32,
// and the saved pass histories are taken from lanes below 32. Garbage in lanes 32..63 never
// reaches a lower lane, since every lag moves data to later times (higher lanes).

#include <type_traits>

#include "device_common.hpp"
#include "kernels.hpp"
#include "tree_regs.hpp"

namespace pirate_amd {

namespace {

using namespace treg;

constexpr int cmin(int a, int b) { return (a < b) ? a : b; }

// Phase-A rows (of 16) of the next tile that rank-7 stage 2 loads during the current tile's
// phase B. Measured on chord16k (stage2_t0, fp32 / fp16 ms): 0 rows 1.10 / 0.87, 4 rows
// 1.01 / 0.88, 8 rows 0.99 / 0.86; 12 rows spilled in the tile loop (fp32 1.14 vs 1.07 for 8
// in an earlier build), all 16 rows 1.31. Rank 8 (which spills already) keeps 4: on a
// chime.yml-shaped workload 0, 2, 4, 8 rows measured 3.70, 3.88, 3.70, 3.73 ms.
constexpr int kStage2Rank7Prefetch = 8;
constexpr int kStage2Rank8Prefetch = 4;

template<int DD>
struct FastDims {
    static constexpr int D = 1 << DD;
    static constexpr int K1 = (DD + 1) / 2, K2 = DD / 2;
    static constexpr int NRA = 1 << K1, NRB = 1 << K2;
    // waves per workgroup: rank 8 runs one sub-tree per wave and phase in 16 waves (128
    // VGPRs each); 8 waves with two sub-trees each (256 VGPRs, 2 waves per SIMD) measured
    // slower (chime.yml-shaped stage 2, 8 beams, float16: 4.29 vs 3.71 ms)
    static constexpr int NW = (DD >= 8) ? 16 : cmin(8, cmin(NRA, NRB));
    static constexpr int SA = NRB / NW, SB = NRA / NW;     // sub-trees per wave in phase A / B
    static constexpr int T2 = NRB * (NRB - 1) / 2;
    static constexpr int SUMLAG = T2 * (NRA * (NRA - 1) / 2);
    static constexpr int NSTATE = NRB * Sub<K1>::NST + NRA * Sub<K2>::NST + SUMLAG;
    static constexpr int LMAX = (NRB - 1) * (NRA - 1);       // largest delay-line lag
    static_assert(LMAX < 256, "delay-line shift handles lags below 256");

    // Delay lines. Row (h, j) holds the last lag(h, j) samples of phase-A output row (h, j)
    // followed by the current tile: phase A writes the tile at a_off(h, j), phase B reads
    // the delayed window at b_off(h, j) (= a_off - lag) and then shifts the line down by 64.
    // Ranks up to 7 use a 2-D layout with a compile-time phase-A offset per (wave, j): row
    // (h, j) starts at h*S + Q(j) and reserves Lm(j) + 64 floats, Lm(j) = (NRB-1)*bitrev(j)
    // the largest lag over h. Rank 8 packs the rows (j-major, each row lag(h, j) + 64
    // floats), which needs 120 KB instead of 176 KB; gfx950 gives one workgroup up to 160 KB.
    static constexpr bool PACKED = (DD >= 8);
    static constexpr int Lm(int j) { return (NRB - 1) * brev_c(j, K1); }
    static constexpr int Q(int j) { return (j == 0) ? 0 : Q(j - 1) + Lm(j - 1) + 64; }
    static constexpr int S = Q(NRA);
    static constexpr int G(int h) { return h * (NRB - 1) - h * (h - 1) / 2; }   // sum of (NRB-1-h') over h' < h
    static constexpr int CQ(int j) { return (j == 0) ? 0 : CQ(j - 1) + brev_c(j - 1, K1) * T2 + 64 * NRB; }
    static constexpr int DL_USED = PACKED ? CQ(NRA) : NRB * S;
    static constexpr int DL_FLOATS = ((DL_USED + 256 + 3) / 4) * 4;   // + padding for the shift reads
    static constexpr int LDS_BYTES = 4 * DL_FLOATS + 2 * 16 * D;   // + two int4 row-descriptor tables
    // rank 7: 62 KB of LDS allows 2 workgroups (16 waves) per CU, so aim for 128 VGPRs;
    // rank 8: one workgroup (16 waves) per CU, also 128 VGPRs
    static constexpr int WAVES_PER_EU = (DD >= 7) ? 4 : 2;

    __host__ __device__ static constexpr int lag_c(int h, int j) { return (NRB - 1 - h) * brev_c(j, K1); }
    // phase-A write offset of row (h, j) (h = the wave's sub-tree, j unrolled)
    __device__ static int a_off(int h, int j)
    {
        if constexpr (PACKED)
            return CQ(j) + 64 * h + brev_c(j, K1) * G(h + 1);
        else
            return h * S + Q(j) + Lm(j);
    }
    // phase-B read offset of row (h, j) (j = the wave's sub-tree, h unrolled), bj = bitrev(j)
    __device__ static int b_off(int h, int j, int bj)
    {
        if constexpr (PACKED)
            return CQ(j) + 64 * h + bj * G(h);
        else
            return Q(j) + h * (S + bj);
    }

    __device__ static int lag(int h, int j) { return (NRB - 1 - h) * brev_c(j, K1); }
    // residue of row (h, j) in the state buffer: rows j-major, h inner, lag(h, j) floats each
    __device__ static int lagsum(int h, int j)
    {
        int sb = 0;
        for (int jj = 0; jj < j; jj++)
            sb += brev_c(jj, K1);
        return T2 * sb + brev_c(j, K1) * (h * (NRB - 1) - h * (h - 1) / 2);
    }
};

// Per-wave view of the phase-B delay lines of sub-tree j.
template<int DD>
struct PhaseB {
    using F = FastDims<DD>;
    int j, bj, cq;     // cq: offset of sub-tree j's rows in the packed layout (unused otherwise)

    __device__ void init(int jj, int lane)
    {
        (void) lane;
        j = jj;
        bj = brev_c(j, F::K1);
        int q = 0;
        for (int k = 0; k < j; k++)
            q += F::PACKED ? (brev_c(k, F::K1) * F::T2 + 64 * F::NRB) : (F::Lm(k) + 64);
        cq = q;
    }

    // start of row (h, j)'s delayed window (h compile-time after unrolling)
    __device__ __forceinline__ int row(int h) const
    {
        if constexpr (F::PACKED)
            return cq + 64 * h + bj * F::G(h);
        else
            return cq + h * (F::S + bj);
    }

    // Row (h, j) at the delayed time of this tile's lane, then shift the line down by 64
    // (positions [64, 64 + lag) -> [0, lag)); each lane copies its own column, in order.
    // Only positions below lag must be right: [lag, lag + 64) is the next tile's phase-A
    // window, so lanes at or above lag may copy garbage (reads past the row stay inside the
    // padded buffer), which keeps the copy branch-free.
    __device__ __forceinline__ void read(float *lbuf, float (&y)[F::NRB], int lane) const
    {
#pragma unroll
        for (int h = 0; h < F::NRB; h++) {
            float *p = lbuf + row(h) + lane;
            const int L = (F::NRB - 1 - h) * bj;
            y[h] = p[0];
            p[0] = p[64];
            if (L > 64)          // wave-uniform
                p[64] = p[128];
            if constexpr (F::LMAX > 128) {
                if (L > 128)
                    p[128] = p[192];
                if (L > 192)
                    p[192] = p[256];
            }
        }
    }

    // Half tile (32 valid lanes): the same read, then a shift down by 32. Here lane t's
    // store lands where lane t - 32 loads, so every load of the wave is issued before any
    // store (LDS instructions of one wave execute in order).
    __device__ __forceinline__ void read_half(float *lbuf, float (&y)[F::NRB], int lane) const
    {
#pragma unroll
        for (int h = 0; h < F::NRB; h++) {
            float *p = lbuf + row(h) + lane;
            const int L = (F::NRB - 1 - h) * bj;
            y[h] = p[0];
            const float s0 = p[32];
            const float s1 = (L > 64) ? p[96] : 0.0f;
            float s2 = 0.0f, s3 = 0.0f;
            if constexpr (F::LMAX > 128) {
                s2 = (L > 128) ? p[160] : 0.0f;
                s3 = (L > 192) ? p[224] : 0.0f;
            }
            __builtin_amdgcn_wave_barrier();
            p[0] = s0;
            if (L > 64)
                p[64] = s1;
            if constexpr (F::LMAX > 128) {
                if (L > 128)
                    p[128] = s2;
                if (L > 192)
                    p[192] = s3;
            }
            __builtin_amdgcn_wave_barrier();
        }
    }

    template<bool LOAD>
    __device__ void residue(float *lbuf, float *res, int lane) const
    {
        for (int h = 0; h < F::NRB; h++) {
            float *p = lbuf + row(h);
            float *r = res + F::lagsum(h, j);
            const int L = F::lag(h, j);
            for (int q = lane; q < L; q += 64) {
                if (LOAD)
                    p[q] = r[q];
                else
                    r[q] = p[q];
            }
        }
    }
};

template<int DD, int K1 = FastDims<DD>::K1>
__device__ __forceinline__ void phase_a_write(float *lbuf, int h, const float (&x)[1 << K1], int lane)
{
    using F = FastDims<DD>;
#pragma unroll
    for (int j = 0; j < F::NRA; j++)
        lbuf[F::a_off(h, j) + lane] = x[j];
}

}  // namespace


long fast_tree_state_nelts(int dd)
{
    switch (dd) {
        case 1: return FastDims<1>::NSTATE;
        case 2: return FastDims<2>::NSTATE;
        case 3: return FastDims<3>::NSTATE;
        case 4: return FastDims<4>::NSTATE;
        case 5: return FastDims<5>::NSTATE;
        case 6: return FastDims<6>::NSTATE;
        case 7: return FastDims<7>::NSTATE;
        case 8: return FastDims<8>::NSTATE;
        default: return 0;
    }
}


// ---------------------------------------------------------------------------------------
// Stage 1: (B, 2^amb trees, 2^dd rows, Tp) -> lag ring rows (tree, final row)
// Stage 2 (final-pass streams only): ring rows (fc, j) at their lag -> peak-finder rows
//
// Per-row addressing is precomputed once per launch into LDS descriptors (int4):
//   ring row (output of stage 1, input of stage 2): x = element offset of the row's position
//   for chunk time 0 of this chunk (row offset + position), y = number of samples before the
//   position wraps (capacity - position), z = capacity. Offsets are 32-bit (checked on host).
//   peak-finder row (stage 2 output): x = element offset of (row, tpad + final lag) or -1.

// HC: the chunk may end in a half tile (Tp a multiple of 32, not of 64); launched only for
// such chunks, so the common case keeps tile-uniform ring-wrap tests and no half-tile code.
template<typename Tst, int DD, bool STAGE2, bool HC, typename Args>
__device__ __forceinline__ void fast_tree_body(const Args &a)
{
    using F = FastDims<DD>;
    using SA_ = Sub<F::K1>;
    using SB_ = Sub<F::K2>;
    constexpr int D = F::D;
    extern __shared__ float lbuf[];
    int4 *idesc = reinterpret_cast<int4 *>(lbuf + F::DL_FLOATS);    // stage 2: input rows
    int4 *odesc = idesc + D;                                          // output rows

    const int lane = threadIdx.x & 63;
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int tr = blockIdx.x, b = blockIdx.y;
    const long gbeam = long(a.beam0) + b;

    // ---- row descriptors
    const long chunk = launch_chunk(a);
    for (int i = threadIdx.x; i < D; i += blockDim.x) {
        if constexpr (!STAGE2) {
            const long rr = long(tr) * D + i;
            const long cap = a.ring.row_cap[rr];
            const long pos = (chunk * a.Tp) % cap;
            odesc[i] = make_int4(int(a.ring.row_off[rr] + pos), int(cap - pos), int(cap), 0);
        } else {
            const long rr = (long(i) << a.amb) + tr;
            const long cap = a.ring.row_cap[rr];
            const long pos = ring_pos(chunk * a.Tp - a.row_lag[rr], cap);
            idesc[i] = make_int4(int(a.ring.row_off[rr] + pos), int(cap - pos), int(cap), 0);
            const int sidx = a.out_stream[i];
            const long dpf_base = long(bitrev_dev(tr, a.amb)) << (DD - a.pf_rank);
            odesc[i] = make_int4((sidx < 0) ? -1 : int((dpf_base * a.M + sidx) * a.rowlen + a.tpad + a.fl[sidx]), 0, 0, 0);
        }
    }

    // ---- state: histories (registers) and delay-line residue (LDS)
    float *st = a.state + gbeam * a.state_bstride + long(tr) * F::NSTATE;
    float *st_b = st + F::NRB * SA_::NST;
    float *st_res = st_b + F::NRA * SB_::NST;

    float ha[F::SA][SA_::NH], hb[F::SB][SB_::NH];
    PhaseB<DD> pb[F::SB];
#pragma unroll
    for (int s = 0; s < F::SA; s++)
        load_hist<F::K1>(ha[s], st + (wid + s * F::NW) * SA_::NST, lane);
#pragma unroll
    for (int s = 0; s < F::SB; s++) {
        const int j = wid + s * F::NW;
        pb[s].init(j, lane);
        load_hist<F::K2>(hb[s], st_b + j * SB_::NST, lane);
        pb[s].template residue<true>(lbuf, st_res, lane);
    }
    __syncthreads();   // descriptors

    const Tst *src;    // stage 1: this tree's input rows; stage 2: the beam's ring
    Tst *dst;          // stage 1: the beam's ring; stage 2: the beam's peak-finder rows
    if constexpr (!STAGE2) {
        src = reinterpret_cast<const Tst *>(a.in) + long(b) * a.in_bstride + long(tr) * D * a.Tp;
        dst = reinterpret_cast<Tst *>(a.ring.base) + gbeam * a.ring.bstride;
    } else {
        src = reinterpret_cast<const Tst *>(a.ring.base) + gbeam * a.ring.bstride;
        dst = reinterpret_cast<Tst *>(a.sb) + gbeam * a.sb_bstride;
    }

    // phase-A input rows of tile t0 (issued one tile ahead); PART: a half tile, whose lanes
    // past the chunk end read zeros
    // Stage 2 keeps them as loaded (raw_st) and converts where phase A uses them: converted
    // at the load, its 16-bit variant waited for each of a tile's row loads before issuing the
    // next (one temporary register for all). Stage 1 converts at the load: kept raw, the
    // compiler packs two rows per register and waits for each pair instead.
    using LdT = typename std::conditional<STAGE2, raw_st<Tst>, float>::type;
    LdT xa[F::SA][F::NRA];
    // Stage 1 (and stage 2 below rank 7) issues all of the next tile's phase-A loads right
    // after this tile's phase A, so they are in flight during phase B; rank-7 stage 2 would
    // spill with all of them (1.31 vs 1.19 ms on chord16k) and prefetches the first NPF rows.
    constexpr int NPF = (!STAGE2 || DD <= 6) ? F::NRA : (DD == 7) ? kStage2Rank7Prefetch : kStage2Rank8Prefetch;
    // rows [L0, L1) of tile t0
    auto load_a = [&](int t0, auto part, auto l0, auto l1) {
        constexpr bool PART = decltype(part)::value;
        constexpr int L0 = decltype(l0)::value, L1 = decltype(l1)::value;
        const bool live = !PART || (t0 + lane < a.Tp);
#pragma unroll
        for (int s = 0; s < F::SA; s++) {
            const int h = wid + s * F::NW;
#pragma unroll
            for (int l = L0; l < L1; l++) {
                if constexpr (!STAGE2) {
                    xa[s][l] = live ? ld_st(src + long(h * F::NRA + l) * a.Tp + t0 + lane) : 0.0f;
                } else {
                    const int4 d = idesc[h * F::NRA + l];
                    const int tl = t0 + lane;
                    xa[s][l] = live ? ld_raw(src + (d.x + tl - ((tl >= d.y) ? d.z : 0))) : LdT(0);
                }
            }
        }
    };
    using Full = std::false_type;
    using Half = std::true_type;
    using R0 = std::integral_constant<int, 0>;
    using RP = std::integral_constant<int, NPF>;
    using RN = std::integral_constant<int, F::NRA>;
    const int nfull = a.Tp >> 6;
    const bool has_half = HC && (a.Tp & 63) != 0;

    auto tile_step = [&](int t0, auto half) {
        constexpr bool HALF = decltype(half)::value;
        if constexpr (HALF)                   // a half tile is never prefetched
            load_a(t0, half, R0{}, RN{});
        else if constexpr (NPF < F::NRA)
            load_a(t0, half, RP{}, RN{});
#pragma unroll
        for (int s = 0; s < F::SA; s++) {
            if constexpr (STAGE2) {
                float x[F::NRA];
#pragma unroll
                for (int l = 0; l < F::NRA; l++)
                    x[l] = from_raw<Tst>(xa[s][l]);
                sub_passes<F::K1>(x, ha[s], lane);
                phase_a_write<DD>(lbuf, wid + s * F::NW, x, lane);
            } else {
                sub_passes<F::K1>(xa[s], ha[s], lane);
                phase_a_write<DD>(lbuf, wid + s * F::NW, xa[s], lane);
            }
            if constexpr (F::SA > 1 && DD >= 8)
                __builtin_amdgcn_sched_barrier(0);    // one sub-tree at a time (register pressure)
        }
        __syncthreads();
        if constexpr (NPF > 0 && !HALF) {
            if (t0 + 128 <= a.Tp)
                load_a(t0 + 64, Full{}, R0{}, RP{});
        }

        const bool live = !HALF || (t0 + lane < a.Tp);
        // Stage-1 ring positions are multiples of 64 when the chunk is, so a tile never
        // wraps and the wrap test uses the tile start; with half tiles it is made per lane.
        const int tw = HC ? (t0 + lane) : t0;
#pragma unroll
        for (int s = 0; s < F::SB; s++) {
            float y[F::NRB];
            if constexpr (HALF)
                pb[s].read_half(lbuf, y, lane);
            else
                pb[s].read(lbuf, y, lane);
            sub_passes<F::K2>(y, hb[s], lane);
#pragma unroll
            for (int hp = 0; hp < F::NRB; hp++) {
                const int4 d = odesc[hp * F::NRA + pb[s].j];
                if constexpr (!STAGE2) {
                    if (live)
                        st_st(dst + (d.x + t0 - ((tw >= d.y) ? d.z : 0) + lane), y[hp]);
                } else {
                    if (d.x >= 0 && live)
                        st_st(dst + (long(d.x) + t0 + lane), y[hp]);
                }
            }
            if constexpr (F::SB > 1 && DD >= 8)
                __builtin_amdgcn_sched_barrier(0);
        }
        __syncthreads();
    };

    if constexpr (NPF > 0) {
        if (nfull > 0)
            load_a(0, Full{}, R0{}, RP{});
    }
    for (int it = 0; it < nfull; it++)
        tile_step(it * 64, Full{});
    if constexpr (HC) {
        if (has_half)
            tile_step(nfull * 64, Half{});
    }

    const int hi = has_half ? (a.Tp & 63) : 64;
#pragma unroll
    for (int s = 0; s < F::SA; s++)
        save_hist<F::K1>(ha[s], st + (wid + s * F::NW) * SA_::NST, lane, hi);
#pragma unroll
    for (int s = 0; s < F::SB; s++) {
        save_hist<F::K2>(hb[s], st_b + pb[s].j * SB_::NST, lane, hi);
        pb[s].template residue<false>(lbuf, st_res, lane);
    }
}

template<typename Tst, int DD, bool HC>
__global__ void __launch_bounds__(FastDims<DD>::NW * 64) __attribute__((amdgpu_waves_per_eu(FastDims<DD>::WAVES_PER_EU)))
k_stage1_fast(Stage1Args a)
{
    fast_tree_body<Tst, DD, false, HC>(a);
}

template<typename Tst, int DD, bool HC>
__global__ void __launch_bounds__(FastDims<DD>::NW * 64) __attribute__((amdgpu_waves_per_eu(FastDims<DD>::WAVES_PER_EU)))
k_stage2_fast(Stage2Args a)
{
    fast_tree_body<Tst, DD, true, HC>(a);
}


// Dynamic LDS above 64 KB (rank 8) must be allowed per kernel; gfx950 allows 160 KB.
static void allow_lds(const void *fn, int bytes)
{
    if (bytes > 65536)
        (void) hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

template<typename Tst>
static void launch_stage1_fast_t(const Stage1Args &a, hipStream_t st)
{
#define PDD_F1(DDV)                                                                          \
    case DDV: {                                                                              \
        using F = FastDims<DDV>;                                                             \
        if (a.Tp & 63) {                                                                     \
            allow_lds((const void *) k_stage1_fast<Tst, DDV, true>, F::LDS_BYTES);             \
            hipLaunchKernelGGL((k_stage1_fast<Tst, DDV, true>), dim3(1 << a.amb, a.B), dim3(F::NW * 64), \
                               F::LDS_BYTES, st, a);                                         \
        } else {                                                                             \
            allow_lds((const void *) k_stage1_fast<Tst, DDV, false>, F::LDS_BYTES);            \
            hipLaunchKernelGGL((k_stage1_fast<Tst, DDV, false>), dim3(1 << a.amb, a.B), dim3(F::NW * 64), \
                               F::LDS_BYTES, st, a);                                         \
        }                                                                                    \
        break;                                                                               \
    }
    switch (a.dd) {
        PDD_F1(1) PDD_F1(2) PDD_F1(3) PDD_F1(4) PDD_F1(5) PDD_F1(6) PDD_F1(7) PDD_F1(8)
        default: break;
    }
#undef PDD_F1
}

template<typename Tst>
static void launch_stage2_fast_t(const Stage2Args &a, hipStream_t st)
{
#define PDD_F2(DDV)                                                                          \
    case DDV: {                                                                              \
        using F = FastDims<DDV>;                                                             \
        if (a.Tp & 63) {                                                                     \
            allow_lds((const void *) k_stage2_fast<Tst, DDV, true>, F::LDS_BYTES);             \
            hipLaunchKernelGGL((k_stage2_fast<Tst, DDV, true>), dim3(1 << a.amb, a.B), dim3(F::NW * 64), \
                               F::LDS_BYTES, st, a);                                         \
        } else {                                                                             \
            allow_lds((const void *) k_stage2_fast<Tst, DDV, false>, F::LDS_BYTES);            \
            hipLaunchKernelGGL((k_stage2_fast<Tst, DDV, false>), dim3(1 << a.amb, a.B), dim3(F::NW * 64), \
                               F::LDS_BYTES, st, a);                                         \
        }                                                                                    \
        break;                                                                               \
    }
    switch (a.dd) {
        PDD_F2(1) PDD_F2(2) PDD_F2(3) PDD_F2(4) PDD_F2(5) PDD_F2(6) PDD_F2(7) PDD_F2(8)
        default: break;
    }
#undef PDD_F2
}

void launch_stage1_fast(Dtype dt, const Stage1Args &a, hipStream_t st)
{
    switch (dt) {
        case Dtype::float32: launch_stage1_fast_t<float>(a, st); break;
        case Dtype::float16: launch_stage1_fast_t<__half>(a, st); break;
        case Dtype::bfloat16: launch_stage1_fast_t<__hip_bfloat16>(a, st); break;
    }
}

// ---------------------------------------------------------------------------------------
// Part A of the rank-8 split stage 2 (k_tree.hip): passes 0..3 of the rank-8 tree, which
// pair rows inside each group of 16 consecutive rows, are a rank-4 tree per group. One wave
// per (tree instance tr, group h) keeps the 16 rows in registers (the register kernel's
// sub-tree passes: the direct tree's operations, bit for bit), reads them from the lag ring
// at their stage-2 lags and writes the tree after pass 3 to the float32 scratch X
// (B, 2^amb, 256, Tp). Pass histories are saved in the same layout as the LDS-tile
// kernels' (8 pairs x (1 + 2 + 4 + 8) floats per group), after the rank-8 split's state.

static constexpr int kSplitHistA = 8 * 15;

template<typename Tst>
__global__ void __launch_bounds__(256)
k_stage2_split_a_reg(Stage2Args a, float *X, long split_nst)
{
    using SA_ = Sub<4>;
    const int lane = threadIdx.x & 63;
    const int grp = blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int tr = grp >> 4, h = grp & 15, b = blockIdx.y;
    const long gbeam = long(a.beam0) + b;
    float *st = a.state + gbeam * a.state_bstride + long(tr) * split_nst + long(h) * kSplitHistA;

    float hh[SA_::NH];
    load_hist<4>(hh, st, lane);

    // per-row ring position of this chunk's time 0 (wave-uniform)
    const Tst *ring = reinterpret_cast<const Tst *>(a.ring.base) + gbeam * a.ring.bstride;
    long rbase[16];
    int rwrap[16], rcap[16];
    const long chunk = launch_chunk(a);
#pragma unroll
    for (int l = 0; l < 16; l++) {
        const long rr = (long(h * 16 + l) << a.amb) + tr;
        const long cap = a.ring.row_cap[rr];
        const long pos = ring_pos(chunk * a.Tp - a.row_lag[rr], cap);
        rbase[l] = a.ring.row_off[rr] + pos;
        rwrap[l] = int(cap - pos);
        rcap[l] = int(cap);
    }
    float *x = X + ((long(b) << a.amb) + tr) * (256L * a.Tp) + long(h) * 16 * a.Tp;

    for (int t0 = 0; t0 < a.Tp; t0 += 64) {
        const int tl = t0 + lane;
        const bool live = tl < a.Tp;
        // Lanes past the chunk end load the last sample instead (in bounds) and take zero:
        // loads under a per-lane condition become branches, and with 16-bit storage the
        // conversion after each branch waited for that load before the next was issued.
        const int tlc = live ? tl : a.Tp - 1;
        raw_st<Tst> vr[16];      // as loaded, converted after all 16 loads are issued
#pragma unroll
        for (int l = 0; l < 16; l++)
            vr[l] = ld_raw(ring + (rbase[l] + tlc - ((tlc >= rwrap[l]) ? rcap[l] : 0)));
        float v[16];
#pragma unroll
        for (int l = 0; l < 16; l++)
            v[l] = live ? from_raw<Tst>(vr[l]) : 0.0f;
        sub_passes<4>(v, hh, lane);
        if (live) {
#pragma unroll
            for (int l = 0; l < 16; l++)
                x[long(l) * a.Tp + tl] = v[l];
        }
    }
    save_hist<4>(hh, st, lane, (a.Tp & 63) ? (a.Tp & 63) : 64);
}

void launch_stage2_split_a_reg(Dtype dt, const Stage2Args &a, float *X, long split_nst, hipStream_t st)
{
    const dim3 grid{unsigned((16 << a.amb) / 4), unsigned(a.B), 1u}, block{256u, 1u, 1u};
    switch (dt) {
        case Dtype::float32: hipLaunchKernelGGL(k_stage2_split_a_reg<float>, grid, block, 0, st, a, X, split_nst); break;
        case Dtype::float16: hipLaunchKernelGGL(k_stage2_split_a_reg<__half>, grid, block, 0, st, a, X, split_nst); break;
        case Dtype::bfloat16:
            hipLaunchKernelGGL(k_stage2_split_a_reg<__hip_bfloat16>, grid, block, 0, st, a, X, split_nst);
            break;
    }
}

void launch_stage2_fast(Dtype dt, const Stage2Args &a, hipStream_t st)
{
    switch (dt) {
        case Dtype::float32: launch_stage2_fast_t<float>(a, st); break;
        case Dtype::float16: launch_stage2_fast_t<__half>(a, st); break;
        case Dtype::bfloat16: launch_stage2_fast_t<__hip_bfloat16>(a, st); break;
    }
}

}  // namespace pirate_amd
