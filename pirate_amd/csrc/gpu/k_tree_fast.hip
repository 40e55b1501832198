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

#include "device_common.hpp"
#include "kernels.hpp"

namespace pirate_amd {

namespace {

constexpr int cmin(int a, int b) { return (a < b) ? a : b; }

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

template<int DD>
struct FastDims {
    static constexpr int K1 = (DD + 1) / 2, K2 = DD / 2;
    static constexpr int NRA = 1 << K1, NRB = 1 << K2;
    static constexpr int NW = cmin(8, cmin(NRA, NRB));     // waves per workgroup
    static constexpr int SA = NRB / NW, SB = NRA / NW;     // sub-trees per wave in phase A / B
    static constexpr int T2 = NRB * (NRB - 1) / 2;
    static constexpr int SUMLAG = T2 * (NRA * (NRA - 1) / 2);
    static constexpr int LDS_FLOATS = SUMLAG + 64 * (1 << DD);
    static constexpr int NSTATE = NRB * Sub<K1>::NST + NRA * Sub<K2>::NST + SUMLAG;

    // delay of phase-A output row (h, j)
    __device__ static int lag(int h, int j) { return (NRB - 1 - h) * brev_c(j, K1); }
    // rows are laid out j-major; row (h, j) holds lag(h, j) + 64 floats; P = sum of earlier lags
    __device__ static int lagsum(int h, int j)
    {
        int sb = 0;
        for (int jj = 0; jj < j; jj++)
            sb += brev_c(jj, K1);
        return T2 * sb + brev_c(j, K1) * (h * (NRB - 1) - h * (h - 1) / 2);
    }
    __device__ static int off(int h, int j) { return lagsum(h, j) + 64 * (j * NRB + h); }
};

// Stream value at time t - L (L = 0..64): cur at lane t - L, or prev (the previous tile)
// at lane 64 + t - L when t < L.
__device__ __forceinline__ float lane_shift(float cur, float prev, int L, int lane)
{
    const float w = (lane >= 64 - L) ? prev : cur;
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

template<int K>
__device__ __forceinline__ void save_hist(const float (&h)[Sub<K>::NH], float *dst, int lane)
{
    constexpr int NP = Sub<K>::NP;
#pragma unroll
    for (int r = 0; r < K; r++) {
        const int n = 1 << r;
#pragma unroll
        for (int q = 0; q < NP; q++)
            if (lane >= 64 - n)
                dst[Sub<K>::hoff(r, q) + lane - (64 - n)] = h[r * NP + q];
    }
}

// Phase-B delay-line rows of sub-tree j: load / save the part that carries over chunks.
template<int DD, bool LOAD>
__device__ __forceinline__ void move_residue(float *lbuf, float *res, int j, int lane)
{
    using F = FastDims<DD>;
    for (int h = 0; h < F::NRB; h++) {
        const int L = F::lag(h, j), P = F::lagsum(h, j), o = F::off(h, j);
        for (int q = lane; q < L; q += 64) {
            if (LOAD)
                lbuf[o + q] = res[P + q];
            else
                res[P + q] = lbuf[o + q];
        }
    }
}

// Phase B input of sub-tree j: row (h, j) at time t is at delay-line position t; then the
// line is shifted down by 64 (positions [64, 64 + lag) -> [0, lag)). In-order LDS access
// within the wave makes the overlapping shift safe (ascending 64-sample blocks).
template<int DD>
__device__ __forceinline__ void phase_b_read(float *lbuf, int j, float (&y)[FastDims<DD>::NRB], int lane)
{
    using F = FastDims<DD>;
#pragma unroll
    for (int h = 0; h < F::NRB; h++) {
        const int o = F::off(h, j), L = F::lag(h, j);
        y[h] = lbuf[o + lane];
        for (int q = lane; q < L; q += 64)
            lbuf[o + q] = lbuf[o + 64 + q];
    }
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
        default: return 0;
    }
}


// ---------------------------------------------------------------------------------------
// Stage 1: (B, 2^amb trees, 2^dd rows, Tp) -> lag ring rows (tree, final row)

template<typename Tst, int DD>
__global__ void __launch_bounds__(FastDims<DD>::NW * 64)
k_stage1_fast(Stage1Args a)
{
    using F = FastDims<DD>;
    using SA_ = Sub<F::K1>;
    using SB_ = Sub<F::K2>;
    extern __shared__ float lbuf[];
    const int lane = threadIdx.x & 63;
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int ia = blockIdx.x, b = blockIdx.y;
    constexpr int D = 1 << DD;

    float *st = a.state + long(a.beam0 + b) * a.state_bstride + long(ia) * F::NSTATE;
    float *st_b = st + F::NRB * SA_::NST;
    float *st_res = st_b + F::NRA * SB_::NST;

    float ha[F::SA][SA_::NH], hb[F::SB][SB_::NH];
#pragma unroll
    for (int s = 0; s < F::SA; s++)
        load_hist<F::K1>(ha[s], st + (wid + s * F::NW) * SA_::NST, lane);
#pragma unroll
    for (int s = 0; s < F::SB; s++) {
        const int j = wid + s * F::NW;
        load_hist<F::K2>(hb[s], st_b + j * SB_::NST, lane);
        move_residue<DD, true>(lbuf, st_res, j, lane);
    }

    const Tst *in = reinterpret_cast<const Tst *>(a.in) + long(b) * a.in_bstride + long(ia) * D * a.Tp;
    Tst *ring = reinterpret_cast<Tst *>(a.ring.base) + long(a.beam0 + b) * a.ring.bstride;

    // ring write positions of this wave's output rows: (chunk*Tp + t0) mod capacity
    int opos[F::SB][F::NRB];
#pragma unroll
    for (int s = 0; s < F::SB; s++)
#pragma unroll
        for (int hp = 0; hp < F::NRB; hp++) {
            const long rr = long(ia) * D + hp * F::NRA + wid + s * F::NW;
            opos[s][hp] = int((a.chunk * a.Tp) % a.ring.row_cap[rr]);
        }

    for (int t0 = 0; t0 < a.Tp; t0 += 64) {
#pragma unroll
        for (int s = 0; s < F::SA; s++) {
            const int h = wid + s * F::NW;
            float x[F::NRA];
#pragma unroll
            for (int l = 0; l < F::NRA; l++)
                x[l] = ld_st(in + long(h * F::NRA + l) * a.Tp + t0 + lane);
            sub_passes<F::K1>(x, ha[s], lane);
#pragma unroll
            for (int j = 0; j < F::NRA; j++)
                lbuf[F::off(h, j) + F::lag(h, j) + lane] = x[j];
        }
        __syncthreads();

#pragma unroll
        for (int s = 0; s < F::SB; s++) {
            const int j = wid + s * F::NW;
            float y[F::NRB];
            phase_b_read<DD>(lbuf, j, y, lane);
            sub_passes<F::K2>(y, hb[s], lane);
#pragma unroll
            for (int hp = 0; hp < F::NRB; hp++) {
                const long rr = long(ia) * D + hp * F::NRA + j;
                const int cap = int(a.ring.row_cap[rr]);
                int p = opos[s][hp];
                st_st(ring + a.ring.row_off[rr] + p + lane, y[hp]);   // capacities are multiples of 64
                p += 64;
                opos[s][hp] = (p >= cap) ? (p - cap) : p;
            }
        }
        __syncthreads();
    }

#pragma unroll
    for (int s = 0; s < F::SA; s++)
        save_hist<F::K1>(ha[s], st + (wid + s * F::NW) * SA_::NST, lane);
#pragma unroll
    for (int s = 0; s < F::SB; s++) {
        const int j = wid + s * F::NW;
        save_hist<F::K2>(hb[s], st_b + j * SB_::NST, lane);
        move_residue<DD, false>(lbuf, st_res, j, lane);
    }
}


// ---------------------------------------------------------------------------------------
// Stage 2 (final-pass streams only): ring rows (fc, j) at their lag -> peak-finder rows

template<typename Tst, int DD>
__global__ void __launch_bounds__(FastDims<DD>::NW * 64)
k_stage2_fast(Stage2Args a)
{
    using F = FastDims<DD>;
    using SA_ = Sub<F::K1>;
    using SB_ = Sub<F::K2>;
    extern __shared__ float lbuf[];
    const int lane = threadIdx.x & 63;
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int jt = blockIdx.x, b = blockIdx.y;

    float *st = a.state + long(a.beam0 + b) * a.state_bstride + long(jt) * F::NSTATE;
    float *st_b = st + F::NRB * SA_::NST;
    float *st_res = st_b + F::NRA * SB_::NST;

    float ha[F::SA][SA_::NH], hb[F::SB][SB_::NH];
#pragma unroll
    for (int s = 0; s < F::SA; s++)
        load_hist<F::K1>(ha[s], st + (wid + s * F::NW) * SA_::NST, lane);
#pragma unroll
    for (int s = 0; s < F::SB; s++) {
        const int j = wid + s * F::NW;
        load_hist<F::K2>(hb[s], st_b + j * SB_::NST, lane);
        move_residue<DD, true>(lbuf, st_res, j, lane);
    }

    const Tst *ring = reinterpret_cast<const Tst *>(a.ring.base) + long(a.beam0 + b) * a.ring.bstride;
    Tst *sbb = reinterpret_cast<Tst *>(a.sb) + long(a.beam0 + b) * a.sb_bstride
               + long(bitrev_dev(jt, a.amb)) * (1 << (DD - a.pf_rank)) * a.M * a.rowlen + a.tpad;

    // ring read positions of this wave's input rows: (chunk*Tp - lag) mod capacity
    int ipos[F::SA][F::NRA];
#pragma unroll
    for (int s = 0; s < F::SA; s++)
#pragma unroll
        for (int l = 0; l < F::NRA; l++) {
            const long rr = (long((wid + s * F::NW) * F::NRA + l) << a.amb) + jt;
            ipos[s][l] = int(ring_pos(a.chunk * a.Tp - a.row_lag[rr], a.ring.row_cap[rr]));
        }

    for (int t0 = 0; t0 < a.Tp; t0 += 64) {
#pragma unroll
        for (int s = 0; s < F::SA; s++) {
            const int h = wid + s * F::NW;
            float x[F::NRA];
#pragma unroll
            for (int l = 0; l < F::NRA; l++) {
                const long rr = (long(h * F::NRA + l) << a.amb) + jt;
                const int cap = int(a.ring.row_cap[rr]);
                int p = ipos[s][l] + lane;
                p = (p >= cap) ? (p - cap) : p;
                x[l] = ld_st(ring + a.ring.row_off[rr] + p);
                int pn = ipos[s][l] + 64;
                ipos[s][l] = (pn >= cap) ? (pn - cap) : pn;
            }
            sub_passes<F::K1>(x, ha[s], lane);
#pragma unroll
            for (int j = 0; j < F::NRA; j++)
                lbuf[F::off(h, j) + F::lag(h, j) + lane] = x[j];
        }
        __syncthreads();

#pragma unroll
        for (int s = 0; s < F::SB; s++) {
            const int j = wid + s * F::NW;
            float y[F::NRB];
            phase_b_read<DD>(lbuf, j, y, lane);
            sub_passes<F::K2>(y, hb[s], lane);
#pragma unroll
            for (int hp = 0; hp < F::NRB; hp++) {
                const int sidx = a.out_stream[hp * F::NRA + j];
                if (sidx >= 0)
                    st_st(sbb + long(sidx) * a.rowlen + t0 + a.fl[sidx] + lane, y[hp]);
            }
        }
        __syncthreads();
    }

#pragma unroll
    for (int s = 0; s < F::SA; s++)
        save_hist<F::K1>(ha[s], st + (wid + s * F::NW) * SA_::NST, lane);
#pragma unroll
    for (int s = 0; s < F::SB; s++) {
        const int j = wid + s * F::NW;
        save_hist<F::K2>(hb[s], st_b + j * SB_::NST, lane);
        move_residue<DD, false>(lbuf, st_res, j, lane);
    }
}


template<typename Tst>
static void launch_stage1_fast_t(const Stage1Args &a, hipStream_t st)
{
#define PDD_F1(DDV)                                                                          \
    case DDV: {                                                                              \
        using F = FastDims<DDV>;                                                             \
        hipLaunchKernelGGL((k_stage1_fast<Tst, DDV>), dim3(1 << a.amb, a.B), dim3(F::NW * 64), \
                           sizeof(float) * F::LDS_FLOATS, st, a);                            \
        break;                                                                               \
    }
    switch (a.dd) {
        PDD_F1(1) PDD_F1(2) PDD_F1(3) PDD_F1(4) PDD_F1(5) PDD_F1(6) PDD_F1(7)
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
        hipLaunchKernelGGL((k_stage2_fast<Tst, DDV>), dim3(1 << a.amb, a.B), dim3(F::NW * 64), \
                           sizeof(float) * F::LDS_FLOATS, st, a);                            \
        break;                                                                               \
    }
    switch (a.dd) {
        PDD_F2(1) PDD_F2(2) PDD_F2(3) PDD_F2(4) PDD_F2(5) PDD_F2(6) PDD_F2(7)
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

void launch_stage2_fast(Dtype dt, const Stage2Args &a, hipStream_t st)
{
    switch (dt) {
        case Dtype::float32: launch_stage2_fast_t<float>(a, st); break;
        case Dtype::float16: launch_stage2_fast_t<__half>(a, st); break;
        case Dtype::bfloat16: launch_stage2_fast_t<__hip_bfloat16>(a, st); break;
    }
}

}  // namespace pirate_amd
