// Peak finding on gfx950 over the subband streams of one output tree.
// Reference semantics: ReferencePeakFindingKernel::apply (src_lib/PeakFindingKernel.cu:252-456):
//
//   boxcars   B_0(e) = x[e-1],  B_l(e) = B_{l-1}(e - 2^(l-1)) + B_{l-1}(e)   (binary-tree sums)
//   candidate (level l, multiplet m, isamp, profile p) ends at e = tout*Dout + (isamp+1)*dt,
//             dt = min(Dcore, 2^l); with x_k = B_l(e - (3-k)*2^l):
//               p=0 (l=0): x3      p=3l+1: x2+x3      p=3l+2: (x1/2 + x2) + x3/2
//               p=3l+3: ((x0/2 + x1) + x2) + x3/2
//   out_max[tout] = max over (l, m, isamp, p) of w[p, n(m)] * y; the first maximum in the loop
//   order (l, m, isamp, p) wins ties; token = t | (p << 8) | (m << 16), t = isamp*dt.
//
// Mapping: one lane per stream (DM row, multiplet m); a block holds 256/M DM rows. Each lane
// walks its stream through time once. Every boxcar level keeps its last values in a register
// window (Wmax and Dout are template parameters, so every shift, sum and candidate of one
// output bin is unrolled at compile time). Each candidate is weighted as it is formed and
// compared with the level's running best (strict '>', candidates visited in (isamp, p)
// order), so the lane ends each bin with the exact first maximum per level. Levels are then
// combined in order (strict '>'), and the multiplets of a DM row are reduced through LDS in
// (level, m) order: the result is the reference's maximum and argmax token, bit for bit.
//
// Samples are staged through LDS in tiles of TL positions (coalesced 4-element loads);
// tiles are aligned to output bins. Positions before the row start are read as zeros,
// which equals the zero initial state of the windows.
//
// The chunk's time axis is split into nseg segments (blockIdx.z), one workgroup each, so
// that short chunks still fill the chip. A segment's lanes start from zero windows and run
// the tpad positions before it unevaluated: every window entry at the segment start is a
// sum over at most the last tpad samples (level l holds boxcar sums of 2^l samples ending at
// most 3*2^l positions back, and 4*2^l <= 2*Wmax <= tpad), so the windows, and hence every
// candidate, equal those of one lane walking the whole chunk. The running variance sums
// depend on the order of the bins, so they use one segment.
//
// Rows are [tpad history | Tp samples | future (final lags)]; at the end of a chunk the
// history and future regions are rotated into place for the next chunk.

#include "device_common.hpp"
#include "kernels.hpp"

namespace pirate_amd {

static constexpr int kPfThreads = 256;
static constexpr int kPfMaxDout = 32;

// Staging tile length (positions) for output time downsampling Dout: whole output bins, at
// least kPfTileMin positions when Dout >= 4, 8 bins below.
static constexpr int kPfTileMin = 16;
__host__ __device__ constexpr int pf_tile_len(int dout)
{
    return (dout >= 4) ? ((dout > kPfTileMin) ? dout : kPfTileMin) : 8 * dout;
}

bool pf_config_supported(int M, int tpad, int nlev, int Dcore, int Dout, int Tp)
{
    (void) nlev;
    if (M < 1 || M > kPfThreads || Dout < 1 || Dout > kPfMaxDout || (Dout & (Dout - 1)))
        return false;
    if (tpad % 4)
        return false;
    const int dc = (Dout < 8) ? Dout : 8;
    return (Dcore == dc) && (Tp % 32 == 0);
}

template<int LW, int LD>
struct PfShape {
    static constexpr int W = 1 << LW;
    static constexpr int NLEV = (LW > 1) ? LW : 1;
    static constexpr int P = 1 + 3 * LW;
    static constexpr int DOUT = 1 << LD;
    static constexpr int DCORE = (DOUT < 8) ? DOUT : 8;
    static constexpr int TL = pf_tile_len(DOUT);                 // tile length (positions)
    static constexpr int NB = TL / DOUT;                          // output bins per tile
    static constexpr int dt(int l) { return ((1 << l) < DCORE) ? (1 << l) : DCORE; }
    static constexpr int sgrid(int l) { return (1 << l) / dt(l); }   // grid points per 2^l samples
    static constexpr int win(int l) { return 3 * sgrid(l) + 1; }
    static constexpr int woff(int l) { return (l == 0) ? 0 : woff(l - 1) + win(l - 1); }
    static constexpr int WTOT = woff(NLEV);
    // Occupancy target: 4 waves/SIMD (128 VGPRs) when the register windows are small,
    // else 2 (the compiler's unrolled schedule would otherwise spill).
    static constexpr int WAVES = (WTOT + P <= 32 && LD >= 2) ? 4 : 2;
};

template<typename S, bool VAR>
struct PfLane {
    float w[S::WTOT];          // per level l: w[woff(l) + i] = B_l(e - i*dt(l)), e = newest end
    float bv[S::NLEV];         // per level: best weighted value in the current bin
    unsigned int bc[S::NLEV];  // per level: its code (see code())
    float wt[S::P];            // weights of the current bin, per profile
    // VAR: per profile, the sum of squared weighted candidates in the current bin (bs) and
    // the chunk's running sum of bs / (nt_out * candidates per bin at that level) (acc),
    // in float64 -- the reference peak finder's optional out_var.
    double bs[VAR ? S::P : 1], acc[VAR ? S::P : 1], wv[S::NLEV];

    template<int P>
    __device__ __forceinline__ void sq(float v)
    {
        if constexpr (VAR)
            bs[P] = bs[P] + double(v) * double(v);
    }

    // End of an output bin: fold the bin's sums into the chunk's (same order as the reference).
    __device__ __forceinline__ void fold_var()
    {
        if constexpr (VAR) {
#pragma unroll
            for (int p = 0; p < S::P; p++) {
                acc[p] = acc[p] + bs[p] * wv[(p == 0) ? 0 : (p - 1) / 3];
                bs[p] = 0.0;
            }
        }
    }

    // Candidate codes within a level. Compact codes isamp*4 + q (q = 0 for p = 0, else
    // p - 3l) stay within the hardware's inline-constant range (<= 64) when Dout <= 16, so the
    // per-candidate selects need no literal moves; token() expands them once per bin.
    static constexpr bool COMPACT = (S::DOUT <= 16);

    template<int L>
    static constexpr unsigned int code(int isamp, int q)
    {
        return COMPACT ? unsigned(4 * isamp + q)
                       : (unsigned(isamp * S::dt(L)) | (unsigned(q ? 3 * L + q : 0) << 8));
    }

    template<int L>
    __device__ __forceinline__ unsigned int token(unsigned int c) const
    {
        if constexpr (COMPACT) {
            const unsigned int q = c & 3u, isamp = c >> 2;
            const unsigned int p = q ? (3u * L + q) : 0u;
            return (isamp * unsigned(S::dt(L))) | (p << 8);
        } else {
            return c;
        }
    }

    template<int L>
    __device__ __forceinline__ void push(float v)
    {
        constexpr int o = S::woff(L), n = S::win(L);
#pragma unroll
        for (int i = n - 1; i > 0; i--)
            w[o + i] = w[o + i - 1];
        w[o] = v;
    }

    template<int L, bool FIRST>
    __device__ __forceinline__ void upd(float v, unsigned int code)
    {
        if constexpr (FIRST) {
            bv[L] = v;
            bc[L] = code;
        } else {
            const bool g = v > bv[L];
            bv[L] = g ? v : bv[L];
            bc[L] = g ? code : bc[L];
        }
    }

    template<int L, int POS>
    __device__ __forceinline__ void candidates()
    {
        constexpr int o = S::woff(L), s = S::sgrid(L);
        constexpr int isamp = (POS + 1) / S::dt(L) - 1;
        constexpr bool first = (isamp == 0);
        const float x3 = w[o], x2 = w[o + s], x1 = w[o + 2 * s], x0 = w[o + 3 * s];
        if constexpr (L == 0) {
            const float v0 = wt[0] * x3;
            upd<0, first>(v0, code<0>(isamp, 0));
            sq<0>(v0);
        }
        if constexpr (S::P > 1) {
            constexpr int p = 3 * L + 1;
            const float h3 = 0.5f * x3;
            const float y1 = x2 + x3;
            const float y2 = (0.5f * x1 + x2) + h3;
            const float y3 = ((0.5f * x0 + x1) + x2) + h3;
            const float v1 = wt[p] * y1, v2 = wt[p + 1] * y2, v3 = wt[p + 2] * y3;
            upd<L, first && (L > 0)>(v1, code<L>(isamp, 1));
            upd<L, false>(v2, code<L>(isamp, 2));
            upd<L, false>(v3, code<L>(isamp, 3));
            sq<p>(v1);
            sq<p + 1>(v2);
            sq<p + 2>(v3);
        }
    }

    template<int L, int POS, bool EVAL>
    __device__ __forceinline__ void levels()
    {
        if constexpr (L < S::NLEV) {
            if constexpr ((POS + 1) % S::dt(L) == 0) {
                if constexpr (L > 0) {
                    constexpr int op = S::woff(L - 1), back = S::sgrid(L - 1);
                    push<L>(w[op + back] + w[op]);
                }
                if constexpr (EVAL)
                    candidates<L, POS>();
                levels<L + 1, POS, EVAL>();
            }
        }
    }

    // One output bin (DOUT positions); xs[POS] is the sample at in-bin position POS.
    template<bool EVAL, int POS = 0>
    __device__ __forceinline__ void bin(const float *xs)
    {
        if constexpr (POS < S::DOUT) {
            push<0>(xs[POS]);
            levels<0, POS, EVAL>();
            bin<EVAL, POS + 1>(xs);
        }
    }

    template<int L = 1>
    __device__ __forceinline__ void combine_from(float &V, unsigned int &C) const
    {
        if constexpr (L < S::NLEV) {
            const bool g = bv[L] > V;
            V = g ? bv[L] : V;
            C = g ? token<L>(bc[L]) : C;
            combine_from<L + 1>(V, C);
        }
    }

    // Levels in order, strict '>': the first level with the bin's maximum; C is its token.
    __device__ __forceinline__ void combine(float &V, unsigned int &C) const
    {
        V = bv[0];
        C = token<0>(bc[0]);
        combine_from<1>(V, C);
    }
};

__device__ __forceinline__ int pf_level_of_token(unsigned int tok)
{
    const int p = int((tok >> 8) & 0xffu);
    return (p == 0) ? 0 : (p - 1) / 3;
}

template<typename Tst, int LW, int LD, bool VAR>
__global__ void __launch_bounds__(kPfThreads) __attribute__((amdgpu_waves_per_eu(VAR ? 2 : PfShape<LW, LD>::WAVES)))
k_pf(PfArgs a)
{
    using S = PfShape<LW, LD>;
    constexpr int TL = S::TL, NB = S::NB, STR = TL + 1, VPR = TL / 4;
    __shared__ float tile[kPfThreads * STR];
    __shared__ float rv[kPfThreads * NB];
    __shared__ unsigned int rt[kPfThreads * NB];

    const int tid = threadIdx.x;
    const int M = a.M, N = a.N;
    const int ndpf = kPfThreads / M;
    const int dpf0 = blockIdx.x * ndpf;
    const int b = blockIdx.y;
    const int ntseg = a.Tp / (TL * a.nseg);      // evaluated tiles of this segment
    const int tile0 = blockIdx.z * ntseg;        // its first tile in the chunk
    const int nd = min(ndpf, a.ndm_out - dpf0);
    const int nstreams = nd * M;
    const bool active = tid < nstreams;
    const int dl = tid / M, m = tid - dl * M;

    Tst *rows = reinterpret_cast<Tst *>(a.sb) + long(a.beam0 + b) * a.sb_bstride + long(dpf0) * M * a.rowlen;
    const int Wds = a.ndm_out / a.ndm_wt;
    const int Tds = a.nt_out / a.nt_wt;
    const float *wrow = a.wt + long(b) * a.wt_bstride + long(active ? (dpf0 + dl) / Wds : 0) * a.nt_wt * S::P * N
                        + (active ? a.m_to_n[m] : 0);
    Tst *outm = reinterpret_cast<Tst *>(a.out_max) + long(b) * a.out_bstride;
    unsigned int *outa = a.out_argmax + long(b) * a.out_bstride;

    PfLane<S, VAR> ln;
#pragma unroll
    for (int i = 0; i < S::WTOT; i++)
        ln.w[i] = 0.0f;
    if constexpr (VAR) {
#pragma unroll
        for (int p = 0; p < S::P; p++)
            ln.bs[p] = ln.acc[p] = 0.0;
#pragma unroll
        for (int l = 0; l < S::NLEV; l++)
            ln.wv[l] = 1.0 / double(a.nt_out * (S::DOUT / S::dt(l)));
    }

    const int nwarm = (a.tpad + TL - 1) / TL;
    const int ntiles = nwarm + ntseg;
    const float *xs = tile + tid * STR;

    for (int it = 0; it < ntiles; it++) {
        const int p0 = a.tpad + (tile0 + it - nwarm) * TL;

        // ---- stage [nstreams][TL] samples (4 per load, VPR loads per row segment); a thread
        // issues all of its loads before it converts and stores them (with the conversion at
        // each load, the 16-bit variants waited for each load before issuing the next)
        rawv_st<Tst, 4> raw[VPR];
#pragma unroll
        for (int k = 0; k < VPR; k++) {
            const int idx = tid + k * kPfThreads;
            const int s = idx / VPR, i = (idx - s * VPR) * 4;
            raw[k] = (s < nstreams && p0 + i >= 0) ? ldv_raw<Tst, 4>(rows + long(s) * a.rowlen + p0 + i)
                                                  : rawv_st<Tst, 4>{};
        }
#pragma unroll
        for (int k = 0; k < VPR; k++) {
            const int idx = tid + k * kPfThreads;
            const int s = idx / VPR, i = (idx - s * VPR) * 4;
            if (s < nstreams)
                from_rawv<Tst, 4>(raw[k], tile + s * STR + i);
        }
        __syncthreads();

        if (it < nwarm) {
            if (active) {
#pragma unroll 1
                for (int j = 0; j < NB; j++)
                    ln.template bin<false>(xs + j * S::DOUT);
            }
        } else {
            const int tb0 = (tile0 + it - nwarm) * NB;
            if (active) {
#pragma unroll 1
                for (int j = 0; j < NB; j++) {
                    const float *wp = wrow + long((tb0 + j) / Tds) * S::P * N;
#pragma unroll
                    for (int p = 0; p < S::P; p++)
                        ln.wt[p] = wp[p * N];
                    ln.template bin<true>(xs + j * S::DOUT);
                    ln.fold_var();
                    float V;
                    unsigned int C;
                    ln.combine(V, C);
                    rv[tid * NB + j] = V;
                    rt[tid * NB + j] = C | (unsigned(m) << 16);
                }
            }
        }
        __syncthreads();

        // ---- reduce over multiplets: first (level, m) in loop order with the largest value
        if (it >= nwarm) {
            const int tb0 = (tile0 + it - nwarm) * NB;
            for (int idx = tid; idx < nd * NB; idx += kPfThreads) {
                const int d = idx / NB, j = idx - d * NB;
                const int s0 = d * M;
                float bv = rv[s0 * NB + j];
                unsigned int bt = rt[s0 * NB + j];
                int bk = pf_level_of_token(bt) * M;
                for (int mm = 1; mm < M; mm++) {
                    const float v = rv[(s0 + mm) * NB + j];
                    const unsigned int tk = rt[(s0 + mm) * NB + j];
                    const int key = pf_level_of_token(tk) * M + mm;
                    if (v > bv || (v == bv && key < bk)) {
                        bv = v;
                        bt = tk;
                        bk = key;
                    }
                }
                const long o = long(dpf0 + d) * a.nt_out + tb0 + j;
                st_st(outm + o, bv);
                outa[o] = bt;
            }
        }
    }

    if constexpr (VAR) {
        if (active) {
            double *ov = a.out_var + long(b) * a.ndm_out * M * S::P + (long(dpf0 + dl) * M + m) * S::P;
#pragma unroll
            for (int p = 0; p < S::P; p++)
                ov[p] = ln.acc[p];
        }
    }

    // ---- rotate rows for the next chunk: positions [Tp, Tp + tpad + flmax) -> [0, ...),
    // i.e. the rows shift down by Tp. All staging loads completed before the last barrier.
    // Done in column blocks of width min(Tp, tpad + flmax) in ascending order, with a
    // barrier between blocks: within a block the positions read ([c0 + Tp, ...)) and
    // written ([c0, c0 + width)) are disjoint, and a later block reads only positions at or
    // above every position an earlier block wrote. Segment 0 rotates; the host chooses
    // segments of at least nwarm*TL + flmax positions, so no other segment reads below
    // tpad + flmax, the positions written here.
    if (blockIdx.z != 0)
        return;
    const int nrot = a.tpad + a.flmax;
    const int wcol = min(a.Tp, nrot);
    for (int c0 = 0; c0 < nrot; c0 += wcol) {
        const int nc = min(wcol, nrot - c0);
        for (int idx = tid; idx < nstreams * nc; idx += kPfThreads) {
            const int s = idx / nc, i = c0 + (idx - s * nc);
            Tst *row = rows + long(s) * a.rowlen;
            const Tst v = row[a.Tp + i];
            row[i] = v;
        }
        if (c0 + wcol < nrot)
            __syncthreads();
    }
}

template<typename Tst, int LW>
static void launch_pf_lw(const PfArgs &a, hipStream_t st, dim3 grid)
{
    dim3 block(kPfThreads);
    switch (a.Dout) {
#define PDD_PF(LDV)                                                                \
        case (1 << LDV):                                                           \
            if (a.out_var)                                                         \
                hipLaunchKernelGGL((k_pf<Tst, LW, LDV, true>), grid, block, 0, st, a);  \
            else                                                                   \
                hipLaunchKernelGGL((k_pf<Tst, LW, LDV, false>), grid, block, 0, st, a); \
            break;
        PDD_PF(0) PDD_PF(1) PDD_PF(2) PDD_PF(3) PDD_PF(4) PDD_PF(5)
#undef PDD_PF
        default: break;
    }
}

// Time segments per chunk: the fewest (a power of two) that give at least a.target_blocks
// workgroups (2048 unless tuned), within segments of whole tiles of at least
// nwarm*TL + flmax positions (see the rotation in k_pf). With out_var, one segment.
static int pf_nseg(const PfArgs &a, long nblocks)
{
    if (a.out_var)
        return 1;
    const int TL = pf_tile_len(a.Dout);
    const int nwarm = (a.tpad + TL - 1) / TL;
    int nseg = 1;
    while (nblocks * nseg < a.target_blocks) {
        const int n2 = 2 * nseg;
        if (a.Tp % (TL * n2) != 0 || a.Tp / n2 < nwarm * TL + a.flmax)
            break;
        nseg = n2;
    }
    return nseg;
}

template<typename Tst>
static void launch_pf_t(const PfArgs &a0, hipStream_t st)
{
    PfArgs a = a0;
    const int ndpf = kPfThreads / a.M;
    const unsigned nbx = unsigned((a.ndm_out + ndpf - 1) / ndpf);
    a.nseg = pf_nseg(a, long(nbx) * a.B);
    dim3 grid{nbx, unsigned(a.B), unsigned(a.nseg)};
    switch ((a.P - 1) / 3) {
        case 0: launch_pf_lw<Tst, 0>(a, st, grid); break;
        case 1: launch_pf_lw<Tst, 1>(a, st, grid); break;
        case 2: launch_pf_lw<Tst, 2>(a, st, grid); break;
        case 3: launch_pf_lw<Tst, 3>(a, st, grid); break;
        case 4: launch_pf_lw<Tst, 4>(a, st, grid); break;
        case 5: launch_pf_lw<Tst, 5>(a, st, grid); break;
        default: break;
    }
}

void launch_pf(Dtype dt, const PfArgs &a, hipStream_t st)
{
    switch (dt) {
        case Dtype::float32: launch_pf_t<float>(a, st); break;
        case Dtype::float16: launch_pf_t<__half>(a, st); break;
        case Dtype::bfloat16: launch_pf_t<__hip_bfloat16>(a, st); break;
    }
}

}  // namespace pirate_amd
