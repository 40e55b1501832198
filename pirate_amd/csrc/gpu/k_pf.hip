// Peak finding on gfx950 over the subband streams of one output tree.
// Reference semantics: ReferencePeakFindingKernel::apply (src_lib/PeakFindingKernel.cu:252-456):
//
//   boxcars   B_0(e) = x[e-1],  B_l(e) = B_{l-1}(e - 2^(l-1)) + B_{l-1}(e)   (binary-tree sums)
//   candidate (level l, multiplet m, isamp, profile p) ends at e = tout*Dout + (isamp+1)*dt,
//             dt = min(Dcore, 2^l); with x_k = B_l(e - (3-k)*2^l):
//               p=0 (l=0): x3      p=3l+1: x2+x3      p=3l+2: (x1/2 + x2) + x3/2
//               p=3l+3: ((x0/2 + x1) + x2) + x3/2
//   out_max[tout] = max over (l, m, isamp, p) of w[p, n(m)] * y,  first maximum in the loop
//   order (l, m, isamp, p) wins ties; token = t | (p << 8) | (m << 16), t = isamp*dt.
//
// Algorithm (bit-exact with the float32 reference, for weights >= 0):
//   1. Per tile, the M streams of one DM row are staged in LDS and every boxcar level is
//      built there with the same binary-tree association as the reference.
//   2. Group maxima G(l,m,bin) = max_p w_p * max_isamp y_p. Rounding is monotonic, so for
//      w_p >= 0 this equals the maximum of the individually rounded products w_p*y.
//   3. Per bin, the first (l, m) in loop order with the largest G is the winning group;
//      a short scan of that group in (isamp, p) order finds the first candidate whose
//      rounded product equals the maximum: that candidate is the reference's argmax.
//
// Rows are [tpad history | Tp samples | future] (see the stage-2 kernel); at the end of a
// chunk the history and future regions are rotated into place for the next chunk.

#include "device_common.hpp"
#include "kernels.hpp"

#include <climits>

namespace pirate_amd {

static constexpr int kPfThreads = 256;
static constexpr int kPfMaxLev = 6;

struct PfGeom {
    int TT;                      // tile length (PF samples)
    int nbins;                   // TT / Dout
    int n[kPfMaxLev];            // entries per level
    int off[kPfMaxLev];          // offset of level l within a stream's level block
    int per_stream;              // floats per stream
};

__host__ __device__ inline PfGeom pf_geom(int TT, int tpad, int nlev, int Dcore, int Dout)
{
    PfGeom g;
    g.TT = TT;
    g.nbins = TT / Dout;
    int o = 0;
    for (int l = 0; l < kPfMaxLev; l++) {
        if (l < nlev) {
            int w = 1 << l, dt = (Dcore < w) ? Dcore : w;
            g.n[l] = (TT + tpad - w) / dt + 1;
            g.off[l] = o;
            o += g.n[l];
        }
        else {
            g.n[l] = 0;
            g.off[l] = o;
        }
    }
    g.per_stream = o;
    return g;
}

static int pf_pick_tt(int M, int tpad, int nlev, int Dcore, int Dout, int Tp)
{
    int best = 0;
    for (int TT = Dout; TT <= Tp && TT <= 512; TT *= 2) {
        if (Tp % TT)
            continue;
        PfGeom g = pf_geom(TT, tpad, nlev, Dcore, Dout);
        long bytes = sizeof(float) * (long(M) * g.per_stream + long(g.nbins) * nlev * M) + 8 * g.nbins;
        if (bytes <= 96 * 1024)
            best = TT;
    }
    return best;
}

int pf_lds_bytes(int M, int tpad, int nlev, int Dcore)
{
    (void) Dcore;
    return int(sizeof(float) * long(M) * 4 * tpad);   // nominal; the kernel picks its own tile
}

// Level-l boxcar ending at PF time e, read from the stream's level block in LDS.
__device__ __forceinline__ float lv_at(const float *blk, const PfGeom &g, int l, int e, int tile_t0, int tpad, int Dcore)
{
    const int w = 1 << l, dt = (Dcore < w) ? Dcore : w;
    const int i = (e - (tile_t0 - tpad + w)) / dt;
    return blk[g.off[l] + i];
}

template<typename Tst>
__global__ void __launch_bounds__(kPfThreads)
k_pf(PfArgs a, PfGeom g)
{
    extern __shared__ float smem[];
    const int M = a.M, nlev = a.nlev, P = a.P, N = a.N;
    float *lv = smem;                                        // [M][per_stream]
    float *Gv = lv + M * g.per_stream;                       // [nbins][nlev][M]
    float *bval = Gv + g.nbins * nlev * M;                   // [nbins]
    int *bkey = reinterpret_cast<int *>(bval + g.nbins);     // [nbins]

    const int dpf = blockIdx.x, b = blockIdx.y, tid = threadIdx.x;
    Tst *rows = reinterpret_cast<Tst *>(a.sb) + long(a.beam0 + b) * a.sb_bstride + long(dpf) * M * a.rowlen;
    const float *wtb = a.wt + long(b) * a.wt_bstride;
    const int Wds = a.ndm_out / a.ndm_wt;
    const int Tds = a.nt_out / a.nt_wt;
    const float *wrow_d = wtb + long(dpf / Wds) * a.nt_wt * P * N;

    for (int t0 = 0; t0 < a.Tp; t0 += g.TT) {
        // ---- stage level 0: samples at PF times [t0 - tpad, t0 + TT) = row positions [t0, t0 + TT + tpad)
        const int n0 = g.n[0];
        for (int idx = tid; idx < M * n0; idx += kPfThreads) {
            const int m = idx / n0, i = idx % n0;
            lv[m * g.per_stream + i] = ld_st(rows + long(m) * a.rowlen + t0 + i);
        }
        __syncthreads();

        // ---- build boxcar levels: B_l(e) = B_{l-1}(e - 2^(l-1)) + B_{l-1}(e)
        for (int l = 1; l < nlev; l++) {
            const int w = 1 << l, dt = min(a.Dcore, w);
            const int wp = w >> 1, dtp = min(a.Dcore, wp);
            const int nl = g.n[l];
            for (int idx = tid; idx < M * nl; idx += kPfThreads) {
                const int m = idx / nl, i = idx % nl;
                float *blk = lv + m * g.per_stream;
                const int e = t0 - a.tpad + w + i * dt;
                const int i1 = (e - wp - (t0 - a.tpad + wp)) / dtp;
                const int i2 = (e - (t0 - a.tpad + wp)) / dtp;
                blk[g.off[l] + i] = blk[g.off[l - 1] + i1] + blk[g.off[l - 1] + i2];
            }
            __syncthreads();
        }

        // ---- group maxima G(l, m, bin)
        const int nitems = g.nbins * nlev * M;
        for (int idx = tid; idx < nitems; idx += kPfThreads) {
            const int m = idx % M;
            const int l = (idx / M) % nlev;
            const int bin = idx / (M * nlev);
            const int tout = t0 / a.Dout + bin;
            const int n = a.m_to_n[m];
            const float *wp = wrow_d + long(tout / Tds) * P * N;
            const float *blk = lv + m * g.per_stream;
            const int w = 1 << l, dt = min(a.Dcore, w);
            const int nsamp = a.Dout / dt;
            float m0 = -INFINITY, m1 = -INFINITY, m2 = -INFINITY, m3 = -INFINITY;
            for (int is = 0; is < nsamp; is++) {
                const int e = tout * a.Dout + (is + 1) * dt;
                const float x3 = lv_at(blk, g, l, e, t0, a.tpad, a.Dcore);
                if (l == 0)
                    m0 = fmaxf(m0, x3);
                if (P > 1) {
                    const float x2 = lv_at(blk, g, l, e - w, t0, a.tpad, a.Dcore);
                    const float x1 = lv_at(blk, g, l, e - 2 * w, t0, a.tpad, a.Dcore);
                    const float x0 = lv_at(blk, g, l, e - 3 * w, t0, a.tpad, a.Dcore);
                    const float y1 = x2 + x3;
                    const float y2 = (0.5f * x1 + x2) + 0.5f * x3;
                    const float y3 = ((0.5f * x0 + x1) + x2) + 0.5f * x3;
                    m1 = fmaxf(m1, y1);
                    m2 = fmaxf(m2, y2);
                    m3 = fmaxf(m3, y3);
                }
            }
            float G = -INFINITY;
            if (l == 0)
                G = fmaxf(G, wp[0 * N + n] * m0);
            if (P > 1) {
                G = fmaxf(G, wp[(3 * l + 1) * N + n] * m1);
                G = fmaxf(G, wp[(3 * l + 2) * N + n] * m2);
                G = fmaxf(G, wp[(3 * l + 3) * N + n] * m3);
            }
            Gv[(bin * nlev + l) * M + m] = G;
        }
        __syncthreads();

        // ---- per bin: winning group (first (l, m) in loop order), then the winning candidate
        for (int bin = tid; bin < g.nbins; bin += kPfThreads) {
            float best = -1.0e30f;
            int key = -1;
            for (int l = 0; l < nlev; l++)
                for (int m = 0; m < M; m++) {
                    const float v = Gv[(bin * nlev + l) * M + m];
                    if (v > best) {
                        best = v;
                        key = l * M + m;
                    }
                }

            const int tout = t0 / a.Dout + bin;
            unsigned int token = ~0u;
            float outv = -1.0e30f;
            if (key >= 0) {
                const int l = key / M, m = key % M;
                const int n = a.m_to_n[m];
                const float *wp = wrow_d + long(tout / Tds) * P * N;
                const float *blk = lv + m * g.per_stream;
                const int w = 1 << l, dt = min(a.Dcore, w);
                const int nsamp = a.Dout / dt;
                bool found = false;
                for (int is = 0; is < nsamp && !found; is++) {
                    const int e = tout * a.Dout + (is + 1) * dt;
                    const float x3 = lv_at(blk, g, l, e, t0, a.tpad, a.Dcore);
                    const unsigned int tk0 = (unsigned(m) << 16) | unsigned(is * dt);
                    if (l == 0) {
                        const float v = wp[0 * N + n] * x3;
                        if (v == best) { found = true; token = tk0; outv = v; }
                    }
                    if (!found && P > 1) {
                        const float x2 = lv_at(blk, g, l, e - w, t0, a.tpad, a.Dcore);
                        const float x1 = lv_at(blk, g, l, e - 2 * w, t0, a.tpad, a.Dcore);
                        const float x0 = lv_at(blk, g, l, e - 3 * w, t0, a.tpad, a.Dcore);
                        const float y[3] = {x2 + x3, (0.5f * x1 + x2) + 0.5f * x3, ((0.5f * x0 + x1) + x2) + 0.5f * x3};
                        for (int q = 0; q < 3 && !found; q++) {
                            const int p = 3 * l + 1 + q;
                            const float v = wp[p * N + n] * y[q];
                            if (v == best) { found = true; token = tk0 | (unsigned(p) << 8); outv = v; }
                        }
                    }
                }
            }
            Tst *om = reinterpret_cast<Tst *>(a.out_max) + long(b) * a.out_bstride + long(dpf) * a.nt_out + tout;
            st_st(om, outv);
            a.out_argmax[long(b) * a.out_bstride + long(dpf) * a.nt_out + tout] = token;
        }
        __syncthreads();
    }

    // ---- rotate each row for the next chunk: history <- last tpad samples, future -> start
    for (int idx = tid; idx < M * (a.tpad + a.flmax); idx += kPfThreads) {
        const int m = idx / (a.tpad + a.flmax), i = idx % (a.tpad + a.flmax);
        Tst *row = rows + long(m) * a.rowlen;
        // positions [Tp, Tp + tpad + flmax) -> [0, tpad + flmax); requires Tp >= tpad + flmax
        const Tst v = row[a.Tp + i];
        row[i] = v;
    }
}

template<typename Tst>
static void launch_pf_t(const PfArgs &a, hipStream_t st)
{
    const int TT = pf_pick_tt(a.M, a.tpad, a.nlev, a.Dcore, a.Dout, a.Tp);
    PfGeom g = pf_geom(TT, a.tpad, a.nlev, a.Dcore, a.Dout);
    size_t shm = sizeof(float) * (size_t(a.M) * g.per_stream + size_t(g.nbins) * a.nlev * a.M) + 8 * g.nbins;
    dim3 grid(a.ndm_out, a.B);
    dim3 block(kPfThreads);
    hipFuncSetAttribute((const void *) k_pf<Tst>, hipFuncAttributeMaxDynamicSharedMemorySize, int(shm));
    hipLaunchKernelGGL(k_pf<Tst>, grid, block, shm, st, a, g);
}

bool pf_config_supported(int M, int tpad, int nlev, int Dcore, int Dout, int Tp)
{
    return (nlev <= kPfMaxLev) && pf_pick_tt(M, tpad, nlev, Dcore, Dout, Tp) > 0;
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
