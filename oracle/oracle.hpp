// ORACLE — test infrastructure only. A CPU restatement of the reference's dedispersion
// chain, used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
// checker. Nothing in the product (pirate_amd/) may include, link or call this code.
//
// Every component is templated on the arithmetic type T (float for parity with the
// reference's float32 CPU path, double for the FP64 precision oracle). Each function
// cites the reference file:line whose semantics it restates. The code is written from
// those semantics, not copied.
//
// Pinning (what checks each part, see tests/test_oracle_pinning.py):
//   RefTree (no subbands)  -> brute-force delay sums (dedispersion_delay) and the
//                             non-incremental whole-array tree; one-hot impulse responses.
//   RefTree (subbands)     -> independent sub-trees per subband (reference test_subbands).
//   RefLagbuf              -> whole-array shift over concatenated chunks.
//   gridding               -> overlap-weight closed form, intensity conservation.
//   lagged downsampling    -> soph0 (plain downsample + one-stage tree) == soph1.
//   peak finder            -> closed-form profiles on impulses, eval_tokens(argmax)==max,
//                             chunked == unchunked.
//   whole dedisperser      -> soph0 == soph1 (bit-exact), pulse recovery.

#ifndef PIRATE_ORACLE_HPP
#define PIRATE_ORACLE_HPP

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>
#ifdef __F16C__
#include <immintrin.h>
#endif

inline float half_bits_to_float(uint16_t h);   // IEEE binary16 -> float (defined below)

namespace oracle {

#define ORC_CHECK(cond, msg)                                                  \
    do {                                                                      \
        if (!(cond)) {                                                        \
            std::ostringstream _s;                                            \
            _s << "oracle: " << msg << " [" #cond "] " << __FILE__ << ":" << __LINE__; \
            throw std::runtime_error(_s.str());                               \
        }                                                                     \
    } while (0)

inline long pow2(long n) { return 1L << n; }
inline bool is_pow2(long n) { return n > 0 && !(n & (n - 1)); }
inline int ilog2(long n)
{
    ORC_CHECK(is_pow2(n), "ilog2 of non-power-of-two " << n);
    int r = 0;
    while ((1L << r) < n) r++;
    return r;
}

// Reference: src_lib/utils.cpp:189-203.
inline long bitrev(long i, int nbits)
{
    long j = 0;
    for (int b = 0; b < nbits; b++) { j = (j << 1) | (i & 1); i >>= 1; }
    return j;
}

// Reference: src_lib/utils.cpp:218-239.
inline long rb_lag(long fc, long dm_brev, int s1_rank, int s2_rank, bool uflag)
{
    long dm = bitrev(dm_brev, s1_rank) + (uflag ? pow2(s1_rank) : 0);
    return (pow2(s2_rank) - 1 - fc) * dm;
}

// Delay (in samples) of input channel 'freq' in output row 'dm_brev' of a rank-'rank'
// tree (reference: src_lib/utils.cpp:287-301).
inline long dedispersion_delay(int rank, long freq, long dm_brev)
{
    long delay = 0, delay0 = 0;
    for (int r = 0; r < rank; r++) {
        long d = (dm_brev & 1) ? (delay0 + 1) : delay0;
        delay += (freq & 1) ? 0 : d;
        delay0 += d;
        dm_brev >>= 1;
        freq >>= 1;
    }
    return delay;
}


// -------------------------------------------------------------------------------------------
// Frequency subbands (reference: src_lib/FrequencySubbands.cpp:32-60, 165-201).

struct Subbands {
    std::vector<long> counts;
    long pf_rank = 0, N = 0, M = 0;
    std::vector<long> m_to_n, m_to_d, n_to_flo, n_to_fhi, n_to_mbase;

    Subbands() : Subbands(std::vector<long>{1}) { }
    explicit Subbands(const std::vector<long> &sc) : counts(sc)
    {
        ORC_CHECK(!sc.empty() && sc.back() == 1, "bad subband_counts");
        pf_rank = long(sc.size()) - 1;
        for (long lev = 0; lev <= pf_rank; lev++) {
            long sp = pow2(std::max(lev - 1, 0L));
            for (long b = 0; b < sc[lev]; b++) {
                n_to_flo.push_back(b * sp);
                n_to_fhi.push_back(b * sp + pow2(lev));
                n_to_mbase.push_back(M);
                for (long d = 0; d < pow2(lev); d++) { m_to_n.push_back(N); m_to_d.push_back(d); }
                M += pow2(lev);
                N++;
            }
        }
    }
    long m_to_flo(long m) const { return n_to_flo[m_to_n[m]]; }
    long m_to_fhi(long m) const { return n_to_fhi[m_to_n[m]]; }

    static std::vector<long> restrict(const std::vector<long> &sc, long et, long new_rank)
    {
        long src_rank = long(sc.size()) - 1;
        long er = std::max(src_rank - et, 0L);
        std::vector<long> early(er + 1);
        early[er] = 1;
        for (long l = 0; l < er; l++) {
            long mx = (l > 0) ? (pow2(er + 1 - l) - 1) : pow2(er);
            early[l] = std::min(sc[l], mx);
        }
        std::vector<long> dst(new_rank + 1);
        dst[new_rank] = 1;
        for (long l = 0; l < new_rank; l++) {
            long s = l - new_rank + er;
            if (s < 0 || early[s] == 0) dst[l] = 0;
            else if (s == 0 && l > 0) dst[l] = 2 * early[s] - 1;
            else if (s > 0 && l == 0) dst[l] = early[s] / 2 + 1;
            else dst[l] = early[s];
        }
        return dst;
    }
};


// -------------------------------------------------------------------------------------------
// Lag buffer: per-row integer delay with state carried across calls.
// Reference: src_lib/ReferenceLagbuf.cpp:71-127.

template<typename T>
struct RefLagbuf {
    std::vector<long> lags;   // one lag per row, rows in row-major order of the lag array
    long ntime = 0;
    std::vector<T> pstate;
    std::vector<T> scratch;

    RefLagbuf() { }
    RefLagbuf(const std::vector<long> &lags_, long ntime_) : lags(lags_), ntime(ntime_)
    {
        long tot = 0;
        for (long l : lags) { ORC_CHECK(l >= 0, "negative lag"); tot += l; }
        pstate.assign(tot, T(0));
        scratch.assign(ntime, T(0));
    }

    // Row r lives at data + row_offset(r), ntime contiguous elements.
    template<typename F>
    void apply(T *data, F row_offset)
    {
        T *ps = pstate.data();
        for (size_t r = 0; r < lags.size(); r++) {
            long lag = lags[r];
            if (lag <= 0)
                continue;
            T *d = data + row_offset(long(r));
            long n = std::min(lag, ntime);
            std::memcpy(scratch.data(), d + (ntime - n), n * sizeof(T));
            std::memmove(d + n, d, (ntime - n) * sizeof(T));
            std::memcpy(d, ps, n * sizeof(T));
            std::memmove(ps, ps + n, (lag - n) * sizeof(T));
            std::memcpy(ps + (lag - n), scratch.data(), n * sizeof(T));
            ps += lag;
        }
    }
};


// -------------------------------------------------------------------------------------------
// Incremental tree dedispersion with subband outputs.
// Reference: src_lib/ReferenceTree.cpp:22-345, include/pirate/ReferenceTree.hpp.
//
// buf: (B, 2^amb, 2^dd, ntime*nspec), dedispersed in place.
// out: (B, 2^(amb+dd-pf_rank), M, ntime*nspec), subband outputs (may be null if M == 1).

template<typename T>
struct RefTree {
    long B = 0, amb_rank = 0, dd_rank = 0, ntime = 0, nspec = 0;
    Subbands fs;
    std::vector<T> pstate;
    std::vector<T> scratch;
    RefLagbuf<T> final_lagbuf;

    RefTree(long B_, long amb_, long dd_, long nt_, long ns_, const std::vector<long> &sc) :
        B(B_), amb_rank(amb_), dd_rank(dd_), ntime(nt_), nspec(ns_), fs(sc)
    {
        ORC_CHECK(fs.pf_rank <= dd_rank, "pf_rank > dd_rank");
        long S = nspec, T_ = ntime;
        scratch.assign(S * (T_ + pow2(dd_rank) + 1), T(0));
        long Dpf = pow2(amb_rank + dd_rank - fs.pf_rank);
        std::vector<long> flags;
        for (long b = 0; b < B; b++)
            for (long dpf = 0; dpf < Dpf; dpf++)
                for (long m = 0; m < fs.M; m++) {
                    long ff = pow2(fs.pf_rank) - fs.m_to_fhi(m);
                    long dd = dpf & (pow2(dd_rank - fs.pf_rank) - 1);
                    flags.push_back(ff * dd * S);
                }
        final_lagbuf = RefLagbuf<T>(flags, T_ * S);
    }

    // dst/src: two rows each (stride ds/ss); dst may equal src. 'lag' in time samples.
    long dedisperse_1d(T *dst, long ds, const T *src, long ss, long ps_used, long lag)
    {
        const T c = T(0.7071067811865476);
        long S = nspec, Tn = ntime;
        if (ps_used + (lag + 1) * S > long(pstate.size()))
            pstate.resize(ps_used + (lag + 1) * S, T(0));
        T *ps = pstate.data() + ps_used;
        T *tmp = scratch.data();
        std::memcpy(tmp, ps, (lag + 1) * S * sizeof(T));
        std::memcpy(tmp + (lag + 1) * S, src, Tn * S * sizeof(T));
        std::memcpy(ps, tmp + Tn * S, (lag + 1) * S * sizeof(T));
        for (long ts = Tn * S - 1; ts >= 0; ts--) {
            T x0 = tmp[ts];
            T x1 = tmp[ts + S];
            T y = src[ss + ts];
            dst[ts] = c * (x1 + y);
            dst[ds + ts] = c * (x0 + y);
        }
        return ps_used + (lag + 1) * S;
    }

    long dedisperse_2d(T *bufp, long bds, T *outp, long ods, long oms, long ps_used)
    {
        long TS = ntime * nspec;
        long pf_rank = fs.pf_rank;
        long pf_ndm = pow2(dd_rank - pf_rank);
        long mcurr = 0;

        for (long r = 0; r <= dd_rank; r++) {
            long ndm_in = pow2(r);

            if (outp && r == dd_rank - pf_rank) {
                long pf_ns = fs.counts[0];
                for (long dm_in = 0; dm_in < ndm_in; dm_in++) {
                    long dmb = bitrev(dm_in, int(r));
                    for (long pfs = 0; pfs < pf_ns; pfs++) {
                        const T *src = bufp + (pfs * ndm_in + dmb) * bds;
                        T *dst = outp + dm_in * ods + pfs * oms;
                        std::memcpy(dst, src, TS * sizeof(T));
                    }
                }
                mcurr += pf_ns;
            }
            if (r >= dd_rank)
                break;

            long pf_level = r + pf_rank - dd_rank + 1;
            if (outp && pf_level > 0) {
                long pf_ns = fs.counts[pf_level];
                long pf_nd2 = pow2(pf_level - 1);
                for (long pf_dm = 0; pf_dm < pf_ndm; pf_dm++)
                    for (long d2 = 0; d2 < pf_nd2; d2++) {
                        long dm_in = pf_dm * pf_nd2 + d2;
                        long dmb = bitrev(dm_in, int(r));
                        for (long pfs = 0; pfs < pf_ns; pfs++) {
                            const T *src = bufp + (pfs * ndm_in + dmb) * bds;
                            long m = mcurr + pfs * (2 * pf_nd2) + 2 * d2;
                            T *dst = outp + pf_dm * ods + m * oms;
                            ps_used = dedisperse_1d(dst, oms, src, ndm_in * bds, ps_used, dm_in);
                        }
                    }
                mcurr += pf_ns * 2 * pf_nd2;
            }

            long nf_out = pow2(dd_rank - r - 1);
            long ndm_out = pow2(r + 1);
            for (long fout = 0; fout < nf_out; fout++)
                for (long dmb = 0; dmb < ndm_in; dmb++) {
                    long dm_in = bitrev(dmb, int(r));
                    T *p = bufp + (fout * ndm_out + dmb) * bds;
                    ps_used = dedisperse_1d(p, ndm_in * bds, p, ndm_in * bds, ps_used, dm_in);
                }
        }
        ORC_CHECK(mcurr == (outp ? fs.M : 0), "multiplet count mismatch");
        return ps_used;
    }

    // Strides in elements; the innermost (time*spec) axis is contiguous.
    void dedisperse(T *buf, long bs_b, long bs_a, long bs_d, T *out, long os_b, long os_d, long os_m)
    {
        ORC_CHECK(out || fs.M == 1, "subband output array required when M > 1");
        long A = pow2(amb_rank);
        long ps_used = 0;
        for (long b = 0; b < B; b++)
            for (long a = 0; a < A; a++) {
                T *bufp = buf + b * bs_b + a * bs_a;
                T *outp = nullptr;
                if (out) {
                    long pfdm = bitrev(a, int(amb_rank)) << (dd_rank - fs.pf_rank);
                    outp = out + b * os_b + pfdm * os_d;
                }
                ps_used = dedisperse_2d(bufp, bs_d, outp, os_d, os_m, ps_used);
            }
        if (out) {
            long Dpf = pow2(amb_rank + dd_rank - fs.pf_rank);
            long M = fs.M;
            final_lagbuf.apply(out, [&](long r) {
                long m = r % M, dpf = (r / M) % Dpf, b = r / (M * Dpf);
                return b * os_b + dpf * os_d + m * os_m;
            });
        }
        ORC_CHECK(ps_used == long(pstate.size()), "pstate size mismatch");
    }
};


// Whole-array (non-incremental) tree, zero history (reference: src_lib/utils.cpp:304-344).
template<typename T>
void dedisperse_non_incremental(T *arr, long nfreq, long ninner, long nspec)
{
    const T c = T(0.7071067811865476);
    int rank = ilog2(nfreq);
    for (int r = 0; r < rank; r++) {
        long pr = pow2(r);
        for (long i = 0; i < nfreq; i += 2 * pr)
            for (long j = 0; j < pr; j++) {
                T *row0 = arr + (i + j) * ninner;
                T *row1 = row0 + pr * ninner;
                long lag1 = bitrev(j, r) * nspec;
                long lag0 = lag1 + nspec;
                for (long k = ninner - 1; k >= 0; k--) {
                    T x0 = (k >= lag0) ? row0[k - lag0] : T(0);
                    T x1 = (k >= lag1) ? row0[k - lag1] : T(0);
                    T y = row1[k];
                    row0[k] = c * (x1 + y);
                    row1[k] = c * (x0 + y);
                }
            }
    }
}


// -------------------------------------------------------------------------------------------
// Tree gridding: out[b,n,t] = sum_f w[n,f] in[b,f,t], w = overlap of [cmap[n+1], cmap[n])
// with [f, f+1). Reference: src_lib/TreeGriddingKernel.cu:62-102.

template<typename T>
void tree_gridding(T *out, const T *in, const double *cmap, long B, long nchan, long nfreq, long ntime)
{
    for (long b = 0; b < B; b++) {
        T *op = out + b * nchan * ntime;
        const T *ip = in + b * nfreq * ntime;
        std::memset(op, 0, nchan * ntime * sizeof(T));
        for (long n = 0; n < nchan; n++) {
            double f0 = cmap[n + 1], f1 = cmap[n];
            long if0 = std::max(long(f0), 0L);
            long if1 = std::min(long(f1) + 1, nfreq);
            for (long f = if0; f < if1; f++) {
                double flo = std::max(f0, double(f));
                double fhi = std::min(f1, double(f) + 1);
                double w = std::max(fhi - flo, 0.0);
                T wt = T(w);
                if (sizeof(T) == sizeof(float))
                    wt = T(float(w));
                for (long t = 0; t < ntime; t++)
                    op[n * ntime + t] += wt * ip[f * ntime + t];
            }
        }
    }
}


// -------------------------------------------------------------------------------------------
// Lagged downsampling for primary trees p >= 1.
// Reference: src_lib/ReferenceLaggedDownsamplingKernel.cpp:27-139.
//
// in: (B, 2^R, ntime) level-0 tree-gridded data.
// outs[p-1]: (B, 2^(R-1), ntime/2^p) for p = 1..npri-1.

template<typename T>
struct RefLDS {
    long B = 0, R = 0, s = 0, npri = 0, ntime = 0;
    RefLagbuf<T> small, large;
    std::unique_ptr<RefLDS<T>> next;

    RefLDS(long B_, long R_, long output_dd_rank, long npri_, long ntime_) :
        B(B_), R(R_), s(output_dd_rank + 1), npri(npri_), ntime(ntime_)
    {
        if (npri <= 1)
            return;
        ORC_CHECK(ntime % 2 == 0, "LDS ntime must be even");
        std::vector<long> sl(B * pow2(R)), ll(B * pow2(R - 1));
        for (long i = 0; i < long(sl.size()); i++)
            sl[i] = (i & 1) ? 0 : 1;
        for (long i = 0; i < B * pow2(R - s); i++)
            for (long j = 0; j < pow2(s - 1); j++)
                ll[i * pow2(s - 1) + j] = pow2(s - 1) - j - 1;
        small = RefLagbuf<T>(sl, ntime / 2);
        large = RefLagbuf<T>(ll, ntime / 2);
        if (npri > 2)
            next.reset(new RefLDS<T>(B, R, output_dd_rank, npri - 1, ntime / 2));
    }

    void apply(const T *in, T *const *outs)
    {
        if (npri <= 1)
            return;
        const T c = T(0.7071067811865476);
        long nr = B * pow2(R), nt2 = ntime / 2;
        std::vector<T> ds(nr * nt2);
        for (long r = 0; r < nr; r++)
            for (long t = 0; t < nt2; t++)
                ds[r * nt2 + t] = c * (in[r * ntime + 2 * t] + in[r * ntime + 2 * t + 1]);
        std::vector<T> ds2(ds);
        small.apply(ds2.data(), [&](long r) { return r * nt2; });
        T *out = outs[0];
        for (long j = 0; j < nr / 2; j++)
            for (long t = 0; t < nt2; t++)
                out[j * nt2 + t] = c * (ds2[(2 * j) * nt2 + t] + ds2[(2 * j + 1) * nt2 + t]);
        large.apply(out, [&](long r) { return r * nt2; });
        if (next)
            next->apply(ds.data(), outs + 1);
    }
};


// -------------------------------------------------------------------------------------------
// Peak finder. Reference: src_lib/PeakFindingKernel.cu:197-557,
// include/pirate/PeakFindingKernel.hpp:20-260.
//
// in:  (B, ndm_out, M, nt_in)
// wt:  (B, ndm_wt, nt_wt, P, N)
// out: (B, ndm_out, nt_out), tokens (t) | (p << 8) | (m << 16)
// var: optional (B, ndm_out, M, P), double

template<typename T>
struct RefPF {
    long B, total_beams, ndm_out, ndm_wt, nt_out, nt_in, nt_wt, Wmax, Dcore;
    Subbands fs;
    long Dout, nbatches, nprofiles, tpad, num_levels;
    long expected_ibatch = 0;
    std::vector<long> tdt, tnt, tiout, tnout, tsout;
    std::vector<std::vector<T>> tmp;     // per level: (B, D, M, tnt[l])
    std::vector<T> pstate;              // (total_beams, D, M, tpad)

    RefPF(const std::vector<long> &sc, long Wmax_, long B_, long total_beams_, long ndm_out_, long ndm_wt_,
          long nt_out_, long nt_in_, long nt_wt_, long Dcore_) :
        B(B_), total_beams(total_beams_), ndm_out(ndm_out_), ndm_wt(ndm_wt_), nt_out(nt_out_), nt_in(nt_in_),
        nt_wt(nt_wt_), Wmax(Wmax_), Dcore(Dcore_), fs(sc)
    {
        ORC_CHECK(is_pow2(Wmax) && total_beams % B == 0 && nt_in % nt_out == 0 && nt_out % nt_wt == 0 &&
                  ndm_out % ndm_wt == 0 && Dcore > 0 && (nt_in / nt_out) % Dcore == 0, "bad peak-finder params");
        nbatches = total_beams / B;
        nprofiles = 3 * ilog2(Wmax) + 1;
        Dout = nt_in / nt_out;
        tpad = std::max(2 * Wmax, 4L);
        num_levels = std::max(ilog2(Wmax), 1);
        pstate.assign(total_beams * ndm_out * fs.M * tpad, T(0));
        for (long l = 0; l < num_levels; l++) {
            long dt = std::min(Dcore, pow2(l));
            ORC_CHECK((nt_in + tpad - pow2(l)) % dt == 0, "level grid");
            long nt = (nt_in + tpad - pow2(l)) / dt + 1;
            tdt.push_back(dt);
            tnt.push_back(nt);
            tnout.push_back(Dout / dt);
            tsout.push_back(pow2(l) / dt);
            tmp.emplace_back(B * ndm_out * fs.M * nt, T(0));
            tiout.push_back((tpad + dt - pow2(l)) / dt);
        }
    }

    static void update(T &maxval, uint32_t &argmax, double *sumsq, T val, uint32_t token)
    {
        argmax = (val > maxval) ? token : argmax;
        maxval = std::max(maxval, val);
        if (sumsq)
            *sumsq += double(val) * double(val);
    }

    void apply(T *out_max, uint32_t *out_argmax, double *out_var, const T *in, const float *wt, long ibatch)
    {
        ORC_CHECK(ibatch == expected_ibatch, "peak finder: batches out of order");
        expected_ibatch = (ibatch + 1) % nbatches;
        long D = ndm_out, M = fs.M, N = fs.N, P = nprofiles;
        long b0 = ibatch * B;
        long t1 = std::min(tpad, nt_in), t0 = tpad - t1;

        for (long b = 0; b < B; b++)
            for (long d = 0; d < D; d++)
                for (long m = 0; m < M; m++) {
                    T *dst = &tmp[0][((b * D + d) * M + m) * tnt[0]];
                    T *ps = &pstate[(((b0 + b) * D + d) * M + m) * tpad];
                    const T *src = &in[((b * D + d) * M + m) * nt_in];
                    for (long t = 0; t < tpad; t++) dst[t] = ps[t];
                    for (long t = 0; t < nt_in; t++) dst[t + tpad] = src[t];
                    for (long t = 0; t < t0; t++) ps[t] = ps[t + nt_in];
                    for (long t = 0; t < t1; t++) ps[t + t0] = src[t + nt_in - t1];
                }

        for (long l = 0; l + 1 < num_levels; l++) {
            long ns = tnt[l], nd = tnt[l + 1];
            long r = tdt[l + 1] / tdt[l];
            long s = pow2(l) / tdt[l];
            ORC_CHECK(r * (nd - 1) + s == ns - 1, "level downsampling grid");
            for (long row = 0; row < B * D * M; row++) {
                T *dst = &tmp[l + 1][row * nd];
                const T *src = &tmp[l][row * ns];
                for (long t = 0; t < nd; t++)
                    dst[t] = src[r * t] + src[r * t + s];
            }
        }

        long Wds = ndm_out / ndm_wt, Tds = nt_out / nt_wt;
        for (long b = 0; b < B; b++)
            for (long d = 0; d < D; d++) {
                double *var_bd = out_var ? &out_var[(b * D + d) * M * P] : nullptr;
                if (var_bd)
                    for (long i = 0; i < M * P; i++) var_bd[i] = 0.0;
                for (long tout = 0; tout < nt_out; tout++) {
                    const float *wp = &wt[(((b * ndm_wt + d / Wds) * nt_wt + tout / Tds) * P) * N];
                    T maxval = T(-1.0e30f);
                    uint32_t argmax = ~0u;
                    for (long l = 0; l < num_levels; l++) {
                        const T *tin = &tmp[l][((b * D + d) * M) * tnt[l]];
                        long mstr = tnt[l], dt = tdt[l], nsamp = tnout[l], S = tsout[l], I = tiout[l];
                        double wvar = 1.0 / double(nt_out * nsamp);
                        for (long m = 0; m < M; m++) {
                            long n = fs.m_to_n[m];
                            T w0 = l ? T(0) : T(wp[n]);
                            T w1 = (P > 1) ? T(wp[(3 * l + 1) * N + n]) : T(0);
                            T w2 = (P > 1) ? T(wp[(3 * l + 2) * N + n]) : T(0);
                            T w3 = (P > 1) ? T(wp[(3 * l + 3) * N + n]) : T(0);
                            double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
                            bool dv = (var_bd != nullptr);
                            for (long is = 0; is < nsamp; is++) {
                                long j = m * mstr + I + tout * nsamp + is;
                                T x0 = tin[j - 3 * S], x1 = tin[j - 2 * S], x2 = tin[j - S], x3 = tin[j];
                                uint32_t tk0 = uint32_t((m << 16) | (is * dt));
                                T y0 = x3;
                                T y1 = x2 + x3;
                                T y2 = T(0.5f) * x1 + x2 + T(0.5f) * x3;
                                T y3 = T(0.5f) * x0 + x1 + x2 + T(0.5f) * x3;
                                if (l == 0)
                                    update(maxval, argmax, dv ? &s0 : nullptr, w0 * y0, tk0);
                                if (P > 1) {
                                    update(maxval, argmax, dv ? &s1 : nullptr, w1 * y1, tk0 | uint32_t((3 * l + 1) << 8));
                                    update(maxval, argmax, dv ? &s2 : nullptr, w2 * y2, tk0 | uint32_t((3 * l + 2) << 8));
                                    update(maxval, argmax, dv ? &s3 : nullptr, w3 * y3, tk0 | uint32_t((3 * l + 3) << 8));
                                }
                            }
                            if (dv) {
                                double *vm = var_bd + m * P;
                                if (l == 0) vm[0] += s0 * wvar;
                                if (P > 1) {
                                    vm[3 * l + 1] += s1 * wvar;
                                    vm[3 * l + 2] += s2 * wvar;
                                    vm[3 * l + 3] += s3 * wvar;
                                }
                            }
                        }
                    }
                    out_max[(b * D + d) * nt_out + tout] = maxval;
                    out_argmax[(b * D + d) * nt_out + tout] = argmax;
                }
            }
    }

    // Re-evaluates the single trial named by each token, from the most recent apply().
    // Reference: src_lib/PeakFindingKernel.cu:460-557.
    void eval_tokens(T *out, const uint32_t *tokens, const float *wt) const
    {
        long D = ndm_out, M = fs.M, N = fs.N, P = nprofiles;
        long Wds = ndm_out / ndm_wt, Tds = nt_out / nt_wt;
        for (long b = 0; b < B; b++)
            for (long d = 0; d < D; d++)
                for (long tout = 0; tout < nt_out; tout++) {
                    uint32_t tk = tokens[(b * D + d) * nt_out + tout];
                    long m = (tk >> 16) & 0xffff, p = (tk >> 8) & 0xff, t = tk & 0xff;
                    ORC_CHECK(m < M && p < P && t < Dout, "bad token " << tk);
                    long l = p ? (p - 1) / 3 : 0, q = p - 3 * l;
                    long dt = tdt[l], is = t / dt;
                    ORC_CHECK(is * dt == t, "bad token time " << tk);
                    long n = fs.m_to_n[m];
                    T w = T(wt[(((b * ndm_wt + d / Wds) * nt_wt + tout / Tds) * P + p) * N + n]);
                    long nsamp = tnout[l], S = tsout[l], I = tiout[l];
                    const T *tin = &tmp[l][((b * D + d) * M + m) * tnt[l]];
                    long j = I + tout * nsamp + is;
                    T x0 = tin[j - 3 * S], x1 = tin[j - 2 * S], x2 = tin[j - S], x3 = tin[j];
                    T y;
                    if (q == 0) y = x3;
                    else if (q == 1) y = x2 + x3;
                    else if (q == 2) y = T(0.5f) * x1 + x2 + T(0.5f) * x3;
                    else y = T(0.5f) * x0 + x1 + x2 + T(0.5f) * x3;
                    out[(b * D + d) * nt_out + tout] = w * y;
                }
    }
};


// Zero-lag autocorrelation of each peak-finding profile (reference: PfVarianceConvolver,
// src_lib/PfVariance.cpp:23-98, evaluated on a single unit sample).
inline std::vector<double> bare_profile_variances(long P)
{
    std::vector<double> v(P);
    for (long p = 0; p < P; p++) {
        if (p == 0) { v[p] = 1.0; continue; }
        long l = (p - 1) / 3, q = p - 3 * l;
        double w = double(pow2(l));
        v[p] = (q == 1) ? 2.0 * w : (q == 2) ? 1.5 * w : 2.5 * w;
    }
    return v;
}


// -------------------------------------------------------------------------------------------
// Dedisperser configuration, derived from config fields exactly as the reference plan does
// (src_lib/DedispersionPlan.cpp:59-156) — restated here independently of pirate_amd's plan.

struct OrcPrimaryTree { long num_early_triggers, max_width, dm_ds, time_ds, wt_dm_ds, wt_time_ds; };

struct OrcTree {
    long ipri, et, amb_rank, dd_rank, nt_ds;
    std::vector<long> subband_counts;
    long M, N, Wmax, dm_ds, time_ds, wt_dm_ds, wt_time_ds, Dcore, nprofiles, ndm_out, ndm_wt, nt_out, nt_wt;
};

struct OrcConfig {
    std::vector<long> zone_nfreq;
    std::vector<double> zone_freq_edges;
    long R = 0, nt_in = 0, beams_per_gpu = 0, beams_per_batch = 0;
    long nelts_per_segment = 32;
    std::vector<long> subband_counts;
    std::vector<OrcPrimaryTree> pts;
    bool reference_dcore = false;   // false: Dcore = min(Dout, 8); true: Dcore = Dout

    long nfreq() const { long s = 0; for (long n : zone_nfreq) s += n; return s; }
    long npri() const { return long(pts.size()); }

    // Reference: src_lib/DedispersionConfig.cpp:52-96, 135-162, 238-250.
    double frequency_to_index(double freq) const
    {
        double fmin = zone_freq_edges.front(), fmax = zone_freq_edges.back();
        freq = std::min(std::max(freq, fmin), fmax);
        double off = 0;
        for (size_t i = 0; i < zone_nfreq.size(); i++) {
            double f0 = zone_freq_edges[i], f1 = zone_freq_edges[i + 1];
            if (freq <= f1) { off += ((freq - f0) / (f1 - f0)) * zone_nfreq[i]; break; }
            off += zone_nfreq[i];
        }
        return std::min(std::max(off, 0.0), double(nfreq()));
    }
    double delay_to_frequency(double delay) const
    {
        double flo = zone_freq_edges.front(), fhi = zone_freq_edges.back();
        double ntree = double(pow2(R));
        delay = std::min(std::max(delay, 0.0), ntree);
        double scale = ntree / (1.0 / (flo * flo) - 1.0 / (fhi * fhi));
        return 1.0 / std::sqrt(delay / scale + 1.0 / (fhi * fhi));
    }
    std::vector<double> channel_map() const
    {
        std::vector<double> c(pow2(R) + 1);
        for (long n = 0; n <= pow2(R); n++)
            c[n] = frequency_to_index(delay_to_frequency(double(n)));
        return c;
    }

    std::vector<OrcTree> make_trees() const
    {
        std::vector<OrcTree> trees;
        for (long ipri = 0; ipri < npri(); ipri++) {
            long prank = ipri ? R - 1 : R;
            long s1dd = prank / 2, s1amb = prank - s1dd;
            const OrcPrimaryTree &pt = pts[ipri];
            for (long et = pt.num_early_triggers; et >= 0; et--) {
                OrcTree t;
                t.ipri = ipri; t.et = et;
                t.amb_rank = s1dd; t.dd_rank = s1amb - et;
                t.nt_ds = nt_in / pow2(ipri);
                long pf_rank = (t.dd_rank + 1) / 2;
                t.subband_counts = Subbands::restrict(subband_counts, et, pf_rank);
                Subbands fs(t.subband_counts);
                t.M = fs.M; t.N = fs.N;
                t.Wmax = pt.max_width;
                t.dm_ds = pt.dm_ds ? pt.dm_ds : pow2(pf_rank);
                t.time_ds = pt.time_ds ? pt.time_ds : t.dm_ds;
                t.wt_dm_ds = pt.wt_dm_ds; t.wt_time_ds = pt.wt_time_ds;
                t.Dcore = reference_dcore ? t.time_ds : std::min(t.time_ds, 8L);
                t.nprofiles = 1 + 3 * ilog2(t.Wmax);
                long tot = t.amb_rank + t.dd_rank;
                t.ndm_out = pow2(tot) / t.dm_ds;
                t.ndm_wt = pow2(tot) / t.wt_dm_ds;
                t.nt_out = t.nt_ds / t.time_ds;
                t.nt_wt = t.nt_ds / t.wt_time_ds;
                trees.push_back(t);
            }
        }
        return trees;
    }
};


// -------------------------------------------------------------------------------------------
// Storage rounding of the 16-bit tiers: x rounded to IEEE binary16 or bfloat16 with
// round-to-nearest-even, as a store to the storage type and a reload would give.

inline uint16_t float_to_half_bits(float f)
{
    uint32_t x;
    std::memcpy(&x, &f, 4);
    const uint32_t sign = (x >> 16) & 0x8000u;
    x &= 0x7fffffffu;
    if (x >= 0x7f800000u)                                  // inf, nan (quiet)
        return uint16_t(sign | 0x7c00u | ((x > 0x7f800000u) ? 0x200u : 0u));
    if (x >= 0x477ff000u)                                  // >= 65520: rounds to inf
        return uint16_t(sign | 0x7c00u);
    if (x >= 0x38800000u) {                                // normal binary16 (>= 2^-14)
        uint32_t h = ((((x >> 23) - 127 + 15)) << 10) | ((x & 0x7fffffu) >> 13);
        const uint32_t rem = x & 0x1fffu;
        if (rem > 0x1000u || (rem == 0x1000u && (h & 1u)))
            h++;                                           // a carry into the exponent is exact
        return uint16_t(sign | h);
    }
    // subnormal binary16: units of 2^-24, nearest-even integer (exact in double)
    float a;
    std::memcpy(&a, &x, 4);
    const double u = double(a) * 16777216.0;
    double r = std::floor(u);
    const double d = u - r;
    if (d > 0.5 || (d == 0.5 && std::fmod(r, 2.0) != 0.0))
        r += 1.0;
    return uint16_t(sign | uint32_t(r));                   // r = 1024 is the smallest normal
}

inline float round_to_half(float x) { return ::half_bits_to_float(float_to_half_bits(x)); }

inline float round_to_bf16(float x)
{
    uint32_t b;
    std::memcpy(&b, &x, 4);
    if ((b & 0x7fffffffu) > 0x7f800000u)
        b = (b | 0x00400000u) & 0xffff0000u;               // nan stays a (quiet) nan
    else
        b = (b + 0x7fffu + ((b >> 16) & 1u)) & 0xffff0000u;
    float r;
    std::memcpy(&r, &b, 4);
    return r;
}

// storage: 0 = float32 (no rounding), 1 = float16, 2 = bfloat16.
template<typename T>
inline void round_to_storage(T *x, long n, int storage)
{
    if (storage == 1) {
        long i = 0;
#ifdef __F16C__
        // The x86 F16C conversion (round-to-nearest-even, immediate rounding mode) gives the
        // same bits as float_to_half_bits; 8 values at a time.
        if constexpr (std::is_same<T, float>::value)
            for (; i + 8 <= n; i += 8) {
                const __m256 v = _mm256_loadu_ps(x + i);
                _mm256_storeu_ps(x + i, _mm256_cvtph_ps(_mm256_cvtps_ph(v, _MM_FROUND_TO_NEAREST_INT | _MM_FROUND_NO_EXC)));
            }
#endif
        for (; i < n; i++) x[i] = T(round_to_half(float(x[i])));
    }
    else if (storage == 2)
        for (long i = 0; i < n; i++) x[i] = T(round_to_bf16(float(x[i])));
}


// -------------------------------------------------------------------------------------------
// Whole dedisperser, sophistication 0 and 1.
// Reference: src_lib/ReferenceDedisperser.cpp:48-450, include/pirate/Dedisperser.hpp:395-482.

template<typename T>
struct RefDedisperser {
    OrcConfig cfg;
    int soph;
    bool tree_domain_input;
    int storage = 0;   // soph 1: round to 16-bit storage where the GPU chain stores (see dedisperse())
    std::vector<OrcTree> trees;
    long B, nbatches, ntrees, npri, R, nt_in, nfreq;
    std::vector<double> cmap;

    std::vector<T> input;                       // (B, nfreq or 2^R, nt_in)
    std::vector<std::vector<float>> wt;         // per tree: (B, ndm_wt, nt_wt, P, N)
    std::vector<std::vector<T>> out_max;        // per tree: (B, ndm_out, nt_out)
    std::vector<std::vector<uint32_t>> out_argmax;
    std::vector<std::vector<double>> out_var;   // per tree: (B, ndm_out, M, P) if enabled
    std::vector<std::unique_ptr<RefPF<T>>> pf;

    // soph 0
    std::vector<std::vector<T>> ds_inputs;      // per primary tree: (B, 2^R, nt_in/2^p)
    std::vector<std::vector<T>> dd_bufs0;       // per tree: (B, 2^ref_rank, nt_ds)
    std::vector<std::vector<T>> sb_bufs;        // per tree: (B, ndm_out*(p>0?2:1), M, nt_ds)
    std::vector<std::unique_ptr<RefTree<T>>> trees0;   // (nbatches * ntrees)

    // soph 1
    std::vector<std::vector<T>> s1_bufs;        // per primary tree: (B, 2^r_p, nt_p)
    std::vector<std::vector<T>> s2_bufs;        // per tree: (B, 2^rank, nt_ds)
    std::vector<std::unique_ptr<RefLDS<T>>> lds; // per batch
    std::vector<std::unique_ptr<RefTree<T>>> s1_trees;   // (nbatches * npri)
    std::vector<std::unique_ptr<RefTree<T>>> s2_trees;   // (nbatches * ntrees)
    std::vector<RefLagbuf<T>> s2_seg_lagbufs;            // (nbatches * ntrees)
    std::vector<RefLagbuf<T>> s2_res_lagbufs;            // (nbatches * ntrees)

    RefDedisperser(const OrcConfig &c, int soph_, bool enable_var, bool tdi = false, int storage_ = 0) :
        cfg(c), soph(soph_), tree_domain_input(tdi), storage(storage_)
    {
        ORC_CHECK(soph == 0 || soph == 1, "sophistication must be 0 or 1");
        ORC_CHECK(storage == 0 || soph == 1, "16-bit storage rounding is modelled for sophistication 1 only");
        trees = cfg.make_trees();
        B = cfg.beams_per_batch;
        nbatches = cfg.beams_per_gpu / B;
        ntrees = long(trees.size());
        npri = cfg.npri();
        R = cfg.R;
        nt_in = cfg.nt_in;
        nfreq = cfg.nfreq();
        cmap = cfg.channel_map();

        input.assign(B * (tdi ? pow2(R) : nfreq) * nt_in, T(0));
        for (auto &t : trees) {
            wt.emplace_back(B * t.ndm_wt * t.nt_wt * t.nprofiles * t.N, 0.0f);
            out_max.emplace_back(B * t.ndm_out * t.nt_out, T(0));
            out_argmax.emplace_back(B * t.ndm_out * t.nt_out, 0u);
            out_var.emplace_back(enable_var ? B * t.ndm_out * t.M * t.nprofiles : 0, 0.0);
            pf.emplace_back(new RefPF<T>(t.subband_counts, t.Wmax, B, cfg.beams_per_gpu, t.ndm_out, t.ndm_wt,
                                         t.nt_out, t.nt_ds, t.nt_wt, t.Dcore));
        }

        if (soph == 0) {
            for (long p = 0; p < npri; p++)
                ds_inputs.emplace_back(B * pow2(R) * (nt_in / pow2(p)), T(0));
            for (long ib = 0; ib < nbatches; ib++)
                for (auto &t : trees) {
                    long rr = t.amb_rank + t.dd_rank + (t.ipri > 0 ? 1 : 0);
                    trees0.emplace_back(new RefTree<T>(B, 0, rr, t.nt_ds, 1, t.subband_counts));
                }
            for (auto &t : trees) {
                long rr = t.amb_rank + t.dd_rank + (t.ipri > 0 ? 1 : 0);
                dd_bufs0.emplace_back(B * pow2(rr) * t.nt_ds, T(0));
                sb_bufs.emplace_back(B * t.ndm_out * (t.ipri > 0 ? 2 : 1) * t.M * t.nt_ds, T(0));
            }
        }
        else {
            for (long p = 0; p < npri; p++) {
                long rp = p ? R - 1 : R;
                s1_bufs.emplace_back(B * pow2(rp) * (nt_in / pow2(p)), T(0));
            }
            for (long ib = 0; ib < nbatches; ib++) {
                long odr = (npri > 1) ? (R - 1) / 2 : 0;
                lds.emplace_back(new RefLDS<T>(B, R, odr, npri, nt_in));
                for (long p = 0; p < npri; p++) {
                    long rp = p ? R - 1 : R;
                    long dd = rp / 2, amb = rp - dd;
                    s1_trees.emplace_back(new RefTree<T>(B, amb, dd, nt_in / pow2(p), 1, {1}));
                }
                for (auto &t : trees)
                    s2_trees.emplace_back(new RefTree<T>(B, t.amb_rank, t.dd_rank, t.nt_ds, 1, t.subband_counts));
            }
            long S = cfg.nelts_per_segment;
            for (long ib = 0; ib < nbatches; ib++)
                for (auto &t : trees) {
                    long rank = t.amb_rank + t.dd_rank;
                    std::vector<long> seg(B * pow2(rank)), res(B * pow2(rank));
                    for (long b = 0; b < B; b++)
                        for (long fc = 0; fc < pow2(t.dd_rank); fc++)
                            for (long dmb = 0; dmb < pow2(t.amb_rank); dmb++) {
                                long lag = rb_lag(fc, dmb, int(t.amb_rank), int(t.dd_rank), t.ipri > 0);
                                // segment lag on row (fc, dmb); residual lag on the transposed
                                // (a=dmb, f=fc) ordering used by the stage-2 kernel.
                                seg[b * pow2(rank) + fc * pow2(t.amb_rank) + dmb] = (lag / S) * S;
                                res[b * pow2(rank) + dmb * pow2(t.dd_rank) + fc] = lag % S;
                            }
                    s2_seg_lagbufs.emplace_back(seg, t.nt_ds);
                    s2_res_lagbufs.emplace_back(res, t.nt_ds);
                    (void) ib;
                }
            for (auto &t : trees)
                s2_bufs.emplace_back(B * pow2(t.amb_rank + t.dd_rank) * t.nt_ds, T(0));
        }
    }

    void grid_or_copy(T *dst)
    {
        if (tree_domain_input)
            std::memcpy(dst, input.data(), B * pow2(R) * nt_in * sizeof(T));
        else
            tree_gridding(dst, input.data(), cmap.data(), B, pow2(R), nfreq, nt_in);
    }

    void run_pf(long itree, T *sb, long ibatch)
    {
        double *var = out_var[itree].empty() ? nullptr : out_var[itree].data();
        pf[itree]->apply(out_max[itree].data(), out_argmax[itree].data(), var, sb, wt[itree].data(), ibatch);
    }

    void dedisperse(long ichunk, long ibatch)
    {
        (void) ichunk;
        ORC_CHECK(ibatch >= 0 && ibatch < nbatches, "ibatch out of range");
        const T c = T(0.7071067811865476);

        if (soph == 0) {
            grid_or_copy(ds_inputs[0].data());
            for (long p = 1; p < npri; p++) {
                long nts = nt_in / pow2(p - 1), ntd = nt_in / pow2(p);
                const T *src = ds_inputs[p - 1].data();
                T *dst = ds_inputs[p].data();
                for (long r = 0; r < B * pow2(R); r++)
                    for (long t = 0; t < ntd; t++)
                        dst[r * ntd + t] = c * (src[r * nts + 2 * t] + src[r * nts + 2 * t + 1]);
            }
            for (long it = 0; it < ntrees; it++) {
                const OrcTree &t = trees[it];
                long rr = t.amb_rank + t.dd_rank + (t.ipri > 0 ? 1 : 0);
                long nt = t.nt_ds;
                const T *in = ds_inputs[t.ipri].data();
                T *dd = dd_bufs0[it].data();
                for (long b = 0; b < B; b++)
                    std::memcpy(dd + b * pow2(rr) * nt, in + b * pow2(R) * nt, pow2(rr) * nt * sizeof(T));
                long ndm_sb = t.ndm_out * (t.ipri > 0 ? 2 : 1);
                T *sb = sb_bufs[it].data();
                trees0[ibatch * ntrees + it]->dedisperse(dd, pow2(rr) * nt, 0, nt, sb, ndm_sb * t.M * nt, t.M * nt, nt);
                if (t.ipri > 0) {
                    // Upper half of the DM rows only.
                    std::vector<T> half(B * t.ndm_out * t.M * nt);
                    for (long b = 0; b < B; b++)
                        std::memcpy(&half[b * t.ndm_out * t.M * nt], sb + (b * ndm_sb + t.ndm_out) * t.M * nt,
                                    t.ndm_out * t.M * nt * sizeof(T));
                    run_pf(it, half.data(), ibatch);
                }
                else
                    run_pf(it, sb, ibatch);
            }
            return;
        }

        // soph 1. With 16-bit storage, values are rounded where the GPU chain stores them:
        // gridded rows, lagged-downsampling outputs, stage-1 outputs (the lag ring), stage-2
        // subband rows (the peak-finder input) and out_max; everything in between is float.
        grid_or_copy(s1_bufs[0].data());
        round_to_storage(s1_bufs[0].data(), long(s1_bufs[0].size()), storage);
        {
            std::vector<T *> outs;
            for (long p = 1; p < npri; p++) outs.push_back(s1_bufs[p].data());
            lds[ibatch]->apply(s1_bufs[0].data(), outs.data());
            for (long p = 1; p < npri; p++)
                round_to_storage(s1_bufs[p].data(), long(s1_bufs[p].size()), storage);
        }
        for (long p = 0; p < npri; p++) {
            long rp = p ? R - 1 : R;
            long dd = rp / 2, amb = rp - dd, nt = nt_in / pow2(p);
            T *buf = s1_bufs[p].data();
            s1_trees[ibatch * npri + p]->dedisperse(buf, pow2(rp) * nt, pow2(dd) * nt, nt, nullptr, 0, 0, 0);
            round_to_storage(buf, long(s1_bufs[p].size()), storage);
            (void) amb;
        }
        for (long it = 0; it < ntrees; it++) {
            const OrcTree &t = trees[it];
            long rank = t.amb_rank + t.dd_rank, nt = t.nt_ds;
            long rp = t.ipri ? R - 1 : R;
            const T *src = s1_bufs[t.ipri].data();
            T *dst = s2_bufs[it].data();
            for (long b = 0; b < B; b++)
                std::memcpy(dst + b * pow2(rank) * nt, src + b * pow2(rp) * nt, pow2(rank) * nt * sizeof(T));
            s2_seg_lagbufs[ibatch * ntrees + it].apply(dst, [&](long r) { return r * nt; });
        }
        for (long it = 0; it < ntrees; it++) {
            const OrcTree &t = trees[it];
            long rank = t.amb_rank + t.dd_rank, nt = t.nt_ds;
            long A = pow2(t.amb_rank), F = pow2(t.dd_rank);
            T *dd = s2_bufs[it].data();
            // Transposed view: element (b, a, f, t) at ((b*F + f)*A + a)*nt + t.
            s2_res_lagbufs[ibatch * ntrees + it].apply(dd, [&](long r) {
                long f = r % F, a = (r / F) % A, b = r / (F * A);
                return ((b * F + f) * A + a) * nt;
            });
            std::vector<T> sb(B * t.ndm_out * t.M * nt);
            s2_trees[ibatch * ntrees + it]->dedisperse(dd, pow2(rank) * nt, nt, A * nt,
                                                      sb.data(), t.ndm_out * t.M * nt, t.M * nt, nt);
            round_to_storage(sb.data(), long(sb.size()), storage);
            run_pf(it, sb.data(), ibatch);
            round_to_storage(out_max[it].data(), long(out_max[it].size()), storage);
        }
    }
};

}  // namespace oracle

// -------------------------------------------------------------------------------------------
// Ring-buffer segment copies driven by octuples (reference: CpuRingbufCopyKernel::apply and its
// helper, src_lib/RingbufCopyKernel.cu:60-123). Octuple i = {src_offset, src_phase, src_len,
// src_nseg, dst_offset, dst_phase, dst_len, dst_nseg}; for beams b = 0..nbeams-1 in order,
// segment src_offset + ((src_phase + iframe + b) % src_len) * src_nseg is copied (memmove) to
// the same formula on the dst side. Octuples are processed in order.

inline void ref_ringbuf_copy(uint8_t *ringbuf, const uint32_t *oct, long noct, long seg_bytes, long nbeams,
                             uint64_t iframe)
{
    for (long i = 0; i < noct; i++) {
        const uint32_t *o = oct + 8 * i;
        uint64_t sph = (uint64_t(o[1]) + iframe) % uint64_t(o[2]);
        uint64_t dph = (uint64_t(o[5]) + iframe) % uint64_t(o[6]);
        for (long b = 0; b < nbeams; b++) {
            const uint64_t s = (uint64_t(o[0]) + sph * uint64_t(o[3])) * uint64_t(seg_bytes);
            const uint64_t d = (uint64_t(o[4]) + dph * uint64_t(o[7])) * uint64_t(seg_bytes);
            std::memmove(ringbuf + d, ringbuf + s, size_t(seg_bytes));
            sph = (sph + 1 == o[2]) ? 0 : sph + 1;
            dph = (dph + 1 == o[6]) ? 0 : dph + 1;
        }
    }
}

#endif


// -------------------------------------------------------------------------------------------
// Dequantization of packed int4 input (reference: ReferenceDequantizationKernel::apply,
// src_lib/GpuDequantizationKernel.cu:326-380; format in include/pirate/GpuDequantizationKernel.hpp:17-43).
// data: (nbeams, nfreq, ntime) int4, two per byte, low nibble = even sample; value -8 is the
// missing-sample sentinel and maps to 0. scoff: (nbeams, nfreq, ntime/256, 2) IEEE binary16
// bit patterns {scale, offset}, converted to float before any arithmetic.
// out = scale * value + offset (float32).

inline float half_bits_to_float(uint16_t h)
{
    // IEEE 754 binary16 -> binary32 (exact): sign, 5-bit exponent (bias 15), 10-bit fraction.
    const uint32_t sign = uint32_t(h >> 15) << 31;
    const int e = (h >> 10) & 0x1f;
    uint32_t frac = h & 0x3ff;
    uint32_t bits;
    if (e == 0x1f)
        bits = sign | 0x7f800000u | (frac << 13);                 // inf / nan
    else if (e != 0)
        bits = sign | (uint32_t(e - 15 + 127) << 23) | (frac << 13);
    else if (frac == 0)
        bits = sign;                                               // signed zero
    else {
        int ee = -14;                                              // subnormal: normalize
        while (!(frac & 0x400)) {
            frac <<= 1;
            ee--;
        }
        bits = sign | (uint32_t(ee + 127) << 23) | ((frac & 0x3ff) << 13);
    }
    float f;
    std::memcpy(&f, &bits, 4);
    return f;
}

inline void ref_dequantize(const uint8_t *data, const uint16_t *scoff, long nbeams, long nfreq, long ntime, float *out)
{
    const long nchunks = ntime / 256;
    for (long b = 0; b < nbeams; b++)
        for (long f = 0; f < nfreq; f++)
            for (long c = 0; c < nchunks; c++) {
                const long pair = (b * nfreq + f) * nchunks + c;
                const float scale = half_bits_to_float(scoff[2 * pair]);
                const float offset = half_bits_to_float(scoff[2 * pair + 1]);
                for (long k = 0; k < 256; k++) {
                    const long i = pair * 256 + k;
                    const int nib = (data[i >> 1] >> (4 * (i & 1))) & 0xf;
                    const int v = (nib >= 8) ? nib - 16 : nib;
                    float x = 0.0f;
                    if (v != -8) {
                        const float p = scale * float(v);
                        x = p + offset;
                    }
                    out[i] = x;
                }
            }
}
