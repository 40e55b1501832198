// Analytic peak-finder variances in closed form; the definition and the approximation are
// stated in pf_variance.hpp.

#include "pf_variance.hpp"
#include "common.hpp"

#include <algorithm>
#include <cmath>

namespace pirate_amd {

std::vector<double> pf_kernel_autocorrelations(long P, long &tlen)
{
    PDD_CHECK(P >= 1 && (P - 1) % 3 == 0, "pf_kernel_autocorrelations: P must be 1 + 3*nlevels, got " << P);
    std::vector<std::vector<double>> h(1, std::vector<double>(1, 1.0));
    for (long l = 0; 3 * l + 1 < P; l++) {
        const long w = 1L << l;
        std::vector<double> k1(2 * w, 1.0), k2(3 * w, 1.0), k3(4 * w, 1.0);
        for (long i = 0; i < w; i++) {
            k2[i] = k2[2 * w + i] = 0.5;
            k3[i] = k3[3 * w + i] = 0.5;
        }
        h.push_back(k1);
        h.push_back(k2);
        h.push_back(k3);
    }
    tlen = long(h.back().size());
    std::vector<double> A(P * tlen, 0.0);
    for (long p = 0; p < P; p++) {
        const long n = long(h[p].size());
        for (long lag = 0; lag < n; lag++) {
            double s = 0.0;
            for (long t = 0; t + lag < n; t++)
                s += h[p][t] * h[p][t + lag];
            A[p * tlen + lag] = s;
        }
    }
    return A;
}


long tree_partial_delay(long j, long d, int k)
{
    // Pass l (0 <= l < k) delays the even child (bit l of j clear) by the row index of the
    // level-l parent plus the new row bit: with the parent's top-l-bit index d >> (k-l) and
    // the child's d >> (k-1-l), that is (d >> (k-1-l)) - (d >> (k-l)).
    long tau = 0;
    for (int l = 0; l < k; l++)
        if (!((j >> l) & 1))
            tau += (d >> (k - 1 - l)) - (d >> (k - l));
    return tau;
}


PfAvarApproximation::PfAvarApproximation(const DedispersionPlan &plan, const double *fv, long nfreq_, long only_tree)
{
    nfreq = plan.nfreq;
    ntrees = plan.ntrees;
    PDD_CHECK(fv, "PfAvarApproximation: null freq_variances");
    PDD_CHECK(nfreq_ == nfreq, "PfAvarApproximation: expected " << nfreq << " freq_variances, got " << nfreq_);
    for (long i = 0; i < nfreq; i++)
        PDD_CHECK(std::isfinite(fv[i]) && fv[i] > 0.0, "PfAvarApproximation: freq_variances must be positive and finite (index " << i << ")");
    PDD_CHECK(only_tree < ntrees, "PfAvarApproximation: tree index out of range");

    const long Rtop = plan.config.toplevel_tree_rank;
    const std::vector<double> cm = plan.config.make_channel_map();
    const long nchan = long(cm.size()) - 1;

    tree_variance.assign(ntrees, {});
    per_tf.assign(ntrees, {});

    for (long itree = 0; itree < ntrees; itree++) {
        if (only_tree >= 0 && itree != only_tree)
            continue;
        const DedispersionTree &tr = plan.trees[itree];
        const FrequencySubbands &fs = tr.frequency_subbands;
        const long r = tr.total_rank();
        const long L = integer_log2(tr.pf.wt_dm_downsampling);
        const long R = fs.pf_rank;
        const long P = tr.nprofiles;
        const long N = fs.N;
        const bool upper = (tr.primary_tree_index > 0);
        PDD_CHECK(R >= 0 && R <= L && L <= r, "PfAvarApproximation: tree " << itree << " needs pf_rank <= log2(wt_dm_downsampling) <= tree rank");
        const int k = int(r - L + (upper ? 1 : 0));
        PDD_CHECK(k <= Rtop && L + k <= Rtop, "PfAvarApproximation: tree " << itree << " level out of range");
        const long nrow = 1L << (r - L);                 // == ndm_wt
        PDD_CHECK(nrow == tr.ndm_wt, "PfAvarApproximation: ndm_wt mismatch");
        const long row0 = upper ? nrow : 0;              // upper half of the 2^k level-k rows
        const long nband = 1L << L;                      // level-k bands used (F < 2^L)
        const long fshift = L - R;                       // level-k band -> pf-rank band
        const double band_norm = std::ldexp(1.0, -int(fshift + k));

        long tlen = 0;
        const std::vector<double> A = pf_kernel_autocorrelations(P, tlen);

        std::vector<double> ptf((1L << R) * nrow * P, 0.0);
        std::vector<double> diag(1L << R, 0.0);           // sum of s_i w^2 2^-k 2^-(L-R) per pf-rank band
        std::vector<long> nlist;
        std::vector<double> wlist;
        std::vector<long> taus;

        // Tree channels overlapping input channel i: cm[n] > i and cm[n+1] < i+1. The
        // channel map is decreasing, so both conditions select a contiguous range.
        for (long i = 0; i < nfreq; i++) {
            long lo = 0, hi = nchan;                      // first n with cm[n+1] < i+1
            while (lo < hi) {
                const long mid = (lo + hi) >> 1;
                if (cm[mid + 1] >= double(i + 1)) lo = mid + 1; else hi = mid;
            }
            const long n_lo = lo;
            lo = 0, hi = nchan;                           // first n with cm[n] <= i
            while (lo < hi) {
                const long mid = (lo + hi) >> 1;
                if (cm[mid] > double(i)) lo = mid + 1; else hi = mid;
            }
            const long n_hi = lo;
            const double si = fv[i];

            long n = n_lo;
            while (n < n_hi) {
                const long F = n >> k;
                const long nend = std::min(n_hi, (F + 1) << k);
                if (F >= nband)
                    break;                                // channels only increase from here
                nlist.clear();
                wlist.clear();
                for (long m = n; m < nend; m++) {
                    const double w = std::min(cm[m], double(i + 1)) - std::max(cm[m + 1], double(i));
                    if (w > 0.0) {
                        nlist.push_back(m & ((1L << k) - 1));
                        wlist.push_back(w);
                    }
                }
                n = nend;
                const long g = long(nlist.size());
                if (g == 0)
                    continue;
                const long f = F >> fshift;
                const double c = si * band_norm;
                double dsum = 0.0;
                for (long a = 0; a < g; a++)
                    dsum += wlist[a] * wlist[a];
                diag[f] += c * dsum;
                if (g < 2)
                    continue;

                // Off-diagonal pairs: 2 w_a w_b A_p(|tau_a - tau_b|), row by row.
                taus.resize(g);
                double *out = &ptf[f * nrow * P];
                for (long D = 0; D < nrow; D++) {
                    const long d = D + row0;
                    for (long a = 0; a < g; a++)
                        taus[a] = tree_partial_delay(nlist[a], d, k);
                    double *o = out + D * P;
                    for (long a = 0; a < g; a++)
                        for (long b = a + 1; b < g; b++) {
                            const long lag = std::labs(taus[a] - taus[b]);
                            if (lag >= tlen)
                                continue;
                            const double cw = 2.0 * c * wlist[a] * wlist[b];
                            for (long p = 0; p < P; p++)
                                o[p] += cw * A[p * tlen + lag];
                        }
                }
            }
        }

        for (long f = 0; f < (1L << R); f++)
            for (long D = 0; D < nrow; D++)
                for (long p = 0; p < P; p++)
                    ptf[(f * nrow + D) * P + p] += diag[f] * A[p * tlen];

        std::vector<double> tv(N * nrow * P, 0.0);
        for (long s = 0; s < N; s++) {
            const long flo = fs.n_to_flo[s], fhi = fs.n_to_fhi[s];
            PDD_CHECK(0 <= flo && flo < fhi && fhi <= (1L << R), "PfAvarApproximation: bad subband range");
            const double inv = 1.0 / double(fhi - flo);
            for (long f = flo; f < fhi; f++)
                for (long x = 0; x < nrow * P; x++)
                    tv[s * nrow * P + x] += inv * ptf[f * nrow * P + x];
        }
        for (long x = 0; x < N * nrow * P; x++)
            PDD_CHECK(tv[x] > 0.0, "PfAvarApproximation: tree " << itree << " has a zero variance (a band without input channels)");
        tree_variance[itree] = std::move(tv);
        per_tf[itree] = std::move(ptf);
    }
}

}  // namespace pirate_amd
