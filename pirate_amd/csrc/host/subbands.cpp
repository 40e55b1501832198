#include "subbands.hpp"
#include "common.hpp"

#include <algorithm>
#include <cmath>
#include <iomanip>

namespace pirate_amd {

static constexpr long max_peak_finding_rank = 4;   // reference: include/pirate/constants.hpp:43


FrequencySubbands::FrequencySubbands() : FrequencySubbands(std::vector<long>{1}) { }


// Reference: src_lib/FrequencySubbands.cpp:32-60.
FrequencySubbands::FrequencySubbands(const std::vector<long> &sc) : subband_counts(sc)
{
    validate_subband_counts(subband_counts);
    pf_rank = long(subband_counts.size()) - 1;

    for (long level = 0; level <= pf_rank; level++) {
        long spacing = pow2(std::max(level - 1, 0L));
        for (long b = 0; b < subband_counts[level]; b++) {
            n_to_flo.push_back(b * spacing);
            n_to_fhi.push_back(b * spacing + pow2(level));
            n_to_mbase.push_back(M);
            for (long d = 0; d < pow2(level); d++) {
                m_to_n.push_back(N);
                m_to_d.push_back(d);
            }
            M += pow2(level);
            N += 1;
        }
    }
}


// Reference: src_lib/FrequencySubbands.cpp:63-84 (f_to_freq equally spaced in freq^-2).
FrequencySubbands::FrequencySubbands(const std::vector<long> &sc, double fmin_, double fmax_) :
    FrequencySubbands(sc)
{
    fmin = fmin_;
    fmax = fmax_;
    PDD_CHECK(fmin > 0 && fmax > fmin, "FrequencySubbands: bad (fmin,fmax)");
    long nf = pow2(pf_rank) + 1;
    double start = std::pow(fmax, -2.0);
    double end = std::pow(fmin, -2.0);
    f_to_freq.resize(nf);
    for (long f = 0; f < nf; f++) {
        double t = double(f) / double(nf - 1);
        f_to_freq[f] = std::pow(start + t * (end - start), -0.5);
    }
}


// Reference: src_lib/FrequencySubbands.cpp:137-160.
void FrequencySubbands::validate_subband_counts(const std::vector<long> &sc)
{
    PDD_CHECK(!sc.empty(), "FrequencySubbands: subband_counts must be nonempty");
    long pf_rank = long(sc.size()) - 1;
    PDD_CHECK(sc[pf_rank] == 1, "FrequencySubbands: subband_counts[pf_rank] must be 1 (full band), got "
              << vec_str(sc));
    if (pf_rank > max_peak_finding_rank)
        throw pdd_error("FrequencySubbands: max allowed pf_rank is " + std::to_string(max_peak_finding_rank));
    for (long level = 0; level <= pf_rank; level++) {
        long max_bands = (level > 0) ? (pow2(pf_rank + 1 - level) - 1) : pow2(pf_rank);
        PDD_CHECK((sc[level] >= 0) && (sc[level] <= max_bands),
                  "FrequencySubbands: subband_counts[" << level << "]=" << sc[level] << " out of range");
    }
}


// Reference: src_lib/FrequencySubbands.cpp:165-201.
std::vector<long> FrequencySubbands::restrict_subband_counts(const std::vector<long> &sc,
                                                             long et_level, long new_pf_rank)
{
    validate_subband_counts(sc);
    PDD_CHECK(et_level >= 0 && new_pf_rank >= 0, "restrict_subband_counts: bad args");

    long src_rank = long(sc.size()) - 1;
    long early_rank = std::max(src_rank - et_level, 0L);

    std::vector<long> early(early_rank + 1);
    early[early_rank] = 1;
    for (long lev = 0; lev < early_rank; lev++) {
        long max_count = (lev > 0) ? (pow2(early_rank + 1 - lev) - 1) : pow2(early_rank);
        early[lev] = std::min(sc[lev], max_count);
    }

    std::vector<long> dst(new_pf_rank + 1);
    dst[new_pf_rank] = 1;
    for (long lev = 0; lev < new_pf_rank; lev++) {
        long src_level = lev - new_pf_rank + early_rank;
        if ((src_level < 0) || (early[src_level] == 0))
            dst[lev] = 0;
        else if ((src_level == 0) && (lev > 0))
            dst[lev] = 2 * early[src_level] - 1;
        else if ((src_level > 0) && (lev == 0))
            dst[lev] = early[src_level] / 2 + 1;
        else
            dst[lev] = early[src_level];
    }

    validate_subband_counts(dst);
    return dst;
}


// Reference: src_lib/FrequencySubbands.cpp:88-134.
FrequencySubbands FrequencySubbands::from_threshold(double fmin, double fmax, double threshold, long pf_rank)
{
    PDD_CHECK(fmin > 0 && fmax > fmin && pf_rank >= 0, "from_threshold: bad args");
    if (pf_rank > max_peak_finding_rank)
        throw pdd_error("FrequencySubbands::from_threshold: max allowed pf_rank is " + std::to_string(max_peak_finding_rank));

    long nf = pow2(pf_rank) + 1;
    double start = std::pow(fmax, -2.0), end = std::pow(fmin, -2.0);
    std::vector<double> f2f(nf);
    for (long f = 0; f < nf; f++)
        f2f[f] = std::pow(start + (double(f) / double(nf - 1)) * (end - start), -0.5);

    std::vector<long> sc(pf_rank + 1, 0);
    for (long level = 0; level <= pf_rank; level++) {
        long max_bands = (level > 0) ? (pow2(pf_rank + 1 - level) - 1) : pow2(pf_rank);
        long s = pow2(std::max(level - 1, 0L));
        for (long b = 0; b < max_bands; b++) {
            long flo = b * s, fhi = b * s + pow2(level);
            double freq_lo = f2f[fhi], freq_hi = f2f[flo];
            if ((level == pf_rank) || ((freq_hi / freq_lo) > (1.0 + threshold)))
                sc[level]++;
        }
    }
    return FrequencySubbands(sc, fmin, fmax);
}


std::string FrequencySubbands::to_string() const
{
    std::ostringstream os;
    os << "FrequencySubbands(subband_counts=" << vec_str(subband_counts) << ")\n";
    for (long n = 0; n < N; n++) {
        long flo = n_to_flo[n], fhi = n_to_fhi[n];
        os << "  n=" << n << ": level=" << integer_log2(fhi - flo)
           << "  (mlo,mhi)=(" << n_to_mbase[n] << "," << (n_to_mbase[n] + fhi - flo) << ")"
           << "  (flo,fhi)=(" << flo << "," << fhi << ")";
        if (!f_to_freq.empty())
            os << std::fixed << std::setprecision(1) << "  (freq_lo,freq_hi)=(" << f_to_freq[fhi] << "," << f_to_freq[flo] << ")";
        os << "\n";
    }
    os << "N=" << N << "\nM=" << M << "\n";
    return os.str();
}

}  // namespace pirate_amd
