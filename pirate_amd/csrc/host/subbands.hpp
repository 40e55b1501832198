// FrequencySubbands: the (pf_level, subband) scheme searched by the peak finder.
// Semantics restated from the reference: include/pirate/FrequencySubbands.hpp:22-80,
// src_lib/FrequencySubbands.cpp:27-190.
//
//   subband_counts has length pf_rank+1; subband_counts[pf_rank] == 1 (full band).
//   Level 0: band s covers coarse-freq channel [s, s+1).
//   Level l>0: band s covers [s*2^(l-1), (s+2)*2^(l-1)) (overlapping, half-stepped).
//   Each band at level l carries 2^l fine DMs; a (band, fine DM) pair is a "multiplet" m.

#ifndef PIRATE_AMD_SUBBANDS_HPP
#define PIRATE_AMD_SUBBANDS_HPP

#include <string>
#include <vector>

namespace pirate_amd {

struct FrequencySubbands {
    FrequencySubbands();   // subband_counts = {1}
    explicit FrequencySubbands(const std::vector<long> &subband_counts);
    FrequencySubbands(const std::vector<long> &subband_counts, double fmin, double fmax);

    std::vector<long> subband_counts;
    long pf_rank = -1;
    long N = 0;   // number of subbands
    long M = 0;   // number of multiplets

    std::vector<long> m_to_n, m_to_d;
    std::vector<long> n_to_flo, n_to_fhi, n_to_mbase;

    std::vector<double> f_to_freq;   // length 2^pf_rank+1, only if fmin/fmax given
    double fmin = 0.0, fmax = 0.0;

    long m_to_flo(long m) const { return n_to_flo.at(m_to_n.at(m)); }
    long m_to_fhi(long m) const { return n_to_fhi.at(m_to_n.at(m)); }

    std::string to_string() const;

    static void validate_subband_counts(const std::vector<long> &subband_counts);
    static std::vector<long> restrict_subband_counts(const std::vector<long> &subband_counts,
                                                     long early_trigger_level, long new_pf_rank);
    static FrequencySubbands from_threshold(double fmin, double fmax, double threshold, long pf_rank);
};

}  // namespace pirate_amd

#endif
