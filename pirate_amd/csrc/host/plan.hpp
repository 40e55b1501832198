// DedispersionPlan: expands a DedispersionConfig into stage-1 trees (one per primary tree)
// and stage-2 output trees (one per (primary tree, early-trigger level)).
// Tree derivation follows the reference: src_lib/DedispersionPlan.cpp:29-156,
// include/pirate/DedispersionTree.hpp:19-70, include/pirate/DedispersionPlan.hpp.
// Argmax decoding follows src_lib/DedispersionPlan.cpp:481-661.

#ifndef PIRATE_AMD_PLAN_HPP
#define PIRATE_AMD_PLAN_HPP

#include "config.hpp"
#include "subbands.hpp"

#include <memory>
#include <string>
#include <vector>

namespace pirate_amd {

struct DedispersionTree {
    int primary_tree_index = -1;
    int early_trigger_level = -1;
    int amb_rank = 0;      // = stage-1 dd_rank of the primary tree
    int dd_rank = 0;       // = stage-1 amb_rank of the primary tree - early_trigger_level
    int nt_ds = 0;         // time samples per chunk after 2^p downsampling

    long total_rank() const { return amb_rank + dd_rank; }

    FrequencySubbands frequency_subbands;
    PrimaryTree pf;        // dm/time downsampling resolved (never 0 here)

    long Dcore = 0;        // peak-finder internal time granularity (argmax token quantum)
    long nprofiles = 0;    // 1 + 3*log2(max_width)
    long ndm_out = 0, ndm_wt = 0, nt_out = 0, nt_wt = 0;

    double dm_min = 0.0, dm_max = 0.0, trigger_frequency = 0.0;
};

struct DedispersionPlan {
    explicit DedispersionPlan(const DedispersionConfig &config, bool gpu_runnable = true);

    DedispersionConfig config;
    bool gpu_runnable = true;

    Dtype dtype = Dtype::float32;
    long nfreq = 0, nt_in = 0, num_primary_trees = 0;
    long beams_per_gpu = 0, beams_per_batch = 0, num_active_batches = 0, nbits = 0;
    int nelts_per_segment = 0, nbytes_per_segment = 0;

    std::vector<long> stage1_dd_rank, stage1_amb_rank;   // length num_primary_trees
    long ntrees = 0;
    std::vector<DedispersionTree> trees;

    std::string to_yaml_string(bool verbose = false) const;

    void decode_argmax(unsigned int token, long itree, long idm_coarse, long itime_coarse,
                       long &fmin, long &fmax, long &tlo, long &thi, long &p) const;
    void decode_argmax2(long itree, long fmin, long fmax, long tlo, long thi, long p,
                        double &freq_lo_MHz, double &freq_hi_MHz, double &dm,
                        double &timestamp_samp, double &width_samp) const;
    std::vector<long> compute_steady_state_it0(long itree) const;
};

}  // namespace pirate_amd

#endif
