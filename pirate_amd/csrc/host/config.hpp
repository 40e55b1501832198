// DedispersionConfig: YAML-backed configuration of the dedispersion chain.
// Field names, defaults and validation follow the reference:
// include/pirate/DedispersionConfig.hpp:21-95, src_lib/DedispersionConfig.cpp:32-717.

#ifndef PIRATE_AMD_CONFIG_HPP
#define PIRATE_AMD_CONFIG_HPP

#include <string>
#include <vector>

namespace pirate_amd {

// Storage dtype of the dedispersion buffers. The reference supports float32 and float16;
// bfloat16 is added here (storage only; arithmetic is float32 on the GPU).
enum class Dtype { float32 = 0, float16 = 1, bfloat16 = 2 };

const char *dtype_str(Dtype d);
Dtype dtype_from_str(const std::string &s);
inline int dtype_nbits(Dtype d) { return (d == Dtype::float32) ? 32 : 16; }

struct PrimaryTree {
    long num_early_triggers = 0;
    long max_width = 0;
    long dm_downsampling = 0;        // 0 = auto (2^pf_rank)
    long time_downsampling = 0;      // 0 = auto (dm_downsampling)
    long wt_dm_downsampling = 0;
    long wt_time_downsampling = 0;
};

struct DedispersionConfig {
    std::vector<long> zone_nfreq;
    std::vector<double> zone_freq_edges;
    double time_sample_ms = 0.0;

    long toplevel_tree_rank = -1;
    long time_samples_per_chunk = 0;
    Dtype dtype = Dtype::float32;

    std::vector<long> frequency_subband_counts;
    std::vector<PrimaryTree> primary_trees;

    long beams_per_gpu = 0;
    long beams_per_batch = 0;
    long num_active_batches = 0;
    long future_write_max_samples = 0;
    long max_gpu_clag = 10000;

    long num_primary_trees() const { return long(primary_trees.size()); }

    void validate() const;

    static DedispersionConfig from_yaml_string(const std::string &text, const std::string &name = "dedispersion_config");
    static DedispersionConfig from_yaml_file(const std::string &filename);
    std::string to_yaml_string(bool verbose = false) const;

    int get_nelts_per_segment() const;   // 128 bytes / element size
    long get_total_nfreq() const;

    double frequency_to_index(double freq) const;
    double index_to_frequency(double index) const;
    double delay_to_frequency(double delay) const;
    double frequency_to_delay(double freq) const;
    double dm_per_unit_delay() const;
    double max_dm_of_all_trees() const;
    long max_width_of_base_tree() const;

    // Length 2^toplevel_tree_rank + 1, monotonically decreasing, values in [0, nfreq].
    std::vector<double> make_channel_map() const;

    // Consistency checks of the frequency/delay maps (reference: DedispersionConfig::test()).
    void test() const;
};

}  // namespace pirate_amd

#endif
