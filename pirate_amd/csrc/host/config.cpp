#include "config.hpp"
#include "common.hpp"
#include "subbands.hpp"
#include "yaml_lite.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iomanip>

namespace pirate_amd {

static constexpr int max_tree_rank = 16;          // reference: include/pirate/constants.hpp:40
static constexpr int max_primary_trees = 7;       // constants.hpp:41
static constexpr int max_pf_width = 32;           // constants.hpp:42
static constexpr double k_dm = 4.148808e6;        // constants.hpp:49 (ms MHz^2 / (pc cm^-3))
static constexpr int bytes_per_segment = 128;     // constants.hpp:22


const char *dtype_str(Dtype d)
{
    switch (d) {
        case Dtype::float32: return "float32";
        case Dtype::float16: return "float16";
        case Dtype::bfloat16: return "bfloat16";
    }
    return "?";
}

Dtype dtype_from_str(const std::string &s)
{
    if (s == "float32" || s == "fp32" || s == "f32")
        return Dtype::float32;
    if (s == "float16" || s == "fp16" || s == "f16")
        return Dtype::float16;
    if (s == "bfloat16" || s == "bf16")
        return Dtype::bfloat16;
    throw pdd_error("DedispersionConfig: invalid dtype '" + s + "' (expected float32, float16 or bfloat16)");
}


// Shortest decimal representation that round-trips.
static std::string fmt_double(double x)
{
    char buf[64];
    if (x == std::floor(x) && std::fabs(x) < 1.0e15) {
        snprintf(buf, sizeof(buf), "%.0f", x);
        return buf;
    }
    for (int prec = 1; prec <= 17; prec++) {
        snprintf(buf, sizeof(buf), "%.*g", prec, x);
        if (std::strtod(buf, nullptr) == x)
            return buf;
    }
    return buf;
}


int DedispersionConfig::get_nelts_per_segment() const
{
    return bytes_per_segment / (dtype_nbits(dtype) / 8);
}


long DedispersionConfig::get_total_nfreq() const
{
    long ret = 0;
    for (long n : zone_nfreq)
        ret += n;
    return ret;
}


// Reference: src_lib/DedispersionConfig.cpp:52-96.
double DedispersionConfig::frequency_to_index(double freq) const
{
    double fmin = zone_freq_edges.front(), fmax = zone_freq_edges.back();
    double eps = 1.0e-5 * (fmax - fmin);
    if ((freq < fmin - eps) || (freq > fmax + eps)) {
        std::ostringstream ss;
        ss << "DedispersionConfig::frequency_to_index(): frequency " << freq << " is out of range [" << fmin << ", " << fmax << "]";
        throw pdd_error(ss.str());
    }
    freq = std::min(std::max(freq, fmin), fmax);

    double channel_offset = 0;
    for (size_t i = 0; i < zone_nfreq.size(); i++) {
        double f0 = zone_freq_edges[i], f1 = zone_freq_edges[i+1];
        if (freq <= f1) {
            channel_offset += ((freq - f0) / (f1 - f0)) * zone_nfreq[i];
            break;
        }
        channel_offset += zone_nfreq[i];
    }
    double tot = double(get_total_nfreq());
    return std::min(std::max(channel_offset, 0.0), tot);
}


// Reference: src_lib/DedispersionConfig.cpp:99-132.
double DedispersionConfig::index_to_frequency(double index) const
{
    double tot = double(get_total_nfreq());
    double eps = 1.0e-5 * tot;
    if ((index < -eps) || (index > tot + eps)) {
        std::ostringstream ss;
        ss << "DedispersionConfig::index_to_frequency(): index " << index << " is out of range [0, " << tot << "]";
        throw pdd_error(ss.str());
    }
    index = std::min(std::max(index, 0.0), tot);

    double off = 0;
    for (size_t i = 0; i < zone_nfreq.size(); i++) {
        double next = off + zone_nfreq[i];
        if (index <= next) {
            double frac = (index - off) / zone_nfreq[i];
            return zone_freq_edges[i] + frac * (zone_freq_edges[i+1] - zone_freq_edges[i]);
        }
        off = next;
    }
    return zone_freq_edges.back();
}


// Reference: src_lib/DedispersionConfig.cpp:135-162. freq = (d/scale + fhi^-2)^(-1/2),
// scale = 2^R / (flo^-2 - fhi^-2).
double DedispersionConfig::delay_to_frequency(double delay) const
{
    double flo = zone_freq_edges.front(), fhi = zone_freq_edges.back();
    double ntree = double(pow2(toplevel_tree_rank));
    double eps = 1.0e-5 * ntree;
    if ((delay < -eps) || (delay > ntree + eps)) {
        std::ostringstream ss;
        ss << "DedispersionConfig::delay_to_frequency(): delay " << delay << " is out of range [0, " << ntree << "]";
        throw pdd_error(ss.str());
    }
    delay = std::min(std::max(delay, 0.0), ntree);
    double scale = ntree / (1.0/(flo*flo) - 1.0/(fhi*fhi));
    return 1.0 / std::sqrt(delay/scale + 1.0/(fhi*fhi));
}


// Reference: src_lib/DedispersionConfig.cpp:165-193.
double DedispersionConfig::frequency_to_delay(double freq) const
{
    double flo = zone_freq_edges.front(), fhi = zone_freq_edges.back();
    double ntree = double(pow2(toplevel_tree_rank));
    double eps = 1.0e-5 * (fhi - flo);
    if ((freq < flo - eps) || (freq > fhi + eps)) {
        std::ostringstream ss;
        ss << "DedispersionConfig::frequency_to_delay(): frequency " << freq << " is out of range [" << flo << ", " << fhi << "]";
        throw pdd_error(ss.str());
    }
    freq = std::min(std::max(freq, flo), fhi);
    double d = 1.0/(freq*freq), dlo = 1.0/(flo*flo), dhi = 1.0/(fhi*fhi);
    return ntree * (d - dhi) / (dlo - dhi);
}


// Reference: src_lib/DedispersionConfig.cpp:195-214.
double DedispersionConfig::dm_per_unit_delay() const
{
    PDD_CHECK(zone_freq_edges.size() >= 2, "dm_per_unit_delay(): zone_freq_edges not set");
    double flo = zone_freq_edges.front(), fhi = zone_freq_edges.back();
    PDD_CHECK(flo > 0.0 && flo < fhi, "dm_per_unit_delay(): bad band edges");
    return time_sample_ms / (k_dm * (1.0/(flo*flo) - 1.0/(fhi*fhi)));
}


double DedispersionConfig::max_dm_of_all_trees() const
{
    PDD_CHECK(num_primary_trees() >= 1, "max_dm_of_all_trees(): no primary trees");
    double dm0 = dm_per_unit_delay() * double(pow2(toplevel_tree_rank));
    return dm0 * double(pow2(num_primary_trees() - 1));
}


long DedispersionConfig::max_width_of_base_tree() const
{
    PDD_CHECK(!primary_trees.empty(), "max_width_of_base_tree(): no primary trees");
    return primary_trees[0].max_width;
}


// Reference: src_lib/DedispersionConfig.cpp:238-250.
std::vector<double> DedispersionConfig::make_channel_map() const
{
    long nchan = pow2(toplevel_tree_rank);
    std::vector<double> cmap(nchan + 1);
    for (long n = 0; n <= nchan; n++)
        cmap[n] = frequency_to_index(delay_to_frequency(double(n)));
    return cmap;
}


// Reference: src_lib/DedispersionConfig.cpp:275-330 (randomized there; a fixed
// low-discrepancy sample here, so that failures reproduce).
void DedispersionConfig::test() const
{
    validate();
    double flo = zone_freq_edges.front(), fhi = zone_freq_edges.back();
    double tot = double(get_total_nfreq());
    double ntree = double(pow2(toplevel_tree_rank));

    double expected_index = 0;
    for (size_t i = 0; i < zone_freq_edges.size(); i++) {
        double f = zone_freq_edges[i];
        PDD_CHECK(std::fabs(frequency_to_index(f) - expected_index) < 1.0e-4 * tot, "test(): frequency_to_index at zone edge");
        PDD_CHECK(std::fabs(index_to_frequency(expected_index) - f) < 1.0e-4 * fhi, "test(): index_to_frequency at zone edge");
        if (i < zone_nfreq.size())
            expected_index += zone_nfreq[i];
    }

    PDD_CHECK(std::fabs(delay_to_frequency(0.0) - fhi) < 1.0e-4 * fhi, "test(): delay_to_frequency(0)");
    PDD_CHECK(std::fabs(delay_to_frequency(ntree) - flo) < 1.0e-4 * fhi, "test(): delay_to_frequency(ntree)");
    PDD_CHECK(std::fabs(frequency_to_delay(fhi)) < 1.0e-4 * ntree, "test(): frequency_to_delay(fhi)");
    PDD_CHECK(std::fabs(frequency_to_delay(flo) - ntree) < 1.0e-4 * ntree, "test(): frequency_to_delay(flo)");

    const double golden = 0.6180339887498949;
    for (int i = 0; i < 64; i++) {
        double u = std::fmod(0.5 + i * golden, 1.0);
        double index = u * tot;
        PDD_CHECK(std::fabs(frequency_to_index(index_to_frequency(index)) - index) < 1.0e-4 * tot, "test(): index round trip");
        double f = flo + u * (fhi - flo);
        PDD_CHECK(std::fabs(index_to_frequency(frequency_to_index(f)) - f) < 1.0e-4 * fhi, "test(): frequency round trip");
        double d = u * ntree;
        PDD_CHECK(std::fabs(frequency_to_delay(delay_to_frequency(d)) - d) < 1.0e-4 * ntree, "test(): delay round trip");
        PDD_CHECK(std::fabs(delay_to_frequency(frequency_to_delay(f)) - f) < 1.0e-4 * fhi, "test(): delay/frequency round trip");
    }
}


// Reference: src_lib/DedispersionConfig.cpp:333-435.
void DedispersionConfig::validate() const
{
    PDD_CHECK(toplevel_tree_rank > 0, "DedispersionConfig: toplevel_tree_rank must be > 0");
    PDD_CHECK(primary_trees.size() > 0, "DedispersionConfig: primary_trees must be nonempty");
    PDD_CHECK(time_samples_per_chunk > 0, "DedispersionConfig: time_samples_per_chunk must be > 0");
    PDD_CHECK(time_sample_ms > 0, "DedispersionConfig: time_sample_ms must be > 0");
    PDD_CHECK(beams_per_gpu > 0, "DedispersionConfig: beams_per_gpu must be > 0");
    PDD_CHECK(beams_per_batch > 0, "DedispersionConfig: beams_per_batch must be > 0");
    PDD_CHECK(num_active_batches > 0, "DedispersionConfig: num_active_batches must be > 0");
    PDD_CHECK(toplevel_tree_rank <= max_tree_rank, "DedispersionConfig: toplevel_tree_rank must be <= " << max_tree_rank);
    PDD_CHECK(num_primary_trees() <= max_primary_trees, "DedispersionConfig: at most " << max_primary_trees << " primary trees");

    PDD_CHECK(zone_nfreq.size() > 0, "DedispersionConfig: zone_nfreq must be nonempty");
    PDD_CHECK(zone_freq_edges.size() == zone_nfreq.size() + 1, "DedispersionConfig: zone_freq_edges must have length len(zone_nfreq)+1");
    for (long n : zone_nfreq)
        PDD_CHECK(n > 0, "DedispersionConfig: zone_nfreq entries must be > 0");
    for (size_t i = 0; i + 1 < zone_freq_edges.size(); i++) {
        PDD_CHECK(zone_freq_edges[i] > 0.0, "DedispersionConfig: zone_freq_edges must be positive");
        PDD_CHECK(zone_freq_edges[i] < zone_freq_edges[i+1], "DedispersionConfig: zone_freq_edges must be increasing");
    }

    long min_nt = long(get_nelts_per_segment()) * pow2(num_primary_trees() - 1);
    if (time_samples_per_chunk % min_nt) {
        std::ostringstream ss;
        ss << "DedispersionConfig: time_samples_per_chunk=" << time_samples_per_chunk
           << " must be a multiple of " << min_nt << " (this value depends on dtype and the number of primary trees)";
        throw pdd_error(ss.str());
    }

    PDD_CHECK(beams_per_gpu % beams_per_batch == 0, "DedispersionConfig: beams_per_gpu must be a multiple of beams_per_batch");
    PDD_CHECK(num_active_batches * beams_per_batch <= beams_per_gpu, "DedispersionConfig: num_active_batches * beams_per_batch must be <= beams_per_gpu");
    PDD_CHECK(max_gpu_clag >= 0, "DedispersionConfig: max_gpu_clag must be >= 0");
    PDD_CHECK(future_write_max_samples >= 0, "DedispersionConfig: future_write_max_samples must be >= 0");

    FrequencySubbands::validate_subband_counts(frequency_subband_counts);

    for (long ipri = 0; ipri < num_primary_trees(); ipri++) {
        const PrimaryTree &pt = primary_trees[ipri];
        long primary_tree_rank = ipri ? (toplevel_tree_rank - 1) : toplevel_tree_rank;
        long stage1_dd_rank = primary_tree_rank / 2;

        PDD_CHECK(pt.num_early_triggers >= 0, "primary_trees[" << ipri << "]: num_early_triggers must be >= 0");
        PDD_CHECK(pt.num_early_triggers <= primary_tree_rank - stage1_dd_rank,
                  "primary_trees[" << ipri << "]: num_early_triggers must be <= " << (primary_tree_rank - stage1_dd_rank));
        PDD_CHECK(pt.max_width > 0 && is_power_of_two(pt.max_width), "primary_trees[" << ipri << "]: max_width must be a power of two");
        PDD_CHECK(pt.max_width <= max_pf_width, "primary_trees[" << ipri << "]: max_width must be <= " << max_pf_width);
        PDD_CHECK(pt.wt_dm_downsampling > 0 && is_power_of_two(pt.wt_dm_downsampling), "primary_trees[" << ipri << "]: wt_dm_downsampling must be a power of two");
        PDD_CHECK(pt.wt_time_downsampling > 0 && is_power_of_two(pt.wt_time_downsampling), "primary_trees[" << ipri << "]: wt_time_downsampling must be a power of two");

        long min_rank = primary_tree_rank - pt.num_early_triggers;
        if (pt.wt_dm_downsampling > pow2(min_rank)) {
            std::ostringstream ss;
            ss << "DedispersionConfig: wt_dm_downsampling[" << ipri << "]=" << pt.wt_dm_downsampling
               << " must be <= " << pow2(min_rank) << ". This upper bound is set by the smallest"
               << " tree at primary_tree_index=" << ipri << " (i.e. the earliest trigger).";
            throw pdd_error(ss.str());
        }
        if (pt.dm_downsampling > 0) {
            PDD_CHECK(is_power_of_two(pt.dm_downsampling), "primary_trees[" << ipri << "]: dm_downsampling must be a power of two");
            PDD_CHECK(pt.wt_dm_downsampling >= pt.dm_downsampling, "primary_trees[" << ipri << "]: wt_dm_downsampling must be >= dm_downsampling");
        }
        if (pt.time_downsampling > 0) {
            PDD_CHECK(is_power_of_two(pt.time_downsampling), "primary_trees[" << ipri << "]: time_downsampling must be a power of two");
            PDD_CHECK(pt.wt_time_downsampling >= pt.time_downsampling, "primary_trees[" << ipri << "]: wt_time_downsampling must be >= time_downsampling");
        }
    }
}


// Reference: src_lib/DedispersionConfig.cpp:651-717.
DedispersionConfig DedispersionConfig::from_yaml_string(const std::string &text, const std::string &name)
{
    YamlView f(YamlNode::parse(text, name), name);
    PDD_CHECK(f.is_map(), name << ": top-level YAML node must be a mapping");

    if (f.has_key("tree_rank"))
        throw pdd_error(name + ": key 'tree_rank' is from an old config syntax -- it has been"
                        " renamed to 'toplevel_tree_rank' (same meaning).");
    for (const char *k : {"peak_finding_params", "early_triggers", "num_downsampling_levels"}) {
        if (f.has_key(k))
            throw pdd_error(name + ": key '" + k + "' is from an old config syntax. The keys"
                            " 'peak_finding_params', 'early_triggers', and 'num_downsampling_levels'"
                            " have been replaced by 'primary_trees'.");
    }

    DedispersionConfig ret;
    ret.zone_nfreq = f["zone_nfreq"].as_long_vector();
    ret.zone_freq_edges = f["zone_freq_edges"].as_double_vector();
    ret.time_sample_ms = f.get_double("time_sample_ms");
    ret.toplevel_tree_rank = f.get_long("toplevel_tree_rank");
    ret.time_samples_per_chunk = f.get_long("time_samples_per_chunk");
    ret.dtype = dtype_from_str(f.get_string("dtype"));
    ret.beams_per_gpu = f.get_long("beams_per_gpu");
    ret.beams_per_batch = f.get_long("beams_per_batch");
    ret.num_active_batches = f.get_long("num_active_batches");
    ret.future_write_max_samples = f.get_long("future_write_max_samples");
    ret.max_gpu_clag = f.get_long("max_gpu_clag", 10000);
    ret.frequency_subband_counts = f["frequency_subband_counts"].as_long_vector();

    YamlView pts = f["primary_trees"];
    for (long i = 0; i < pts.size(); i++) {
        YamlView p = pts[i];
        PrimaryTree pt;
        pt.num_early_triggers = p.get_long("num_early_triggers");
        pt.max_width = p.get_long("max_width");
        pt.dm_downsampling = p.get_long("dm_downsampling", 0);
        pt.time_downsampling = p.get_long("time_downsampling", 0);
        pt.wt_dm_downsampling = p.get_long("wt_dm_downsampling");
        pt.wt_time_downsampling = p.get_long("wt_time_downsampling");
        ret.primary_trees.push_back(pt);
        p.check_for_invalid_keys();
    }

    f.check_for_invalid_keys();
    ret.validate();
    return ret;
}


DedispersionConfig DedispersionConfig::from_yaml_file(const std::string &filename)
{
    std::ifstream in(filename);
    if (!in)
        throw pdd_error("DedispersionConfig::from_yaml(): couldn't open '" + filename + "'");
    std::stringstream ss;
    ss << in.rdbuf();
    return from_yaml_string(ss.str(), filename);
}


// Same key order and flow style as the reference emitter (src_lib/DedispersionConfig.cpp:438-640).
std::string DedispersionConfig::to_yaml_string(bool verbose) const
{
    validate();
    std::ostringstream os;
    auto seq = [](const auto &v) {
        std::ostringstream s;
        s << "[";
        for (size_t i = 0; i < v.size(); i++) {
            if (i) s << ", ";
            if constexpr (std::is_same_v<std::decay_t<decltype(v[0])>, double>)
                s << fmt_double(v[i]);
            else
                s << v[i];
        }
        s << "]";
        return s.str();
    };

    if (verbose) {
        os << "# Frequency channels. The observed band is divided into zones of equal-width channels.\n";
        os << "# Total frequency channels: " << get_total_nfreq() << "\n";
    }
    os << "zone_nfreq: " << seq(zone_nfreq) << "\n";
    os << "zone_freq_edges: " << seq(zone_freq_edges) << "\n";
    os << "time_sample_ms: " << fmt_double(time_sample_ms) << "\n";
    if (verbose) {
        os << "# DM ranges searched by each primary tree:\n";
        for (long p = 0; p < num_primary_trees(); p++) {
            long dlo = (p == 0) ? 0 : pow2(toplevel_tree_rank + p - 1);
            long dhi = pow2(toplevel_tree_rank + p);
            os << std::fixed << std::setprecision(1) << "#   p=" << p << ": " << time_sample_ms * pow2(p)
               << " ms samples, DM range [" << dlo * dm_per_unit_delay() << ", " << dhi * dm_per_unit_delay() << "] pc/cm^3\n";
            os.unsetf(std::ios::fixed);
        }
    }
    os << "toplevel_tree_rank: " << toplevel_tree_rank << "\n";
    os << "time_samples_per_chunk: " << time_samples_per_chunk << "\n";
    os << "dtype: " << dtype_str(dtype) << "\n";
    os << "frequency_subband_counts: " << seq(frequency_subband_counts) << "\n";
    os << "primary_trees:\n";
    for (const PrimaryTree &pt : primary_trees) {
        os << "  - {num_early_triggers: " << pt.num_early_triggers << ", max_width: " << pt.max_width
           << ", dm_downsampling: " << pt.dm_downsampling << ", time_downsampling: " << pt.time_downsampling
           << ", wt_dm_downsampling: " << pt.wt_dm_downsampling << ", wt_time_downsampling: " << pt.wt_time_downsampling << "}\n";
    }
    os << "beams_per_gpu: " << beams_per_gpu << "\n";
    os << "beams_per_batch: " << beams_per_batch << "\n";
    os << "num_active_batches: " << num_active_batches << "\n";
    os << "future_write_max_samples: " << future_write_max_samples << "\n";
    if (max_gpu_clag < 10000)
        os << "max_gpu_clag: " << max_gpu_clag << "\n";
    return os.str();
}

}  // namespace pirate_amd
