#include "plan.hpp"
#include "common.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <iomanip>

namespace pirate_amd {

static std::string fmt_g(double x)
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


DedispersionPlan::DedispersionPlan(const DedispersionConfig &config_, bool gpu_runnable_) :
    config(config_), gpu_runnable(gpu_runnable_)
{
    config.validate();

    dtype = config.dtype;
    nfreq = config.get_total_nfreq();
    nt_in = config.time_samples_per_chunk;
    num_primary_trees = config.num_primary_trees();
    beams_per_gpu = config.beams_per_gpu;
    beams_per_batch = config.beams_per_batch;
    num_active_batches = config.num_active_batches;
    nbits = dtype_nbits(dtype);
    nelts_per_segment = config.get_nelts_per_segment();
    nbytes_per_segment = 128;

    // Reference: src_lib/DedispersionPlan.cpp:59-156.
    for (long ipri = 0; ipri < num_primary_trees; ipri++) {
        long primary_tree_rank = ipri ? (config.toplevel_tree_rank - 1) : config.toplevel_tree_rank;
        long st1_dd = primary_tree_rank / 2;
        long st1_amb = primary_tree_rank - st1_dd;
        stage1_dd_rank.push_back(st1_dd);
        stage1_amb_rank.push_back(st1_amb);

        long num_et = config.primary_trees[ipri].num_early_triggers;

        // Earliest trigger first, then the main (et_level=0) tree.
        for (long et = num_et; et >= 0; et--) {
            DedispersionTree t;
            t.primary_tree_index = int(ipri);
            t.early_trigger_level = int(et);
            t.amb_rank = int(st1_dd);            // amb <-> dd swap between stages
            t.dd_rank = int(st1_amb - et);
            t.nt_ds = int(xdiv(nt_in, pow2(ipri)));
            PDD_CHECK(t.dd_rank >= 1, "DedispersionPlan: stage-2 dd_rank must be >= 1");

            long tot_rank = t.total_rank();
            long pf_rank = (t.dd_rank + 1) / 2;

            long dmax = pow2(config.toplevel_tree_rank - et);
            double fmin = config.delay_to_frequency(double(dmax));
            double fmax = config.zone_freq_edges.back();

            std::vector<long> sc = FrequencySubbands::restrict_subband_counts(config.frequency_subband_counts, et, pf_rank);
            t.frequency_subbands = FrequencySubbands(sc, fmin, fmax);

            t.pf = config.primary_trees[ipri];
            if (t.pf.dm_downsampling == 0)
                t.pf.dm_downsampling = pow2(pf_rank);
            if (t.pf.time_downsampling == 0)
                t.pf.time_downsampling = t.pf.dm_downsampling;

            PDD_CHECK(t.pf.dm_downsampling <= t.pf.wt_dm_downsampling, "DedispersionPlan: dm_downsampling must be <= wt_dm_downsampling");
            PDD_CHECK(t.pf.wt_dm_downsampling <= pow2(tot_rank), "DedispersionPlan: wt_dm_downsampling must be <= 2^tree_rank");
            PDD_CHECK(t.pf.time_downsampling <= t.pf.wt_time_downsampling, "DedispersionPlan: time_downsampling must be <= wt_time_downsampling");
            PDD_CHECK(t.pf.wt_time_downsampling <= t.nt_ds, "DedispersionPlan: wt_time_downsampling must be <= nt_ds");

            // The subband outputs of the stage-2 tree have 2^(tot_rank - pf_rank) DM rows,
            // which the peak finder consumes 1:1 (reference: ReferenceTree::dedisperse()
            // output shape vs PeakFindingKernelParams::ndm_out).
            PDD_CHECK(t.pf.dm_downsampling == pow2(pf_rank),
                      "DedispersionPlan: tree (p=" << ipri << ", et=" << et << ") requires dm_downsampling == 2^pf_rank = " << pow2(pf_rank));

            t.nprofiles = 1 + 3 * integer_log2(t.pf.max_width);
            t.ndm_out = xdiv(pow2(tot_rank), t.pf.dm_downsampling);
            t.ndm_wt = xdiv(pow2(tot_rank), t.pf.wt_dm_downsampling);
            t.nt_out = xdiv(t.nt_ds, t.pf.time_downsampling);
            t.nt_wt = xdiv(t.nt_ds, t.pf.wt_time_downsampling);

            // Dcore: argmax time quantum of the peak finder. The GPU peak finder of this
            // build uses Dcore = min(time_downsampling, 8) (the reference's default rule for
            // generated kernels, makefile_helper.py:316-321); non-GPU plans use the
            // reference's fallback Dcore = time_downsampling.
            t.Dcore = gpu_runnable ? std::min(t.pf.time_downsampling, 8L) : t.pf.time_downsampling;

            double dm0 = config.dm_per_unit_delay() * double(pow2(config.toplevel_tree_rank));
            t.dm_min = dm0 * ((ipri > 0) ? double(pow2(ipri - 1)) : 0.0);
            t.dm_max = dm0 * double(pow2(ipri));
            t.trigger_frequency = fmin;

            // Peak-finder shape requirements (reference: PeakFindingKernelParams::validate()).
            PDD_CHECK(t.nt_out % t.nt_wt == 0, "DedispersionPlan: nt_out must be a multiple of nt_wt");
            PDD_CHECK((t.nt_ds / t.nt_out) % t.Dcore == 0, "DedispersionPlan: Dcore must divide nt_in/nt_out");

            trees.push_back(t);
        }
    }
    ntrees = long(trees.size());
}


std::string DedispersionPlan::to_yaml_string(bool verbose) const
{
    std::ostringstream os;
    os << "dtype: " << dtype_str(dtype) << "\n";
    os << "nfreq: " << nfreq << "\n";
    os << "nt_in: " << nt_in << "\n";
    os << "toplevel_tree_rank: " << config.toplevel_tree_rank << "\n";
    os << "num_primary_trees: " << num_primary_trees << "\n";
    os << "beams_per_gpu: " << beams_per_gpu << "\n";
    os << "beams_per_batch: " << beams_per_batch << "\n";
    os << "num_active_batches: " << num_active_batches << "\n";
    os << "stage1_dd_rank: " << vec_str(stage1_dd_rank) << "\n";
    os << "stage1_amb_rank: " << vec_str(stage1_amb_rank) << "\n";
    os << "ntrees: " << ntrees << "\n";
    os << "trees:\n";
    for (long i = 0; i < ntrees; i++) {
        const DedispersionTree &t = trees[i];
        os << "  - tree_index: " << i << "\n";
        os << "    ndm_out: " << t.ndm_out << "\n";
        os << "    nt_out: " << t.nt_out << "\n";
        os << "    dm_min: " << fmt_g(t.dm_min) << "\n";
        os << "    dm_max: " << fmt_g(t.dm_max) << "\n";
        os << "    trigger_frequency: " << fmt_g(t.trigger_frequency) << "\n";
        os << "    primary_tree_index: " << t.primary_tree_index << "\n";
        os << "    early_trigger_level: " << t.early_trigger_level << "\n";
        os << "    amb_rank: " << t.amb_rank << "\n";
        os << "    dd_rank: " << t.dd_rank << "\n";
        os << "    nt_ds: " << t.nt_ds << "\n";
        os << "    max_width: " << t.pf.max_width << "\n";
        os << "    dm_downsampling: " << t.pf.dm_downsampling << "\n";
        os << "    time_downsampling: " << t.pf.time_downsampling << "\n";
        os << "    Dcore: " << t.Dcore << "\n";
        os << "    nprofiles: " << t.nprofiles << "\n";
        os << "    wt_dm_downsampling: " << t.pf.wt_dm_downsampling << "\n";
        os << "    wt_time_downsampling: " << t.pf.wt_time_downsampling << "\n";
        os << "    ndm_wt: " << t.ndm_wt << "\n";
        os << "    nt_wt: " << t.nt_wt << "\n";
        if (verbose) {
            os << "    # " << t.frequency_subbands.N << " frequency subband(s), M=" << t.frequency_subbands.M << " multiplets\n";
        }
        os << "    frequency_subband_counts: " << vec_str(t.frequency_subbands.subband_counts) << "\n";
    }
    return os.str();
}


// Reference: src_lib/DedispersionPlan.cpp:481-547.
void DedispersionPlan::decode_argmax(unsigned int token, long itree, long idm_coarse, long itime_coarse,
                                     long &fmin, long &fmax, long &tlo, long &thi, long &p) const
{
    PDD_CHECK(itree >= 0 && itree < ntrees, "decode_argmax: itree out of range");
    const DedispersionTree &tr = trees[itree];
    const FrequencySubbands &fs = tr.frequency_subbands;
    PDD_CHECK(idm_coarse >= 0 && idm_coarse < tr.ndm_out, "decode_argmax: idm_coarse out of range");
    PDD_CHECK(itime_coarse >= 0 && itime_coarse < tr.nt_out, "decode_argmax: itime_coarse out of range");

    long Dout = xdiv(tr.nt_ds, tr.nt_out);
    long Dcore = tr.Dcore;

    long m = (token >> 16) & 0xffff;
    p = (token >> 8) & 0xff;
    long t = token & 0xff;
    PDD_CHECK(m < fs.M, "decode_argmax: bad token (m out of range)");
    PDD_CHECK(p < tr.nprofiles, "decode_argmax: bad token (p out of range)");
    PDD_CHECK(t < Dout, "decode_argmax: bad token (t out of range)");

    long lpf = p ? ((p - 1) / 3) : 0;
    long dt = std::min(Dcore, pow2(lpf));
    PDD_CHECK(t % dt == 0, "decode_argmax: bad token (t not a multiple of dt)");

    long n = fs.m_to_n[m];
    long dfine = fs.m_to_d[m];
    long flo = fs.n_to_flo[n], fhi = fs.n_to_fhi[n];
    long lsb = integer_log2(fhi - flo);

    long ipri = tr.primary_tree_index;
    long rr = tr.total_rank() + ((ipri > 0) ? 1 : 0);
    PDD_CHECK(rr == config.toplevel_tree_rank - tr.early_trigger_level, "decode_argmax: internal rank mismatch");

    long G = pow2(rr - fs.pf_rank);
    fmin = flo * G;
    fmax = fhi * G - 1;

    long dhi = idm_coarse + ((ipri > 0) ? tr.ndm_out : 0);
    long Tpf = itime_coarse * Dout + t + dt - 1;
    long thi_ds = Tpf - (pow2(fs.pf_rank) - fhi) * dhi;
    long tlo_ds = thi_ds - (dhi * pow2(lsb) + dfine);

    thi = (thi_ds + 1) << ipri;
    tlo = (tlo_ds + 1) << ipri;
}


// Reference: src_lib/DedispersionPlan.cpp:551-615.
void DedispersionPlan::decode_argmax2(long itree, long fmin, long fmax, long tlo, long thi, long p,
                                      double &freq_lo_MHz, double &freq_hi_MHz, double &dm,
                                      double &timestamp_samp, double &width_samp) const
{
    PDD_CHECK(itree >= 0 && itree < ntrees, "decode_argmax2: itree out of range");
    long ntree = pow2(config.toplevel_tree_rank);
    long ipri = trees[itree].primary_tree_index;
    PDD_CHECK(fmin >= 0 && fmin < fmax && fmax < ntree, "decode_argmax2: bad (fmin, fmax)");
    PDD_CHECK(tlo <= thi, "decode_argmax2: tlo > thi");
    PDD_CHECK(p >= 0 && p < trees[itree].nprofiles, "decode_argmax2: p out of range");

    double dslope = double(thi - tlo) / double(fmax - fmin);
    long pdiv = p / 3, pmod = p - 3 * pdiv;
    double pf_width, pf_shift;
    if (p == 0) {
        pf_width = 1.0 * (1 << ipri);
        pf_shift = 0.5 * (1 << ipri);
    }
    else if (pmod == 1) {
        pf_width = 1.0 * (1 << (ipri + pdiv + 1));
        pf_shift = 0.5 * (1 << (ipri + pdiv + 1));
    }
    else if (pmod == 2) {
        pf_width = 2.0 * (1 << (ipri + pdiv));
        pf_shift = 1.5 * (1 << (ipri + pdiv));
    }
    else {
        pf_width = 3.0 * (1 << (ipri + pdiv - 1));
        pf_shift = 2.0 * (1 << (ipri + pdiv - 1));
    }

    freq_lo_MHz = config.delay_to_frequency(double(fmax + 1));
    freq_hi_MHz = config.delay_to_frequency(double(fmin));
    dm = dslope * double(ntree) * config.dm_per_unit_delay();
    timestamp_samp = double(thi) + dslope * (double(ntree) - 0.5 - double(fmax)) - pf_shift;
    width_samp = pf_width;
}


// Reference: src_lib/DedispersionPlan.cpp:618-661.
std::vector<long> DedispersionPlan::compute_steady_state_it0(long itree) const
{
    PDD_CHECK(itree >= 0 && itree < ntrees, "compute_steady_state_it0: itree out of range");
    const DedispersionTree &tr = trees[itree];
    long p = tr.primary_tree_index, e = tr.early_trigger_level;
    long T_ds = tr.pf.time_downsampling, D_ds = tr.pf.dm_downsampling, Wmax = tr.pf.max_width;
    long d_lo = (p > 0) ? pow2(config.toplevel_tree_rank + p - 1) : 0;
    long d0 = xdiv(d_lo, pow2(e + p));
    std::vector<long> ret(tr.ndm_out);
    for (long idm = 0; idm < tr.ndm_out; idm++) {
        long dmax = d0 + (idm + 1) * D_ds - 1;
        ret[idm] = (dmax + 4 * Wmax + T_ds - 1) / T_ds;
    }
    return ret;
}

}  // namespace pirate_amd
