// Analytic white-noise variances of the peak finder's outputs (the reference's
// PfAvarApproximation: include/pirate/PfVariance.hpp:84-104, src_lib/PfVariance.cpp:238-366,
// pirate_frb/slow_avar/PfVariance.py:465-503), used to turn the peak finder's output into an
// SNR (GpuDedisperser::fill_analytic_weights, src_lib/GpuDedisperser.cpp:1162-1238).
//
// The reference evaluates the quantity with sparse tiles propagated through the tree
// (src_lib/SparseTile.cpp). Here it is evaluated in closed form:
//
//   Input channel i carries white noise of variance s_i. Tree gridding spreads it over tree
//   channels n with overlap weights w[n,i]. After k passes of the tree recursion
//   (out[F,2d+b,t] = (in[2F,d,t-d-b] + in[2F+1,d,t]) / sqrt(2)), tree channel n sits in
//   coarse band F = n >> k at local index j = n & (2^k-1), and reaches row d (natural DM
//   order, 0 <= d < 2^k) of that band as one sample of amplitude 2^(-k/2), delayed by
//   tau_k(j, d). Peak-finding profile p filters a row with kernel h_p, so for white noise
//
//     V(F, d, p) = sum_i s_i 2^-k sum_{a,b in F} w[a,i] w[b,i] A_p(|tau(j_a,d) - tau(j_b,d)|),
//
//   with A_p(lag) = sum_t h_p[t] h_p[t+lag] (zero for lag >= len(h_p)).
//
// The approximation (the reference's): a tree of total rank r whose weights are downsampled
// 2^L-fold in DM stops the recursion at level k = r - L of the top-level tree (k = r - L + 1
// for primary trees p > 0, which take the upper half of the level-k rows, still at full time
// resolution), treats the first 2^L coarse bands as independent, and sums their variances
// with weight 2^-(L-R) into 2^R pf-rank bands (R = the tree's pf_rank); subband n's variance
// is the mean over the pf-rank bands [n_to_flo[n], n_to_fhi[n]).

#ifndef PIRATE_AMD_PF_VARIANCE_HPP
#define PIRATE_AMD_PF_VARIANCE_HPP

#include "plan.hpp"

#include <vector>

namespace pirate_amd {

// Autocorrelations A_p(lag) of the first P peak-finding kernels (p = 0: one sample; then for
// each level l with w = 2^l: ones(2w); [1/2]*w + [1]*w + [1/2]*w; [1/2]*w + [1]*2w + [1/2]*w),
// as a (P, tlen) row-major table, tlen = length of the longest of the P kernels.
// Reference: PfVarianceConvolver::PfVarianceConvolver (src_lib/PfVariance.cpp:23-75).
std::vector<double> pf_kernel_autocorrelations(long P, long &tlen);

// Delay of local channel j in natural-order row d after k tree passes (see above).
long tree_partial_delay(long j, long d, int k);

struct PfAvarApproximation {
    long nfreq = 0, ntrees = 0;

    // Per tree: (N, ndm_wt, nprofiles) float64, the shape fill_host_weights() takes.
    std::vector<std::vector<double>> tree_variance;
    // Per tree: (2^pf_rank, ndm_wt, nprofiles) float64, the frequency-summed pf-rank bands.
    std::vector<std::vector<double>> per_tf;

    // freq_variances: length plan.nfreq, positive and finite. only_tree >= 0 computes that
    // tree alone (the other entries stay empty).
    PfAvarApproximation(const DedispersionPlan &plan, const double *freq_variances, long nfreq,
                        long only_tree = -1);
};

}  // namespace pirate_amd

#endif
