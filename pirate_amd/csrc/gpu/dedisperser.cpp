#include "dedisperser.hpp"
#include "scorer.hpp"
#include "../host/common.hpp"
#include "../host/pf_variance.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace pirate_amd {

#define HIP_CHECK(expr)                                                                   \
    do {                                                                                  \
        hipError_t _e = (expr);                                                           \
        if (_e != hipSuccess)                                                             \
            throw pdd_error(std::string("HIP error: ") + hipGetErrorString(_e) + " in " + #expr); \
    } while (0)

static constexpr int kGpuMaxDD = 8;          // register-tree kernels (stage 1, final-pass stage 2)
static constexpr int kGpuMaxDDGeneral = 7;   // LDS-tile stage 2 (subband searches)
static constexpr long kLdsBytesPerCU = 160 * 1024;


GpuDedisperser::GpuDedisperser(std::shared_ptr<DedispersionPlan> plan_, const GpuDedisperserParams &params_) :
    plan(plan_), params(params_)
{
    PDD_CHECK(plan, "GpuDedisperser: null plan");
    PDD_CHECK(plan->gpu_runnable, "GpuDedisperser: plan was built with gpu_runnable=false");
    dtype = plan->dtype;
    elt_bytes = dtype_nbits(dtype) / 8;
    nfreq = plan->nfreq;
    nt_in = plan->nt_in;
    total_beams = plan->beams_per_gpu;
    beams_per_batch = plan->beams_per_batch;
    nbatches = total_beams / beams_per_batch;
    nstreams = plan->num_active_batches;
    ntrees = plan->ntrees;

    // Defaults and lower bounds as the reference's (src_lib/GpuDedisperser.cpp:116-123).
    nbatches_out = params.nbatches_out ? params.nbatches_out : nstreams;
    nbatches_wt = params.nbatches_wt ? params.nbatches_wt : nstreams;
    num_consumers = params.num_consumers;
    PDD_CHECK(nbatches_out >= nstreams, "GpuDedisperser: nbatches_out (" << nbatches_out
              << ") must be >= num_active_batches (" << nstreams << ")");
    PDD_CHECK(nbatches_wt >= nstreams, "GpuDedisperser: nbatches_wt (" << nbatches_wt
              << ") must be >= num_active_batches (" << nstreams << ")");
    PDD_CHECK(num_consumers >= 1, "GpuDedisperser: num_consumers must be >= 1");
    PDD_CHECK(params.initial_chunk >= 0, "GpuDedisperser: initial_chunk must be >= 0");

    curr_output_acquire.assign(num_consumers, 0);
    curr_output_release.assign(num_consumers, 0);

    build_tables();
    validate_gpu_limits();
}


GpuDedisperser::~GpuDedisperser()
{
    // Copies of spilled ring rows may still be in flight on the copy streams.
    for (hipStream_t x : {d2h_stream, h2d_stream})
        if (x)
            (void) hipStreamSynchronize(x);
    for (hipStream_t x : compute_streams)
        (void) hipStreamSynchronize(x);
    for (auto &bs : branch_streams)
        for (size_t p = 1; p < bs.size(); p++)
            if (bs[p])
                (void) hipStreamSynchronize(bs[p]);
    auto freeit = [](DeviceArray &a) { if (a.ptr) (void) hipFree(a.ptr); a.ptr = nullptr; };
    freeit(input_slots);
    for (auto &a : grid_bufs)
        freeit(a);
    for (auto &v : lds_outs)
        for (auto &a : v)
            freeit(a);
    freeit(lds_state);
    freeit(grid_rows);
    freeit(grid_w);
    for (void *h : host_ring)
        if (h)
            (void) hipHostFree(h);
    for (auto *v : {&rings, &ring_row_off, &ring_row_cap, &ring1_row_off, &ring1_row_cap, &spill_rows,
                    &stage2_lags, &sb, &d_main_emit, &d_stag,
                    &d_lvl0, &d_fl, &d_m_to_n, &d_stag_begin, &d_stag_goff, &d_out_stream, &s1_state, &s2_state, &weights, &out_var})
        for (auto &a : *v)
            freeit(a);
    freeit(scorer_params);
    freeit(scorer_origins);
    freeit(scorer_out);
    freeit(out_region);     // out_max / out_argmax are views into it
    freeit(split_x);
    for (auto &t : timed_pending) {
        timing_event_pool.push_back(t.start);
        timing_event_pool.push_back(t.stop);
    }
    for (hipStream_t x : {d2h_stream, h2d_stream})
        if (x)
            (void) hipStreamDestroy(x);
    for (auto &bs : branch_streams)
        for (size_t p = 1; p < bs.size(); p++)
            if (bs[p])
                (void) hipStreamDestroy(bs[p]);
    if (ev_split_free)
        (void) hipEventDestroy(ev_split_free);
    for (auto &v : ev_join)
        for (hipEvent_t e : v)
            (void) hipEventDestroy(e);
    for (auto *v : {&ev_input_free, &ev_input_ready, &ev_output_ready, &ev_output_released, &timing_event_pool,
                    &ev_spill_s1, &ev_spill_d2h, &ev_spill_h2d, &ev_spill_scatter, &ev_fork, &ev_batch_done})
        for (hipEvent_t e : *v)
            (void) hipEventDestroy(e);
    for (hipStream_t x : compute_streams)
        (void) hipStreamDestroy(x);
}


// Subband-stream tables of each stage-2 tree. The loops restate the subband extraction
// of ReferenceTree::dedisperse_2d (src_lib/ReferenceTree.cpp:142-287) for one tree
// instance: level-0 copies, main-pair outputs (even subband index), staggered extra
// butterflies (odd subband index), multiplet numbering, and the final lags
// (ReferenceTree.cpp:42-54).
void GpuDedisperser::build_tables()
{
    tables.resize(ntrees);
    for (long it = 0; it < ntrees; it++) {
        const DedispersionTree &t = plan->trees[it];
        const FrequencySubbands &fs = t.frequency_subbands;
        TreeTables &T = tables[it];
        const int dd = t.dd_rank, pf_rank = int(fs.pf_rank);
        const int NP = 1 << (dd - 1), M = int(fs.M);
        const int pf_ndm = 1 << (dd - pf_rank);
        T.M = M;
        T.N = int(fs.N);
        T.main_emit.assign(dd * NP, -1);
        T.fl.assign(pf_ndm * M, 0);
        long mcurr = 0;
        int hoff = 0;
        for (int r = 0; r <= dd; r++) {
            const int ndm_in = 1 << r;
            if (r == dd - pf_rank) {
                const int pf_ns = int(fs.subband_counts[0]);
                for (int dm_in = 0; dm_in < ndm_in; dm_in++)
                    for (int pfs = 0; pfs < pf_ns; pfs++) {
                        T.lvl0.push_back(pfs * ndm_in + int(bit_reverse(dm_in, r)));
                        T.lvl0.push_back(dm_in * M + pfs);
                    }
                mcurr += pf_ns;
            }
            if (r >= dd)
                break;
            const int pf_level = r + pf_rank - dd + 1;
            if (pf_level <= 0)
                continue;
            const int pf_ns = int(fs.subband_counts[pf_level]);
            const int pf_nd2 = 1 << (pf_level - 1);
            for (int pf_dm = 0; pf_dm < pf_ndm; pf_dm++)
                for (int d2 = 0; d2 < pf_nd2; d2++) {
                    const int dm_in = pf_dm * pf_nd2 + d2;
                    const int dmb = int(bit_reverse(dm_in, r));
                    for (int pfs = 0; pfs < pf_ns; pfs++) {
                        const int m = int(mcurr) + pfs * 2 * pf_nd2 + 2 * d2;
                        const int s0 = pf_dm * M + m;
                        if ((pfs & 1) == 0) {
                            const int p = (pfs / 2) * ndm_in + dmb;
                            T.main_emit[r * NP + p] = s0;
                        }
                        else {
                            int e[6] = {r, pfs * ndm_in + dmb, (pfs + 1) * ndm_in + dmb, dm_in, s0, hoff};
                            T.stag.insert(T.stag.end(), e, e + 6);
                            hoff += ndm_in;
                        }
                    }
                }
            mcurr += long(pf_ns) * 2 * pf_nd2;
        }
        PDD_CHECK(mcurr == M, "GpuDedisperser: internal error: multiplet count mismatch");
        T.nstag = int(T.stag.size() / 6);
        // Entries were generated pass by pass: record each pass's range.
        T.stag_begin.assign(dd + 1, 0);
        for (int r = 0, k = 0; r <= dd; r++) {
            while (k < T.nstag && T.stag[6 * k] < r)
                k++;
            T.stag_begin[r] = k;
        }
        T.nlvl0 = int(T.lvl0.size() / 2);
        T.stag_hist_nelts = hoff;
        if (dd == 8) {
            // Rank-8 split: each group of rows with equal low 4 bits keeps only the histories
            // of its own staggered pairs (group = row0 & 15), packed.
            std::vector<int> gtot(16, 0);
            T.stag_goff.assign(T.nstag, 0);
            for (int k = 0; k < T.nstag; k++) {
                const int g = T.stag[6 * k + 1] & 15;
                T.stag_goff[k] = gtot[g];
                gtot[g] += 1 << T.stag[6 * k];
            }
            T.stag_hist_g = *std::max_element(gtot.begin(), gtot.end());
        }

        // Register-tree path: every stream is a final-pass output (main pairs of the last
        // pass, or level-0 copies after the last pass), each row feeding at most one stream.
        T.out_stream.assign(pow2(dd), -1);
        T.fast = (T.nstag == 0);
        for (int r = 0; r < dd && T.fast; r++)
            for (int p = 0; p < NP; p++) {
                const int s0 = T.main_emit[r * NP + p];
                if (s0 < 0)
                    continue;
                if (r != dd - 1 || T.out_stream[p] >= 0 || T.out_stream[p + NP] >= 0)
                    T.fast = false;
                else {
                    T.out_stream[p] = s0;
                    T.out_stream[p + NP] = s0 + 1;
                }
            }
        if (T.nlvl0 > 0 && dd - pf_rank != dd)
            T.fast = false;
        for (int k = 0; k < T.nlvl0 && T.fast; k++) {
            const int row = T.lvl0[2 * k], s0 = T.lvl0[2 * k + 1];
            if (T.out_stream[row] >= 0)
                T.fast = false;
            else
                T.out_stream[row] = s0;
        }
        T.split = !T.fast && dd == 8;

        for (int pf_dm = 0; pf_dm < pf_ndm; pf_dm++)
            for (int m = 0; m < M; m++) {
                int v = int((pow2(pf_rank) - fs.m_to_fhi(m)) * pf_dm);
                T.fl[pf_dm * M + m] = v;
                T.flmax = std::max(T.flmax, v);
            }
        for (int m = 0; m < M; m++)
            T.m_to_n.push_back(int(fs.m_to_n[m]));

        T.tpad = int(std::max(2 * t.pf.max_width, 4L));
        T.nlev = std::max(integer_log2(t.pf.max_width), 1);
        // Rows start on 128-byte lines and the buffer is offset by 'lead' elements so that
        // chunk sample 0 (position tpad) of every row starts a line: the stage-2 tile stores
        // and the peak finder's loads then cover whole lines.
        const int line_elts = int(128 / elt_bytes);
        T.lead = int(align_up(T.tpad, line_elts)) - T.tpad;
        T.rowlen = int(align_up(T.tpad + t.nt_ds + T.flmax, line_elts));
    }
}


void GpuDedisperser::validate_gpu_limits()
{
    const long R = plan->config.toplevel_tree_rank;
    const long npri = plan->num_primary_trees;
    PDD_CHECK(npri <= 7, "GpuDedisperser: at most 7 primary trees");
    PDD_CHECK(nt_in % 4 == 0, "GpuDedisperser: nt_in must be a multiple of 4");
    for (long p = 0; p < npri; p++) {
        long Tp = nt_in >> p;
        PDD_CHECK(Tp % 32 == 0, "GpuDedisperser: pirate_amd needs nt_in/2^p to be a multiple of 32 for every primary tree p (nt_in="
                  << nt_in << ", p=" << p << ")");
        PDD_CHECK(plan->stage1_dd_rank[p] <= kGpuMaxDD && plan->stage1_dd_rank[p] >= 1,
                  "GpuDedisperser: stage-1 dd_rank " << plan->stage1_dd_rank[p] << " not supported by pirate_amd (max " << kGpuMaxDD << ")");
        if (p > 0) {
            long s = plan->stage1_dd_rank[1] + 1;
            PDD_CHECK(pow2(s - 1) <= Tp, "GpuDedisperser: lagged-downsampling lag exceeds the chunk at p=" << p);
        }
    }
    (void) R;
    for (long it = 0; it < ntrees; it++) {
        const DedispersionTree &t = plan->trees[it];
        const TreeTables &T = tables[it];
        PDD_CHECK(t.dd_rank <= kGpuMaxDD, "GpuDedisperser: stage-2 dd_rank " << t.dd_rank << " not supported by pirate_amd (max " << kGpuMaxDD << ")");
        if (!T.fast && t.dd_rank <= kGpuMaxDDGeneral)
            PDD_CHECK(stage2_lds_bytes(t.dd_rank, T.stag_hist_nelts) <= kLdsBytesPerCU,
                      "GpuDedisperser: tree " << it << " needs too much LDS (subband history)");
        if (!T.fast && t.dd_rank == 8) {
            PDD_CHECK(t.frequency_subbands.pf_rank <= 4, "GpuDedisperser: tree " << it << ": rank-8 split needs pf_rank <= 4");
            PDD_CHECK(split_b_lds_bytes(T.stag_hist_g) <= kLdsBytesPerCU,
                      "GpuDedisperser: tree " << it << " needs too much LDS (subband history)");
        }
        PDD_CHECK(t.pf.time_downsampling <= 512, "GpuDedisperser: time_downsampling too large");
        PDD_CHECK(pf_config_supported(T.M, T.tpad, T.nlev, int(t.Dcore), int(t.pf.time_downsampling), t.nt_ds),
                  "GpuDedisperser: tree " << it << ": peak-finder shape not supported (M=" << T.M << ")");
    }
}


DeviceArray GpuDedisperser::dalloc(size_t nbytes)
{
    DeviceArray a;
    a.nbytes = nbytes;
    if (nbytes == 0)
        return a;
    HIP_CHECK(hipMalloc(&a.ptr, nbytes));
    HIP_CHECK(hipMemset(a.ptr, 0, nbytes));
    gpu_bytes_allocated += nbytes;
    return a;
}


void GpuDedisperser::allocate()
{
    std::lock_guard<std::mutex> lk(mutex);
    PDD_CHECK(!is_allocated, "GpuDedisperser::allocate() called twice");
    HIP_CHECK(hipSetDevice(params.device));
    compute_streams.assign(nstreams, nullptr);
    for (auto &x : compute_streams)
        HIP_CHECK(hipStreamCreateWithFlags(&x, hipStreamNonBlocking));
    compute_stream = compute_streams[0];

    const long B = beams_per_batch, R = plan->config.toplevel_tree_rank, npri = plan->num_primary_trees;
    const long N = pow2(R);

    auto up_int = [&](const std::vector<int> &v) {
        DeviceArray a = dalloc(std::max<size_t>(v.size(), 1) * sizeof(int));
        if (!v.empty())
            HIP_CHECK(hipMemcpy(a.ptr, v.data(), v.size() * sizeof(int), hipMemcpyHostToDevice));
        return a;
    };
    auto up_long = [&](const std::vector<long> &v) {
        DeviceArray a = dalloc(std::max<size_t>(v.size(), 1) * sizeof(long));
        if (!v.empty())
            HIP_CHECK(hipMemcpy(a.ptr, v.data(), v.size() * sizeof(long), hipMemcpyHostToDevice));
        return a;
    };

    input_slots = dalloc(size_t(nstreams) * B * nfreq * nt_in * elt_bytes);
    grid_bufs.resize(nstreams);
    for (auto &a : grid_bufs)
        a = dalloc(size_t(B) * N * nt_in * elt_bytes);

    // Gridding CSR: per tree channel n, ascending input channels f in [if0, if1) with
    // overlap weights (reference: src_lib/TreeGriddingKernel.cu:77-100).
    {
        std::vector<double> cmap = plan->config.make_channel_map();
        std::vector<int> rows(3 * N);
        std::vector<float> w;
        for (long n = 0; n < N; n++) {
            double f0 = cmap[n + 1], f1 = cmap[n];
            long if0 = std::max(long(f0), 0L);
            long if1 = std::min(long(f1) + 1, nfreq);
            rows[n] = int(if0);
            rows[N + n] = int(std::max(if1 - if0, 0L));
            rows[2 * N + n] = int(w.size());
            for (long f = if0; f < if1; f++) {
                double flo = std::max(f0, double(f)), fhi = std::min(f1, double(f) + 1);
                w.push_back(float(std::max(fhi - flo, 0.0)));
            }
        }
        grid_rows = up_int(rows);
        grid_w = dalloc(std::max<size_t>(w.size(), 1) * sizeof(float));
        HIP_CHECK(hipMemcpy(grid_w.ptr, w.data(), w.size() * sizeof(float), hipMemcpyHostToDevice));
    }

    lds_outs.assign(nstreams, std::vector<DeviceArray>(npri));
    for (auto &v : lds_outs)
        for (long p = 1; p < npri; p++)
            v[p] = dalloc(size_t(B) * (N / 2) * (nt_in >> p) * elt_bytes);
    if (npri > 1) {
        long s = plan->stage1_dd_rank[1] + 1;
        lds_state_bstride = lds_state_nelts(int(R), int(npri), int(s));
        lds_state = dalloc(size_t(total_beams) * lds_state_bstride * sizeof(float));
    }

    // Lag ring: per primary tree p and stage-1 output row (a1, j), a circular row whose
    // capacity covers the largest stage-2 lag of that row plus one chunk; rows whose
    // smallest lag spans more than max_gpu_clag chunks keep k whole chunks of it in host
    // memory (see SpillGroup in the header).
    rings.resize(npri);
    ring_row_off.resize(npri);
    ring_row_cap.resize(npri);
    ring1_row_off.resize(npri);
    ring1_row_cap.resize(npri);
    ring_bstride.resize(npri);
    spill_groups.assign(npri, {});
    spill_k.assign(npri, {});
    nspill.assign(npri, 0);
    outbox_off.assign(npri, 0);
    inbox_off.assign(npri, 0);
    host_bstride.assign(npri, 0);
    spill_rows.resize(npri);
    host_ring.assign(npri, nullptr);
    s1_state.resize(npri);
    s1_state_bstride.resize(npri);
    const long max_gpu_clag = plan->config.max_gpu_clag;
    for (long p = 0; p < npri; p++) {
        const int dd1 = int(plan->stage1_dd_rank[p]), amb1 = int(plan->stage1_amb_rank[p]);
        const long Tp = nt_in >> p;
        const long nrows = pow2(dd1 + amb1);
        std::vector<long> off(nrows), cap(nrows), kr(nrows, 0);
        long tot = 0;
        for (long a1 = 0; a1 < pow2(amb1); a1++)
            for (long j = 0; j < pow2(dd1); j++) {
                long maxlag = 0, minlag = -1;
                for (long it = 0; it < ntrees; it++) {
                    const DedispersionTree &t = plan->trees[it];
                    if (t.primary_tree_index != p || a1 >= pow2(t.dd_rank))
                        continue;
                    const long lag = rb_lag(a1, j, t.amb_rank, t.dd_rank, p > 0);
                    maxlag = std::max(maxlag, lag);
                    minlag = (minlag < 0) ? lag : std::min(minlag, lag);
                }
                long r = a1 * pow2(dd1) + j;
                if (minlag >= 0 && max_gpu_clag < (1L << 40))
                    kr[r] = std::max(0L, minlag / Tp - max_gpu_clag);
                cap[r] = align_up(maxlag - kr[r] * Tp + Tp, 64);
                off[r] = tot;
                tot += cap[r];
            }
        // Spilled rows in outbox order: grouped by k, ascending.
        std::vector<long> order;
        for (long r = 0; r < nrows; r++)
            if (kr[r] > 0)
                order.push_back(r);
        std::stable_sort(order.begin(), order.end(), [&](long x, long y) { return kr[x] < kr[y]; });
        nspill[p] = long(order.size());
        outbox_off[p] = align_up(tot, 64);
        inbox_off[p] = outbox_off[p] + nspill[p] * Tp;
        std::vector<long> off1(off), cap1(cap);
        std::vector<int> srows(order.size());
        long hoff = 0;
        for (size_t i = 0; i < order.size(); i++) {
            const long r = order[i];
            srows[i] = int(r);
            off1[r] = outbox_off[p] + long(i) * Tp;
            cap1[r] = Tp;
            if (spill_groups[p].empty() || spill_groups[p].back().k != kr[r])
                spill_groups[p].push_back(SpillGroup{kr[r], long(i), 0, 0});
            spill_groups[p].back().nrows++;
        }
        for (SpillGroup &g : spill_groups[p]) {
            g.host_off = hoff;
            hoff += (g.k + 1) * g.nrows * Tp;
        }
        host_bstride[p] = align_up(hoff, 64);
        spill_k[p] = kr;
        ring_bstride[p] = align_up(inbox_off[p] + nspill[p] * Tp, 64);
        PDD_CHECK(ring_bstride[p] < (1L << 31), "GpuDedisperser: lag ring of primary tree " << p
                  << " exceeds 2^31 elements per beam (tree kernels use 32-bit row offsets)");
        rings[p] = dalloc(size_t(total_beams) * ring_bstride[p] * elt_bytes);
        ring_row_off[p] = up_long(off);
        ring_row_cap[p] = up_long(cap);
        ring1_row_off[p] = up_long(off1);
        ring1_row_cap[p] = up_long(cap1);
        spill_rows[p] = up_int(srows);
        if (nspill[p] > 0) {
            const size_t hb = size_t(total_beams) * host_bstride[p] * elt_bytes;
            HIP_CHECK(hipHostMalloc(&host_ring[p], hb, hipHostMallocDefault));
            std::memset(host_ring[p], 0, hb);
            host_bytes_allocated += hb;
        }
        s1_state_bstride[p] = pow2(amb1) * fast_tree_state_nelts(int(dd1));
        s1_state[p] = dalloc(size_t(total_beams) * s1_state_bstride[p] * sizeof(float));
    }

    size_t split_nelts = 0;   // float32 scratch of rank-8 split stage-2 trees
    stage2_lags.resize(ntrees);
    sb.resize(ntrees);
    d_main_emit.resize(ntrees);
    d_stag.resize(ntrees);
    d_lvl0.resize(ntrees);
    d_fl.resize(ntrees);
    d_m_to_n.resize(ntrees);
    d_stag_begin.resize(ntrees);
    d_stag_goff.resize(ntrees);
    d_out_stream.resize(ntrees);
    s2_state.resize(ntrees);
    s2_state_bstride.resize(ntrees);
    weights.resize(ntrees);
    out_max.resize(ntrees);
    out_argmax.resize(ntrees);
    out_var.resize(ntrees);
    for (long it = 0; it < ntrees; it++) {
        const DedispersionTree &t = plan->trees[it];
        const TreeTables &T = tables[it];
        std::vector<int> lags(pow2(t.dd_rank + t.amb_rank));
        const std::vector<long> &kr = spill_k[t.primary_tree_index];
        for (long fc = 0; fc < pow2(t.dd_rank); fc++)
            for (long j = 0; j < pow2(t.amb_rank); j++) {
                const long rr = fc * pow2(t.amb_rank) + j;
                lags[rr] = int(rb_lag(fc, j, t.amb_rank, t.dd_rank, t.primary_tree_index > 0) - kr[rr] * t.nt_ds);
            }
        stage2_lags[it] = up_int(lags);
        // Peak-finder input rows hold per-beam history (and final-lag "future" samples)
        // across chunks: one set of rows per beam of the GPU, not per batch.
        PDD_CHECK(t.ndm_out * T.M * T.rowlen < (1L << 31), "GpuDedisperser: peak-finder rows of tree " << it
                  << " exceed 2^31 elements per beam");
        sb[it] = dalloc((size_t(total_beams) * t.ndm_out * T.M * T.rowlen + T.lead) * elt_bytes);
        d_main_emit[it] = up_int(T.main_emit);
        d_stag[it] = up_int(T.stag);
        d_lvl0[it] = up_int(T.lvl0);
        d_fl[it] = up_int(T.fl);
        d_m_to_n[it] = up_int(T.m_to_n);
        d_stag_begin[it] = up_int(T.stag_begin);
        d_stag_goff[it] = up_int(T.stag_goff);
        d_out_stream[it] = up_int(T.out_stream);
        s2_state_bstride[it] = pow2(t.amb_rank) * (T.fast ? fast_tree_state_nelts(t.dd_rank)
                                                   : T.split ? split_state_nelts_per_tree(T.stag_hist_g)
                                                             : stage2_state_nelts_per_tree(t.dd_rank, T.stag_hist_nelts));
        if (T.split)
            split_nelts = std::max(split_nelts, size_t(B) * pow2(t.amb_rank) * 256 * t.nt_ds);
        s2_state[it] = dalloc(size_t(total_beams) * s2_state_bstride[it] * sizeof(float));
        weights[it] = dalloc(size_t(nbatches_wt) * B * t.ndm_wt * t.nt_wt * t.nprofiles * T.N * sizeof(float));
        // out_max / out_argmax: views into the output region (below)
        out_max[it].nbytes = size_t(nbatches_out) * B * t.ndm_out * t.nt_out * elt_bytes;
        out_argmax[it].nbytes = size_t(nbatches_out) * B * t.ndm_out * t.nt_out * sizeof(unsigned int);
        if (params.enable_variances)
            out_var[it] = dalloc(size_t(nbatches_out) * B * t.ndm_out * T.M * t.nprofiles * sizeof(double));
    }

    if (split_nelts > 0)
        split_x = dalloc(split_nelts * sizeof(float));

    // Output ring: every tree's out_max and out_argmax in one device allocation (256-byte
    // aligned sub-arrays, tree-major), so that one IPC handle exports all of it to a
    // consumer process (the reference shares its output ring the same way:
    // FrbServer.cpp:1583-1612, FrbGrouper.cpp:565-581).
    {
        size_t tot = 0;
        out_offsets.assign(2 * ntrees, 0);
        for (long it = 0; it < ntrees; it++) {
            out_offsets[2 * it] = tot;
            tot = size_t(align_up(long(tot + out_max[it].nbytes), 256));
            out_offsets[2 * it + 1] = tot;
            tot = size_t(align_up(long(tot + out_argmax[it].nbytes), 256));
        }
        out_region = dalloc(tot);
        for (long it = 0; it < ntrees; it++) {
            out_max[it].ptr = (char *) out_region.ptr + out_offsets[2 * it];
            out_argmax[it].ptr = (char *) out_region.ptr + out_offsets[2 * it + 1];
        }
    }

    auto mkev = [](std::vector<hipEvent_t> &v, long n) {
        v.resize(n);
        for (auto &e : v)
            HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    };
    mkev(ev_spill_s1, npri * nbatches);
    mkev(ev_spill_d2h, npri * nbatches);
    mkev(ev_spill_h2d, npri * nbatches);
    mkev(ev_spill_scatter, npri * nbatches);
    if (std::any_of(nspill.begin(), nspill.end(), [](long n) { return n > 0; })) {
        HIP_CHECK(hipStreamCreateWithFlags(&d2h_stream, hipStreamNonBlocking));
        HIP_CHECK(hipStreamCreateWithFlags(&h2d_stream, hipStreamNonBlocking));
    }
    branch_streams.assign(nstreams, std::vector<hipStream_t>(npri, nullptr));
    ev_join.resize(nstreams);
    for (long s = 0; s < nstreams; s++) {
        for (long p = 1; p < npri; p++)
            HIP_CHECK(hipStreamCreateWithFlags(&branch_streams[s][p], hipStreamNonBlocking));
        mkev(ev_join[s], npri);
    }
    mkev(ev_fork, nstreams);
    mkev(ev_batch_done, nbatches);
    HIP_CHECK(hipEventCreateWithFlags(&ev_split_free, hipEventDisableTiming));
    mkev(ev_input_free, nstreams);
    mkev(ev_input_ready, nstreams);
    mkev(ev_output_ready, nbatches_out);
    mkev(ev_output_released, nbatches_out * num_consumers);
    // Every slot starts out "free": record once so that first waits return immediately.
    for (auto &e : ev_input_free) HIP_CHECK(hipEventRecord(e, compute_stream));
    for (auto &e : ev_output_released) HIP_CHECK(hipEventRecord(e, compute_stream));
    for (auto &e : ev_batch_done) HIP_CHECK(hipEventRecord(e, compute_stream));
    HIP_CHECK(hipStreamSynchronize(compute_stream));

    is_allocated = true;
}


size_t GpuDedisperser::spill_bytes_per_batch() const
{
    size_t n = 0;
    for (size_t p = 0; p < nspill.size(); p++)
        n += size_t(beams_per_batch) * nspill[p] * (nt_in >> p) * elt_bytes;
    return n;
}


void GpuDedisperser::set_tuning(const std::string &name, long value)
{
    std::lock_guard<std::mutex> lk(mutex);
    if (name == "pf_target_blocks") {
        PDD_CHECK(value >= 1 && value <= (1L << 20), "set_tuning(): pf_target_blocks must be in [1, 2^20]");
        tuning.pf_target_blocks = value;
    }
    else if (name == "grid_xcd_run") {
        PDD_CHECK(value >= 0 && value <= 64, "set_tuning(): grid_xcd_run must be in [0, 64]");
        tuning.grid_xcd_run = int(value);
    }
    else if (name == "fork") {
        PDD_CHECK(value >= -1 && value <= 2, "set_tuning(): fork must be -1, 0, 1 or 2");
        tuning.fork = int(value);
    }
    else
        throw pdd_error("set_tuning(): unknown launch parameter '" + name + "'");
}

long GpuDedisperser::get_tuning(const std::string &name) const
{
    if (name == "pf_target_blocks")
        return tuning.pf_target_blocks;
    if (name == "grid_xcd_run")
        return tuning.grid_xcd_run;
    if (name == "fork")
        return tuning.fork;
    throw pdd_error("get_tuning(): unknown launch parameter '" + name + "'");
}


void *GpuDedisperser::input_slot(long slot) const
{
    PDD_CHECK(is_allocated && slot >= 0 && slot < nstreams, "input_slot(): bad slot or not allocated");
    return (char *) input_slots.ptr + size_t(slot) * beams_per_batch * nfreq * nt_in * elt_bytes;
}

void *GpuDedisperser::output_slot(long itree, long oslot, int which) const
{
    PDD_CHECK(is_allocated && itree >= 0 && itree < ntrees && oslot >= 0 && oslot < nbatches_out && (which == 0 || which == 1),
              "output_slot(): bad tree, slot or array, or not allocated");
    const DedispersionTree &t = plan->trees[itree];
    const size_t n = size_t(beams_per_batch) * t.ndm_out * t.nt_out;
    return which ? (void *) ((unsigned int *) out_argmax[itree].ptr + oslot * n)
                 : (void *) ((char *) out_max[itree].ptr + oslot * n * elt_bytes);
}


void GpuDedisperser::launch_chain(const void *input, long ichunk, long ibatch, long oslot, long wslot, hipStream_t st,
                                  long istream, const long *chunk_dev)
{
    const long B = beams_per_batch, R = plan->config.toplevel_tree_rank, npri = plan->num_primary_trees;
    const long N = pow2(R);
    const int beam0 = int(ibatch * B);
    PDD_CHECK(istream >= 0 && istream < nstreams, "launch_chain(): stream index out of range");
    PDD_CHECK(!chunk_dev || std::all_of(nspill.begin(), nspill.end(), [](long n) { return n == 0; }),
              "launch_chain(): a device chunk counter (graph replay) needs a plan without host spill");
    const DeviceArray &grid_buf = grid_bufs[istream];
    const std::vector<DeviceArray> &lds_out = lds_outs[istream];
    const std::vector<hipStream_t> &bstreams = branch_streams[istream];
    hipEvent_t ev_fork_s = ev_fork[istream];
    last_stream = istream;

    {
        GridLdsArgs g;
        g.in = input;
        g.in_bstride = nfreq * nt_in;
        g.grid = grid_buf.ptr;
        g.grid_bstride = N * nt_in;
        g.B = int(B);
        g.nchan = int(N);
        g.nfreq = int(nfreq);
        g.ntime = int(nt_in);
        g.row_f0 = (const int *) grid_rows.ptr;
        g.row_cnt = g.row_f0 + N;
        g.row_woff = g.row_f0 + 2 * N;
        g.w = (const float *) grid_w.ptr;
        for (long p = 0; p < kMaxPrimaryTrees; p++) {
            g.out[p] = (p > 0 && p < npri) ? lds_out[p].ptr : nullptr;
            g.out_bstride[p] = (p > 0 && p < npri) ? (N / 2) * (nt_in >> p) : 0;
        }
        g.state = (float *) lds_state.ptr;
        g.state_bstride = lds_state_bstride;
        g.npri = int(npri);
        g.xcd_run = tuning.grid_xcd_run;
        g.s = (npri > 1) ? int(plan->stage1_dd_rank[1] + 1) : 1;
        g.beam0 = beam0;
        g.chunk = ichunk;
        g.chunk_dev = chunk_dev;
        hipEvent_t ev;
        timed_begin("grid_lds", st, ev);
        launch_grid_lds(dtype, g, st);
        timed_end("grid_lds", st, ev);
    }
    for (long p = 0; p < npri; p++)
        if (nspill[p] > 0)
            spill_h2d(p, ichunk, ibatch, beam0);
    // Primary tree p's stage 1, spill scatter, stage-2 trees and peak finders run on branch
    // stream p (0: st); st waits for every branch at the end. The branches p >= 1 start
    // after the gridding, or, when tree 0's stage 2 is the register-tree kernel, after it:
    // stage 1 and that stage 2 run near the HBM bandwidth on their own (chord16k fp32:
    // 5.57 ms/step serial, 5.66 with the branches from the gridding on), while the
    // peak finders and the subband stage-2 kernels are bound by VALU or LDS latency and
    // leave bandwidth to share (chime_sb1-shaped, fp16: 15.68 -> 15.11 ms).
    // tuning.fork == 2: every branch on 'st' in order (one kernel at a time, for per-kernel timing).
    const bool serial = (tuning.fork == 2);
    const bool late_fork = (tuning.fork < 0) ? tables[0].fast : (tuning.fork == 1);
    if (!late_fork && !serial)
        HIP_CHECK(hipEventRecord(ev_fork_s, st));
    for (long p = 0; p < npri; p++) {
        hipStream_t sp = (p && !serial) ? bstreams[p] : st;
        if (sp != st)
            HIP_CHECK(hipStreamWaitEvent(sp, ev_fork_s, 0));
        Stage1Args s;
        const long Tp = nt_in >> p;
        s.in = p ? lds_out[p].ptr : grid_buf.ptr;
        s.in_bstride = p ? (N / 2) * Tp : N * nt_in;
        s.ring.base = rings[p].ptr;
        s.ring.row_off = (const long *) ring1_row_off[p].ptr;     // spilled rows: the outbox
        s.ring.row_cap = (const long *) ring1_row_cap[p].ptr;
        s.ring.bstride = ring_bstride[p];
        s.state = (float *) s1_state[p].ptr;
        s.state_bstride = s1_state_bstride[p];
        s.B = int(B);
        s.amb = int(plan->stage1_amb_rank[p]);
        s.dd = int(plan->stage1_dd_rank[p]);
        s.Tp = int(Tp);
        s.beam0 = beam0;
        s.chunk = ichunk;
        s.chunk_dev = chunk_dev;
        hipEvent_t ev;
        const std::string nm = "stage1_p" + std::to_string(p);
        timed_begin(nm, sp, ev);
        if (nspill[p] > 0)     // the outbox is free once the previous chunk's copy has read it
            HIP_CHECK(hipStreamWaitEvent(sp, ev_spill_d2h[p * nbatches + ibatch], 0));
        launch_stage1_fast(dtype, s, sp);
        timed_end(nm, sp, ev);
        if (nspill[p] > 0) {
            spill_d2h(p, ichunk, ibatch, beam0, sp);
            spill_scatter(p, ichunk, ibatch, beam0, sp);
        }
        for (long it = 0; it < ntrees; it++) {
            const DedispersionTree &t = plan->trees[it];
            const TreeTables &T = tables[it];
            if (t.primary_tree_index != p)
                continue;
            Stage2Args s2;
            s2.ring.base = rings[p].ptr;
            s2.ring.row_off = (const long *) ring_row_off[p].ptr;
            s2.ring.row_cap = (const long *) ring_row_cap[p].ptr;
            s2.ring.bstride = ring_bstride[p];
            s2.row_lag = (const int *) stage2_lags[it].ptr;
            s2.sb = (char *) sb[it].ptr + size_t(T.lead) * elt_bytes;
            s2.sb_bstride = t.ndm_out * T.M * T.rowlen;
            s2.rowlen = T.rowlen;
            s2.tpad = T.tpad;
            s2.main_emit = (const int *) d_main_emit[it].ptr;
            s2.stag = (const int *) d_stag[it].ptr;
            s2.lvl0 = (const int *) d_lvl0[it].ptr;
            s2.fl = (const int *) d_fl[it].ptr;
            s2.nstag = T.nstag;
            s2.nlvl0 = T.nlvl0;
            s2.stag_hist_nelts = T.split ? T.stag_hist_g : T.stag_hist_nelts;
            s2.stag_goff = (const int *) d_stag_goff[it].ptr;
            s2.stag_begin = (const int *) d_stag_begin[it].ptr;
            s2.M = T.M;
            s2.pf_rank = int(t.frequency_subbands.pf_rank);
            s2.state = (float *) s2_state[it].ptr;
            s2.state_bstride = s2_state_bstride[it];
            s2.B = int(B);
            s2.amb = t.amb_rank;
            s2.dd = t.dd_rank;
            s2.Tp = t.nt_ds;
            s2.beam0 = beam0;
            s2.amb1_dd1 = t.amb_rank;
            s2.chunk = ichunk;
            s2.chunk_dev = chunk_dev;
            s2.out_stream = (const int *) d_out_stream[it].ptr;
            hipEvent_t ev2;
            const std::string nm2 = "stage2_t" + std::to_string(it);
            timed_begin(nm2, sp, ev2);
            if (T.fast)
                launch_stage2_fast(dtype, s2, sp);
            else if (T.split) {
                // the float32 scratch is shared by every rank-8 split tree
                HIP_CHECK(hipStreamWaitEvent(sp, ev_split_free, 0));
                launch_stage2_split(dtype, s2, (float *) split_x.ptr, sp);
                HIP_CHECK(hipEventRecord(ev_split_free, sp));
            }
            else
                launch_stage2(dtype, s2, sp);
            timed_end(nm2, sp, ev2);
            if (late_fork && !serial && it == 0)
                HIP_CHECK(hipEventRecord(ev_fork_s, st));
            if (it == scorer_tree) {
                // the learned scorer reads this chunk's rows before the peak finder rotates them
                scorer::TileSource src;
                src.base = (char *) sb[it].ptr + (size_t(T.lead + T.tpad) + size_t(beam0) * s2.sb_bstride) * elt_bytes;
                src.dtype = (dtype == Dtype::float32) ? scorer::kFloat32 : (dtype == Dtype::float16) ? scorer::kFloat16
                                                                                                     : scorer::kBfloat16;
                src.nbeams = B;
                src.beam_stride = s2.sb_bstride;
                src.nrows = long(t.ndm_out) * T.M;
                src.row_stride = T.rowlen;
                src.t_lo = 0;
                src.t_hi = t.nt_ds;
                src.origins = (const int *) scorer_origins.ptr;
                const long ntiles = B * scorer_ntr * scorer_ntt;
                hipEvent_t ev3;
                timed_begin("scorer_t" + std::to_string(it), sp, ev3);
                scorer::launch_forward((const float *) scorer_params.ptr, src, ntiles,
                                       (float *) scorer_out.ptr + oslot * ntiles, sp);
                timed_end("scorer_t" + std::to_string(it), sp, ev3);
            }

            PfArgs f;
            f.sb = (char *) sb[it].ptr + size_t(T.lead) * elt_bytes;
            f.sb_bstride = s2.sb_bstride;
            f.rowlen = T.rowlen;
            f.tpad = T.tpad;
            f.flmax = T.flmax;
            f.m_to_n = (const int *) d_m_to_n[it].ptr;
            const long wslot_n = B * t.ndm_wt * t.nt_wt * t.nprofiles * T.N;
            f.wt = (const float *) weights[it].ptr + wslot * wslot_n;
            f.wt_bstride = t.ndm_wt * t.nt_wt * t.nprofiles * T.N;
            const long oslot_n = B * t.ndm_out * t.nt_out;
            f.out_max = (char *) out_max[it].ptr + oslot * oslot_n * elt_bytes;
            f.out_argmax = (unsigned int *) out_argmax[it].ptr + oslot * oslot_n;
            f.out_bstride = t.ndm_out * t.nt_out;
            f.B = int(B);
            f.ndm_out = int(t.ndm_out);
            f.M = T.M;
            f.N = T.N;
            f.P = int(t.nprofiles);
            f.nlev = T.nlev;
            f.Tp = t.nt_ds;
            f.Dout = int(t.pf.time_downsampling);
            f.Dcore = int(t.Dcore);
            f.ndm_wt = int(t.ndm_wt);
            f.nt_wt = int(t.nt_wt);
            f.nt_out = int(t.nt_out);
            f.beam0 = beam0;
            f.target_blocks = tuning.pf_target_blocks;
            if (params.enable_variances)
                f.out_var = (double *) out_var[it].ptr + oslot * B * t.ndm_out * T.M * t.nprofiles;
            const std::string nmp = "pf_t" + std::to_string(it);
            timed_begin(nmp, sp, ev2);
            launch_pf(dtype, f, sp);
            timed_end(nmp, sp, ev2);
        }
        if (sp != st) {
            HIP_CHECK(hipEventRecord(ev_join[istream][p], sp));
            HIP_CHECK(hipStreamWaitEvent(st, ev_join[istream][p], 0));
        }
    }
    HIP_CHECK(hipGetLastError());
}


// Host spill of primary tree p for (chunk, batch), on two copy streams so that both PCIe
// directions run at once and overlap the kernels:
//   spill_h2d   (issued first): host slots ((chunk - k) mod (k+1)) -> inbox, after the
//               previous chunk's device-to-host copy of this batch (which wrote those
//               slots, k >= 1) and its scatter (which read the inbox);
//   spill_d2h   (after stage 1): outbox -> host slot (chunk mod (k+1));
//   spill_scatter (on the compute stream, before stage 2): inbox -> device ring rows at
//               ring time chunk * Tp.
// Stage 1 waits for the previous chunk's device-to-host copy before it refills the outbox.
// A slot not yet written holds zeros (the stream before time 0).
void GpuDedisperser::spill_copy(long p, long ichunk, int beam0, bool to_host)
{
    const long Tp = nt_in >> p, B = beams_per_batch;
    const size_t rb_pitch = size_t(ring_bstride[p]) * elt_bytes, hb_pitch = size_t(host_bstride[p]) * elt_bytes;
    char *ring_b = (char *) rings[p].ptr + size_t(beam0) * rb_pitch;
    char *host_b = (char *) host_ring[p] + size_t(beam0) * hb_pitch;
    for (const SpillGroup &g : spill_groups[p]) {
        const size_t width = size_t(g.nrows) * Tp * elt_bytes;
        if (to_host) {
            char *dev = ring_b + size_t(outbox_off[p] + g.row0 * Tp) * elt_bytes;
            char *host = host_b + size_t(g.host_off + (ichunk % (g.k + 1)) * g.nrows * Tp) * elt_bytes;
            HIP_CHECK(hipMemcpy2DAsync(host, hb_pitch, dev, rb_pitch, width, B, hipMemcpyDeviceToHost, d2h_stream));
        } else {
            const long slot = (((ichunk - g.k) % (g.k + 1)) + (g.k + 1)) % (g.k + 1);
            char *dev = ring_b + size_t(inbox_off[p] + g.row0 * Tp) * elt_bytes;
            char *host = host_b + size_t(g.host_off + slot * g.nrows * Tp) * elt_bytes;
            HIP_CHECK(hipMemcpy2DAsync(dev, rb_pitch, host, hb_pitch, width, B, hipMemcpyHostToDevice, h2d_stream));
        }
    }
}

void GpuDedisperser::spill_h2d(long p, long ichunk, long ibatch, int beam0)
{
    const long e = p * nbatches + ibatch;
    HIP_CHECK(hipStreamWaitEvent(h2d_stream, ev_spill_d2h[e], 0));
    HIP_CHECK(hipStreamWaitEvent(h2d_stream, ev_spill_scatter[e], 0));
    spill_copy(p, ichunk, beam0, false);
    HIP_CHECK(hipEventRecord(ev_spill_h2d[e], h2d_stream));
}

void GpuDedisperser::spill_d2h(long p, long ichunk, long ibatch, int beam0, hipStream_t st)
{
    const long e = p * nbatches + ibatch;
    HIP_CHECK(hipEventRecord(ev_spill_s1[e], st));
    HIP_CHECK(hipStreamWaitEvent(d2h_stream, ev_spill_s1[e], 0));
    spill_copy(p, ichunk, beam0, true);
    HIP_CHECK(hipEventRecord(ev_spill_d2h[e], d2h_stream));
}

void GpuDedisperser::spill_scatter(long p, long ichunk, long ibatch, int beam0, hipStream_t st)
{
    const long e = p * nbatches + ibatch;
    HIP_CHECK(hipStreamWaitEvent(st, ev_spill_h2d[e], 0));
    hipEvent_t ev;
    timed_begin("spill_p" + std::to_string(p), st, ev);
    launch_spill_scatter(rings[p].ptr, ring_bstride[p], int(elt_bytes), inbox_off[p], (const int *) spill_rows[p].ptr,
                         nspill[p], (const long *) ring_row_off[p].ptr, (const long *) ring_row_cap[p].ptr, nt_in >> p,
                         ichunk, beam0, int(beams_per_batch), st);
    timed_end("spill_p" + std::to_string(p), st, ev);
    HIP_CHECK(hipEventRecord(ev_spill_scatter[e], st));
}


void GpuDedisperser::check_stopped_locked()
{
    if (stopped)
        throw pdd_error("GpuDedisperser was stopped: " + stop_reason);
}


void GpuDedisperser::stop(const std::string &reason)
{
    std::lock_guard<std::mutex> lk(mutex);
    if (!stopped) {
        stopped = true;
        stop_reason = reason.empty() ? "stop() called" : reason;
    }
    cv.notify_all();
}


bool GpuDedisperser::is_stopped()
{
    std::lock_guard<std::mutex> lk(mutex);
    return stopped;
}


// Reference: src_lib/GpuDedisperser.cpp:742-823 (cursor checks; slot = seq_id % nstreams;
// the caller's stream waits until the slot's previous contents have been consumed).
void *GpuDedisperser::acquire_input(long seq_id, hipStream_t stream)
{
    std::unique_lock<std::mutex> lk(mutex);
    check_stopped_locked();
    PDD_CHECK(is_allocated, "GpuDedisperser::acquire_input(): allocate() has not been called");
    PDD_CHECK(!group_member, "GpuDedisperser::acquire_input(): this dedisperser is driven by a GpuDedisperserGroup");
    if (seq_id != curr_input_acquire || curr_input_acquire != curr_input_release) {
        std::ostringstream ss;
        ss << "GpuDedisperser::acquire_input(): expected seq_id=" << curr_input_release << ", got " << seq_id
           << " (acquire/release must alternate, in order)";
        stopped = true;
        stop_reason = ss.str();
        cv.notify_all();
        throw pdd_error(ss.str());
    }
    const long slot = seq_id % nstreams;
    HIP_CHECK(hipStreamWaitEvent(stream, ev_input_free[slot], 0));
    curr_input_acquire++;
    return (char *) input_slots.ptr + size_t(slot) * beams_per_batch * nfreq * nt_in * elt_bytes;
}


// Reference: src_lib/GpuDedisperser.cpp:826-855, 990-1105. Back-pressure: the host blocks
// until every consumer has released output seq_id - nbatches_out.
void GpuDedisperser::release_input_and_launch(long seq_id, hipStream_t stream)
{
    std::unique_lock<std::mutex> lk(mutex);
    check_stopped_locked();
    if (seq_id != curr_input_release || curr_input_acquire != curr_input_release + 1) {
        std::ostringstream ss;
        ss << "GpuDedisperser::release_input_and_launch(): expected seq_id=" << curr_input_release << ", got " << seq_id;
        stopped = true;
        stop_reason = ss.str();
        cv.notify_all();
        throw pdd_error(ss.str());
    }
    const long oslot = seq_id % nbatches_out;
    cv.wait(lk, [&] {
        if (stopped)
            return true;
        for (long c = 0; c < num_consumers; c++)
            if (curr_output_release[c] < seq_id - nbatches_out + 1)
                return false;
        return true;
    });
    check_stopped_locked();

    // Input slot s runs on compute stream s (reference: src_lib/GpuDedisperser.cpp:998); the
    // batch's previous chunk, which may have run on another stream, must have finished
    // (its tree states, lag ring and peak-finder rows carry over; :1010-1013).
    const long slot = seq_id % nstreams;
    const long ichunk = seq_id / nbatches, ibatch = seq_id % nbatches;
    hipStream_t cs = compute_streams[slot];
    HIP_CHECK(hipEventRecord(ev_input_ready[slot], stream));
    HIP_CHECK(hipStreamWaitEvent(cs, ev_input_ready[slot], 0));
    HIP_CHECK(hipStreamWaitEvent(cs, ev_batch_done[ibatch], 0));
    for (long c = 0; c < num_consumers; c++)
        HIP_CHECK(hipStreamWaitEvent(cs, ev_output_released[oslot * num_consumers + c], 0));

    const void *in = (char *) input_slots.ptr + size_t(slot) * beams_per_batch * nfreq * nt_in * elt_bytes;
    launch_chain(in, ichunk, ibatch, oslot, seq_id % nbatches_wt, cs, slot);

    HIP_CHECK(hipEventRecord(ev_batch_done[ibatch], cs));
    HIP_CHECK(hipEventRecord(ev_input_free[slot], cs));
    HIP_CHECK(hipEventRecord(ev_output_ready[oslot], cs));
    curr_input_release++;
    curr_launched = seq_id + 1;
    cv.notify_all();
}


// Reference: src_lib/GpuDedisperser.cpp:858-935.
void GpuDedisperser::acquire_output(long consumer, long seq_id, hipStream_t stream, bool sync,
                                    std::vector<void *> &omax, std::vector<unsigned int *> &oarg)
{
    std::unique_lock<std::mutex> lk(mutex);
    check_stopped_locked();
    PDD_CHECK(consumer >= 0 && consumer < num_consumers, "acquire_output(): consumer out of range");
    PDD_CHECK(!group_member, "GpuDedisperser::acquire_output(): this dedisperser is driven by a GpuDedisperserGroup");
    long &acq = curr_output_acquire[consumer];
    long &rel = curr_output_release[consumer];
    bool ok = (seq_id == acq) && (params.synchronous ? (acq == rel) : true);
    if (!ok) {
        std::ostringstream ss;
        ss << "GpuDedisperser::acquire_output(): consumer " << consumer << " expected seq_id=" << acq << ", got " << seq_id;
        stopped = true;
        stop_reason = ss.str();
        cv.notify_all();
        throw pdd_error(ss.str());
    }
    cv.wait(lk, [&] { return stopped || curr_launched > seq_id; });
    check_stopped_locked();

    const long oslot = seq_id % nbatches_out;
    HIP_CHECK(hipStreamWaitEvent(stream, ev_output_ready[oslot], 0));
    if (sync)
        HIP_CHECK(hipEventSynchronize(ev_output_ready[oslot]));
    acq++;

    omax.resize(ntrees);
    oarg.resize(ntrees);
    for (long it = 0; it < ntrees; it++) {
        const DedispersionTree &t = plan->trees[it];
        const long n = beams_per_batch * t.ndm_out * t.nt_out;
        omax[it] = (char *) out_max[it].ptr + oslot * n * elt_bytes;
        oarg[it] = (unsigned int *) out_argmax[it].ptr + oslot * n;
    }
}


void GpuDedisperser::output_variances(long seq_id, std::vector<double *> &ov)
{
    std::lock_guard<std::mutex> lk(mutex);
    check_stopped_locked();
    PDD_CHECK(params.enable_variances, "output_variances(): the dedisperser was created without enable_variances");
    bool held = false;
    for (long c = 0; c < num_consumers; c++)
        held = held || (seq_id >= curr_output_release[c] && seq_id < curr_output_acquire[c]);
    PDD_CHECK(held, "output_variances(): seq_id " << seq_id << " is not an acquired output");
    const long oslot = seq_id % nbatches_out;
    ov.resize(ntrees);
    for (long it = 0; it < ntrees; it++) {
        const DedispersionTree &t = plan->trees[it];
        ov[it] = (double *) out_var[it].ptr + oslot * beams_per_batch * t.ndm_out * tables[it].M * t.nprofiles;
    }
}


void GpuDedisperser::release_output(long consumer, long seq_id, hipStream_t stream)
{
    std::unique_lock<std::mutex> lk(mutex);
    check_stopped_locked();
    PDD_CHECK(consumer >= 0 && consumer < num_consumers, "release_output(): consumer out of range");
    long &acq = curr_output_acquire[consumer];
    long &rel = curr_output_release[consumer];
    if (seq_id != rel || rel >= acq) {
        std::ostringstream ss;
        ss << "GpuDedisperser::release_output(): consumer " << consumer << " expected seq_id=" << rel << ", got " << seq_id;
        stopped = true;
        stop_reason = ss.str();
        cv.notify_all();
        throw pdd_error(ss.str());
    }
    const long oslot = seq_id % nbatches_out;
    HIP_CHECK(hipEventRecord(ev_output_released[oslot * num_consumers + consumer], stream));
    rel++;
    cv.notify_all();
}


// Reference: GpuDedisperser::fill_all_weights (src_lib/GpuDedisperser.cpp:1257-1279).
// The peak finder here takes the per-group maximum before weighting, which equals the
// maximum of the weighted candidates only for non-negative weights; negative or non-finite
// weights are rejected.
void GpuDedisperser::fill_all_weights(long itree, const float *host, long nelts)
{
    std::lock_guard<std::mutex> lk(mutex);
    check_stopped_locked();
    PDD_CHECK(is_allocated, "fill_all_weights(): allocate() has not been called");
    PDD_CHECK(itree >= 0 && itree < ntrees, "fill_all_weights(): itree out of range");
    const DedispersionTree &t = plan->trees[itree];
    long expect = nbatches_wt * beams_per_batch * t.ndm_wt * t.nt_wt * t.nprofiles * tables[itree].N;
    PDD_CHECK(nelts == expect, "fill_all_weights(): expected " << expect << " weights (nbatches_wt, B, ndm_wt, nt_wt, P, N), got " << nelts);
    for (long i = 0; i < nelts; i++)
        PDD_CHECK(std::isfinite(host[i]) && host[i] >= 0.0f, "fill_all_weights(): weights must be finite and >= 0 (index " << i << ")");
    HIP_CHECK(hipMemcpy(weights[itree].ptr, host, nelts * sizeof(float), hipMemcpyHostToDevice));
}


// Reference: GpuDedisperser::fill_analytic_weights (src_lib/GpuDedisperser.cpp:1162-1238) and
// PeakFindingKernelParams::fill_host_weights with variances (src_lib/PeakFindingKernel.cu:
// 606-691, randomize=false): w[b, d, t, p, n] = 1/sqrt(float(variance[n, d, p])) in every
// beam and weight slot.
void GpuDedisperser::fill_analytic_weights(const double *freq_variances, long nfreq_in)
{
    {
        std::lock_guard<std::mutex> lk(mutex);
        check_stopped_locked();
        PDD_CHECK(is_allocated, "fill_analytic_weights(): allocate() has not been called");
    }
    const PfAvarApproximation avar(*plan, freq_variances, nfreq_in);
    for (long itree = 0; itree < ntrees; itree++) {
        const DedispersionTree &t = plan->trees[itree];
        const long D = t.ndm_wt, T = t.nt_wt, P = t.nprofiles, N = t.frequency_subbands.N;
        const std::vector<double> &v = avar.tree_variance[itree];     // (N, D, P)
        std::vector<float> base(D * P * N);                            // (D, P, N)
        for (long n = 0; n < N; n++)
            for (long d = 0; d < D; d++)
                for (long p = 0; p < P; p++)
                    base[(d * P + p) * N + n] = 1.0f / std::sqrt(float(v[(n * D + d) * P + p]));
        const long nslot = nbatches_wt * beams_per_batch;
        std::vector<float> host(nslot * D * T * P * N);
        float *o = host.data();
        for (long s = 0; s < nslot; s++)
            for (long d = 0; d < D; d++)
                for (long it = 0; it < T; it++, o += P * N)
                    std::memcpy(o, &base[d * P * N], P * N * sizeof(float));
        fill_all_weights(itree, host.data(), long(host.size()));
    }
}


hipEvent_t GpuDedisperser::timing_event()
{
    if (!timing_event_pool.empty()) {
        hipEvent_t e = timing_event_pool.back();
        timing_event_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    HIP_CHECK(hipEventCreate(&e));
    return e;
}

void GpuDedisperser::timed_begin(const std::string &, hipStream_t st, hipEvent_t &ev)
{
    ev = nullptr;
    if (!timing)
        return;
    ev = timing_event();
    HIP_CHECK(hipEventRecord(ev, st));
}

void GpuDedisperser::timed_end(const std::string &name, hipStream_t st, hipEvent_t ev)
{
    if (!timing || !ev)
        return;
    hipEvent_t e2 = timing_event();
    HIP_CHECK(hipEventRecord(e2, st));
    timed_pending.push_back({name, ev, e2});
    if (timed_pending.size() > 4096)
        collect_timing_locked();
}

void GpuDedisperser::collect_timing_locked()
{
    for (TimedLaunch &t : timed_pending) {
        HIP_CHECK(hipEventSynchronize(t.stop));
        float ms = 0.0f;
        HIP_CHECK(hipEventElapsedTime(&ms, t.start, t.stop));
        bool found = false;
        for (auto &kv : timing_totals)
            if (kv.first == t.name) {
                kv.second.first++;
                kv.second.second += ms;
                found = true;
            }
        if (!found)
            timing_totals.push_back({t.name, {1, double(ms)}});
        timing_event_pool.push_back(t.start);
        timing_event_pool.push_back(t.stop);
    }
    timed_pending.clear();
}

void GpuDedisperser::set_kernel_timing(bool enable)
{
    std::lock_guard<std::mutex> lk(mutex);
    collect_timing_locked();
    timing = enable;
    timing_totals.clear();
}

std::string GpuDedisperser::kernel_timing_report()
{
    std::lock_guard<std::mutex> lk(mutex);
    collect_timing_locked();
    std::ostringstream os;
    for (auto &kv : timing_totals)
        os << kv.first << " " << kv.second.first << " " << kv.second.second << "\n";
    return os.str();
}


void GpuDedisperser::output_ipc_handle(void *handle, size_t *nbytes)
{
    PDD_CHECK(is_allocated, "output_ipc_handle(): allocate() has not been called");
    PDD_CHECK(out_region.nbytes < (size_t(1) << 31),
              "output_ipc_handle(): the output ring is " << out_region.nbytes << " bytes; pirate_amd exports at most 2 GiB");
    static_assert(sizeof(hipIpcMemHandle_t) == 64, "IPC handle size");
    HIP_CHECK(hipSetDevice(params.device));
    hipIpcMemHandle_t h;
    HIP_CHECK(hipIpcGetMemHandle(&h, out_region.ptr));
    std::memcpy(handle, &h, sizeof(h));
    *nbytes = out_region.nbytes;
}

size_t GpuDedisperser::output_array_offset(long itree, int which) const
{
    PDD_CHECK(is_allocated, "output_array_offset(): allocate() has not been called");
    PDD_CHECK(itree >= 0 && itree < ntrees && (which == 0 || which == 1), "output_array_offset(): bad tree or array");
    return out_offsets[2 * itree + which];
}


size_t GpuDedisperser::debug_copy(const std::string &name, void *host, size_t maxbytes)
{
    std::lock_guard<std::mutex> lk(mutex);
    PDD_CHECK(is_allocated, "debug_copy(): not allocated");
    for (hipStream_t x : compute_streams)
        HIP_CHECK(hipStreamSynchronize(x));
    const DeviceArray *a = nullptr;
    if (name == "grid")
        a = &grid_bufs[last_stream];
    else if (name.rfind("lds", 0) == 0)
        a = &lds_outs[last_stream].at(std::stol(name.substr(3)));
    else if (name.rfind("sb", 0) == 0)
        a = &sb.at(std::stol(name.substr(2)));
    else if (name.rfind("ring", 0) == 0)
        a = &rings.at(std::stol(name.substr(4)));
    PDD_CHECK(a, "debug_copy(): unknown buffer '" << name << "'");
    if (host) {
        PDD_CHECK(maxbytes >= a->nbytes, "debug_copy(): host buffer too small");
        HIP_CHECK(hipMemcpy(host, a->ptr, a->nbytes, hipMemcpyDeviceToHost));
    }
    return a->nbytes;
}

void GpuDedisperser::set_scorer(long itree, const float *params_host, long nparams)
{
    std::lock_guard<std::mutex> lk(mutex);
    PDD_CHECK(is_allocated, "set_scorer(): not allocated");
    PDD_CHECK(itree >= 0 && itree < ntrees, "set_scorer(): itree out of range");
    PDD_CHECK(params_host && nparams == scorer::kNParams, "set_scorer(): need " << scorer::kNParams << " parameters");
    const DedispersionTree &t = plan->trees[itree];
    const TreeTables &T = tables[itree];
    const long nrows = long(t.ndm_out) * T.M;
    PDD_CHECK(T.N == 1 && T.flmax == 0, "set_scorer(): tree " << itree << " has a subband search; the scorer reads "
              "full-band DM trials");
    PDD_CHECK(nrows % scorer::kC0 == 0 && t.nt_ds % scorer::kT == 0, "set_scorer(): tree " << itree << " has " << nrows
              << " rows and " << t.nt_ds << " samples per chunk; tiles need multiples of 64 and 256");
    for (hipStream_t x : compute_streams)
        HIP_CHECK(hipStreamSynchronize(x));
    const long B = beams_per_batch, ntr = nrows / scorer::kC0, ntt = t.nt_ds / scorer::kT;
    if (!scorer_params.ptr)
        scorer_params = dalloc(size_t(scorer::kNParams) * sizeof(float));
    HIP_CHECK(hipMemcpy(scorer_params.ptr, params_host, size_t(scorer::kNParams) * sizeof(float), hipMemcpyHostToDevice));
    if (scorer_tree != itree) {
        std::vector<int> o;
        o.reserve(size_t(B * ntr * ntt * 3));
        for (long b = 0; b < B; b++)
            for (long r = 0; r < ntr; r++)
                for (long c = 0; c < ntt; c++) {
                    o.push_back(int(b));
                    o.push_back(int(r * scorer::kC0));
                    o.push_back(int(c * scorer::kT));
                }
        for (DeviceArray *a : {&scorer_origins, &scorer_out})
            if (a->ptr) {
                HIP_CHECK(hipFree(a->ptr));
                gpu_bytes_allocated -= a->nbytes;
                a->ptr = nullptr;
            }
        scorer_origins = dalloc(o.size() * sizeof(int));
        HIP_CHECK(hipMemcpy(scorer_origins.ptr, o.data(), o.size() * sizeof(int), hipMemcpyHostToDevice));
        scorer_out = dalloc(size_t(nbatches_out) * B * ntr * ntt * sizeof(float));
        scorer_ntr = ntr;
        scorer_ntt = ntt;
        scorer_tree = itree;
    }
}

void GpuDedisperser::output_scores(long seq_id, float *&scores, long &ntiles_dm, long &ntiles_t)
{
    std::lock_guard<std::mutex> lk(mutex);
    check_stopped_locked();
    PDD_CHECK(scorer_tree >= 0, "output_scores(): no scorer set (set_scorer)");
    bool held = false;
    for (long c = 0; c < num_consumers; c++)
        held = held || (seq_id >= curr_output_release[c] && seq_id < curr_output_acquire[c]);
    PDD_CHECK(held, "output_scores(): seq_id " << seq_id << " is not an acquired output");
    const long oslot = seq_id % nbatches_out;
    scores = (float *) scorer_out.ptr + oslot * beams_per_batch * scorer_ntr * scorer_ntt;
    ntiles_dm = scorer_ntr;
    ntiles_t = scorer_ntt;
}

void GpuDedisperser::tree_rows(long itree, void *&base, long &nrows, long &beam_stride, long &row_stride, long &t_lo,
                               long &t_hi)
{
    std::lock_guard<std::mutex> lk(mutex);
    PDD_CHECK(is_allocated, "tree_rows(): not allocated");
    PDD_CHECK(itree >= 0 && itree < ntrees, "tree_rows(): itree out of range");
    const TreeTables &T = tables[itree];
    const DedispersionTree &t = plan->trees[itree];
    base = (char *) sb[itree].ptr + size_t(T.lead + T.tpad) * elt_bytes;
    nrows = long(t.ndm_out) * T.M;
    row_stride = T.rowlen;
    beam_stride = nrows * T.rowlen;
    // the peak finder ends the chunk by moving its last tpad + flmax samples to the row
    // start, so [flmax, nt_ds) is what still holds this chunk's samples
    t_lo = T.flmax;
    t_hi = t.nt_ds;
}

}  // namespace pirate_amd
