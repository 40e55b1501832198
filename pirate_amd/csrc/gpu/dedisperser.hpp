// GpuDedisperser: the MI355X dedispersion chain behind the reference's acquire/release API.
// Reference interface: include/pirate/Dedisperser.hpp:45-392 (GpuDedisperser),
// src_lib/GpuDedisperser.cpp:742-984 (acquire_input / release_input_and_launch_dd_kernels /
// acquire_output / release_output cursors, back-pressure and stop semantics).
//
// One compute stream per active batch (num_active_batches of them, as the reference's
// stream pool: include/pirate/CudaStreamPool.hpp:20-50) runs, per (chunk, beam batch):
//   tree gridding -> lagged downsampling -> stage-1 trees -> lag ring ->
//   stage-2 trees + subband streams -> peak finding -> output ring slot.
// The caller's streams synchronize with it through events only.

#ifndef PIRATE_AMD_DEDISPERSER_HPP
#define PIRATE_AMD_DEDISPERSER_HPP

#include "../host/plan.hpp"
#include "kernels.hpp"

#include <condition_variable>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

namespace pirate_amd {

struct GpuDedisperserParams {
    long nbatches_out = 0;     // output ring slots (0 = num_active_batches)
    long nbatches_wt = 0;      // weight ring slots (0 = num_active_batches)
    long num_consumers = 1;
    int device = 0;
    bool synchronous = false;
    long initial_chunk = 0;
    bool enable_variances = false;   // peak finder also writes out_var (float64) per output slot
};

// Launch shapes that change no result (only speed), settable between chunks; the values
// below are the defaults, the autotuner's table overrides them per workload.
struct LaunchTuning {
    long pf_target_blocks = 2048;   // peak finder: fewest time segments giving this many workgroups
    int grid_xcd_run = 8;           // gridding: consecutive channel pairs per XCD (0: launch order)
    int fork = -1;                  // primary trees p >= 1 start: -1 auto, 0 after the gridding,
                                    // 1 after tree 0's stage 2, 2 serial (no branch streams)
};

struct DeviceArray {
    void *ptr = nullptr;
    size_t nbytes = 0;
};

class GpuDedisperser {
public:
    GpuDedisperser(std::shared_ptr<DedispersionPlan> plan, const GpuDedisperserParams &params);
    ~GpuDedisperser();
    GpuDedisperser(const GpuDedisperser &) = delete;
    GpuDedisperser &operator=(const GpuDedisperser &) = delete;

    std::shared_ptr<DedispersionPlan> plan;
    GpuDedisperserParams params;
    LaunchTuning tuning;
    // Set when a GpuDedisperserGroup drives this dedisperser: its own acquire/release entry
    // points are then refused (the group launches its chains).
    bool group_member = false;

    // name in {"pf_target_blocks", "grid_xcd_run", "fork"}; throws on a bad name or value.
    void set_tuning(const std::string &name, long value);
    long get_tuning(const std::string &name) const;

    Dtype dtype;
    long nfreq, nt_in, total_beams, beams_per_batch, nbatches, nstreams, ntrees;
    long nbatches_out, nbatches_wt, num_consumers;
    size_t elt_bytes;

    void allocate();
    bool is_allocated = false;
    // PCIe bytes per beam batch of the host spill (each direction; 0 without max_gpu_clag).
    size_t spill_bytes_per_batch() const;
    size_t gpu_bytes_allocated = 0;
    size_t host_bytes_allocated = 0;  // pinned host memory of spilled lag-ring rows

    // Returns the device address of the (B, nfreq, nt_in) input slot for seq_id.
    void *acquire_input(long seq_id, hipStream_t stream);
    void release_input_and_launch(long seq_id, hipStream_t stream);

    // out_max[itree], out_argmax[itree]: device addresses of (B, ndm_out, nt_out) arrays.
    void acquire_output(long consumer, long seq_id, hipStream_t stream, bool sync,
                        std::vector<void *> &out_max, std::vector<unsigned int *> &out_argmax);
    void release_output(long consumer, long seq_id, hipStream_t stream);

    // out_var[itree]: device address of the (B, ndm_out, M, nprofiles) float64 array of an
    // acquired (not yet released) seq_id; requires enable_variances.
    void output_variances(long seq_id, std::vector<double *> &out_var);

    // Learned peak scorer on tree itree's rows, run on each chunk right after the tree's
    // stage 2: the tree's rows (DM trials, with one full-band subband) are cut into tiles of
    // 64 rows x 256 chunk samples on a fixed grid, and their logits go to an output ring of
    // (nbatches_out, B, nrows / 64, nt_ds / 256) floats. params: host copy of the scorer's
    // parameters (copied to the device). Needs allocate(); applies from the next launch.
    void set_scorer(long itree, const float *params, long nparams);
    // the (B, ntiles_dm, ntiles_t) logits of an acquired seq_id, and the tile counts
    void output_scores(long seq_id, float *&scores, long &ntiles_dm, long &ntiles_t);

    // host: (nbatches_wt, B, ndm_wt, nt_wt, P, N) float32, all entries finite and >= 0.
    void fill_all_weights(long itree, const float *host, long nelts);

    // Weights 1/sqrt(analytic variance) for every tree and weight slot, so that the peak
    // finder's output is an SNR for white-noise input with these per-channel variances
    // (length nfreq, positive).
    void fill_analytic_weights(const double *freq_variances, long nfreq);

    void stop(const std::string &reason);
    bool is_stopped();

    // Launches the chain for (ichunk, ibatch) on 'stream' without the acquire/release
    // protocol; the input is read from 'input' (device, (B, nfreq, nt_in), storage dtype)
    // and the outputs go to output slot 'oslot'. 'istream' selects the set of per-stream
    // intermediate buffers and branch streams (< num_active_batches). With 'chunk_dev' set,
    // the kernels read the chunk index from that device address instead of 'ichunk' (for a
    // captured HIP graph that is replayed for every chunk).
    void launch_chain(const void *input, long ichunk, long ibatch, long oslot, long wslot, hipStream_t stream,
                      long istream = 0, const long *chunk_dev = nullptr);

    // Device address of input slot 'slot' (< num_active_batches) and of tree itree's output
    // slot 'oslot' (which: 0 = out_max, 1 = out_argmax), for callers that drive launch_chain.
    void *input_slot(long slot) const;
    void *output_slot(long itree, long oslot, int which) const;

    // Output ring export to a consumer process: a 64-byte hipIpcMemHandle_t of the device
    // allocation that holds every tree's out_max and out_argmax ring (nbatches_out slots of
    // (B, ndm_out, nt_out)), its size, and each array's byte offset in it (which: 0 =
    // out_max, 1 = out_argmax).
    void output_ipc_handle(void *handle, size_t *nbytes);
    size_t output_array_offset(long itree, int which) const;

    // Copies an internal buffer to the host (tests): name in {"grid", "lds<p>", "sb<t>"}.
    size_t debug_copy(const std::string &name, void *host, size_t maxbytes);
    // Peak-finder input rows of tree itree (the stage-2 output): row r = idm_out * M + m of global
    // beam b, chunk sample t at base + (b * beam_stride + r * row_stride + t) * elt_bytes. After
    // the chunk's peak finding (i.e. once its output is acquired) samples flmax <= t < nt_ds of
    // the latest chunk of the beam's batch are there (flmax = 0 without subband final lags).
    void tree_rows(long itree, void *&base, long &nrows, long &beam_stride, long &row_stride, long &t_lo,
                   long &t_hi);

    // Per-kernel timing with HIP events recorded on the compute stream around every kernel
    // launch (kernel names: gridding, lds, stage1_p<p>, stage2_t<i>, pf_t<i>). The report
    // lists "name launches total_ms" lines, accumulated since timing was enabled.
    void set_kernel_timing(bool enable);
    bool kernel_timing_enabled() const { return timing; }
    std::string kernel_timing_report();

    // Input slot s (seq_id % num_active_batches) is processed on compute_streams[s], with its
    // own gridded and downsampled buffers, so that active batches run concurrently; the next
    // chunk of a batch waits for the previous one (ev_batch_done) when it runs on another
    // stream. compute_stream is compute_streams[0].
    hipStream_t compute_stream = nullptr;
    std::vector<hipStream_t> compute_streams;
    // Primary trees p >= 1 run their stage 1, stage 2 and peak finding on branch_streams[s][p]
    // (branch 0 is the calling stream), forked after the gridding and joined before the chain
    // returns, so that the primary trees' kernels overlap.
    std::vector<std::vector<hipStream_t>> branch_streams;

    // Per-tree stage-2 tables (host copies, exposed for tests).
    struct TreeTables {
        std::vector<int> main_emit, stag, lvl0, fl, m_to_n, stag_begin;
        std::vector<int> out_stream;   // register-tree path: final row -> stream (-1: none)
        bool fast = false;             // stage 2 runs on the register-tree kernel
        bool split = false;            // rank-8 subband search: two-launch split kernel
        int nstag = 0, nlvl0 = 0, stag_hist_nelts = 0, flmax = 0, tpad = 0, nlev = 0, rowlen = 0, M = 0, N = 0;
        int lead = 0;                  // element offset of the peak-finder rows in their buffer
        std::vector<int> stag_goff;    // rank 8: staggered history offset within its row group
        int stag_hist_g = 0;           // rank 8: staggered history floats per row group
    };
    std::vector<TreeTables> tables;

private:
    std::mutex mutex;
    std::condition_variable cv;
    bool stopped = false;
    std::string stop_reason;

    long curr_input_acquire = 0, curr_input_release = 0;
    std::vector<long> curr_output_acquire, curr_output_release;

    DeviceArray input_slots;          // (nstreams, B, nfreq, nt_in)
    std::vector<DeviceArray> grid_bufs;               // [s] (B, 2^R, nt_in)
    std::vector<std::vector<DeviceArray>> lds_outs;   // [s][p] (B, 2^(R-1), nt_in >> p), p >= 1
    long last_stream = 0;             // stream of the latest launch (debug_copy of "grid", "lds<p>")
    std::vector<DeviceArray> rings;   // [p]
    std::vector<DeviceArray> ring_row_off, ring_row_cap;   // [p] device (long), stage-2 (read) side
    std::vector<long> ring_bstride;

    // Host spill of lag-ring rows (config max_gpu_clag). Row r of primary tree p whose
    // smallest stage-2 lag spans more than max_gpu_clag chunks is delayed by k_r whole chunks
    // in pinned host memory: stage 1 writes the row's chunk to an outbox at the end of the
    // beam's ring region, the outbox goes to host slot (chunk mod (k_r+1)), the slot written
    // k_r chunks earlier comes back into an inbox beside it, and a scatter kernel appends it
    // to the row's (shorter) device ring, which stage 2 then reads with lag - k_r * Tp. Rows
    // are grouped by k_r, so each group is one 2-D copy per direction per batch. The two
    // directions run on their own streams, ordered by events per (primary tree, batch).
    struct SpillGroup { long k, row0, nrows, host_off; };
    std::vector<std::vector<SpillGroup>> spill_groups;       // [p]
    std::vector<std::vector<long>> spill_k;                  // [p][ring row] host delay in chunks
    std::vector<long> nspill, outbox_off, inbox_off, host_bstride;   // [p]
    hipStream_t d2h_stream = nullptr, h2d_stream = nullptr;
    std::vector<hipEvent_t> ev_spill_s1, ev_spill_d2h, ev_spill_h2d, ev_spill_scatter;   // [p * nbatches + batch]
    hipEvent_t ev_split_free = nullptr;
    std::vector<hipEvent_t> ev_fork;                          // [s]
    std::vector<std::vector<hipEvent_t>> ev_join;             // [s][p]
    std::vector<hipEvent_t> ev_batch_done;                    // [batch]: latest chain of the batch
    std::vector<DeviceArray> spill_rows;                     // [p] device int: ring row of each outbox row
    std::vector<DeviceArray> ring1_row_off, ring1_row_cap;   // [p] stage-1 (write) side
    std::vector<void *> host_ring;                           // [p] pinned, (total_beams, host_bstride)
    std::vector<DeviceArray> stage2_lags;    // [itree] device int
    std::vector<DeviceArray> sb;      // [itree] (B, ndm_out, M, rowlen)
    std::vector<DeviceArray> d_main_emit, d_stag, d_lvl0, d_fl, d_m_to_n, d_stag_begin, d_stag_goff, d_out_stream;
    DeviceArray lds_state;  long lds_state_bstride = 0;
    std::vector<DeviceArray> s1_state; std::vector<long> s1_state_bstride;
    std::vector<DeviceArray> s2_state; std::vector<long> s2_state_bstride;
    std::vector<DeviceArray> weights; // [itree] (nbatches_wt, B, ndm_wt, nt_wt, P, N) float
    std::vector<DeviceArray> out_max, out_argmax;   // [itree] (nbatches_out, B, ndm_out, nt_out), views into out_region
    DeviceArray out_region;
    DeviceArray split_x;              // float32 scratch of the rank-8 split stage 2
    std::vector<size_t> out_offsets;                // [2 * itree + which]
    std::vector<DeviceArray> out_var; // [itree] (nbatches_out, B, ndm_out, M, P) float64, if enabled
    // learned scorer (set_scorer): tree, parameters, tile origins of one batch, logits ring
    long scorer_tree = -1, scorer_ntr = 0, scorer_ntt = 0;
    DeviceArray scorer_params, scorer_origins, scorer_out;
    DeviceArray grid_rows;            // int (3*nchan) f0, cnt, woff
    DeviceArray grid_w;               // float

    std::vector<hipEvent_t> ev_input_free;        // [slot]
    std::vector<hipEvent_t> ev_input_ready;       // [slot]
    std::vector<hipEvent_t> ev_output_ready;      // [oslot]
    std::vector<hipEvent_t> ev_output_released;   // [oslot * num_consumers + c]
    long curr_launched = 0;

    bool timing = false;
    struct TimedLaunch { std::string name; hipEvent_t start, stop; };
    std::vector<TimedLaunch> timed_pending;
    std::vector<hipEvent_t> timing_event_pool;
    std::vector<std::pair<std::string, std::pair<long, double>>> timing_totals;   // name -> (launches, ms)
    hipEvent_t timing_event();
    void timed_begin(const std::string &name, hipStream_t st, hipEvent_t &ev);
    void timed_end(const std::string &name, hipStream_t st, hipEvent_t ev);
    void collect_timing_locked();

    DeviceArray dalloc(size_t nbytes);
    void check_stopped_locked();
    void spill_copy(long p, long ichunk, int beam0, bool to_host);
    void spill_h2d(long p, long ichunk, long ibatch, int beam0);
    void spill_d2h(long p, long ichunk, long ibatch, int beam0, hipStream_t st);
    void spill_scatter(long p, long ichunk, long ibatch, int beam0, hipStream_t st);
    void build_tables();
    void validate_gpu_limits();
};

}  // namespace pirate_amd

#endif
