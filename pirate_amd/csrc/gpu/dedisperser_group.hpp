// GpuDedisperserGroup: beam groups of different shapes (channel counts, tree ranks, numbers of
// primary trees and downsampling levels), each with its own DedispersionPlan, processed as
// one chunk pipeline on one GPU.
//
// The reference batches homogeneous beams only (beams_per_batch beams of one plan per launch
// sequence, include/pirate/DedispersionConfig.hpp:81-84, src_lib/GpuDedisperser.cpp:988-1105).
// Here every member is a GpuDedisperser whose whole beam set is one batch; per chunk, every
// member's chain (gridding + lagged downsampling, stage-1 trees, stage-2 trees, peak finding)
// is issued on a branch stream of its own, forked from and joined into one launch stream, so
// the members' kernels share the device at once. The kernels can read the chunk index from a
// device counter that the chunk's last node advances (for replay of a captured HIP graph, one
// per (input slot, output slot, weight slot) phase); graph replay is refused
// (see dedisperser_group.cpp), and the launches are issued one by one.
//
// Per-chunk protocol (one producer, one consumer, chunks in order):
//   acquire_input(c) -> per-member input slots (B_m, F_m, T) to fill on the caller's stream
//   release_input_and_launch(c)
//   acquire_output(c) -> member outputs ready on the caller's stream
//   release_output(c)
// release_input_and_launch(c) blocks until output c - nbatches_out has been released.

#ifndef PIRATE_AMD_DEDISPERSER_GROUP_HPP
#define PIRATE_AMD_DEDISPERSER_GROUP_HPP

#include "dedisperser.hpp"

#include <condition_variable>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

namespace pirate_amd {

struct GpuDedisperserGroupParams {
    long nslots = 2;         // input slots per member (the caller fills chunk c+1 while c runs)
    long nbatches_out = 2;   // output slots per member
    int device = 0;
    bool use_graph = false;  // true (HIP graph replay) is refused: see dedisperser_group.cpp
};

class GpuDedisperserGroup {
public:
    GpuDedisperserGroup(const std::vector<std::shared_ptr<DedispersionPlan>> &plans, const GpuDedisperserGroupParams &params);
    ~GpuDedisperserGroup();
    GpuDedisperserGroup(const GpuDedisperserGroup &) = delete;
    GpuDedisperserGroup &operator=(const GpuDedisperserGroup &) = delete;

    GpuDedisperserGroupParams params;
    std::vector<std::shared_ptr<GpuDedisperser>> members;
    long nslots = 0, nbatches_out = 0, nbatches_wt = 0;
    long period = 0;          // graphs per phase cycle: lcm(nslots, nbatches_out, nbatches_wt)

    void allocate();
    bool is_allocated = false;
    size_t gpu_bytes() const;

    void acquire_input(long ichunk, hipStream_t stream, std::vector<void *> &slots);
    void release_input_and_launch(long ichunk, hipStream_t stream);
    void acquire_output(long ichunk, hipStream_t stream, bool sync);
    void release_output(long ichunk, hipStream_t stream);
    void stop(const std::string &reason);

    // Captured graphs are dropped (and re-captured on next use) when a launch shape changes.
    void set_use_graph(bool on);
    void set_tuning(long imember, const std::string &name, long value);

    // Kernel nodes of one captured chunk graph (0 before the first graph launch).
    long graph_nodes() const { return graph_nnodes; }

private:
    std::mutex mutex;
    std::condition_variable cv;
    bool stopped = false;
    std::string stop_reason;
    long curr_input_acquire = 0, curr_input_release = 0, curr_output_acquire = 0, curr_output_release = 0;

    hipStream_t launch_stream = nullptr;            // graph launches / eager fork point
    std::vector<hipStream_t> member_streams;        // [m] branch origin of member m
    long *chunk_counter = nullptr;                  // device: chunk index read by the kernels
    std::vector<DeviceArray> inputs;                // [m] (nslots, B_m, F_m, T) input slots
    std::vector<hipGraphExec_t> graphs;             // [phase]
    long graph_nnodes = 0;
    hipEvent_t ev_fork = nullptr;
    std::vector<hipEvent_t> ev_join;                // [m]
    std::vector<hipEvent_t> ev_input_free, ev_input_ready;          // [slot]
    std::vector<hipEvent_t> ev_output_ready, ev_output_released;    // [oslot]

    void *input_slot(size_t m, long slot) const;
    void check_stopped_locked();
    void fail_locked(const std::string &msg);
    void issue_chunk(long ichunk, hipStream_t st, bool capture);    // every member's chain
    void drop_graphs();
};

}  // namespace pirate_amd

#endif
