// GpuDedisperserGroup: heterogeneous beam groups as one chunk pipeline (one HIP graph launch
// per chunk). See dedisperser_group.hpp.

#include "dedisperser_group.hpp"
#include "../host/common.hpp"

#include <hip/hip_runtime.h>

#include <numeric>
#include <sstream>

namespace pirate_amd {

void launch_advance_counter(long *counter, hipStream_t st);   // k_group.hip

#define HIP_CHECK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) \
    throw pdd_error(std::string("HIP error: ") + hipGetErrorString(e_) + " in " #x); } while (0)

// Graph replay is switched off: its first run on MI355X ended in a segmentation fault inside
// release_input_and_launch (the capture / instantiate / launch of the members' chains), whose
// cause has not been identified. The chunk's launches are issued one by one.
static void refuse_graph()
{
    throw pdd_error("GpuDedisperserGroup: HIP graph replay is disabled in pirate_amd (a segmentation fault in "
                    "its capture path is not yet explained); use use_graph = false");
}


GpuDedisperserGroup::GpuDedisperserGroup(const std::vector<std::shared_ptr<DedispersionPlan>> &plans,
                                         const GpuDedisperserGroupParams &params_) :
    params(params_)
{
    PDD_CHECK(!plans.empty(), "GpuDedisperserGroup: no plans");
    const DedispersionPlan &p0 = *plans[0];
    for (size_t m = 0; m < plans.size(); m++) {
        const DedispersionPlan &p = *plans[m];
        PDD_CHECK(p.beams_per_batch == p.beams_per_gpu, "GpuDedisperserGroup: member " << m
                  << " must hold its beams in one batch (beams_per_batch == beams_per_gpu)");
        PDD_CHECK(p.nt_in == p0.nt_in && p.config.time_sample_ms == p0.config.time_sample_ms,
                  "GpuDedisperserGroup: every member needs the same chunk (time_samples_per_chunk, time_sample_ms)");
    }
    PDD_CHECK(params.nslots >= 1 && params.nbatches_out >= 1, "GpuDedisperserGroup: nslots and nbatches_out must be >= 1");
    if (params.use_graph)
        refuse_graph();
    nslots = params.nslots;
    for (size_t m = 0; m < plans.size(); m++) {
        GpuDedisperserParams dp;
        dp.nbatches_out = params.nbatches_out;
        dp.device = params.device;
        auto dd = std::make_shared<GpuDedisperser>(plans[m], dp);
        dd->group_member = true;
        members.push_back(dd);
    }
    nbatches_out = members[0]->nbatches_out;
    nbatches_wt = members[0]->nbatches_wt;
    period = std::lcm(std::lcm(nslots, nbatches_out), nbatches_wt);
}


GpuDedisperserGroup::~GpuDedisperserGroup()
{
    if (launch_stream)
        (void) hipStreamSynchronize(launch_stream);
    for (hipGraphExec_t g : graphs)
        if (g)
            (void) hipGraphExecDestroy(g);
    if (chunk_counter)
        (void) hipFree(chunk_counter);
    for (DeviceArray &a : inputs)
        if (a.ptr)
            (void) hipFree(a.ptr);
    for (hipStream_t s : member_streams)
        (void) hipStreamDestroy(s);
    if (launch_stream)
        (void) hipStreamDestroy(launch_stream);
    if (ev_fork)
        (void) hipEventDestroy(ev_fork);
    for (auto *v : {&ev_join, &ev_input_free, &ev_input_ready, &ev_output_ready, &ev_output_released})
        for (hipEvent_t e : *v)
            (void) hipEventDestroy(e);
    members.clear();
}


void GpuDedisperserGroup::allocate()
{
    std::lock_guard<std::mutex> lk(mutex);
    PDD_CHECK(!is_allocated, "GpuDedisperserGroup::allocate() called twice");
    for (size_t m = 0; m < members.size(); m++) {
        members[m]->allocate();
        PDD_CHECK(members[m]->spill_bytes_per_batch() == 0, "GpuDedisperserGroup: member " << m
                  << " spills lag-ring rows to the host (max_gpu_clag); graph replay needs a device-resident ring");
    }
    HIP_CHECK(hipSetDevice(params.device));
    HIP_CHECK(hipStreamCreateWithFlags(&launch_stream, hipStreamNonBlocking));
    member_streams.assign(members.size(), nullptr);
    for (auto &s : member_streams)
        HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    HIP_CHECK(hipMalloc(&chunk_counter, 256));
    inputs.resize(members.size());
    for (size_t m = 0; m < members.size(); m++) {
        const GpuDedisperser &d = *members[m];
        inputs[m].nbytes = size_t(nslots) * d.beams_per_batch * d.nfreq * d.nt_in * d.elt_bytes;
        HIP_CHECK(hipMalloc(&inputs[m].ptr, inputs[m].nbytes));
        HIP_CHECK(hipMemset(inputs[m].ptr, 0, inputs[m].nbytes));
    }
    const long zero = 0;
    HIP_CHECK(hipMemcpy(chunk_counter, &zero, sizeof(zero), hipMemcpyHostToDevice));
    auto mkev = [](std::vector<hipEvent_t> &v, long n) {
        v.resize(n);
        for (auto &e : v)
            HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    };
    HIP_CHECK(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
    mkev(ev_join, long(members.size()));
    mkev(ev_input_free, nslots);
    mkev(ev_input_ready, nslots);
    mkev(ev_output_ready, nbatches_out);
    mkev(ev_output_released, nbatches_out);
    for (auto &e : ev_input_free) HIP_CHECK(hipEventRecord(e, launch_stream));
    for (auto &e : ev_output_released) HIP_CHECK(hipEventRecord(e, launch_stream));
    HIP_CHECK(hipStreamSynchronize(launch_stream));
    graphs.assign(period, nullptr);
    is_allocated = true;
}


size_t GpuDedisperserGroup::gpu_bytes() const
{
    size_t n = 0;
    for (auto &m : members)
        n += m->gpu_bytes_allocated;
    for (const DeviceArray &a : inputs)
        n += a.nbytes;
    return n;
}


void GpuDedisperserGroup::check_stopped_locked()
{
    if (stopped)
        throw pdd_error("GpuDedisperserGroup was stopped: " + stop_reason);
}

void GpuDedisperserGroup::fail_locked(const std::string &msg)
{
    stopped = true;
    stop_reason = msg;
    cv.notify_all();
    throw pdd_error(msg);
}

void GpuDedisperserGroup::stop(const std::string &reason)
{
    std::lock_guard<std::mutex> lk(mutex);
    if (!stopped) {
        stopped = true;
        stop_reason = reason.empty() ? "stop() called" : reason;
    }
    cv.notify_all();
}


void GpuDedisperserGroup::drop_graphs()
{
    if (launch_stream)
        HIP_CHECK(hipStreamSynchronize(launch_stream));
    for (hipGraphExec_t &g : graphs)
        if (g) {
            HIP_CHECK(hipGraphExecDestroy(g));
            g = nullptr;
        }
}

void GpuDedisperserGroup::set_use_graph(bool on)
{
    std::lock_guard<std::mutex> lk(mutex);
    if (on)
        refuse_graph();
    params.use_graph = on;
    drop_graphs();
}

void GpuDedisperserGroup::set_tuning(long imember, const std::string &name, long value)
{
    std::lock_guard<std::mutex> lk(mutex);
    PDD_CHECK(imember >= 0 && imember < long(members.size()), "set_tuning(): member index out of range");
    members[imember]->set_tuning(name, value);
    drop_graphs();
}


void *GpuDedisperserGroup::input_slot(size_t m, long slot) const
{
    const GpuDedisperser &d = *members[m];
    return (char *) inputs[m].ptr + size_t(slot) * d.beams_per_batch * d.nfreq * d.nt_in * d.elt_bytes;
}


void GpuDedisperserGroup::acquire_input(long ichunk, hipStream_t stream, std::vector<void *> &slots)
{
    std::lock_guard<std::mutex> lk(mutex);
    check_stopped_locked();
    PDD_CHECK(is_allocated, "GpuDedisperserGroup::acquire_input(): allocate() has not been called");
    if (ichunk != curr_input_acquire || curr_input_acquire != curr_input_release) {
        std::ostringstream ss;
        ss << "GpuDedisperserGroup::acquire_input(): expected ichunk=" << curr_input_release << ", got " << ichunk
           << " (acquire/release must alternate, in order)";
        fail_locked(ss.str());
    }
    const long slot = ichunk % nslots;
    HIP_CHECK(hipStreamWaitEvent(stream, ev_input_free[slot], 0));
    slots.resize(members.size());
    for (size_t m = 0; m < members.size(); m++)
        slots[m] = input_slot(m, slot);
    curr_input_acquire++;
}


// Every member's chain for chunk 'ichunk' as branches forked from and joined into 'st',
// then the chunk counter's advance. With 'capture' the launches go into the graph being
// captured on 'st' and the kernels read the chunk index from the counter.
void GpuDedisperserGroup::issue_chunk(long ichunk, hipStream_t st, bool capture)
{
    const long slot = ichunk % nslots, oslot = ichunk % nbatches_out, wslot = ichunk % nbatches_wt;
    HIP_CHECK(hipEventRecord(ev_fork, st));
    for (size_t m = 0; m < members.size(); m++) {
        hipStream_t ms = member_streams[m];
        HIP_CHECK(hipStreamWaitEvent(ms, ev_fork, 0));
        members[m]->launch_chain(input_slot(m, slot), ichunk, 0, oslot, wslot, ms, 0,
                                 capture ? chunk_counter : nullptr);
        HIP_CHECK(hipEventRecord(ev_join[m], ms));
        HIP_CHECK(hipStreamWaitEvent(st, ev_join[m], 0));
    }
    launch_advance_counter(chunk_counter, st);
}


void GpuDedisperserGroup::release_input_and_launch(long ichunk, hipStream_t stream)
{
    std::unique_lock<std::mutex> lk(mutex);
    check_stopped_locked();
    if (ichunk != curr_input_release || curr_input_acquire != curr_input_release + 1) {
        std::ostringstream ss;
        ss << "GpuDedisperserGroup::release_input_and_launch(): expected ichunk=" << curr_input_release << ", got " << ichunk;
        fail_locked(ss.str());
    }
    cv.wait(lk, [&] { return stopped || curr_output_release >= ichunk - nbatches_out + 1; });
    check_stopped_locked();

    const long slot = ichunk % nslots, oslot = ichunk % nbatches_out;
    HIP_CHECK(hipEventRecord(ev_input_ready[slot], stream));
    HIP_CHECK(hipStreamWaitEvent(launch_stream, ev_input_ready[slot], 0));
    HIP_CHECK(hipStreamWaitEvent(launch_stream, ev_output_released[oslot], 0));

    bool timing = false;
    for (auto &m : members)
        timing = timing || m->kernel_timing_enabled();
    try {
        if (params.use_graph && !timing) {
            hipGraphExec_t &exec = graphs[ichunk % period];
            if (!exec) {
                // Captured on its own stream so that nothing already queued on launch_stream
                // becomes part of the graph.
                hipStream_t cs;
                HIP_CHECK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
                hipGraph_t g = nullptr;
                HIP_CHECK(hipStreamBeginCapture(cs, hipStreamCaptureModeRelaxed));
                try {
                    issue_chunk(ichunk, cs, true);
                }
                catch (...) {
                    (void) hipStreamEndCapture(cs, &g);
                    if (g)
                        (void) hipGraphDestroy(g);
                    (void) hipStreamDestroy(cs);
                    throw;
                }
                HIP_CHECK(hipStreamEndCapture(cs, &g));
                size_t nn = 0;
                HIP_CHECK(hipGraphGetNodes(g, nullptr, &nn));
                graph_nnodes = long(nn);
                HIP_CHECK(hipGraphInstantiate(&exec, g, nullptr, nullptr, 0));
                HIP_CHECK(hipGraphDestroy(g));
                HIP_CHECK(hipStreamDestroy(cs));
            }
            HIP_CHECK(hipGraphLaunch(exec, launch_stream));
        }
        else
            issue_chunk(ichunk, launch_stream, false);
    }
    catch (std::exception &e) {
        fail_locked(e.what());
    }
    HIP_CHECK(hipEventRecord(ev_input_free[slot], launch_stream));
    HIP_CHECK(hipEventRecord(ev_output_ready[oslot], launch_stream));
    curr_input_release++;
    cv.notify_all();
}


void GpuDedisperserGroup::acquire_output(long ichunk, hipStream_t stream, bool sync)
{
    std::unique_lock<std::mutex> lk(mutex);
    check_stopped_locked();
    if (ichunk != curr_output_acquire || curr_output_acquire != curr_output_release) {
        std::ostringstream ss;
        ss << "GpuDedisperserGroup::acquire_output(): expected ichunk=" << curr_output_acquire << ", got " << ichunk;
        fail_locked(ss.str());
    }
    cv.wait(lk, [&] { return stopped || curr_input_release > ichunk; });
    check_stopped_locked();
    const long oslot = ichunk % nbatches_out;
    HIP_CHECK(hipStreamWaitEvent(stream, ev_output_ready[oslot], 0));
    if (sync)
        HIP_CHECK(hipEventSynchronize(ev_output_ready[oslot]));
    curr_output_acquire++;
}


void GpuDedisperserGroup::release_output(long ichunk, hipStream_t stream)
{
    std::lock_guard<std::mutex> lk(mutex);
    check_stopped_locked();
    if (ichunk != curr_output_release || curr_output_release >= curr_output_acquire) {
        std::ostringstream ss;
        ss << "GpuDedisperserGroup::release_output(): expected ichunk=" << curr_output_release << ", got " << ichunk;
        fail_locked(ss.str());
    }
    HIP_CHECK(hipEventRecord(ev_output_released[ichunk % nbatches_out], stream));
    curr_output_release++;
    cv.notify_all();
}

}  // namespace pirate_amd
