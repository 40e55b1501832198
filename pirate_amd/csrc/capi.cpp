// extern "C" entry points of libpirate_amd.so (declared in include/pirate_amd.h).

#include "../../include/pirate_amd.h"
#include "host/common.hpp"
#include "host/config.hpp"
#include "host/plan.hpp"
#include "host/pf_variance.hpp"
#include "gpu/dedisperser.hpp"
#include "gpu/dedisperser_group.hpp"
#include "gpu/kernels.hpp"
#include "gpu/scorer.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>

using namespace pirate_amd;

struct pdd_config { DedispersionConfig c; };
struct pdd_plan { std::shared_ptr<DedispersionPlan> p; };
struct pdd_dedisperser { std::shared_ptr<GpuDedisperser> d; };
struct pdd_group { std::unique_ptr<GpuDedisperserGroup> g; };

static thread_local std::string g_last_error;

static int fail(const std::string &msg)
{
    g_last_error = msg;
    return -1;
}

#define PDD_TRY(body)                                                     \
    try {                                                                 \
        body;                                                             \
        return 0;                                                         \
    }                                                                     \
    catch (std::exception &e) { return fail(e.what()); }                  \
    catch (...) { return fail("unknown C++ exception"); }

// Dedisperser entry points: a throw stops the dedisperser (first error latches).
#define PDD_TRY_DD(dd, body)                                              \
    try {                                                                 \
        if (!(dd) || !(dd)->d) return fail("null dedisperser handle");    \
        body;                                                             \
        return 0;                                                         \
    }                                                                     \
    catch (std::exception &e) { (dd)->d->stop(e.what()); return fail(e.what()); } \
    catch (...) { (dd)->d->stop("unknown C++ exception"); return fail("unknown C++ exception"); }

static void hip_check(hipError_t e, const char *what)
{
    if (e != hipSuccess)
        throw pdd_error(std::string("HIP error: ") + hipGetErrorString(e) + " in " + what);
}

static int copy_string(const std::string &s, char *buf, size_t buflen, size_t *needed)
{
    if (needed)
        *needed = s.size() + 1;
    if (buf && buflen > 0) {
        size_t n = std::min(buflen - 1, s.size());
        std::memcpy(buf, s.data(), n);
        buf[n] = 0;
    }
    return 0;
}

extern "C" {

const char *pdd_last_error(void) { return g_last_error.c_str(); }
int pdd_abi_version(void) { return 4; }

int pdd_config_from_yaml_file(const char *filename, pdd_config **out)
{
    PDD_TRY({
        PDD_CHECK(filename && out, "null argument");
        *out = new pdd_config{DedispersionConfig::from_yaml_file(filename)};
    });
}

int pdd_config_from_yaml_string(const char *text, pdd_config **out)
{
    PDD_TRY({
        PDD_CHECK(text && out, "null argument");
        *out = new pdd_config{DedispersionConfig::from_yaml_string(text)};
    });
}

void pdd_config_destroy(pdd_config *c) { delete c; }

int pdd_config_to_yaml(const pdd_config *c, int verbose, char *buf, size_t buflen, size_t *needed)
{
    PDD_TRY({
        PDD_CHECK(c, "null config");
        copy_string(c->c.to_yaml_string(verbose != 0), buf, buflen, needed);
    });
}

int pdd_config_make_channel_map(const pdd_config *c, double *out, size_t n)
{
    PDD_TRY({
        PDD_CHECK(c && out, "null argument");
        std::vector<double> m = c->c.make_channel_map();
        PDD_CHECK(n >= m.size(), "pdd_config_make_channel_map: output too small (need " << m.size() << ")");
        std::copy(m.begin(), m.end(), out);
    });
}

int pdd_config_eval(const pdd_config *c, const char *what, double x, double *out)
{
    PDD_TRY({
        PDD_CHECK(c && what && out, "null argument");
        std::string w(what);
        if (w == "frequency_to_index") *out = c->c.frequency_to_index(x);
        else if (w == "index_to_frequency") *out = c->c.index_to_frequency(x);
        else if (w == "delay_to_frequency") *out = c->c.delay_to_frequency(x);
        else if (w == "frequency_to_delay") *out = c->c.frequency_to_delay(x);
        else if (w == "dm_per_unit_delay") *out = c->c.dm_per_unit_delay();
        else if (w == "max_dm_of_all_trees") *out = c->c.max_dm_of_all_trees();
        else throw pdd_error("pdd_config_eval: unknown function '" + w + "'");
    });
}

int pdd_config_test(const pdd_config *c)
{
    PDD_TRY({ PDD_CHECK(c, "null config"); c->c.test(); });
}

int pdd_plan_create(const pdd_config *c, int gpu_runnable, pdd_plan **out)
{
    PDD_TRY({
        PDD_CHECK(c && out, "null argument");
        *out = new pdd_plan{std::make_shared<DedispersionPlan>(c->c, gpu_runnable != 0)};
    });
}

void pdd_plan_destroy(pdd_plan *p) { delete p; }

int pdd_plan_to_yaml(const pdd_plan *p, int verbose, char *buf, size_t buflen, size_t *needed)
{
    PDD_TRY({
        PDD_CHECK(p, "null plan");
        copy_string(p->p->to_yaml_string(verbose != 0), buf, buflen, needed);
    });
}

int pdd_plan_ntrees(const pdd_plan *p, int64_t *n)
{
    PDD_TRY({ PDD_CHECK(p && n, "null argument"); *n = p->p->ntrees; });
}

int pdd_plan_tree_info(const pdd_plan *p, int64_t itree, pdd_tree_info *o)
{
    PDD_TRY({
        PDD_CHECK(p && o, "null argument");
        PDD_CHECK(itree >= 0 && itree < p->p->ntrees, "itree out of range");
        const DedispersionTree &t = p->p->trees[itree];
        o->primary_tree_index = t.primary_tree_index;
        o->early_trigger_level = t.early_trigger_level;
        o->amb_rank = t.amb_rank;
        o->dd_rank = t.dd_rank;
        o->nt_ds = t.nt_ds;
        o->max_width = t.pf.max_width;
        o->dm_downsampling = t.pf.dm_downsampling;
        o->time_downsampling = t.pf.time_downsampling;
        o->wt_dm_downsampling = t.pf.wt_dm_downsampling;
        o->wt_time_downsampling = t.pf.wt_time_downsampling;
        o->Dcore = t.Dcore;
        o->nprofiles = t.nprofiles;
        o->ndm_out = t.ndm_out;
        o->ndm_wt = t.ndm_wt;
        o->nt_out = t.nt_out;
        o->nt_wt = t.nt_wt;
        o->nsubbands_N = t.frequency_subbands.N;
        o->nmultiplets_M = t.frequency_subbands.M;
        o->pf_rank = t.frequency_subbands.pf_rank;
        o->dm_min = t.dm_min;
        o->dm_max = t.dm_max;
        o->trigger_frequency = t.trigger_frequency;
    });
}

int pdd_plan_decode_argmax(const pdd_plan *p, uint32_t token, int64_t itree, int64_t idm, int64_t it, int64_t *out5)
{
    PDD_TRY({
        PDD_CHECK(p && out5, "null argument");
        long v[5];
        p->p->decode_argmax(token, itree, idm, it, v[0], v[1], v[2], v[3], v[4]);
        for (int i = 0; i < 5; i++) out5[i] = v[i];
    });
}

int pdd_plan_decode_argmax_batch(const pdd_plan *p, int64_t itree, const uint32_t *tokens, int64_t n, int64_t *out)
{
    PDD_TRY({
        PDD_CHECK(p && (n == 0 || (tokens && out)), "null argument");
        PDD_CHECK(itree >= 0 && itree < p->p->ntrees, "pdd_plan_decode_argmax_batch: itree out of range");
        const long nd = p->p->trees[itree].ndm_out;
        const long nt = p->p->trees[itree].nt_out;
        PDD_CHECK(n >= 0 && n % (nd * nt) == 0, "pdd_plan_decode_argmax_batch: n must be a multiple of ndm_out * nt_out = " << nd * nt);
        for (int64_t i = 0; i < n; i++) {
            long v[5];
            p->p->decode_argmax(tokens[i], itree, (i / nt) % nd, i % nt, v[0], v[1], v[2], v[3], v[4]);
            std::copy(v, v + 5, out + 5 * i);
        }
    });
}

int pdd_plan_decode_argmax2(const pdd_plan *p, int64_t itree, int64_t fmin, int64_t fmax, int64_t tlo, int64_t thi,
                            int64_t pp, double *out5)
{
    PDD_TRY({
        PDD_CHECK(p && out5, "null argument");
        p->p->decode_argmax2(itree, fmin, fmax, tlo, thi, pp, out5[0], out5[1], out5[2], out5[3], out5[4]);
    });
}

int pdd_plan_steady_state_it0(const pdd_plan *p, int64_t itree, int64_t *out, size_t n)
{
    PDD_TRY({
        PDD_CHECK(p && out, "null argument");
        std::vector<long> v = p->p->compute_steady_state_it0(itree);
        PDD_CHECK(n >= v.size(), "pdd_plan_steady_state_it0: output too small");
        std::copy(v.begin(), v.end(), out);
    });
}

int pdd_plan_pf_variances(const pdd_plan *p, const double *freq_variances, int64_t nfreq, int64_t itree,
                          double *tree_variance, int64_t tv_nelts, double *per_tf, int64_t ptf_nelts)
{
    PDD_TRY({
        PDD_CHECK(p && freq_variances && tree_variance, "null argument");
        PDD_CHECK(itree >= 0 && itree < p->p->ntrees, "pdd_plan_pf_variances: itree out of range");
        PfAvarApproximation avar(*p->p, freq_variances, nfreq, itree);
        const std::vector<double> &tv = avar.tree_variance[itree];
        const std::vector<double> &pt = avar.per_tf[itree];
        PDD_CHECK(tv_nelts == int64_t(tv.size()), "pdd_plan_pf_variances: tree_variance needs " << tv.size() << " elements (N, ndm_wt, nprofiles), got " << tv_nelts);
        std::copy(tv.begin(), tv.end(), tree_variance);
        if (per_tf) {
            PDD_CHECK(ptf_nelts == int64_t(pt.size()), "pdd_plan_pf_variances: per_tf needs " << pt.size() << " elements (2^pf_rank, ndm_wt, nprofiles), got " << ptf_nelts);
            std::copy(pt.begin(), pt.end(), per_tf);
        }
    });
}

int pdd_dedisperser_create(const pdd_plan *p, const pdd_dedisperser_params *prm, pdd_dedisperser **out)
{
    PDD_TRY({
        PDD_CHECK(p && out, "null argument");
        GpuDedisperserParams gp;
        if (prm) {
            gp.nbatches_out = prm->nbatches_out;
            gp.nbatches_wt = prm->nbatches_wt;
            gp.num_consumers = prm->num_consumers ? prm->num_consumers : 1;
            gp.device = prm->device;
            gp.synchronous = prm->synchronous != 0;
            gp.initial_chunk = prm->initial_chunk;
            gp.enable_variances = prm->enable_variances != 0;
        }
        *out = new pdd_dedisperser{std::make_shared<GpuDedisperser>(p->p, gp)};
    });
}

void pdd_dedisperser_destroy(pdd_dedisperser *dd) { delete dd; }

int pdd_dedisperser_allocate(pdd_dedisperser *dd) { PDD_TRY_DD(dd, dd->d->allocate()); }

int pdd_dedisperser_gpu_bytes(const pdd_dedisperser *dd, int64_t *nbytes)
{
    PDD_TRY({ PDD_CHECK(dd && nbytes, "null argument"); *nbytes = int64_t(dd->d->gpu_bytes_allocated); });
}

int pdd_dedisperser_host_bytes(const pdd_dedisperser *dd, int64_t *nbytes)
{
    PDD_TRY({ PDD_CHECK(dd && nbytes, "null argument"); *nbytes = int64_t(dd->d->host_bytes_allocated); });
}

int pdd_output_ipc_handle(pdd_dedisperser *dd, void *handle, int64_t *nbytes)
{
    PDD_TRY({
        PDD_CHECK(dd && handle && nbytes, "null argument");
        size_t n = 0;
        dd->d->output_ipc_handle(handle, &n);
        *nbytes = int64_t(n);
    });
}

int pdd_output_array_offset(const pdd_dedisperser *dd, int64_t itree, int which, int64_t *byte_offset)
{
    PDD_TRY({
        PDD_CHECK(dd && byte_offset, "null argument");
        *byte_offset = int64_t(dd->d->output_array_offset(itree, which));
    });
}

int pdd_ipc_open(const void *handle, int device, void **base)
{
    PDD_TRY({
        PDD_CHECK(handle && base, "null argument");
        hipIpcMemHandle_t h;
        std::memcpy(&h, handle, sizeof(h));
        hip_check(hipSetDevice(device), "hipSetDevice");
        hip_check(hipIpcOpenMemHandle(base, h, hipIpcMemLazyEnablePeerAccess), "hipIpcOpenMemHandle");
    });
}

int pdd_ipc_close(void *base, int device)
{
    PDD_TRY({
        PDD_CHECK(base, "null argument");
        hip_check(hipSetDevice(device), "hipSetDevice");
        // work queued on this device may still read the mapping
        hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        hip_check(hipIpcCloseMemHandle(base), "hipIpcCloseMemHandle");
    });
}

int pdd_acquire_input(pdd_dedisperser *dd, int64_t seq_id, void *stream, void **slot)
{
    PDD_TRY_DD(dd, {
        PDD_CHECK(slot, "null slot pointer");
        *slot = dd->d->acquire_input(seq_id, (hipStream_t) stream);
    });
}

int pdd_release_input_and_launch(pdd_dedisperser *dd, int64_t seq_id, void *stream)
{
    PDD_TRY_DD(dd, dd->d->release_input_and_launch(seq_id, (hipStream_t) stream));
}

int pdd_acquire_output(pdd_dedisperser *dd, int64_t consumer, int64_t seq_id, void *stream, int sync,
                       void **out_max, uint32_t **out_argmax, int64_t ntrees)
{
    PDD_TRY_DD(dd, {
        PDD_CHECK(out_max && out_argmax, "null output arrays");
        PDD_CHECK(ntrees == dd->d->ntrees, "pdd_acquire_output: ntrees mismatch");
        std::vector<void *> om;
        std::vector<unsigned int *> oa;
        dd->d->acquire_output(consumer, seq_id, (hipStream_t) stream, sync != 0, om, oa);
        for (int64_t i = 0; i < ntrees; i++) {
            out_max[i] = om[i];
            out_argmax[i] = oa[i];
        }
    });
}

int pdd_release_output(pdd_dedisperser *dd, int64_t consumer, int64_t seq_id, void *stream)
{
    PDD_TRY_DD(dd, dd->d->release_output(consumer, seq_id, (hipStream_t) stream));
}

int pdd_output_variances(pdd_dedisperser *dd, int64_t seq_id, double **out_var, int64_t ntrees)
{
    PDD_TRY_DD(dd, {
        PDD_CHECK(out_var, "null out_var");
        PDD_CHECK(ntrees == dd->d->ntrees, "pdd_output_variances: ntrees mismatch");
        std::vector<double *> ov;
        dd->d->output_variances(seq_id, ov);
        std::copy(ov.begin(), ov.end(), out_var);
    });
}

int pdd_fill_all_weights(pdd_dedisperser *dd, int64_t itree, const float *w, int64_t nelts)
{
    PDD_TRY_DD(dd, {
        PDD_CHECK(w, "null weights");
        dd->d->fill_all_weights(itree, w, nelts);
    });
}

int pdd_fill_analytic_weights(pdd_dedisperser *dd, const double *freq_variances, int64_t nfreq)
{
    PDD_TRY_DD(dd, {
        PDD_CHECK(freq_variances, "null freq_variances");
        dd->d->fill_analytic_weights(freq_variances, nfreq);
    });
}

int pdd_stop(pdd_dedisperser *dd, const char *reason)
{
    PDD_TRY({
        PDD_CHECK(dd && dd->d, "null dedisperser handle");
        dd->d->stop(reason ? reason : "");
    });
}

int pdd_kernel_timing(pdd_dedisperser *dd, int enable)
{
    PDD_TRY_DD(dd, dd->d->set_kernel_timing(enable != 0));
}

int pdd_kernel_timing_report(pdd_dedisperser *dd, char *buf, size_t buflen, size_t *needed)
{
    PDD_TRY_DD(dd, copy_string(dd->d->kernel_timing_report(), buf, buflen, needed));
}

int pdd_debug_copy(pdd_dedisperser *dd, const char *name, void *host, size_t maxbytes, size_t *nbytes)
{
    PDD_TRY({
        PDD_CHECK(dd && dd->d && name, "null argument");
        size_t n = dd->d->debug_copy(name, host, maxbytes);
        if (nbytes)
            *nbytes = n;
    });
}

int pdd_dequantize(const char *out_dtype, int64_t nbeams, int64_t nfreq, int64_t ntime, const void *data,
                   const void *scales_offsets, void *out, void *stream)
{
    PDD_TRY({
        PDD_CHECK(out_dtype && data && scales_offsets && out, "pdd_dequantize: null argument");
        PDD_CHECK(nbeams > 0 && nfreq > 0 && ntime > 0, "pdd_dequantize: shape must be positive");
        PDD_CHECK(ntime % 256 == 0, "pdd_dequantize: ntime must be a multiple of 256 (got " << ntime << ")");
        const Dtype dt = dtype_from_str(out_dtype);
        PDD_CHECK((reinterpret_cast<uintptr_t>(out) % 16) == 0 && (reinterpret_cast<uintptr_t>(data) % 4) == 0
                      && (reinterpret_cast<uintptr_t>(scales_offsets) % 4) == 0,
                  "pdd_dequantize: misaligned pointer (out: 16 bytes, data and scales_offsets: 4 bytes)");
        launch_dequant(dt, data, scales_offsets, out, nbeams * nfreq * ntime, (hipStream_t) stream);
        hipError_t e = hipGetLastError();
        PDD_CHECK(e == hipSuccess, "pdd_dequantize: launch failed: " << hipGetErrorString(e));
    });
}


struct pdd_ringbuf_copy {
    long total_beams = 0, beams_per_batch = 0, seg_bytes = 0, noct = 0;
    long max_seg = -1;             // largest segment index any octuple addresses
    uint32_t *d_oct = nullptr;     // device copy of the octuples
};

int pdd_ringbuf_copy_create(int64_t total_beams, int64_t beams_per_batch, int64_t nbytes_per_segment,
                            const uint32_t *octuples, int64_t n_uint32, pdd_ringbuf_copy **out)
{
    PDD_TRY({
        PDD_CHECK(out, "pdd_ringbuf_copy_create: null output handle");
        PDD_CHECK(total_beams > 0 && beams_per_batch > 0 && total_beams % beams_per_batch == 0,
                  "pdd_ringbuf_copy_create: total_beams must be a positive multiple of beams_per_batch");
        PDD_CHECK(nbytes_per_segment == 128 || nbytes_per_segment == 256,
                  "pdd_ringbuf_copy_create: nbytes_per_segment must be 128 or 256 (got " << nbytes_per_segment << ")");
        PDD_CHECK(n_uint32 >= 0 && n_uint32 % 8 == 0, "pdd_ringbuf_copy_create: octuples must hold 8*N uint32");
        PDD_CHECK(n_uint32 == 0 || octuples, "pdd_ringbuf_copy_create: null octuples");
        std::unique_ptr<pdd_ringbuf_copy> k(new pdd_ringbuf_copy);
        k->total_beams = total_beams;
        k->beams_per_batch = beams_per_batch;
        k->seg_bytes = nbytes_per_segment;
        k->noct = n_uint32 / 8;
        for (long i = 0; i < k->noct; i++)
            for (int side = 0; side < 2; side++) {
                const uint32_t *q = octuples + 8 * i + 4 * side;
                PDD_CHECK(q[2] > 0, "pdd_ringbuf_copy_create: octuple " << i << " has a zero frame count");
                k->max_seg = std::max(k->max_seg, long(q[0]) + long(q[2] - 1) * long(q[3]));
            }
        if (k->noct > 0) {
            hip_check(hipMalloc(&k->d_oct, n_uint32 * sizeof(uint32_t)), "hipMalloc(octuples)");
            const hipError_t e = hipMemcpy(k->d_oct, octuples, n_uint32 * sizeof(uint32_t), hipMemcpyHostToDevice);
            if (e != hipSuccess)
                (void) hipFree(k->d_oct);
            hip_check(e, "hipMemcpy(octuples)");
        }
        *out = k.release();
    });
}

void pdd_ringbuf_copy_destroy(pdd_ringbuf_copy *k)
{
    if (k) {
        if (k->d_oct)
            (void) hipFree(k->d_oct);
        delete k;
    }
}

int pdd_ringbuf_copy_launch(pdd_ringbuf_copy *k, void *ringbuf, int64_t ringbuf_nbytes, int64_t ichunk,
                            int64_t ibatch, void *stream)
{
    PDD_TRY({
        PDD_CHECK(k && ringbuf, "pdd_ringbuf_copy_launch: null argument");
        PDD_CHECK((reinterpret_cast<uintptr_t>(ringbuf) % 16) == 0, "pdd_ringbuf_copy_launch: ringbuf must be 16-byte aligned");
        PDD_CHECK(ichunk >= 0 && ibatch >= 0 && ibatch * k->beams_per_batch < k->total_beams,
                  "pdd_ringbuf_copy_launch: bad (ichunk, ibatch) = (" << ichunk << ", " << ibatch << ")");
        PDD_CHECK((k->max_seg + 1) * k->seg_bytes <= ringbuf_nbytes,
                  "pdd_ringbuf_copy_launch: octuples address segment " << k->max_seg << ", beyond the "
                  << ringbuf_nbytes << "-byte ring buffer");
        const unsigned long long iframe = (unsigned long long) (ichunk * k->total_beams + ibatch * k->beams_per_batch);
        launch_ringbuf_copy(ringbuf, k->d_oct, k->noct, k->seg_bytes, int(k->beams_per_batch), iframe, (hipStream_t) stream);
        hipError_t e = hipGetLastError();
        PDD_CHECK(e == hipSuccess, "pdd_ringbuf_copy_launch: launch failed: " << hipGetErrorString(e));
    });
}

int pdd_dedisperser_spill_bytes(const pdd_dedisperser *dd, int64_t *nbytes)
{
    PDD_TRY({
        PDD_CHECK(dd && dd->d && nbytes, "null argument");
        PDD_CHECK(dd->d->is_allocated, "pdd_dedisperser_spill_bytes: allocate() has not been called");
        *nbytes = int64_t(dd->d->spill_bytes_per_batch());
    });
}

int pdd_dedisperser_set_tuning(pdd_dedisperser *dd, const char *name, int64_t value)
{
    PDD_TRY({
        PDD_CHECK(dd && dd->d && name, "null argument");
        dd->d->set_tuning(name, value);
    });
}

int pdd_dedisperser_get_tuning(const pdd_dedisperser *dd, const char *name, int64_t *value)
{
    PDD_TRY({
        PDD_CHECK(dd && dd->d && name && value, "null argument");
        *value = dd->d->get_tuning(name);
    });
}


// ---- GpuDedisperserGroup ----

#define PDD_TRY_G(gp, body)                                               \
    try {                                                                 \
        if (!(gp) || !(gp)->g) return fail("null group handle");          \
        body;                                                             \
        return 0;                                                         \
    }                                                                     \
    catch (std::exception &e) { (gp)->g->stop(e.what()); return fail(e.what()); } \
    catch (...) { (gp)->g->stop("unknown C++ exception"); return fail("unknown C++ exception"); }

int pdd_group_create(const pdd_plan *const *plans, int64_t nplans, const pdd_group_params *prm, pdd_group **out)
{
    PDD_TRY({
        PDD_CHECK(plans && nplans > 0 && out, "null argument");
        std::vector<std::shared_ptr<DedispersionPlan>> v;
        for (int64_t i = 0; i < nplans; i++) {
            PDD_CHECK(plans[i], "pdd_group_create: null plan " << i);
            v.push_back(plans[i]->p);
        }
        GpuDedisperserGroupParams gp;
        if (prm) {
            gp.nslots = prm->nslots ? prm->nslots : 2;
            gp.nbatches_out = prm->nbatches_out ? prm->nbatches_out : 2;
            gp.device = prm->device;
            gp.use_graph = prm->use_graph != 0;
        }
        *out = new pdd_group{std::make_unique<GpuDedisperserGroup>(v, gp)};
    });
}

void pdd_group_destroy(pdd_group *g) { delete g; }

int pdd_group_allocate(pdd_group *g) { PDD_TRY_G(g, g->g->allocate()); }

int pdd_group_member(pdd_group *g, int64_t imember, pdd_dedisperser **dd)
{
    PDD_TRY({
        PDD_CHECK(g && g->g && dd, "null argument");
        PDD_CHECK(imember >= 0 && imember < int64_t(g->g->members.size()), "pdd_group_member: index out of range");
        *dd = new pdd_dedisperser{g->g->members[imember]};
    });
}

int pdd_group_gpu_bytes(const pdd_group *g, int64_t *nbytes)
{
    PDD_TRY({
        PDD_CHECK(g && g->g && nbytes, "null argument");
        *nbytes = int64_t(g->g->gpu_bytes());
    });
}

int pdd_group_acquire_input(pdd_group *g, int64_t ichunk, void *stream, void **slots, int64_t nmembers)
{
    PDD_TRY_G(g, {
        PDD_CHECK(slots && nmembers == int64_t(g->g->members.size()), "pdd_group_acquire_input: need one slot pointer per member");
        std::vector<void *> v;
        g->g->acquire_input(ichunk, (hipStream_t) stream, v);
        for (int64_t m = 0; m < nmembers; m++)
            slots[m] = v[m];
    });
}

int pdd_group_release_input_and_launch(pdd_group *g, int64_t ichunk, void *stream)
{
    PDD_TRY_G(g, g->g->release_input_and_launch(ichunk, (hipStream_t) stream));
}

int pdd_group_acquire_output(pdd_group *g, int64_t ichunk, void *stream, int sync, void **out_max,
                             void **out_argmax, int64_t n)
{
    PDD_TRY_G(g, {
        int64_t tot = 0;
        for (auto &m : g->g->members)
            tot += m->ntrees;
        PDD_CHECK(out_max && out_argmax && n == tot, "pdd_group_acquire_output: need " << tot
                  << " (member, tree) pointers, got " << n);
        g->g->acquire_output(ichunk, (hipStream_t) stream, sync != 0);
        const long oslot = ichunk % g->g->nbatches_out;
        int64_t k = 0;
        for (auto &m : g->g->members)
            for (long it = 0; it < m->ntrees; it++, k++) {
                out_max[k] = m->output_slot(it, oslot, 0);
                out_argmax[k] = m->output_slot(it, oslot, 1);
            }
    });
}

int pdd_group_release_output(pdd_group *g, int64_t ichunk, void *stream)
{
    PDD_TRY_G(g, g->g->release_output(ichunk, (hipStream_t) stream));
}

int pdd_group_set_use_graph(pdd_group *g, int use_graph)
{
    PDD_TRY_G(g, g->g->set_use_graph(use_graph != 0));
}

int pdd_group_set_tuning(pdd_group *g, int64_t imember, const char *name, int64_t value)
{
    PDD_TRY({
        PDD_CHECK(g && g->g && name, "null argument");
        g->g->set_tuning(imember, name, value);
    });
}

int pdd_group_graph_nodes(const pdd_group *g, int64_t *n)
{
    PDD_TRY({
        PDD_CHECK(g && g->g && n, "null argument");
        *n = g->g->graph_nodes();
    });
}

int pdd_group_stop(pdd_group *g, const char *reason)
{
    PDD_TRY({
        PDD_CHECK(g && g->g, "null group handle");
        g->g->stop(reason ? reason : "");
    });
}


// ---- learned peak scorer ----

static scorer::TileSource tile_source(const pdd_tile_source *s, const char *who)
{
    PDD_CHECK(s && s->base, who << ": null tile source");
    PDD_CHECK(s->dtype >= 0 && s->dtype <= 2, who << ": tile dtype must be 0 (float32), 1 (float16) or 2 (bfloat16)");
    PDD_CHECK(s->nbeams >= 0 && s->nrows >= 0 && s->beam_stride >= 0 && s->row_stride >= 0 && s->t_lo <= s->t_hi,
              who << ": bad tile source extents");
    scorer::TileSource t;
    t.base = s->base;
    t.dtype = s->dtype;
    t.nbeams = s->nbeams;
    t.beam_stride = s->beam_stride;
    t.nrows = s->nrows;
    t.row_stride = s->row_stride;
    t.t_lo = s->t_lo;
    t.t_hi = s->t_hi;
    t.origins = s->origins;
    return t;
}

static void launch_check(const char *who)
{
    hipError_t e = hipGetLastError();
    PDD_CHECK(e == hipSuccess, who << ": launch failed: " << hipGetErrorString(e));
}

static constexpr int kScorerMaxParts = scorer::kScratchParts;   // same partial count as the no-allocation variant

int64_t pdd_scorer_nparams(void) { return scorer::kNParams; }

int pdd_scorer_forward(const float *params, const pdd_tile_source *src, int64_t ntiles, float *logits, void *stream)
{
    PDD_TRY({
        PDD_CHECK(params && (logits || ntiles == 0), "pdd_scorer_forward: null argument");
        PDD_CHECK(ntiles >= 0, "pdd_scorer_forward: ntiles < 0");
        const scorer::TileSource t = tile_source(src, "pdd_scorer_forward");
        scorer::launch_forward(params, t, ntiles, logits, (hipStream_t) stream);
        launch_check("pdd_scorer_forward");
    });
}

int pdd_scorer_grad(const float *params, const pdd_tile_source *src, const float *labels, int64_t ntiles,
                    float *grad, float *logits, float *loss_terms, void *stream)
{
    PDD_TRY({
        PDD_CHECK(params && grad, "pdd_scorer_grad: null argument");
        PDD_CHECK(ntiles >= 0, "pdd_scorer_grad: ntiles < 0");
        PDD_CHECK(ntiles == 0 || (labels && logits && loss_terms), "pdd_scorer_grad: null argument");
        const scorer::TileSource t = tile_source(src, "pdd_scorer_grad");
        hipStream_t st = (hipStream_t) stream;
        const int nparts = scorer::train_nparts(ntiles, kScorerMaxParts);
        float *partials = nullptr;
        if (nparts > 0)
            hip_check(hipMallocAsync((void **) &partials, size_t(nparts) * scorer::kNParams * sizeof(float), st),
                      "pdd_scorer_grad workspace");
        scorer::launch_train(params, t, labels, ntiles, ntiles ? float(1.0 / double(ntiles)) : 0.0f, logits,
                             loss_terms, partials, nparts, st);
        scorer::launch_reduce(partials, nparts, grad, st);
        launch_check("pdd_scorer_grad");
        if (partials)
            hip_check(hipFreeAsync(partials, st), "pdd_scorer_grad workspace");
    });
}

int pdd_scorer_grad_noalloc(const float *params, const pdd_tile_source *src, const float *labels, int64_t ntiles,
                            float *grad, float *logits, float *loss_terms, void *stream)
{
    PDD_TRY({
        PDD_CHECK(params && grad, "pdd_scorer_grad_noalloc: null argument");
        PDD_CHECK(ntiles >= 0, "pdd_scorer_grad_noalloc: ntiles < 0");
        PDD_CHECK(ntiles == 0 || (labels && logits && loss_terms), "pdd_scorer_grad_noalloc: null argument");
        const scorer::TileSource t = tile_source(src, "pdd_scorer_grad_noalloc");
        scorer::grad_noalloc(params, t, labels, ntiles, ntiles ? float(1.0 / double(ntiles)) : 0.0f, grad, logits,
                             loss_terms, (hipStream_t) stream);
        launch_check("pdd_scorer_grad_noalloc");
    });
}

int pdd_scorer_adam(float *params, float *m, float *v, const float *grad, int64_t step, float lr, float beta1,
                    float beta2, float eps, void *stream)
{
    PDD_TRY({
        PDD_CHECK(params && m && v && grad, "pdd_scorer_adam: null argument");
        PDD_CHECK(step >= 1, "pdd_scorer_adam: step counts from 1");
        PDD_CHECK(beta1 >= 0 && beta1 < 1 && beta2 >= 0 && beta2 < 1 && eps > 0, "pdd_scorer_adam: bad hyperparameters");
        const float bc1 = float(1.0 - std::pow(double(beta1), double(step)));
        const float bc2 = float(1.0 - std::pow(double(beta2), double(step)));
        scorer::launch_adam(params, m, v, grad, scorer::kNParams, lr, beta1, beta2, eps, bc1, bc2, (hipStream_t) stream);
        launch_check("pdd_scorer_adam");
    });
}

int pdd_dedisperser_set_scorer(pdd_dedisperser *dd, int64_t itree, const float *params, int64_t nparams)
{
    PDD_TRY_DD(dd, dd->d->set_scorer(itree, params, nparams));
}

int pdd_output_scores(pdd_dedisperser *dd, int64_t seq_id, float **scores, int64_t *ntiles_dm, int64_t *ntiles_t)
{
    PDD_TRY_DD(dd, {
        PDD_CHECK(scores && ntiles_dm && ntiles_t, "null argument");
        long a = 0;
        long b = 0;
        dd->d->output_scores(seq_id, *scores, a, b);
        *ntiles_dm = a;
        *ntiles_t = b;
    });
}

int pdd_dedisperser_tree_rows(pdd_dedisperser *dd, int64_t itree, pdd_tile_source *src)
{
    PDD_TRY({
        PDD_CHECK(dd && dd->d && src, "null argument");
        void *base = nullptr;
        long nrows = 0;
        long bstride = 0;
        long rstride = 0;
        long t_lo = 0;
        long t_hi = 0;
        dd->d->tree_rows(itree, base, nrows, bstride, rstride, t_lo, t_hi);
        src->base = base;
        src->dtype = dd->d->dtype == Dtype::float32 ? 0 : (dd->d->dtype == Dtype::float16 ? 1 : 2);
        src->nbeams = dd->d->total_beams;
        src->beam_stride = bstride;
        src->nrows = nrows;
        src->row_stride = rstride;
        src->t_lo = t_lo;
        src->t_hi = t_hi;
        src->origins = nullptr;
    });
}

}  // extern "C"
