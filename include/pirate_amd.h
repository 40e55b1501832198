/*
 * pirate_amd.h — C-ABI of the MI355X (gfx950) dedispersion chain.
 *
 * Plain C: opaque handles, plain pointers and sizes; HIP streams are passed as void*
 * (a hipStream_t, or NULL for the null stream). Every function returns 0 on success and a
 * nonzero status on failure, with the message available from pdd_last_error() (thread-local).
 * A dedisperser that has failed (or was stopped) latches the first error: every later call
 * on it fails with that message (reference "stoppable" semantics,
 * src_lib/GpuDedisperser.cpp:361-372, notes/stoppable_class.md).
 *
 * Each entry point names the reference interface it replaces. The reference exposes these
 * through pybind11 (src_pybind11/pirate_pybind11.cpp); this library exposes them as C, and
 * the Python package pirate_amd binds them with ctypes.
 */

#ifndef PIRATE_AMD_H
#define PIRATE_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pdd_config pdd_config;
typedef struct pdd_plan pdd_plan;
typedef struct pdd_dedisperser pdd_dedisperser;
typedef struct pdd_ringbuf_copy pdd_ringbuf_copy;

/* Last error message of the calling thread ("" if none). */
const char *pdd_last_error(void);

/* ABI version of this library. */
int pdd_abi_version(void);

/* ---- DedispersionConfig (include/pirate/DedispersionConfig.hpp:21-260) ---- */

/* DedispersionConfig::from_yaml(filename), src_lib/DedispersionConfig.cpp:643-717 (file or text). */
int pdd_config_from_yaml_file(const char *filename, pdd_config **out);
int pdd_config_from_yaml_string(const char *yaml_text, pdd_config **out);
void pdd_config_destroy(pdd_config *config);

/* DedispersionConfig::to_yaml_string(verbose), DedispersionConfig.cpp:635-640. Writes up to
 * buflen bytes (NUL-terminated) and the required size (including NUL) to *needed. */
int pdd_config_to_yaml(const pdd_config *config, int verbose, char *buf, size_t buflen, size_t *needed);

/* DedispersionConfig::make_channel_map(), DedispersionConfig.cpp:238-250; out has 2^R+1 doubles. */
int pdd_config_make_channel_map(const pdd_config *config, double *out, size_t n);

/* DedispersionConfig::{frequency_to_index, index_to_frequency, delay_to_frequency,
 * frequency_to_delay, dm_per_unit_delay, max_dm_of_all_trees}, DedispersionConfig.cpp:52-232.
 * 'what' names the function; x is its argument (ignored for the last two). */
int pdd_config_eval(const pdd_config *config, const char *what, double x, double *out);

/* DedispersionConfig::test(), DedispersionConfig.cpp:275-330. */
int pdd_config_test(const pdd_config *config);

/* ---- DedispersionPlan (include/pirate/DedispersionPlan.hpp) ---- */

/* DedispersionPlan(config, Params{gpu_runnable}), src_lib/DedispersionPlan.cpp:29-365. */
int pdd_plan_create(const pdd_config *config, int gpu_runnable, pdd_plan **out);
void pdd_plan_destroy(pdd_plan *plan);

/* DedispersionPlan::to_yaml_string(verbose), DedispersionPlan.cpp:664-920. */
int pdd_plan_to_yaml(const pdd_plan *plan, int verbose, char *buf, size_t buflen, size_t *needed);

/* DedispersionTree fields, include/pirate/DedispersionTree.hpp:19-70. */
typedef struct {
    int64_t primary_tree_index, early_trigger_level, amb_rank, dd_rank, nt_ds;
    int64_t max_width, dm_downsampling, time_downsampling, wt_dm_downsampling, wt_time_downsampling;
    int64_t Dcore, nprofiles, ndm_out, ndm_wt, nt_out, nt_wt;
    int64_t nsubbands_N, nmultiplets_M, pf_rank;
    double dm_min, dm_max, trigger_frequency;
} pdd_tree_info;

int pdd_plan_ntrees(const pdd_plan *plan, int64_t *ntrees);
int pdd_plan_tree_info(const pdd_plan *plan, int64_t itree, pdd_tree_info *out);

/* DedispersionPlan::decode_argmax(), DedispersionPlan.cpp:481-547.
 * out5 = {fmin, fmax, tlo, thi, p}. */
int pdd_plan_decode_argmax(const pdd_plan *plan, uint32_t token, int64_t itree, int64_t idm_coarse,
                           int64_t itime_coarse, int64_t *out5);

/* Batch form of decode_argmax (the vectorized decoding the reference's Python binding offers,
 * src_pybind11/pirate_pybind11.cpp:356-415): tokens is an array of n uint32 whose last two axes
 * are (ndm_out, nt_out) of tree itree (n a multiple of ndm_out * nt_out); element i is decoded
 * at idm_coarse = (i / nt_out) % ndm_out, itime_coarse = i % nt_out into out[5*i .. 5*i+4]
 * = {fmin, fmax, tlo, thi, p}. */
int pdd_plan_decode_argmax_batch(const pdd_plan *plan, int64_t itree, const uint32_t *tokens, int64_t n, int64_t *out);

/* DedispersionPlan::decode_argmax2(), DedispersionPlan.cpp:551-615.
 * out5 = {freq_lo_MHz, freq_hi_MHz, dm, timestamp_samp, width_samp}. */
int pdd_plan_decode_argmax2(const pdd_plan *plan, int64_t itree, int64_t fmin, int64_t fmax,
                            int64_t tlo, int64_t thi, int64_t p, double *out5);

/* DedispersionPlan::compute_steady_state_it0(itree), DedispersionPlan.cpp:618-661. */
int pdd_plan_steady_state_it0(const pdd_plan *plan, int64_t itree, int64_t *out, size_t n);

/* PfAvarApproximation(plan, freq_variances).tree_variance[itree] and .per_tf[itree]
 * (include/pirate/PfVariance.hpp:84-104, src_lib/PfVariance.cpp:238-366): analytic
 * white-noise variances of tree itree's peak-finder outputs. freq_variances: float64
 * (nfreq,), positive. tree_variance: float64 (N, ndm_wt, nprofiles). per_tf (may be NULL):
 * float64 (2^pf_rank, ndm_wt, nprofiles), the frequency-summed pf-rank bands. */
int pdd_plan_pf_variances(const pdd_plan *plan, const double *freq_variances, int64_t nfreq, int64_t itree,
                          double *tree_variance, int64_t tv_nelts, double *per_tf, int64_t ptf_nelts);

/* ---- GpuDedisperser (include/pirate/Dedisperser.hpp:45-392) ---- */

/* GpuDedisperser::Params, Dedisperser.hpp:47-86 (stream_pool replaced by an internal
 * compute stream; cuda_device_id -> device). Zero means "default" for the counts:
 * nbatches_out and nbatches_wt default to num_active_batches and must not be below it
 * (GpuDedisperser.cpp:116-123).
 * enable_variances: the peak finder also writes, per output slot and tree, the float64
 * (B, ndm_out, M, nprofiles) mean square of each (multiplet, profile)'s weighted candidates
 * over the chunk -- the out_var of the reference peak finder (ReferenceDedisperserBase::Params::
 * enable_variances, Dedisperser.hpp:425, 475-478; PeakFindingKernel.hpp:163-168), which the
 * reference's GpuDedisperser does not offer. Read it with pdd_output_variances. */
typedef struct {
    int64_t nbatches_out;
    int64_t nbatches_wt;
    int64_t num_consumers;
    int32_t device;
    int32_t synchronous;
    int64_t initial_chunk;
    int32_t enable_variances;
} pdd_dedisperser_params;

/* GpuDedisperser::create(params), src_lib/GpuDedisperser.cpp:90-372. */
int pdd_dedisperser_create(const pdd_plan *plan, const pdd_dedisperser_params *params, pdd_dedisperser **out);
void pdd_dedisperser_destroy(pdd_dedisperser *dd);

/* GpuDedisperser::allocate(gpu_allocator, host_allocator), GpuDedisperser.cpp:375-467
 * (this library allocates its own device memory). */
int pdd_dedisperser_allocate(pdd_dedisperser *dd);

/* Bytes of device memory held by the dedisperser. */
int pdd_dedisperser_gpu_bytes(const pdd_dedisperser *dd, int64_t *nbytes);

/* Bytes of pinned host memory held for lag-ring rows spilled to the host: rows whose
 * smallest stage-2 lag spans more than config max_gpu_clag chunks (the reference's host
 * ring zones, MegaRingbuf.hpp:45-58, GpuDedisperser.cpp:628-707); 0 when nothing spills. */
int pdd_dedisperser_host_bytes(const pdd_dedisperser *dd, int64_t *nbytes);

/* GpuDedisperser::acquire_input(seq_id, stream), GpuDedisperser.cpp:742-823.
 * *slot receives the device address of a contiguous (beams_per_batch, nfreq, nt_in) array
 * of the config dtype, valid until the matching release. */
int pdd_acquire_input(pdd_dedisperser *dd, int64_t seq_id, void *stream, void **slot);

/* GpuDedisperser::release_input_and_launch_dd_kernels(seq_id, stream), GpuDedisperser.cpp:826-855. */
int pdd_release_input_and_launch(pdd_dedisperser *dd, int64_t seq_id, void *stream);

/* GpuDedisperser::acquire_output(consumer, seq_id, stream, sync), GpuDedisperser.cpp:858-935.
 * out_max[itree] / out_argmax[itree] (arrays of length ntrees) receive device addresses of
 * contiguous (beams_per_batch, ndm_out, nt_out) arrays (config dtype / uint32). */
int pdd_acquire_output(pdd_dedisperser *dd, int64_t consumer, int64_t seq_id, void *stream, int sync,
                       void **out_max, uint32_t **out_argmax, int64_t ntrees);

/* GpuDedisperser::release_output(consumer, seq_id, stream), GpuDedisperser.cpp:938-984. */
int pdd_release_output(pdd_dedisperser *dd, int64_t consumer, int64_t seq_id, void *stream);

/* out_var of an acquired, not yet released, output (requires enable_variances):
 * out_var[itree] receives the device address of a (beams_per_batch, ndm_out, M, nprofiles)
 * float64 array, valid until the release. No reference counterpart on the GPU. */
int pdd_output_variances(pdd_dedisperser *dd, int64_t seq_id, double **out_var, int64_t ntrees);

/* GpuDedisperser::fill_all_weights(itree, weights), GpuDedisperser.cpp:1257-1279.
 * host_weights: float32 (nbatches_wt, beams_per_batch, ndm_wt, nt_wt, nprofiles, N), >= 0. */
int pdd_fill_all_weights(pdd_dedisperser *dd, int64_t itree, const float *host_weights, int64_t nelts);

/* GpuDedisperser::fill_analytic_weights(freq_variances), Dedisperser.hpp:244,
 * GpuDedisperser.cpp:1162-1238: weights 1/sqrt(analytic variance) for every tree, beam and
 * weight slot, so out_max is an SNR. freq_variances: float64 (nfreq,), positive. */
int pdd_fill_analytic_weights(pdd_dedisperser *dd, const double *freq_variances, int64_t nfreq);

/* GpuDedisperser::stop(e), GpuDedisperser.cpp:520-577. */
int pdd_stop(pdd_dedisperser *dd, const char *reason);

/* ---- output sharing with a consumer process ---- */

/* Output-ring export (the reference's IPC handshake of FrbServer::_fill_handshake,
 * src_lib/FrbServer.cpp:1583-1612): handle (64 bytes, a hipIpcMemHandle_t) of the one device
 * allocation that holds every tree's out_max and out_argmax ring, and its size. Each array
 * is (nbatches_out * beams_per_batch, ndm_out, nt_out) of the config dtype (out_max) or
 * uint32 (out_argmax), contiguous, at the byte offset pdd_output_array_offset gives
 * (which: 0 = out_max, 1 = out_argmax); output seq_id lives in slot seq_id % nbatches_out.
 * The producer and consumer agree on released seq_ids out of band, as the reference's
 * grouper protocol does. Regions of 2 GiB or more are refused. */
int pdd_output_ipc_handle(pdd_dedisperser *dd, void *handle, int64_t *nbytes);
int pdd_output_array_offset(const pdd_dedisperser *dd, int64_t itree, int which, int64_t *byte_offset);

/* Consumer side (FrbGrouper.cpp:565-581): maps a handle from another process of this node
 * on HIP device 'device'; *base receives its device address. pdd_ipc_close waits for the
 * device's queued work (which may read the mapping), then unmaps it. */
int pdd_ipc_open(const void *handle, int device, void **base);
int pdd_ipc_close(void *base, int device);

/* ---- input stage ---- */

/* GpuDequantizationKernel::launch (include/pirate/GpuDequantizationKernel.hpp:56-65; semantics of
 * ReferenceDequantizationKernel::apply, src_lib/GpuDequantizationKernel.cu:326-380).
 * data: (nbeams, nfreq, ntime) int4, two per byte, low nibble = even sample, -8 = missing (-> 0);
 * scales_offsets: (nbeams, nfreq, ntime/256, 2) float16 {scale, offset}; out: (nbeams, nfreq, ntime)
 * of out_dtype ("float32", "float16" or "bfloat16"); all device pointers, ntime % 256 == 0.
 * out = scale * value + offset computed in float32 (exactly the reference CPU kernel's value),
 * rounded to 16 bits for 16-bit outputs. Asynchronous on 'stream'. */
int pdd_dequantize(const char *out_dtype, int64_t nbeams, int64_t nfreq, int64_t ntime, const void *data,
                   const void *scales_offsets, void *out, void *stream);

/* ---- ring-buffer segment copies ---- */

/* GpuRingbufCopyKernel(params) + allocate() (include/pirate/RingbufCopyKernel.hpp:15-60,
 * src_lib/RingbufCopyKernel.cu:23-45, 128-158). octuples: host array of 8*N uint32, N rows of
 * {src_offset, src_phase, src_len, src_nseg, dst_offset, dst_phase, dst_len, dst_nseg}
 * (offsets in segments; len = frames in the zone; nseg = segments per frame), copied to the
 * device. nbytes_per_segment: 128 or 256 (nelts_per_segment x element size).
 * total_beams must be a multiple of beams_per_batch. */
int pdd_ringbuf_copy_create(int64_t total_beams, int64_t beams_per_batch, int64_t nbytes_per_segment,
                            const uint32_t *octuples, int64_t n_uint32, pdd_ringbuf_copy **out);
void pdd_ringbuf_copy_destroy(pdd_ringbuf_copy *k);

/* GpuRingbufCopyKernel::launch(ringbuf, ichunk, ibatch, stream), RingbufCopyKernel.cu:197-220
 * (semantics of CpuRingbufCopyKernel::apply, :60-123): for iframe = ichunk * total_beams +
 * ibatch * beams_per_batch and beams b < beams_per_batch, copies segment
 * src_offset + ((src_phase + iframe + b) % src_len) * src_nseg to the same formula on the dst
 * side. ringbuf: device pointer (16-byte aligned) to ringbuf_nbytes bytes; every addressed
 * segment is checked against ringbuf_nbytes before launch. Asynchronous on 'stream'. */
int pdd_ringbuf_copy_launch(pdd_ringbuf_copy *k, void *ringbuf, int64_t ringbuf_nbytes, int64_t ichunk,
                            int64_t ibatch, void *stream);

/* PCIe bytes per beam batch moved by the host spill of lag-ring rows, in each direction (0
 * without max_gpu_clag): the memcpy_h2g / memcpy_g2h entries of the reference's
 * ResourceTracker (include/pirate/ResourceTracker.hpp:14-75, GpuDedisperser.cpp:183-190). */
int pdd_dedisperser_spill_bytes(const pdd_dedisperser *dd, int64_t *nbytes);

/* ---- launch shapes ---- */

/* Launch shapes that change no output, only speed (no reference counterpart: the reference
 * fixes its launch shapes in its kernel generators). name: "pf_target_blocks" (the peak
 * finder splits each chunk into the fewest power-of-two time segments giving at least this
 * many workgroups; default 2048), "grid_xcd_run" (gridding workgroups per XCD run of adjacent
 * channel pairs; 0 = launch order; default 8), "fork" (primary trees p >= 1 start after the
 * gridding = 0, after tree 0's stage 2 = 1, automatic = -1, all kernels in order on one stream
 * = 2, for per-kernel timing). Takes effect at the next launch. */
int pdd_dedisperser_set_tuning(pdd_dedisperser *dd, const char *name, int64_t value);
int pdd_dedisperser_get_tuning(const pdd_dedisperser *dd, const char *name, int64_t *value);

/* ---- heterogeneous beam groups, one HIP graph launch per chunk ---- */

/* GpuDedisperserGroup (no reference counterpart: the reference batches homogeneous beams, its
 * beams_per_batch / num_active_batches of DedispersionConfig.hpp:81-84 and the launch loop of
 * GpuDedisperser.cpp:988-1105). Each plan is one beam group (its own channels, tree rank,
 * primary trees, downsampling levels), all its beams in one batch (beams_per_batch ==
 * beams_per_gpu), every plan with the same chunk (time_samples_per_chunk, time_sample_ms),
 * none spilling to the host (no max_gpu_clag). The group owns nslots input slots per member
 * (the caller fills chunk c+1 while chunk c runs). Per chunk, every member's chain runs on a
 * branch stream of its own, forked from and joined into one launch stream, so that the
 * members' kernels run at once. use_graph must be 0: HIP graph replay of the chunk is refused
 * in this library (its capture path ended in a segmentation fault on MI355X, cause not yet
 * identified); pdd_group_create and pdd_group_set_use_graph fail for use_graph != 0. */
typedef struct pdd_group pdd_group;
typedef struct {
    int64_t nslots;          /* input slots per member (0 = 2) */
    int64_t nbatches_out;    /* output slots per member (0 = 2) */
    int32_t device;
    int32_t use_graph;       /* must be 0 (see above) */
} pdd_group_params;

int pdd_group_create(const pdd_plan *const *plans, int64_t nplans, const pdd_group_params *params, pdd_group **out);
void pdd_group_destroy(pdd_group *g);
int pdd_group_allocate(pdd_group *g);
int pdd_group_gpu_bytes(const pdd_group *g, int64_t *nbytes);

/* A new handle to member imember's dedisperser (release it with pdd_dedisperser_destroy; the
 * group keeps the dedisperser alive). Use it for weights (pdd_fill_all_weights,
 * pdd_fill_analytic_weights), timing and launch shapes; its acquire/release calls are refused. */
int pdd_group_member(pdd_group *g, int64_t imember, pdd_dedisperser **dd);

/* The chunk protocol: chunks 0, 1, 2, ... in order; acquire_input / release_input_and_launch
 * alternate, as do acquire_output / release_output. slots[m]: member m's (B_m, F_m, T) input
 * slot (device, storage dtype), writable on 'stream' after the call. release_input_and_launch
 * blocks until output ichunk - nbatches_out has been released. acquire_output fills
 * out_max[k] / out_argmax[k] for k running over (member, tree) in order, each
 * (B_m, ndm_out, nt_out); they are ready on 'stream' (and on the host with sync != 0) and
 * valid until release_output. A failing call stops the group; later calls fail with the first
 * error. */
int pdd_group_acquire_input(pdd_group *g, int64_t ichunk, void *stream, void **slots, int64_t nmembers);
int pdd_group_release_input_and_launch(pdd_group *g, int64_t ichunk, void *stream);
int pdd_group_acquire_output(pdd_group *g, int64_t ichunk, void *stream, int sync, void **out_max,
                             void **out_argmax, int64_t n);
int pdd_group_release_output(pdd_group *g, int64_t ichunk, void *stream);
int pdd_group_stop(pdd_group *g, const char *reason);

/* Graph or one-by-one launches, and a member's launch shapes (see pdd_dedisperser_set_tuning);
 * both drop the captured graphs, which are re-captured at the next launch. */
int pdd_group_set_use_graph(pdd_group *g, int use_graph);
int pdd_group_set_tuning(pdd_group *g, int64_t imember, const char *name, int64_t value);
/* Nodes of the latest captured chunk graph (0 before the first graph launch). */
int pdd_group_graph_nodes(const pdd_group *g, int64_t *n);

/* Kernel timing (reference counterpart: the per-kernel KernelTimer loops and ResourceTracker
 * rates of GpuDedisperser::time, src_lib/GpuDedisperser.cpp:1573-1690). When enabled, HIP
 * events bracket every kernel launch on the compute stream; the report holds one line
 * "name launches total_ms" per kernel (grid_lds, stage1_p<p>, stage2_t<i>, pf_t<i>). */
int pdd_kernel_timing(pdd_dedisperser *dd, int enable);
int pdd_kernel_timing_report(pdd_dedisperser *dd, char *buf, size_t buflen, size_t *needed);

/* Test hook (no reference counterpart): copies an internal device buffer to the host.
 * name: "grid", "lds<p>", "sb<itree>", "ring<p>". With host == NULL only *nbytes is set. */
int pdd_debug_copy(pdd_dedisperser *dd, const char *name, void *host, size_t maxbytes, size_t *nbytes);

/* ---- learned peak scorer ----
 *
 * No counterpart in the reference: a small bf16 1-D convolutional network over (DM, time) tiles
 * of the dedispersed output (64 rows x 256 samples), run on MFMA, that stands beside the boxcar
 * statistic of the peak finder (PeakFindingKernel, src_lib/PeakFindingKernel.cu:252-456), which
 * stays the fallback and the baseline of its completeness. Network, parameter order and rounding:
 * pirate_amd/csrc/gpu/scorer.hpp. All pointers are device pointers; all calls are asynchronous
 * on 'stream' (a hipStream_t, NULL = default stream). */

/* Tiles: element (beam, row, t) at base + beam*beam_stride + row*row_stride + t (elements of
 * dtype 0 = float32, 1 = float16, 2 = bfloat16), valid for beam < nbeams, row < nrows,
 * t_lo <= t < t_hi. origins: (ntiles, 3) int32 (beam, row0, t0), or NULL for tile i = (i, 0, 0).
 * A tile outside the valid range gets a NaN logit and no gradient. */
typedef struct {
    const void *base;
    int32_t dtype;
    int64_t nbeams, beam_stride, nrows, row_stride, t_lo, t_hi;
    const int32_t *origins;
} pdd_tile_source;

/* Number of float32 parameters (13905). */
int64_t pdd_scorer_nparams(void);

/* logits[i] of tile i (ntiles floats). */
int pdd_scorer_forward(const float *params, const pdd_tile_source *src, int64_t ntiles, float *logits, void *stream);

/* Gradient of the mean binary cross-entropy over the ntiles tiles with labels[i] in {0, 1}:
 * grad (nparams floats), plus logits[i] and loss_terms[i] = softplus(logit) - label*logit. The
 * per-work-group partial gradients live in a workspace that this call allocates and frees in
 * stream order (hipMallocAsync / hipFreeAsync); their sum is formed in a fixed order, so the
 * result does not depend on timing. */
int pdd_scorer_grad(const float *params, const pdd_tile_source *src, const float *labels, int64_t ntiles,
                    float *grad, float *logits, float *loss_terms, void *stream);

/* pdd_scorer_grad without allocation. No counterpart in the reference: for callers that
 * cannot allocate device memory (code under HIP graph capture, serving loops that forbid
 * hipMalloc). Same arguments and results, bit for bit, as pdd_scorer_grad; it calls no
 * hipMalloc / hipFree. The per-work-group partial gradients live in a file-scope device array
 * of 512 x 13905 floats (28.5 MB per GPU, held until the process ends), which covers any
 * ntiles (a work group accumulates several tiles). Calls that overlap on one GPU race on that
 * array: serialize them (one stream, or events). */
int pdd_scorer_grad_noalloc(const float *params, const pdd_tile_source *src, const float *labels, int64_t ntiles,
                            float *grad, float *logits, float *loss_terms, void *stream);

/* One Adam step (step counted from 1) on params, with moment estimates m and v (nparams floats
 * each, zero before step 1). */
int pdd_scorer_adam(float *params, float *m, float *v, const float *grad, int64_t step, float lr, float beta1,
                    float beta2, float eps, void *stream);

/* Tile source over the peak-finder input rows of tree itree (its stage-2 output): rows
 * idm_out * M + m of every global beam, t = chunk sample of the tree, t_lo <= t < t_hi = chunk
 * samples (t_lo = the largest subband final lag, 0 without subbands: the peak finder moves the
 * samples before it). With a single full-band subband (M = 2^pf_rank multiplets), row r is DM
 * trial r of the tree. The rows hold the latest chunk of the beam's batch: read them after
 * acquire_output of that chunk and before the batch's next launch. */
int pdd_dedisperser_tree_rows(pdd_dedisperser *dd, int64_t itree, pdd_tile_source *src);

/* The learned scorer inside the chain: from the next launch on, each chunk's rows of tree
 * itree (full-band DM trials; rows a multiple of 64, chunk samples a multiple of 256) are cut
 * into tiles of 64 rows x 256 samples on a fixed grid and scored right after the tree's stage
 * 2, on the same stream. params: the scorer's pdd_scorer_nparams() float32 parameters in host
 * memory (copied). Calling it again replaces the parameters. */
int pdd_dedisperser_set_scorer(pdd_dedisperser *dd, int64_t itree, const float *params, int64_t nparams);

/* The (B, ntiles_dm, ntiles_t) float32 logits of an acquired output seq_id (tile (b, i, j)
 * covers rows 64 i .. 64 i + 63 and samples 256 j .. 256 j + 255 of beam b). */
int pdd_output_scores(pdd_dedisperser *dd, int64_t seq_id, float **scores, int64_t *ntiles_dm, int64_t *ntiles_t);

#ifdef __cplusplus
}
#endif

#endif /* PIRATE_AMD_H */
