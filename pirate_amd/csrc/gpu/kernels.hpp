// Host-visible argument structs and launchers for the gfx950 dedispersion kernels.
// All pointers are device pointers; strides and offsets are in elements.

#ifndef PIRATE_AMD_KERNELS_HPP
#define PIRATE_AMD_KERNELS_HPP

#include <hip/hip_runtime.h>
#include "../host/config.hpp"

namespace pirate_amd {

// ---- tree gridding fused with lagged downsampling ----
// grid[b,n,t] = sum_f w[n,f] in[b,f,t] (CSR weights, ascending f), n < nchan = 2^R, and for
// primary trees p = 1..npri-1: out[p] (B, 2^(R-1), ntime >> p).
static constexpr int kMaxPrimaryTrees = 8;

struct GridLdsArgs {
    const void *in;                      // (B, nfreq, ntime) input slot
    long in_bstride;
    void *grid;                          // (B, nchan, ntime) gridded rows (primary tree 0 input)
    long grid_bstride;
    int B, nchan, nfreq, ntime;
    const int *row_f0;                   // (nchan) first input channel of tree channel n
    const int *row_cnt;                  // (nchan) number of input channels
    const int *row_woff;                 // (nchan) offset into w
    const float *w;                      // overlap weights
    void *out[kMaxPrimaryTrees];         // out[p], p >= 1
    long out_bstride[kMaxPrimaryTrees];
    float *state;                        // per global beam, see lds_state_nelts()
    long state_bstride;
    int npri, s;                         // s = (stage-1 dd_rank of p>=1) + 1
    int xcd_run = 8;                     // consecutive channel pairs per XCD (0: plain order)
    int beam0;                           // global index of the batch's first beam
    long chunk;                          // chunk index (selects the state parity)
    const long *chunk_dev = nullptr;     // if set, the chunk index is read here (graph replay)
};
long lds_state_nelts(int R, int npri, int s);   // per beam
void launch_grid_lds(Dtype dt, const GridLdsArgs &a, hipStream_t st);


// ---- stage-1 tree: (B, 2^amb, 2^dd, Tp) -> lag ring rows (a, dm_brev) ----
struct RingView {
    void *base;               // ring storage of one primary tree
    const long *row_off;      // (nrows) element offset of each row within a beam
    const long *row_cap;      // (nrows) capacity (elements) of each row
    long bstride;             // elements per beam
};

struct Stage1Args {
    const void *in;
    long in_bstride;
    RingView ring;
    float *state;             // per (global beam, a): tree pass history
    long state_bstride;       // per beam
    int B, amb, dd, Tp, beam0;
    long chunk;               // chunk index (ring time = chunk*Tp + t)
    const long *chunk_dev = nullptr;   // if set, the chunk index is read here (graph replay)
};


// ---- stage-2 tree + subband extraction -> peak-finder input rows ----
// Per-tree device tables built on the host (see dedisperser.cpp):
//   main_emit[r*(D/2) + p]: first stream (pf_dm*M + m) emitted by main pair p of pass r, or -1
//   stag[k*6 + 0..5]: (pass r, row0, row1, lag, stream s0, hist offset) staggered butterflies
//   lvl0[k*2 + 0..1]: (row, stream) copies of the pass (dd - pf_rank) input
//   fl[s]: final lag of stream s (applied as a write offset)
struct Stage2Args {
    RingView ring;            // ring of the tree's primary tree
    const int *row_lag;       // (2^dd rows fc, 2^amb j) -> lag, indexed fc*2^amb + j
    void *sb;                 // (B, ndm_out, M, rowlen) peak-finder input rows
    long sb_bstride;
    int rowlen, tpad;
    const int *main_emit, *stag, *lvl0, *fl;
    int nstag, nlvl0, stag_hist_nelts;
    const int *stag_begin;    // (dd+1): staggered entries of pass r are [stag_begin[r], stag_begin[r+1])
    int M, pf_rank;
    float *state;             // per (global beam, j): pass history + staggered history
    long state_bstride;
    int B, amb, dd, Tp, beam0, amb1_dd1;   // amb1_dd1: stage-1 dd rank (ring row = fc*2^amb1_dd1 + j)
    long chunk;
    const long *chunk_dev = nullptr;   // if set, the chunk index is read here (graph replay)
    const int *out_stream;    // register-tree path: (2^dd) final row -> stream, or -1
    const int *stag_goff;     // rank-8 split: staggered history offset within its row group
};
long stage2_state_nelts_per_tree(int dd, int stag_hist_nelts);
int stage2_lds_bytes(int dd, int stag_hist_nelts);
void launch_stage2(Dtype dt, const Stage2Args &a, hipStream_t st);
// Rank-8 stage 2 with subband searches: two launches through a float32 scratch X of
// (B, 2^amb, 256, Tp) floats; state per tree instance: split_state_nelts_per_tree(...).
long split_state_nelts_per_tree(int stag_hist_nelts);
int split_b_lds_bytes(int stag_hist_nelts);
void launch_stage2_split(Dtype dt, const Stage2Args &a, float *X, hipStream_t st);
// its part A (passes 0..3) on the register-tree code, one wave per 16-row group
void launch_stage2_split_a_reg(Dtype dt, const Stage2Args &a, float *X, long split_nst, hipStream_t st);

// ---- register-resident trees (phase A / LDS delay lines / phase B), bit-identical results ----
// Stage 1 always; stage 2 when every stream is a final-pass output (no staggered pairs, no
// earlier emissions). State per tree instance: fast_tree_state_nelts(dd) floats.
long fast_tree_state_nelts(int dd);
void launch_stage1_fast(Dtype dt, const Stage1Args &a, hipStream_t st);
void launch_stage2_fast(Dtype dt, const Stage2Args &a, hipStream_t st);


// ---- int4 dequantization: out = scale * value + offset per 256-sample (beam, freq) block ----
// data: nsamples int4 (two per byte, low nibble first); scoff: float16 {scale, offset} pairs.
void launch_dequant(Dtype dt, const void *data, const void *scoff, void *out, long nsamples, hipStream_t st);


// ---- ring-buffer segment copies driven by octuples (device array of 8*noct uint32) ----
// Host spill: appends outbox row i (Tp elements of elt_bytes) of each beam of the batch to
// device ring row rows[i] at ring time chunk * Tp.
void launch_spill_scatter(void *ring, long bstride, int elt_bytes, long outbox_off, const int *rows, long nrows,
                          const long *row_off, const long *row_cap, long Tp, long chunk, int beam0, int B, hipStream_t st);
void launch_ringbuf_copy(void *ringbuf, const uint32_t *octuples, long noct, long seg_bytes, int nbeams,
                         unsigned long long iframe, hipStream_t st);


// ---- peak finder over (B, ndm_out, M, rowlen) rows -> out_max, out_argmax ----
struct PfArgs {
    void *sb;                 // rows [tpad history | Tp | future (max final lag)]
    long sb_bstride;
    int rowlen, tpad, flmax;
    const int *m_to_n;        // (M)
    const float *wt;          // (B, ndm_wt, nt_wt, P, N) non-negative weights
    long wt_bstride;
    void *out_max;            // (B, ndm_out, nt_out), storage dtype
    unsigned int *out_argmax; // (B, ndm_out, nt_out)
    long out_bstride;
    int B, ndm_out, M, N, P, nlev, Tp, Dout, Dcore, ndm_wt, nt_wt, nt_out;
    int beam0;                // global index of the batch's first beam (sb rows are per global beam)
    int nseg = 1;             // time segments per chunk (set by launch_pf)
    long target_blocks = 2048;   // launch_pf picks the fewest segments giving this many workgroups
    double *out_var = nullptr;   // optional (B, ndm_out, M, P) float64: per-chunk mean square of each
                                 // (multiplet, profile)'s weighted candidates
};
bool pf_config_supported(int M, int tpad, int nlev, int Dcore, int Dout, int Tp);
void launch_pf(Dtype dt, const PfArgs &a, hipStream_t st);

}  // namespace pirate_amd

#endif
