// Learned peak scorer: a 3-layer 1-D convolutional network over (DM, time) tiles of the
// dedispersed output, run on gfx950 bf16 MFMA, with its training kernels.
//
// The reference has no learned scorer: this network stands beside the boxcar peak statistic
// of src_lib/PeakFindingKernel.cu:252-456, which stays the fallback and the baseline of its
// completeness.
//
// Network (one tile X: 64 DM rows x 256 samples; "same" convolutions, zero outside 0..255):
//   Y_l[co, t] = b_l[co] + sum_{ci, k} W_l[co, ci, k] A_{l-1}[ci, t + k - 4]
//   A_0 = bf16(X), A_1 = bf16(relu(Y_1)), A_2 = bf16(relu(Y_2)), A_3 = relu(Y_3)   (float32)
//   16 channels, 9 taps, Cin = 64 / 16 / 16; conv weights rounded to bf16, float32 accumulation
//   pmax[c] = max_t A_3[c, t] (first argmax tau[c]), pmean[c] = mean_t A_3[c, t]
//   logit = c0 + sum_c u[c] pmax[c] + v[c] pmean[c]
// Training minimizes the mean binary cross-entropy; the backward signals dY_3, dY_2, dY_1 are
// rounded to bf16 (after the relu masks) before they enter the matrix products.
//
// Parameter vector (float32): W1 (16,64,9) | W2 (16,16,9) | W3 (16,16,9) | b1 | b2 | b3 | u | v | c0

#ifndef PIRATE_AMD_SCORER_HPP
#define PIRATE_AMD_SCORER_HPP

#include <hip/hip_runtime.h>

namespace pirate_amd {
namespace scorer {

constexpr int kC0 = 64, kC = 16, kK = 9, kT = 256;
constexpr int oW1 = 0;
constexpr int oW2 = oW1 + kC * kC0 * kK;
constexpr int oW3 = oW2 + kC * kC * kK;
constexpr int ob1 = oW3 + kC * kC * kK;
constexpr int ob2 = ob1 + kC;
constexpr int ob3 = ob2 + kC;
constexpr int ou = ob3 + kC;
constexpr int ov = ou + kC;
constexpr int oc0 = ov + kC;
constexpr int kNParams = oc0 + 1;       // 13905

// Where the tiles are read: element (beam, row, t) at base + beam*beam_stride + row*row_stride + t,
// valid for beam < nbeams, row < nrows, t_lo <= t < t_hi. Tile i covers rows row0..row0+63 and
// samples t0..t0+255 of beam, with (beam, row0, t0) = origins[3i..3i+2]; with origins == nullptr
// tile i is (i, 0, 0). A tile outside the valid range is not read: its logit is NaN and it
// contributes nothing to a gradient.
enum SrcDtype { kFloat32 = 0, kFloat16 = 1, kBfloat16 = 2 };
struct TileSource {
    const void *base;
    int dtype;
    long nbeams, beam_stride, nrows, row_stride, t_lo, t_hi;
    const int *origins;
};

// Work-group count of the training kernel for ntiles tiles (= number of partial gradients).
int train_nparts(long ntiles, int max_parts);

void launch_forward(const float *params, const TileSource &src, long ntiles, float *logits, hipStream_t st);

// Per-tile logits and loss terms softplus(logit) - y*logit; partial gradients (nparts, kNParams)
// of (mean over ntiles) loss, with 1/ntiles passed as inv_n; nparts = train_nparts(ntiles, max_parts).
void launch_train(const float *params, const TileSource &src, const float *labels, long ntiles, float inv_n,
                  float *logits, float *loss_terms, float *partials, int nparts, hipStream_t st);

// grad[p] = sum_{j < nparts} partials[j, p], in order of j.
void launch_reduce(const float *partials, int nparts, float *grad, hipStream_t st);

// The gradient above without allocation: partial gradients in a file-scope device array of
// kScratchParts x kNParams floats per GPU (28.5 MB), whose address comes from
// hipGetSymbolAddress on the first call on each device. Identical results to
// launch_train + launch_reduce with nparts = train_nparts(ntiles, kScratchParts); overlapping
// calls on one GPU race on the array.
constexpr int kScratchParts = 512;
void grad_noalloc(const float *params, const TileSource &src, const float *labels, long ntiles, float inv_n,
                  float *grad, float *logits, float *loss_terms, hipStream_t st);

// Adam (Kingma & Ba 2015, Algorithm 1) on n float32 parameters; bc1 = 1 - beta1^step,
// bc2 = 1 - beta2^step.
void launch_adam(float *params, float *m, float *v, const float *grad, long n, float lr, float beta1, float beta2,
                 float eps, float bc1, float bc2, hipStream_t st);

}  // namespace scorer
}  // namespace pirate_amd

#endif
