// Learned peak scorer kernels on gfx950: bf16 MFMA (v_mfma_f32_16x16x16_bf16) forward and
// backward passes of the 3-layer 1-D conv net described in scorer.hpp, a deterministic
// reduction of per-work-group partial gradients, and Adam.
//
// Work decomposition. One work group (4 waves) runs one tile at a time (grid-stride over the
// tiles). Every convolution is a 16-row matrix product on MFMA 16x16x16 blocks:
//   forward  Y_l[co, t]  = sum_{k, ci} W_l[co, ci, k] A_{l-1}[ci, t+k-4]   M = co, N = t, K = (k, ci)
//   backward dA[ci, t]   = sum_{k, co} W[co, ci, k] dY[co, t-k+4]          M = ci, N = t, K = (k, co)
//            dW[co,ci,k] = sum_t dY[co, t] A[ci, t+k-4]                     M = co, N = ci, K = t
// Wave w owns output samples 64w..64w+63 (four 16-sample column blocks) of every forward and
// data-gradient product, t-range 64w..64w+63 of the dW2 and dW3 products and input channels
// 16w..16w+15 of the dW1 product. Activations live in LDS as [t + 4][channel] bf16 rows with a
// zero halo of 4 samples on each side, so a B operand (four consecutive channels at one sample)
// is one 8-byte read, and the taps t+k-4 never leave the buffer. dW accumulators stay in
// registers across all tiles of a work group; the group writes one partial gradient at the end.
//
// MFMA 16x16x16 bf16 lane layout (lane l, q = l / 16, j < 4, r < 4):
//   A[m = l % 16][k = 4q + j], B[k = 4q + j][n = l % 16], D[m = 4q + r][n = l % 16].

#include "device_common.hpp"
#include "scorer.hpp"

#include <mutex>
#include <stdexcept>
#include <string>

namespace pirate_amd {
namespace scorer {

typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short ushort_t;

constexpr int kPad = 4;
constexpr int kTT = kT + 2 * kPad;        // 264 rows: samples -4 .. 259
constexpr int kXS = 72;                   // X row stride (ushorts): 64 channels + 8
constexpr int kAS = 20;                   // 16-channel row stride (ushorts): 16 + 4
constexpr int kXBytes = kTT * kXS * 2;    // 38016
constexpr int kABytes = kTT * kAS * 2;    // 10560
// weight fragments (one s16x4 per lane): W1 (k, cb) k*4+cb | W2 k | W3 k | W2^T k | W3^T k
constexpr int fW1 = 0, fW2 = 36, fW3 = 45, fW2T = 54, fW3T = 63, kNFrag = 72;
constexpr int kNFragFwd = 54;
constexpr int kThreads = 256;

__device__ __forceinline__ ushort_t f2bf(float x)
{
    // round to nearest even; NaN stays NaN
    const unsigned u = __float_as_uint(x);
    if ((u & 0x7fffffffu) > 0x7f800000u)
        return ushort_t((u >> 16) | 0x40u);
    return ushort_t((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

__device__ __forceinline__ float bf2f(ushort_t h) { return __uint_as_float(unsigned(h) << 16); }

__device__ __forceinline__ f32x4 mfma16(s16x4 a, s16x4 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ s16x4 ld4(const ushort_t *p) { return *reinterpret_cast<const s16x4 *>(p); }

__device__ __forceinline__ s16x4 pack4(ushort_t a, ushort_t b, ushort_t c, ushort_t d)
{
    s16x4 v;
    v[0] = short(a); v[1] = short(b); v[2] = short(c); v[3] = short(d);
    return v;
}

__device__ __forceinline__ ushort_t src_bf16(const void *base, int dtype, long idx)
{
    if (dtype == kFloat32)
        return f2bf(static_cast<const float *>(base)[idx]);
    if (dtype == kFloat16)
        return f2bf(__half2float(static_cast<const __half *>(base)[idx]));
    return static_cast<const ushort_t *>(base)[idx];
}

// Tile i's element offset of (row0, t0) and whether the tile lies in the valid range.
__device__ __forceinline__ bool tile_origin(const TileSource &s, long i, long &off)
{
    long beam = i, row0 = 0, t0 = 0;
    if (s.origins) {
        beam = s.origins[3 * i];
        row0 = s.origins[3 * i + 1];
        t0 = s.origins[3 * i + 2];
    }
    if (beam < 0 || beam >= s.nbeams || row0 < 0 || row0 + kC0 > s.nrows || t0 < s.t_lo || t0 + kT > s.t_hi)
        return false;
    off = beam * s.beam_stride + row0 * s.row_stride + t0;
    return true;
}

// fp32 parameters -> bf16 MFMA fragments in LDS (nfrag = kNFragFwd or kNFrag).
__device__ void fill_fragments(const float *P, s16x4 *Wf, int nfrag)
{
    for (int e = threadIdx.x; e < nfrag * 64; e += kThreads) {
        const int f = e >> 6, l = e & 63, m = l & 15, q = l >> 4;
        ushort_t h[4];
        for (int j = 0; j < 4; j++) {
            float x;
            if (f < fW2) {                       // W1[co = m][ci = 16 cb + 4q + j][k]
                const int k = f >> 2, cb = f & 3;
                x = P[oW1 + (m * kC0 + 16 * cb + 4 * q + j) * kK + k];
            } else if (f < fW3) {                // W2[co = m][ci = 4q + j][k]
                x = P[oW2 + (m * kC + 4 * q + j) * kK + (f - fW2)];
            } else if (f < fW2T) {
                x = P[oW3 + (m * kC + 4 * q + j) * kK + (f - fW3)];
            } else if (f < fW3T) {               // W2[co = 4q + j][ci = m][k]  (A operand of dA1)
                x = P[oW2 + ((4 * q + j) * kC + m) * kK + (f - fW2T)];
            } else {
                x = P[oW3 + ((4 * q + j) * kC + m) * kK + (f - fW3T)];
            }
            h[j] = f2bf(x);
        }
        Wf[e] = pack4(h[0], h[1], h[2], h[3]);
    }
}

__device__ void zero_halo(ushort_t *buf, int stride)
{
    for (int e = threadIdx.x; e < 2 * kPad * stride; e += kThreads) {
        const int r = e / stride, c = e - r * stride;
        buf[(r < kPad ? r : kT + r) * stride + c] = 0;
    }
}

// X[ci, t] (bf16) -> Xt[(t + 4) * kXS + ci]; thread t loads sample t of four rows at a time.
__device__ void load_tile(const TileSource &s, long off, ushort_t *Xt)
{
    const int t = threadIdx.x;
    for (int cq = 0; cq < kC0 / 4; cq++) {
        ushort_t h[4];
        for (int j = 0; j < 4; j++)
            h[j] = src_bf16(s.base, s.dtype, off + long(4 * cq + j) * s.row_stride + t);
        *reinterpret_cast<s16x4 *>(Xt + (t + kPad) * kXS + 4 * cq) = pack4(h[0], h[1], h[2], h[3]);
    }
}

// relu(acc + bias) rounded to bf16 -> out[(t + 4) * kAS + 4q .. 4q+3]
__device__ __forceinline__ void store_act(ushort_t *out, int t, int q, f32x4 acc, const float *bias)
{
    ushort_t h[4];
    for (int r = 0; r < 4; r++)
        h[r] = f2bf(fmaxf(acc[r] + bias[r], 0.0f));
    *reinterpret_cast<s16x4 *>(out + (t + kPad) * kAS + 4 * q) = pack4(h[0], h[1], h[2], h[3]);
}

// Layers 1..3 of the wave's four column blocks; y3[i][r] = Y_3[co = 4q + r][t = 16(4w + i) + l % 16].
// Barriers: the caller has synchronized after load_tile; A1t and A2t are complete on return.
__device__ void forward_layers(const ushort_t *Xt, ushort_t *A1t, ushort_t *A2t, const s16x4 *Wf,
                               const float *bias1, const float *bias2, const float *bias3, float y3[4][4])
{
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6, m = l & 15, q = l >> 4;
    for (int i = 0; i < 4; i++) {
        const int t0 = 16 * (4 * w + i);
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        for (int k = 0; k < kK; k++) {
            const ushort_t *row = Xt + (t0 + m + k) * kXS + 4 * q;
#pragma unroll
            for (int cb = 0; cb < 4; cb++)
                acc = mfma16(Wf[(fW1 + 4 * k + cb) * 64 + l], ld4(row + 16 * cb), acc);
        }
        store_act(A1t, t0 + m, q, acc, bias1);
    }
    __syncthreads();
    for (int i = 0; i < 4; i++) {
        const int t0 = 16 * (4 * w + i);
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int k = 0; k < kK; k++)
            acc = mfma16(Wf[(fW2 + k) * 64 + l], ld4(A1t + (t0 + m + k) * kAS + 4 * q), acc);
        store_act(A2t, t0 + m, q, acc, bias2);
    }
    __syncthreads();
    for (int i = 0; i < 4; i++) {
        const int t0 = 16 * (4 * w + i);
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int k = 0; k < kK; k++)
            acc = mfma16(Wf[(fW3 + k) * 64 + l], ld4(A2t + (t0 + m + k) * kAS + 4 * q), acc);
        for (int r = 0; r < 4; r++)
            y3[i][r] = acc[r] + bias3[r];
    }
}

// Pooling of A_3 = relu(Y_3) over the tile: pmax, first argmax tau, pmean per channel, and the
// logit. red: LDS scratch of 4 waves x 16 channels x 3 words; hd: 16 pmax | 16 pmean | 16 tau | logit.
__device__ void pool_and_logit(const float y3[4][4], const float *P, float *red, float *hd)
{
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6, m = l & 15, q = l >> 4;
    for (int r = 0; r < 4; r++) {
        float best = fmaxf(y3[0][r], 0.0f), sum = best;
        int bidx = 16 * (4 * w) + m;
        for (int i = 1; i < 4; i++) {
            const float a = fmaxf(y3[i][r], 0.0f);
            if (a > best) {
                best = a;
                bidx = 16 * (4 * w + i) + m;
            }
            sum += a;
        }
        for (int off = 8; off >= 1; off >>= 1) {
            const float bo = __shfl_xor(best, off);
            const int io = __shfl_xor(bidx, off);
            const float so = __shfl_xor(sum, off);
            if (bo > best || (bo == best && io < bidx)) {
                best = bo;
                bidx = io;
            }
            sum += so;
        }
        if (m == 0) {
            float *e = red + (w * kC + 4 * q + r) * 3;
            e[0] = best;
            e[1] = __int_as_float(bidx);
            e[2] = sum;
        }
    }
    __syncthreads();
    if (threadIdx.x < kC) {
        const int c = threadIdx.x;
        float best = red[c * 3], sum = red[c * 3 + 2];
        int bidx = __float_as_int(red[c * 3 + 1]);
        for (int ww = 1; ww < 4; ww++) {
            const float *e = red + (ww * kC + c) * 3;
            if (e[0] > best) {            // later waves hold later samples: ties keep the first
                best = e[0];
                bidx = __float_as_int(e[1]);
            }
            sum += e[2];
        }
        hd[c] = best;
        hd[kC + c] = sum * (1.0f / kT);
        hd[2 * kC + c] = __int_as_float(bidx);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float z = P[oc0];
        for (int c = 0; c < kC; c++)
            z += P[ou + c] * hd[c] + P[ov + c] * hd[kC + c];
        hd[3 * kC] = z;
    }
    __syncthreads();
}

__device__ __forceinline__ void load_bias(const float *P, int o, int q, float *b)
{
    for (int r = 0; r < 4; r++)
        b[r] = P[o + 4 * q + r];
}


__global__ void __launch_bounds__(kThreads)
k_scorer_forward(const float *P, TileSource src, long ntiles, float *logits)
{
    __shared__ __align__(16) unsigned char smem[kXBytes + 2 * kABytes + kNFragFwd * 64 * 8 + 4 * kC * 3 * 4 + 128 * 4];
    ushort_t *Xt = reinterpret_cast<ushort_t *>(smem);
    ushort_t *A1t = reinterpret_cast<ushort_t *>(smem + kXBytes);
    ushort_t *A2t = reinterpret_cast<ushort_t *>(smem + kXBytes + kABytes);
    s16x4 *Wf = reinterpret_cast<s16x4 *>(smem + kXBytes + 2 * kABytes);
    float *red = reinterpret_cast<float *>(smem + kXBytes + 2 * kABytes + kNFragFwd * 64 * 8);
    float *hd = red + 4 * kC * 3;

    const int q = (threadIdx.x & 63) >> 4;
    float bias1[4], bias2[4], bias3[4];
    load_bias(P, ob1, q, bias1);
    load_bias(P, ob2, q, bias2);
    load_bias(P, ob3, q, bias3);
    fill_fragments(P, Wf, kNFragFwd);
    zero_halo(Xt, kXS);
    zero_halo(A1t, kAS);
    zero_halo(A2t, kAS);

    for (long n = blockIdx.x; n < ntiles; n += gridDim.x) {
        long off = 0;
        if (!tile_origin(src, n, off)) {            // uniform over the work group
            if (threadIdx.x == 0)
                logits[n] = __int_as_float(0x7fc00000);
            continue;
        }
        __syncthreads();                            // previous tile's readers of Xt / A1t / A2t are done
        load_tile(src, off, Xt);
        __syncthreads();
        float y3[4][4];
        forward_layers(Xt, A1t, A2t, Wf, bias1, bias2, bias3, y3);
        pool_and_logit(y3, P, red, hd);
        if (threadIdx.x == 0)
            logits[n] = hd[3 * kC];
    }
}


__global__ void __launch_bounds__(kThreads)
k_scorer_train(const float *P, TileSource src, const float *labels, long ntiles, float inv_n, float *logits,
               float *loss_terms, float *partials)
{
    __shared__ __align__(16) unsigned char smem[kXBytes + 4 * kABytes + kNFrag * 64 * 8 + 4 * kC * 3 * 4 + 128 * 4];
    ushort_t *Xt = reinterpret_cast<ushort_t *>(smem);
    ushort_t *A1t = reinterpret_cast<ushort_t *>(smem + kXBytes);
    ushort_t *A2t = reinterpret_cast<ushort_t *>(smem + kXBytes + kABytes);
    ushort_t *dYa = reinterpret_cast<ushort_t *>(smem + kXBytes + 2 * kABytes);
    ushort_t *dYb = reinterpret_cast<ushort_t *>(smem + kXBytes + 3 * kABytes);
    s16x4 *Wf = reinterpret_cast<s16x4 *>(smem + kXBytes + 4 * kABytes);
    float *red = reinterpret_cast<float *>(smem + kXBytes + 4 * kABytes + kNFrag * 64 * 8);
    float *hd = red + 4 * kC * 3;   // [0,16) pmax [16,32) pmean [32,48) tau, 48 logit, 49 g, [64,80) g*u, [80,96) g*v/T

    const int l = threadIdx.x & 63, w = threadIdx.x >> 6, m = l & 15, q = l >> 4;
    float bias1[4], bias2[4], bias3[4];
    load_bias(P, ob1, q, bias1);
    load_bias(P, ob2, q, bias2);
    load_bias(P, ob3, q, bias3);
    fill_fragments(P, Wf, kNFrag);
    zero_halo(Xt, kXS);
    zero_halo(A1t, kAS);
    zero_halo(A2t, kAS);
    zero_halo(dYa, kAS);
    zero_halo(dYb, kAS);

    const f32x4 z4 = {0.0f, 0.0f, 0.0f, 0.0f};
    f32x4 acc1[kK], acc2[kK], acc3[kK];        // dW1 (ci block w), dW2 / dW3 (t-range of wave w)
    for (int k = 0; k < kK; k++)
        acc1[k] = acc2[k] = acc3[k] = z4;
    float db1[4] = {0, 0, 0, 0}, db2[4] = {0, 0, 0, 0}, db3[4] = {0, 0, 0, 0};
    float du = 0.0f, dv = 0.0f, dc0 = 0.0f;    // threads 0..15 (channel = thread), thread 0 (c0)

    for (long n = blockIdx.x; n < ntiles; n += gridDim.x) {
        long off = 0;
        if (!tile_origin(src, n, off)) {
            if (threadIdx.x == 0) {
                logits[n] = __int_as_float(0x7fc00000);
                loss_terms[n] = __int_as_float(0x7fc00000);
            }
            continue;
        }
        __syncthreads();
        load_tile(src, off, Xt);
        __syncthreads();
        float y3[4][4];
        forward_layers(Xt, A1t, A2t, Wf, bias1, bias2, bias3, y3);
        pool_and_logit(y3, P, red, hd);

        // head: g = d(mean loss)/d(logit)
        if (threadIdx.x == 0) {
            const float z = hd[3 * kC], y = labels[n];
            const float g = (1.0f / (1.0f + expf(-z)) - y) * inv_n;
            logits[n] = z;
            loss_terms[n] = fmaxf(z, 0.0f) + log1pf(expf(-fabsf(z))) - y * z;
            hd[3 * kC + 1] = g;
            dc0 += g;
        }
        __syncthreads();
        const float g = hd[3 * kC + 1];
        if (threadIdx.x < kC) {
            const int c = threadIdx.x;
            du += g * hd[c];
            dv += g * hd[kC + c];
            hd[4 * kC + c] = g * P[ou + c];
            hd[5 * kC + c] = g * P[ov + c] * (1.0f / kT);
        }
        __syncthreads();

        // dY3 = bf16(relu'(Y3) * dA3), dA3[c, t] = g v[c] / T + [t == tau[c]] g u[c]
        for (int i = 0; i < 4; i++) {
            const int t = 16 * (4 * w + i) + m;
            ushort_t h[4];
            for (int r = 0; r < 4; r++) {
                const int c = 4 * q + r;
                float d = hd[5 * kC + c];
                if (t == __float_as_int(hd[2 * kC + c]))
                    d += hd[4 * kC + c];
                h[r] = (y3[i][r] > 0.0f) ? f2bf(d) : ushort_t(0);
                db3[r] += bf2f(h[r]);
            }
            *reinterpret_cast<s16x4 *>(dYa + (t + kPad) * kAS + 4 * q) = pack4(h[0], h[1], h[2], h[3]);
        }
        __syncthreads();

        // dW3 over the wave's t-range; dA2 -> dY2 over its column blocks
        for (int s = 0; s < 4; s++) {
            const int ts = 64 * w + 16 * s + 4 * q;
            const s16x4 a = pack4(dYa[(ts + 0 + kPad) * kAS + m], dYa[(ts + 1 + kPad) * kAS + m],
                                  dYa[(ts + 2 + kPad) * kAS + m], dYa[(ts + 3 + kPad) * kAS + m]);
            for (int k = 0; k < kK; k++) {
                const s16x4 b = pack4(A2t[(ts + 0 + k) * kAS + m], A2t[(ts + 1 + k) * kAS + m],
                                      A2t[(ts + 2 + k) * kAS + m], A2t[(ts + 3 + k) * kAS + m]);
                acc3[k] = mfma16(a, b, acc3[k]);
            }
        }
        for (int i = 0; i < 4; i++) {
            const int t0 = 16 * (4 * w + i);
            f32x4 acc = z4;
#pragma unroll
            for (int k = 0; k < kK; k++)
                acc = mfma16(Wf[(fW3T + k) * 64 + l], ld4(dYa + (t0 + m - k + 2 * kPad) * kAS + 4 * q), acc);
            const s16x4 a2 = ld4(A2t + (t0 + m + kPad) * kAS + 4 * q);
            ushort_t h[4];
            for (int r = 0; r < 4; r++) {
                h[r] = (ushort_t(a2[r]) != 0 && !(ushort_t(a2[r]) & 0x8000u)) ? f2bf(acc[r]) : ushort_t(0);
                db2[r] += bf2f(h[r]);
            }
            *reinterpret_cast<s16x4 *>(dYb + (t0 + m + kPad) * kAS + 4 * q) = pack4(h[0], h[1], h[2], h[3]);
        }
        __syncthreads();

        // dW2; dA1 -> dY1 (into dYa: dY3 is no longer read after the barrier above)
        for (int s = 0; s < 4; s++) {
            const int ts = 64 * w + 16 * s + 4 * q;
            const s16x4 a = pack4(dYb[(ts + 0 + kPad) * kAS + m], dYb[(ts + 1 + kPad) * kAS + m],
                                  dYb[(ts + 2 + kPad) * kAS + m], dYb[(ts + 3 + kPad) * kAS + m]);
            for (int k = 0; k < kK; k++) {
                const s16x4 b = pack4(A1t[(ts + 0 + k) * kAS + m], A1t[(ts + 1 + k) * kAS + m],
                                      A1t[(ts + 2 + k) * kAS + m], A1t[(ts + 3 + k) * kAS + m]);
                acc2[k] = mfma16(a, b, acc2[k]);
            }
        }
        for (int i = 0; i < 4; i++) {
            const int t0 = 16 * (4 * w + i);
            f32x4 acc = z4;
#pragma unroll
            for (int k = 0; k < kK; k++)
                acc = mfma16(Wf[(fW2T + k) * 64 + l], ld4(dYb + (t0 + m - k + 2 * kPad) * kAS + 4 * q), acc);
            const s16x4 a1 = ld4(A1t + (t0 + m + kPad) * kAS + 4 * q);
            ushort_t h[4];
            for (int r = 0; r < 4; r++) {
                h[r] = (ushort_t(a1[r]) != 0 && !(ushort_t(a1[r]) & 0x8000u)) ? f2bf(acc[r]) : ushort_t(0);
                db1[r] += bf2f(h[r]);
            }
            *reinterpret_cast<s16x4 *>(dYa + (t0 + m + kPad) * kAS + 4 * q) = pack4(h[0], h[1], h[2], h[3]);
        }
        __syncthreads();

        // dW1 for input channels 16w .. 16w+15, all samples
        for (int s = 0; s < 16; s++) {
            const int ts = 16 * s + 4 * q;
            const s16x4 a = pack4(dYa[(ts + 0 + kPad) * kAS + m], dYa[(ts + 1 + kPad) * kAS + m],
                                  dYa[(ts + 2 + kPad) * kAS + m], dYa[(ts + 3 + kPad) * kAS + m]);
            for (int k = 0; k < kK; k++) {
                const int ci = 16 * w + m;
                const s16x4 b = pack4(Xt[(ts + 0 + k) * kXS + ci], Xt[(ts + 1 + k) * kXS + ci],
                                      Xt[(ts + 2 + k) * kXS + ci], Xt[(ts + 3 + k) * kXS + ci]);
                acc1[k] = mfma16(a, b, acc1[k]);
            }
        }
    }

    // ---- partial gradient of this work group ----
    __syncthreads();
    float *out = partials + long(blockIdx.x) * kNParams;
    // dW1: complete per wave (ci block w); lane holds co = 4q + r, ci = 16w + m
    for (int k = 0; k < kK; k++)
        for (int r = 0; r < 4; r++)
            out[oW1 + ((4 * q + r) * kC0 + 16 * w + m) * kK + k] = acc1[k][r];
    // dW2, dW3: sum of the four waves' t-ranges, in wave order, through LDS (the activation
    // buffers are free now)
    float *S = reinterpret_cast<float *>(smem);        // (4 waves, 2 layers, 9 taps, 16 co, 16 ci)
    for (int k = 0; k < kK; k++)
        for (int r = 0; r < 4; r++) {
            const int e = (k * kC + 4 * q + r) * kC + m;
            S[(w * 2 + 0) * kK * kC * kC + e] = acc2[k][r];
            S[(w * 2 + 1) * kK * kC * kC + e] = acc3[k][r];
        }
    // bias gradients: lanes of one 16-lane group share channels 4q..4q+3
    for (int r = 0; r < 4; r++)
        for (int o = 8; o >= 1; o >>= 1) {
            db1[r] += __shfl_xor(db1[r], o);
            db2[r] += __shfl_xor(db2[r], o);
            db3[r] += __shfl_xor(db3[r], o);
        }
    float *B = S + 4 * 2 * kK * kC * kC;                 // (4 waves, 3 layers, 16 channels)
    if (m == 0)
        for (int r = 0; r < 4; r++) {
            B[(w * 3 + 0) * kC + 4 * q + r] = db1[r];
            B[(w * 3 + 1) * kC + 4 * q + r] = db2[r];
            B[(w * 3 + 2) * kC + 4 * q + r] = db3[r];
        }
    __syncthreads();
    for (int e = threadIdx.x; e < kK * kC * kC; e += kThreads) {
        const int k = e / (kC * kC), co = (e / kC) % kC, ci = e % kC;
        float s2 = 0.0f, s3 = 0.0f;
        for (int ww = 0; ww < 4; ww++) {
            s2 += S[(ww * 2 + 0) * kK * kC * kC + e];
            s3 += S[(ww * 2 + 1) * kK * kC * kC + e];
        }
        out[oW2 + (co * kC + ci) * kK + k] = s2;
        out[oW3 + (co * kC + ci) * kK + k] = s3;
    }
    if (threadIdx.x < 3 * kC) {
        const int layer = threadIdx.x / kC, c = threadIdx.x % kC;
        float s = 0.0f;
        for (int ww = 0; ww < 4; ww++)
            s += B[(ww * 3 + layer) * kC + c];
        out[ob1 + layer * kC + c] = s;
    }
    if (threadIdx.x < kC) {
        out[ou + threadIdx.x] = du;
        out[ov + threadIdx.x] = dv;
    }
    if (threadIdx.x == 0)
        out[oc0] = dc0;
}


__global__ void __launch_bounds__(256)
k_scorer_reduce(const float *partials, int nparts, float *grad)
{
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= kNParams)
        return;
    float s = 0.0f;
    for (int j = 0; j < nparts; j++)
        s += partials[long(j) * kNParams + p];
    grad[p] = s;
}


__global__ void __launch_bounds__(256)
k_scorer_adam(float *params, float *m, float *v, const float *grad, long n, float lr, float beta1, float beta2,
              float eps, float bc1, float bc2)
{
    const long i = long(blockIdx.x) * 256 + threadIdx.x;
    if (i >= n)
        return;
    const float g = grad[i];
    const float mi = beta1 * m[i] + (1.0f - beta1) * g;
    const float vi = beta2 * v[i] + (1.0f - beta2) * g * g;
    m[i] = mi;
    v[i] = vi;
    params[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}


int train_nparts(long ntiles, int max_parts)
{
    if (ntiles <= 0)
        return 0;
    return int(ntiles < long(max_parts) ? ntiles : long(max_parts));
}

static int forward_blocks(long ntiles)
{
    // one resident work group per CU (LDS-limited); 4 rounds of 256 CUs, then grid-stride
    return int(ntiles < 1024 ? ntiles : 1024);
}

void launch_forward(const float *params, const TileSource &src, long ntiles, float *logits, hipStream_t st)
{
    if (ntiles <= 0)
        return;
    hipLaunchKernelGGL(k_scorer_forward, dim3(forward_blocks(ntiles)), dim3(kThreads), 0, st, params, src, ntiles,
                       logits);
}

void launch_train(const float *params, const TileSource &src, const float *labels, long ntiles, float inv_n,
                  float *logits, float *loss_terms, float *partials, int nparts, hipStream_t st)
{
    if (nparts > 0)
        hipLaunchKernelGGL(k_scorer_train, dim3(nparts), dim3(kThreads), 0, st, params, src, labels, ntiles, inv_n,
                           logits, loss_terms, partials);
}

void launch_reduce(const float *partials, int nparts, float *grad, hipStream_t st)
{
    hipLaunchKernelGGL(k_scorer_reduce, dim3((kNParams + 255) / 256), dim3(256), 0, st, partials, nparts, grad);
}

void launch_adam(float *params, float *m, float *v, const float *grad, long n, float lr, float beta1, float beta2,
                 float eps, float bc1, float bc2, hipStream_t st)
{
    if (n <= 0)
        return;
    hipLaunchKernelGGL(k_scorer_adam, dim3((n + 255) / 256), dim3(256), 0, st, params, m, v, grad, n, lr, beta1, beta2,
                       eps, bc1, bc2);
}


// ---- gradient without allocation ----
// The partial gradients of pdd_scorer_grad_noalloc live in this file-scope device array (one
// copy per GPU, held until the process ends): kScratchParts work groups of kNParams floats,
// the same partial count as the allocating gradient's, so both give identical results. Every
// call overwrites all the partials it reduces, so nothing needs clearing; calls that overlap
// on one GPU race on the array.
static __device__ __align__(256) float scratch[size_t(kScratchParts) * kNParams];

static constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mutex;
static float *g_scratch_ptr[kMaxDevices];        // [device]: address of that device's array

// The array's address on the current device, looked up once per device (with that device
// current) by hipGetSymbolAddress; host code uses only this address, never the array's name.
static float *scratch_address()
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess)
        throw std::runtime_error(std::string("scorer scratch: hipGetDevice: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices)
        throw std::runtime_error("scorer scratch: device index " + std::to_string(dev) + " out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            throw std::runtime_error(std::string("scorer scratch: hipGetSymbolAddress: ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<float *>(ptr);
    }
    return g_scratch_ptr[dev];
}

void grad_noalloc(const float *params, const TileSource &src, const float *labels, long ntiles, float inv_n,
                  float *grad, float *logits, float *loss_terms, hipStream_t st)
{
    float *partials = scratch_address();
    const int nparts = train_nparts(ntiles, kScratchParts);
    if (nparts > kScratchParts)
        throw std::runtime_error("scorer gradient without allocation: " + std::to_string(nparts)
                                 + " partial gradients exceed the " + std::to_string(kScratchParts) + " of the scratch array");
    launch_train(params, src, labels, ntiles, inv_n, logits, loss_terms, partials, nparts, st);
    launch_reduce(partials, nparts, grad, st);
}

}  // namespace scorer
}  // namespace pirate_amd
