// Device-side helpers shared by the dedispersion kernels (gfx950, wave64).
//
// Storage types: float, __half, __hip_bfloat16. All arithmetic is float32; the
// 16-bit types are storage only (converted with round-to-nearest-even on store).
//
// The st*_st helpers store to global memory with the nontemporal hint: every array they
// write (gridded rows, lag rings, peak-finder rows, outputs) is streamed once per chunk and
// far larger than the caches, so keeping its lines in L2 only evicts lines that are still to
// be read. Measured on MI355X (chord16k float32 chain): gridding 1.16 -> 1.10 ms, stage2_t0
// 1.04 -> 1.01 ms, step 5.33 -> 5.20 ms. The same hint on the loads is slower (stage2_t0
// 1.12 ms): neighbouring tiles of a ring row share lines. st_st_cached is the plain store.
// Every kernel is compiled with -ffp-contract=off: each multiply and add rounds
// separately, which makes the float32 path bit-identical to the CPU oracle.

#ifndef PIRATE_AMD_DEVICE_COMMON_HPP
#define PIRATE_AMD_DEVICE_COMMON_HPP

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>

#include <type_traits>

namespace pirate_amd {

static constexpr float kRsqrt2 = 0.7071067811865476f;

// Chunk index of a launch: read from device memory when the launch belongs to a replayed HIP
// graph (whose last node advances the counter), else the launch argument.
template <typename A>
__device__ __forceinline__ long launch_chunk(const A &a)
{
    return a.chunk_dev ? *a.chunk_dev : a.chunk;
}

__device__ __forceinline__ float ld_st(const float *p) { return *p; }
__device__ __forceinline__ float ld_st(const __half *p) { return __half2float(*p); }
__device__ __forceinline__ float ld_st(const __hip_bfloat16 *p) { return __bfloat162float(*p); }

// A storage element as loaded, before conversion: the float itself, or the 16 storage bits
// zero-extended to 32. Kernels that hold loaded rows across other work keep them in this form
// and convert at first use, so that each load writes the register that holds the row and the
// compiler needs no temporary per load: with the conversion right at the load, the 16-bit
// rank-7 stage 2 reused one temporary for all of its rows and waited for each load before
// issuing the next.
template<typename Tst>
using raw_st = typename std::conditional<sizeof(Tst) == 4, float, unsigned>::type;

__device__ __forceinline__ float ld_raw(const float *p) { return *p; }
__device__ __forceinline__ unsigned ld_raw(const __half *p) { return *reinterpret_cast<const unsigned short *>(p); }
__device__ __forceinline__ unsigned ld_raw(const __hip_bfloat16 *p) { return *reinterpret_cast<const unsigned short *>(p); }

template<typename Tst> __device__ __forceinline__ float from_raw(raw_st<Tst> v);
template<> __device__ __forceinline__ float from_raw<float>(float v) { return v; }
template<> __device__ __forceinline__ float from_raw<__half>(unsigned v) { return __half2float(__ushort_as_half((unsigned short) v)); }
template<> __device__ __forceinline__ float from_raw<__hip_bfloat16>(unsigned v) { return __uint_as_float(v << 16); }

// The value as computed in float32. Without this the compiler may fuse the multiply that
// produced x into the float16 conversion (v_fma_mix*_f16), rounding once from the exact
// product to float16 instead of rounding to float32 and then to float16 -- a different
// value in rare cases. The empty constraint emits no instruction.
__device__ __forceinline__ float as_f32(float x)
{
    asm volatile("" : "+v"(x));
    return x;
}

// Storage bits of x in a 16-bit type (round-to-nearest-even).
__device__ __forceinline__ unsigned short st_bits(const __half *, float x) { return __half_as_ushort(__float2half(as_f32(x))); }
__device__ __forceinline__ unsigned short st_bits(const __hip_bfloat16 *, float x) { return __bfloat16_as_ushort(__float2bfloat16(x)); }

// x rounded to the storage type and back (as if stored and re-read).
template<typename Tst>
__device__ __forceinline__ float round_st(float x)
{
    if constexpr (sizeof(Tst) == 4)
        return x;
    else {
        const unsigned short u = st_bits(static_cast<const Tst *>(nullptr), x);
        Tst q;
        __builtin_memcpy(&q, &u, 2);
        return ld_st(&q);
    }
}

// One element to global memory, without the nontemporal hint: for rows whose stores start at
// unaligned offsets, so that a line that one wave writes in part stays in L2 until the wave
// holding the rest of it writes it (the lagged-downsampling outputs: 1.10 -> 1.08 ms for the
// chord16k gridding kernel).
__device__ __forceinline__ void st_st_cached(float *p, float x) { *p = x; }
__device__ __forceinline__ void st_st_cached(__half *p, float x) { *p = __float2half(as_f32(x)); }
__device__ __forceinline__ void st_st_cached(__hip_bfloat16 *p, float x) { *p = __float2bfloat16(x); }

// One element to global memory (nontemporal).
__device__ __forceinline__ void st_st(float *p, float x) { __builtin_nontemporal_store(x, p); }
__device__ __forceinline__ void st_st(__half *p, float x)
{
    __builtin_nontemporal_store(st_bits(p, x), reinterpret_cast<unsigned short *>(p));
}
__device__ __forceinline__ void st_st(__hip_bfloat16 *p, float x)
{
    __builtin_nontemporal_store(st_bits(p, x), reinterpret_cast<unsigned short *>(p));
}

typedef float f32x4_nt __attribute__((ext_vector_type(4)));
typedef unsigned u32x2_nt __attribute__((ext_vector_type(2)));
typedef unsigned u32x4_nt __attribute__((ext_vector_type(4)));

// Four consecutive storage elements from float (round-to-nearest-even for 16-bit types),
// to global memory (nontemporal).
__device__ __forceinline__ void st4_st(float *p, float4 v)
{
    const f32x4_nt t = {v.x, v.y, v.z, v.w};
    __builtin_nontemporal_store(t, reinterpret_cast<f32x4_nt *>(p));
}

template<typename T16>
__device__ __forceinline__ void st4_st16(T16 *p, float4 v)
{
    const unsigned a = st_bits(p, v.x), b = st_bits(p, v.y), c = st_bits(p, v.z), d = st_bits(p, v.w);
    const u32x2_nt t = {a | (b << 16), c | (d << 16)};
    __builtin_nontemporal_store(t, reinterpret_cast<u32x2_nt *>(p));
}

__device__ __forceinline__ void st4_st(__half *p, float4 v) { st4_st16(p, v); }
__device__ __forceinline__ void st4_st(__hip_bfloat16 *p, float4 v) { st4_st16(p, v); }

// V consecutive storage elements <-> floats: V = 4 (16-byte float or 8-byte 16-bit access) or,
// for 16-bit types, V = 8 (one 16-byte access). Stores are nontemporal. Loads come in two
// steps, as for raw_st: ldv_raw (the bits as loaded) and from_rawv (to floats, at first use).
template<typename Tst, int V>
using rawv_st = typename std::conditional<sizeof(Tst) == 4, float4,
                                          typename std::conditional<V == 8, uint4, uint2>::type>::type;

template<typename Tst, int V>
__device__ __forceinline__ rawv_st<Tst, V> ldv_raw(const Tst *p)
{
    static_assert(V == 4 || (V == 8 && sizeof(Tst) == 2), "4-wide access, or 8-wide for 16-bit types");
    return *reinterpret_cast<const rawv_st<Tst, V> *>(p);
}

template<typename Tst, int V>
__device__ __forceinline__ void from_rawv(const rawv_st<Tst, V> &r, float *x)
{
    if constexpr (sizeof(Tst) == 4) {
        x[0] = r.x; x[1] = r.y; x[2] = r.z; x[3] = r.w;
    } else {
        unsigned w[V / 2];
        w[0] = r.x;
        w[1] = r.y;
        if constexpr (V == 8) {
            w[2] = r.z;
            w[3] = r.w;
        }
#pragma unroll
        for (int k = 0; k < V / 2; k++) {
            x[2 * k] = from_raw<Tst>(w[k] & 0xffffu);
            x[2 * k + 1] = from_raw<Tst>(w[k] >> 16);
        }
    }
}

template<typename Tst, int V>
__device__ __forceinline__ void stv_st(Tst *p, const float *x)
{
    if constexpr (V == 4) {
        st4_st(p, make_float4(x[0], x[1], x[2], x[3]));
    } else {
        static_assert(V == 8 && sizeof(Tst) == 2, "8-wide access is for 16-bit types");
        u32x4_nt t;
#pragma unroll
        for (int k = 0; k < 4; k++)
            t[k] = unsigned(st_bits(p, x[2 * k])) | (unsigned(st_bits(p, x[2 * k + 1])) << 16);
        __builtin_nontemporal_store(t, reinterpret_cast<u32x4_nt *>(p));
    }
}

// Read-only table entry through the constant address space: with a wave-uniform index the
// compiler issues a scalar load (scalar cache) instead of a vector memory round trip. For
// tables that no kernel writes while it runs.
__device__ __forceinline__ int ld_table(const int *p, int i)
{
    return ((const __attribute__((address_space(4))) int *) p)[i];
}

__device__ __forceinline__ int bitrev_dev(int i, int nbits)
{
    return (nbits > 0) ? int(__brev(unsigned(i)) >> (32 - nbits)) : 0;
}

// Position of time tau in a circular ring row of capacity C (tau may be negative).
__device__ __forceinline__ long ring_pos(long tau, long C)
{
    long r = tau % C;
    return (r < 0) ? (r + C) : r;
}

}  // namespace pirate_amd

#endif
