// Device-side helpers shared by the dedispersion kernels (gfx950, wave64).
//
// Storage types: float, __half, __hip_bfloat16. All arithmetic is float32; the
// 16-bit types are storage only (converted with round-to-nearest-even on store).
// Every kernel is compiled with -ffp-contract=off: each multiply and add rounds
// separately, which makes the float32 path bit-identical to the CPU oracle.

#ifndef PIRATE_AMD_DEVICE_COMMON_HPP
#define PIRATE_AMD_DEVICE_COMMON_HPP

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>

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

// The value as computed in float32. Without this the compiler may fuse the multiply that
// produced x into the float16 conversion (v_fma_mix*_f16), rounding once from the exact
// product to float16 instead of rounding to float32 and then to float16 -- a different
// value in rare cases. The empty constraint emits no instruction.
__device__ __forceinline__ float as_f32(float x)
{
    asm volatile("" : "+v"(x));
    return x;
}

__device__ __forceinline__ void st_st(float *p, float x) { *p = x; }
__device__ __forceinline__ void st_st(__half *p, float x) { *p = __float2half(as_f32(x)); }
__device__ __forceinline__ void st_st(__hip_bfloat16 *p, float x) { *p = __float2bfloat16(x); }

// Four consecutive storage elements (16-byte aligned for float, 8-byte for 16-bit types).
__device__ __forceinline__ float4 ld4_st(const float *p) { return *reinterpret_cast<const float4 *>(p); }

__device__ __forceinline__ float4 ld4_st(const __half *p)
{
    const uint2 u = *reinterpret_cast<const uint2 *>(p);
    return make_float4(__half2float(__ushort_as_half((unsigned short) (u.x & 0xffffu))),
                       __half2float(__ushort_as_half((unsigned short) (u.x >> 16))),
                       __half2float(__ushort_as_half((unsigned short) (u.y & 0xffffu))),
                       __half2float(__ushort_as_half((unsigned short) (u.y >> 16))));
}

__device__ __forceinline__ float4 ld4_st(const __hip_bfloat16 *p)
{
    // bfloat16 -> float is exact: the 16 bits are the high half of the float.
    const uint2 u = *reinterpret_cast<const uint2 *>(p);
    return make_float4(__uint_as_float(u.x << 16), __uint_as_float(u.x & 0xffff0000u),
                       __uint_as_float(u.y << 16), __uint_as_float(u.y & 0xffff0000u));
}

// Four consecutive storage elements from float (round-to-nearest-even for 16-bit types).
__device__ __forceinline__ void st4_st(float *p, float4 v) { *reinterpret_cast<float4 *>(p) = v; }

__device__ __forceinline__ void st4_st(__half *p, float4 v)
{
    const unsigned a = __half_as_ushort(__float2half(as_f32(v.x))), b = __half_as_ushort(__float2half(as_f32(v.y)));
    const unsigned c = __half_as_ushort(__float2half(as_f32(v.z))), d = __half_as_ushort(__float2half(as_f32(v.w)));
    *reinterpret_cast<uint2 *>(p) = make_uint2(a | (b << 16), c | (d << 16));
}

__device__ __forceinline__ void st4_st(__hip_bfloat16 *p, float4 v)
{
    const unsigned a = __bfloat16_as_ushort(__float2bfloat16(v.x)), b = __bfloat16_as_ushort(__float2bfloat16(v.y));
    const unsigned c = __bfloat16_as_ushort(__float2bfloat16(v.z)), d = __bfloat16_as_ushort(__float2bfloat16(v.w));
    *reinterpret_cast<uint2 *>(p) = make_uint2(a | (b << 16), c | (d << 16));
}

// V consecutive storage elements <-> floats: V = 4 (16-byte float or 8-byte 16-bit access) or,
// for 16-bit types, V = 8 (one 16-byte access).
template<typename Tst, int V>
__device__ __forceinline__ void ldv_st(const Tst *p, float *x)
{
    if constexpr (V == 4) {
        const float4 v = ld4_st(p);
        x[0] = v.x; x[1] = v.y; x[2] = v.z; x[3] = v.w;
    } else {
        static_assert(V == 8 && sizeof(Tst) == 2, "8-wide access is for 16-bit types");
        const uint4 u = *reinterpret_cast<const uint4 *>(p);
        const unsigned w[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
        for (int k = 0; k < 4; k++) {
            Tst lo, hi;
            const unsigned short l16 = (unsigned short) (w[k] & 0xffffu), h16 = (unsigned short) (w[k] >> 16);
            __builtin_memcpy(&lo, &l16, 2);
            __builtin_memcpy(&hi, &h16, 2);
            x[2 * k] = ld_st(&lo);
            x[2 * k + 1] = ld_st(&hi);
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
        unsigned w[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            Tst lo, hi;
            st_st(&lo, x[2 * k]);
            st_st(&hi, x[2 * k + 1]);
            unsigned short l16, h16;
            __builtin_memcpy(&l16, &lo, 2);
            __builtin_memcpy(&h16, &hi, 2);
            w[k] = unsigned(l16) | (unsigned(h16) << 16);
        }
        *reinterpret_cast<uint4 *>(p) = make_uint4(w[0], w[1], w[2], w[3]);
    }
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
