// ksh_devutil.hpp -- small gfx950 device helpers shared by the band kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace ksh {

typedef unsigned int ksh_u32x2 __attribute__((ext_vector_type(2)));

// broadcast of lane `src`'s value (v_readlane: the result is wave-uniform)
__device__ __forceinline__ double readlane_f64(double v, int src)
{
    int lo = __builtin_amdgcn_readlane(__double2loint(v), src);
    int hi = __builtin_amdgcn_readlane(__double2hiint(v), src);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ float readlane_f32(float v, int src)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
}
__device__ __forceinline__ double bcast(double v, int src) { return readlane_f64(v, src); }
__device__ __forceinline__ float bcast(float v, int src) { return readlane_f32(v, src); }

// workgroup barrier that orders LDS only: global loads in flight (register prefetches private to
// their thread) are not waited for
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// value the compiler cannot see through: keeps per-lane comparisons inside a loop instead of
// hoisted (as one lane mask per comparison) out of it
__device__ __forceinline__ int opaque(int v)
{
    asm volatile("" : "+v"(v));
    return v;
}

// the same for a wave-uniform value (kept in an SGPR): values derived from it are recomputed where
// they are used instead of hoisted out of a loop, each into a scalar register of its own
__device__ __forceinline__ int opaque_s(int v)
{
    asm volatile("" : "+s"(v));
    return v;
}

// Band reads and writes go through a buffer resource covering exactly the band (bytes < 2 GiB,
// checked by the launchers).  A masked-off access is sent to offset 0x80000000, beyond the
// resource: the buffer unit returns 0 for such a read and drops such a write.  Masked accesses
// thus stay straight-line code with many loads in flight, and no mask stays live across a load.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t band_rsrc(const void *base, unsigned bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)bytes, 0x00020000);
}

constexpr unsigned kMaskedOff = 0x80000000u;

template <typename T>
__device__ __forceinline__ T load_if(bool ok, __amdgpu_buffer_rsrc_t rs, unsigned byte_off)
{
    const int off = (int)(ok ? byte_off : kMaskedOff);
    if constexpr (sizeof(T) == 8)
        return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rs, off, 0, 0));
    else
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, off, 0, 0));
}

template <typename T>
__device__ __forceinline__ void store_if(bool ok, __amdgpu_buffer_rsrc_t rs, unsigned byte_off, T v)
{
    const int off = (int)(ok ? byte_off : kMaskedOff);
    if constexpr (sizeof(T) == 8)
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(ksh_u32x2, v), rs, off, 0, 0);
    else
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), rs, off, 0, 0);
}

// buffer access at a per-lane offset (kMaskedOff for a masked-off lane) plus a wave-uniform
// offset carried in an SGPR
template <typename T>
__device__ __forceinline__ T load_vs(__amdgpu_buffer_rsrc_t rs, unsigned voff, unsigned soff)
{
    if constexpr (sizeof(T) == 8)
        return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rs, (int)voff, (int)soff, 0));
    else
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, (int)voff, (int)soff, 0));
}

template <typename T>
__device__ __forceinline__ void store_vs(__amdgpu_buffer_rsrc_t rs, unsigned voff, unsigned soff, T v)
{
    if constexpr (sizeof(T) == 8)
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(ksh_u32x2, v), rs, (int)voff, (int)soff, 0);
    else
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), rs, (int)voff, (int)soff, 0);
}

}  // namespace ksh
