// Address-space views of generic pointers (gfx950).
#pragma once

#include <hip/hip_runtime.h>

namespace cm {

// Global-memory view of a generic pointer.  A pointer that reaches a kernel
// through a table in memory, or a parameter of a non-inlined device function,
// is generic: its loads and stores are issued as flat instructions, which
// count against the LDS wait counter as well as the vector-memory one, so a
// wait for an LDS result (or the wait before a barrier) also waits for every
// flat load or store in flight, and a prefetch no longer overlaps the work
// before it.  Every buffer these kernels address lives in device global
// memory, so their accesses go through this view as global loads and stores.
#if defined(__HIP_DEVICE_COMPILE__)
template <typename T> using gmem = __attribute__((address_space(1))) T;
template <typename T> using lmem = __attribute__((address_space(3))) T;
#else
// host pass: device functions are parsed there but never compiled for the GPU
template <typename T> using gmem = T;
template <typename T> using lmem = T;
#endif
template <typename T> __device__ __forceinline__ gmem<T> *as_global(T *p) { return (gmem<T> *)p; }

// Float atomic add to global memory without return (the hardware add, as
// unsafeAtomicAdd), issued as a global atomic.
__device__ __forceinline__ void global_add(float *p, float v) { (void)__builtin_amdgcn_global_atomic_fadd_f32(as_global(p), v); }

// Store with the nontemporal hint, for a pass output that the next pass reads
// only after the whole array has been written (the cache lines would be
// evicted long before a re-read): measured 2% faster per 512^3 measurement
// in the y, fused x and c2r passes and the deposit's mesh write-out (slower
// in the r2c z pass and for the partial-line record stores of the binning,
// which keep plain stores).
__device__ __forceinline__ void store_stream(float2 *p, float2 v) {
    typedef float v2f __attribute__((ext_vector_type(2)));
    __builtin_nontemporal_store(v2f{v.x, v.y}, (gmem<v2f> *)p);
}

__device__ __forceinline__ void store_stream(float *p, float v) { __builtin_nontemporal_store(v, (gmem<float> *)p); }
// Load with the nontemporal hint, for a pass input read once.
__device__ __forceinline__ float2 load_stream(const float2 *p) {
    typedef float v2f __attribute__((ext_vector_type(2)));
    const v2f v = __builtin_nontemporal_load((const gmem<v2f> *)p);
    return float2{v.x, v.y};
}

// Input of a spectrum pass of an N^3 mesh: nontemporal from N = 512 up, whose
// arrays (0.54 GB) exceed the memory-side cache (512^3, all five passes:
// 2.42 -> 2.37 ms per measurement; r2c 0.227 -> 0.218, c2r 0.229 -> 0.203,
// fused x 0.414 -> 0.397, forward y 0.218 -> 0.226 ms), plain below: a 256^3
// spectrum (68 MB) is read back from that cache, and the hint took the
// generator's 256^3 column passes from 0.045 to 0.080 ms.
template <int N>
__device__ __forceinline__ float2 load_pass(const float2 *p) {
    if constexpr (N >= 512)
        return load_stream(p);
    else
        return *p;
}

// LDS view of a generic pointer that points into the workgroup's shared
// memory: ds_ instructions instead of flat ones where the compiler cannot see
// that a pointer (a parameter of a non-inlined function) addresses LDS.
template <typename T> __device__ __forceinline__ lmem<T> *as_local(T *p) { return (lmem<T> *)p; }

// Raw buffer resource over the n bytes at p (word 3: 32-bit data format, as
// gfx9 raw buffers take it).  Buffer instructions take a 32-bit per-lane
// offset beside the uniform base, so an array of fewer than 4 GiB is addressed
// without 64-bit address arithmetic per lane; an access at or beyond n reads
// zero and a store there is dropped.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_rsrc(const void *p, unsigned n) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)n, 0x00020000);
}

}  // namespace cm
