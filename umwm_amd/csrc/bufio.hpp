// Memory access helpers of the fused kernels.
//
// Every per-point stream of the i-contiguous layout is read and written
// through raw buffer instructions: a row of ldi values (one (o, p) plane of
// a spectrum, one frequency of a table, one per-point array) becomes a
// buffer resource held in scalar registers, and each lane addresses it with
// a 32-bit byte offset (its own point, or a neighbour's) computed once per
// thread.  The address arithmetic of a row change is then scalar work, the
// loads carry no 64-bit vector address math, and they are counted by
// vmcnt alone, so a load issued ahead of its use is waited for exactly.
// An offset past the row's end reads 0 and drops stores.
//
// Tables indexed by a wave-uniform direction or frequency (cth, sth, f ...)
// are read through the constant address space so they become scalar loads.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace umwm {

using rsrc_t = __amdgpu_buffer_rsrc_t;

// Word 3 of a gfx9-family buffer descriptor: 32-bit data format, no
// swizzle, raw (untyped) addressing.
constexpr int kRsrcWord3 = 0x00020000;

__device__ __forceinline__ rsrc_t make_rsrc(const void* base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes,
                                           kRsrcWord3);
}

// One row of `n` values of type T starting at `base + row * n` (a row may
// span several planes of ldi values, addressed by a scalar offset).
template <typename T>
__device__ __forceinline__ rsrc_t row_rsrc(const T* base, size_t row, size_t n) {
  return make_rsrc(base + row * n, (uint32_t)(n * sizeof(T)));
}

// Loads and stores with a per-lane byte offset `voff` and a wave-uniform
// byte offset `soff` (held in a scalar register).
template <typename T>
__device__ __forceinline__ T bld(rsrc_t r, uint32_t voff, uint32_t soff = 0);
template <>
__device__ __forceinline__ float bld<float>(rsrc_t r, uint32_t voff, uint32_t soff) {
  return __builtin_bit_cast(float,
                            __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
}
template <>
__device__ __forceinline__ double bld<double>(rsrc_t r, uint32_t voff, uint32_t soff) {
  return __builtin_bit_cast(double,
                            __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, 0));
}
template <>
__device__ __forceinline__ int32_t bld<int32_t>(rsrc_t r, uint32_t voff, uint32_t soff) {
  return (int32_t)__builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0);
}

__device__ __forceinline__ void bst(rsrc_t r, uint32_t voff, float v, uint32_t soff = 0) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, (int)voff,
                                        (int)soff, 0);
}
__device__ __forceinline__ void bst(rsrc_t r, uint32_t voff, double v, uint32_t soff = 0) {
  using u2 = __attribute__((ext_vector_type(2))) uint32_t;
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), r, (int)voff, (int)soff, 0);
}
__device__ __forceinline__ void bst(rsrc_t r, uint32_t voff, int32_t v, uint32_t soff = 0) {
  __builtin_amdgcn_raw_buffer_store_b32((uint32_t)v, r, (int)voff, (int)soff, 0);
}

// Streaming variants for spectrum rows a kernel touches once: the
// cache-policy operand of the gfx940-family buffer instructions takes nt
// (bit 1, non-temporal) and sc1 (bit 4, device-scope write-through).  k_advect
// writes E^(n+1) with sc1|nt, so its 1 GB of stores leaves no dirty lines in
// the L2s for the end-of-kernel write-back, and reads E^n, which nothing
// reads again in the substep, with nt.
constexpr int kCpolNt = 2, kCpolSc1 = 16;
template <typename T>
__device__ __forceinline__ T bld_nt(rsrc_t r, uint32_t voff, uint32_t soff = 0);
template <>
__device__ __forceinline__ float bld_nt<float>(rsrc_t r, uint32_t voff, uint32_t soff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff,
                                                                        kCpolNt));
}
template <>
__device__ __forceinline__ double bld_nt<double>(rsrc_t r, uint32_t voff, uint32_t soff) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff,
                                                                         kCpolNt));
}
__device__ __forceinline__ void bst_stream(rsrc_t r, uint32_t voff, float v, uint32_t soff = 0) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, (int)voff,
                                        (int)soff, kCpolSc1 | kCpolNt);
}
__device__ __forceinline__ void bst_stream(rsrc_t r, uint32_t voff, double v, uint32_t soff = 0) {
  using u2 = __attribute__((ext_vector_type(2))) uint32_t;
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), r, (int)voff, (int)soff,
                                        kCpolSc1 | kCpolNt);
}

// 16-byte accesses.  The builtins' unsigned vectors are always bit-cast as
// a whole to a typed vector: extracting single elements of the builtin's
// own return value does not reliably keep the wide load.
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));

template <typename T>
struct Vec16;
template <>
struct Vec16<float> {
  using type = f32x4;
  static constexpr int n = 4;
};
template <>
struct Vec16<double> {
  using type = f64x2;
  static constexpr int n = 2;
};

template <typename T>
__device__ __forceinline__ typename Vec16<T>::type bld16(rsrc_t r, uint32_t voff,
                                                         uint32_t soff = 0) {
  return __builtin_bit_cast(typename Vec16<T>::type,
                            __builtin_amdgcn_raw_buffer_load_b128(r, (int)voff, (int)soff, 0));
}

template <typename V>
__device__ __forceinline__ void bst16(rsrc_t r, uint32_t voff, V v, uint32_t soff = 0) {
  static_assert(sizeof(V) == 16, "16-byte store");
  __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, v), r, (int)voff, (int)soff,
                                         0);
}

// N consecutive values of T starting at byte offset voff (N * sizeof(T) a
// multiple of 16).
template <typename T, int N>
__device__ __forceinline__ void bldn(rsrc_t r, uint32_t voff, T (&out)[N]) {
  constexpr int k = Vec16<T>::n;
  static_assert(N % k == 0, "whole 16-byte chunks");
#pragma unroll
  for (int c = 0; c < N / k; ++c) {
    const auto v = bld16<T>(r, voff + 16u * (uint32_t)c);
#pragma unroll
    for (int i = 0; i < k; ++i) out[c * k + i] = v[i];
  }
}

// Scalar (constant address space) load of a wave-uniform element.
template <typename T>
__device__ __forceinline__ T cld(const T* p, int idx) {
  return ((const __attribute__((address_space(4))) T*)p)[idx];
}

// Four consecutive wave-uniform values (one scalar load).
template <typename T>
__device__ __forceinline__ void cld4(const T* p, int idx, T& a, T& b, T& c, T& d) {
  typedef T v4 __attribute__((ext_vector_type(4)));
  const v4 v = ((const __attribute__((address_space(4))) v4*)p)[idx];
  a = v[0];
  b = v[1];
  c = v[2];
  d = v[3];
}

// A bfloat16 value (raw 16-bit buffer load) widened to float.
__device__ __forceinline__ float bld_bf16(rsrc_t r, uint32_t voff, uint32_t soff = 0) {
  const unsigned short h = __builtin_amdgcn_raw_buffer_load_b16(r, (int)voff, (int)soff, 0);
  return __builtin_bit_cast(float, (uint32_t)h << 16);
}

// The raw bits of a bfloat16 value, zero-extended, for loads issued ahead
// of their use: widening at the load would need the data at once, and with
// few registers free the compiler then waits for every load in flight
// after each one.  bf16_bits_to_f widens at the use.
__device__ __forceinline__ uint32_t bld_u16(rsrc_t r, uint32_t voff, uint32_t soff = 0) {
  return __builtin_amdgcn_raw_buffer_load_b16(r, (int)voff, (int)soff, 0);
}
__device__ __forceinline__ float bf16_bits_to_f(uint32_t h) {
  return __builtin_bit_cast(float, h << 16);
}

// A lane offset past every descriptor's range: the hardware range check
// drops a store made with it (descriptors cover less than 2^31 bytes), so
// stores can be predicated per lane without a divergent branch.
constexpr uint32_t kDropOffset = 0x80000000u;

// Wave index of the calling thread as a wave-uniform value.
__device__ __forceinline__ int wave_id() {
  return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
}

}  // namespace umwm
