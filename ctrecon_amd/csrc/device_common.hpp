// Device helpers shared by the HIP translation units of the library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace ctr {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;

// BF16 storage: the upper 16 bits of an FP32 value, held as uint16_t.  Values are
// widened to FP32 on load and rounded to nearest even on store (v_cvt_pk_bf16_f32);
// all arithmetic is FP32 or wider.
__device__ __forceinline__ float ld(const float* p, size_t i) { return p[i]; }
__device__ __forceinline__ float ld(const uint16_t* p, size_t i) {
  return __uint_as_float((uint32_t)p[i] << 16);
}
__device__ __forceinline__ void st(float* p, size_t i, float v) { p[i] = v; }
__device__ __forceinline__ uint16_t to_bf16(float v) {
  return __builtin_bit_cast(uint16_t, (__bf16)v);
}
__device__ __forceinline__ void st(uint16_t* p, size_t i, float v) { p[i] = to_bf16(v); }

// BF16 with a fixed per-element dither, for the solvers' shadow of the iterate: the 16
// dropped bits are raised by r(i) in [0, 2^16) before truncation, so element i rounds up
// with probability (dropped bits) / 2^16 over a region.  Round to nearest gives every
// voxel of a flat region the same error, which a projection sums; the dithered errors
// average out along a ray.  r(i) is the upper half of a well-mixed 32-bit hash of the
// element index (the "lowbias32" integer finalizer published by C. Wellons): a linear
// sequence in the index would leave the errors correlated along some ray directions.
// |error| < 1 BF16 ulp.
__device__ __forceinline__ uint32_t mix32(uint32_t h) {
  h ^= h >> 16;
  h *= 0x7feb352du;
  h ^= h >> 15;
  h *= 0x846ca68bu;
  h ^= h >> 16;
  return h;
}

__device__ __forceinline__ uint16_t to_bf16_dither(float v, size_t i) {
  const uint32_t r = mix32((uint32_t)i) >> 16;
  return (uint16_t)((__float_as_uint(v) + r) >> 16);
}

// Global-memory view of a pointer.  Pointers read from memory (the descriptors of the
// batched kernels) carry no address space, and the compiler then emits flat accesses,
// which wait on both the scalar and the vector memory counters; kernel arguments are
// known to be global.  The projector bodies access their operands through gp() so both
// cases compile to global loads and stores.
#define CTR_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ CTR_GLOBAL T* gp(T* p) {
  return (CTR_GLOBAL T*)p;
}
// Element i of an array of 4-byte-word structs (ViewP, float2, float4, double2) read
// through a global pointer.
template <typename T>
__device__ __forceinline__ T gload(const T* p, size_t i) {
  static_assert(sizeof(T) % 4 == 0, "gload: 4-byte words");
  const CTR_GLOBAL uint32_t* w = (const CTR_GLOBAL uint32_t*)p + i * (sizeof(T) / 4);
  T v;
  uint32_t* o = reinterpret_cast<uint32_t*>(&v);
#pragma unroll
  for (int k = 0; k < (int)(sizeof(T) / 4); ++k) o[k] = w[k];
  return v;
}
__device__ __forceinline__ void st(CTR_GLOBAL float* p, size_t i, float v) { p[i] = v; }
__device__ __forceinline__ void st(CTR_GLOBAL uint16_t* p, size_t i, float v) {
  p[i] = to_bf16(v);
}

// Raw buffer view of a float or BF16 array: 32-bit byte offsets, and reads past
// num_records return 0 instead of touching memory outside the array.
template <typename T>
struct Buf;
//
// tap3(o, t1, t2, ...): elements o, o + t1, o + t2 with t1, t2 in {0, 1, 2} (taps of one
// footprint along a contiguous axis; an unused tap repeats element o).
template <>
struct Buf<float> {
  __amdgpu_buffer_rsrc_t r;
  __device__ __forceinline__ Buf(const float* p, uint32_t n)
      : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(p), (short)0, (int)(n * 4u),
                                            0x00020000)) {}
  __device__ __forceinline__ float operator[](int i) const {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, i * 4, 0, 0));
  }
  // element i + j, j wave-uniform: its byte offset goes in the loads' scalar offset, which
  // the range check does not include, so i + j must lie inside the array
  __device__ __forceinline__ float at(int i, int j) const {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, i * 4, j * 4, 0));
  }
  // at() as stored bits, widened by widen() where the value is used (a load carried
  // across a loop then stays in flight until its use)
  __device__ __forceinline__ uint32_t at_bits(int i, int j) const {
    return __builtin_amdgcn_raw_buffer_load_b32(r, i * 4, j * 4, 0);
  }
  static __device__ __forceinline__ float widen(uint32_t b) { return __uint_as_float(b); }
  __device__ __forceinline__ void tap3(int o, int t1, int t2, float& a0, float& a1,
                                       float& a2) const {
    a0 = (*this)[o];
    a1 = (*this)[o + t1];
    a2 = (*this)[o + t2];
  }
};
template <>
struct Buf<uint16_t> {
  __amdgpu_buffer_rsrc_t r;
  __device__ __forceinline__ Buf(const uint16_t* p, uint32_t n)
      : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<uint16_t*>(p), (short)0, (int)(n * 2u),
                                            0x00020000)) {}
  __device__ __forceinline__ float operator[](int i) const {
    return __uint_as_float((uint32_t)__builtin_amdgcn_raw_buffer_load_b16(r, i * 2, 0, 0) << 16);
  }
  __device__ __forceinline__ float at(int i, int j) const {
    return __uint_as_float((uint32_t)__builtin_amdgcn_raw_buffer_load_b16(r, i * 2, j * 2, 0)
                           << 16);
  }
  __device__ __forceinline__ uint32_t at_bits(int i, int j) const {
    return __builtin_amdgcn_raw_buffer_load_b16(r, i * 2, j * 2, 0);
  }
  static __device__ __forceinline__ float widen(uint32_t b) { return __uint_as_float(b << 16); }
  // the 8 bytes of elements i .. i + 3 (i even, the array 4-byte aligned)
  __device__ __forceinline__ uint2 load8b(int i) const {
    return __builtin_bit_cast(uint2, __builtin_amdgcn_raw_buffer_load_b64(r, i * 2, 0, 0));
  }
  // 16-bit loads (two aligned 32-bit loads with half-word selects measured slower: the
  // projectors are VALU-bound)
  __device__ __forceinline__ void tap3(int o, int t1, int t2, float& a0, float& a1,
                                       float& a2) const {
    a0 = (*this)[o];
    a1 = (*this)[o + t1];
    a2 = (*this)[o + t2];
  }
};
using BufF32 = Buf<float>;

// Overlap weights of [lo, hi] (lo < hi, clipped to the array) with the cells
// c0 = floor(lo + .5) ... c0 + n - 1: first, interior (1) and last.
__device__ __forceinline__ void edge_weights(float lo, float hi, int c0, int n, float& wfirst,
                                             float& wlast) {
  const float e0 = (float)c0 + 0.5f;  // upper edge of the first cell
  wfirst = (n == 1) ? hi - lo : e0 - lo;
  wlast = hi - ((float)(c0 + n - 1) - 0.5f);
}

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own L2):
// remap the flat workgroup id so that every XCD works through one contiguous range
// of logical blocks (neighbouring detector columns of the same view).
__device__ __forceinline__ int xcd_logical_block(int hw, int G) {
  const int q = G >> 3, r = G & 7, xcd = hw & 7, loc = hw >> 3;
  return (xcd < r) ? xcd * (q + 1) + loc : r * (q + 1) + (xcd - r) * q + loc;
}

// asin for the fan angles of the reference back projector.  On |x| <= 0.5 (fan
// half-angles up to 30 degrees): x + x^3/6 in FP64 plus the remainder x^5 Q(x^2) in
// FP32, Q of degree 5 fitted by Chebyshev least squares to
// (asin(x) - x - x^3/6) / x^5 on [0, 0.25]; the remainder is below 3e-3, so the
// result is within 3.4e-10 of asin (FP32 coefficients take no registers on gfx9,
// FP64 ones would).  Wider arguments take the library asin.  The argument is
// uniform across a wave where it is used, so the branch does not diverge.
__device__ __forceinline__ double asin_fan(double x) {
  if (fabs(x) > 0.5) return asin(x);
  const double t = x * x;
  const double x3 = x * t;
  const float tf = (float)t;
  float q = 0.027716955170035362f;
  q = fmaf(q, tf, 0.012915842235088348f);
  q = fmaf(q, tf, 0.022988272830843925f);
  q = fmaf(q, tf, 0.030344299972057343f);
  q = fmaf(q, tf, 0.04464368522167206f);
  q = fmaf(q, tf, 0.07499999552965164f);
  return fma(x3, 1.0 / 6.0, x) + (double)((float)(x3 * t) * q);
}

// Double-float (hi + lo) value: an FP32 pair carrying about 48 bits.
struct DF {
  float hi, lo;
};

// Error-free sum of two floats (Knuth's TwoSum): s + e == a + b exactly.
__device__ __forceinline__ void two_sum(float a, float b, float& s, float& e) {
  s = a + b;
  const float bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

// Compensated addition of two double-float values.
__device__ __forceinline__ DF df_add(DF a, DF b) {
  float s, e;
  two_sum(a.hi, b.hi, s, e);
  e += a.lo + b.lo;
  DF r;
  r.hi = s + e;
  r.lo = e - (r.hi - s);
  return r;
}

}  // namespace ctr
