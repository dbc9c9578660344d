// Shared device definitions for the guanaco gfx950 kernels.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace gx {

// Workgroup barrier that orders every LDS access before it against every
// access after it. The empty asm statements with a memory clobber stop the
// compiler from moving loads or stores across the barrier: hipcc has sunk
// LDS loads below __syncthreads() (into a following loop) so that a fast wave
// overwrote the data a slower wave had still to read.
__device__ __forceinline__ void lds_barrier() {
  asm volatile("" ::: "memory");
  __syncthreads();
  asm volatile("" ::: "memory");
}

// Spectrum columns per block in the column-blocked half-spectrum layout
// F[t][cb][y][ci] (kx = cb * CB + ci). A block of CB columns of four
// consecutive rows is one 128-byte segment.
constexpr int CB = 4;

// Storage type of the filtered projections ("slabs").
enum Store : int { kF32 = 0, kF16 = 1, kBF16 = 2 };

// bfloat16 as raw bits, round-to-nearest-even from float.
struct bf16_t {
  uint16_t u;
};

__host__ __device__ __forceinline__ uint16_t f32_to_bf16_bits(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu))
    return (uint16_t)((u >> 16) | 0x40u);  // quiet NaN
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

template <int S>
struct StoreT;
template <>
struct StoreT<kF32> {
  using T = float;
  __device__ __forceinline__ static T pack(float f) { return f; }
  __device__ __forceinline__ static float unpack(T v) { return v; }
};
template <>
struct StoreT<kF16> {
  using T = __half;
  __device__ __forceinline__ static T pack(float f) { return __float2half(f); }
  __device__ __forceinline__ static float unpack(T v) {
    return __half2float(v);
  }
};
// gfx950 converts to bfloat16 in hardware (v_cvt_pk_bf16_f32, two values per
// instruction), rounding to nearest even: the same bits as f32_to_bf16_bits
// for every finite value.
__device__ __forceinline__ uint32_t bf16x2_bits(float a, float b) {
  typedef float f32x2 __attribute__((ext_vector_type(2)));
  typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
  const bf16x2 r = __builtin_convertvector((f32x2){a, b}, bf16x2);
  uint32_t u;
  __builtin_memcpy(&u, &r, 4);
  return u;
}

template <>
struct StoreT<kBF16> {
  using T = bf16_t;
  __device__ __forceinline__ static T pack(float f) {
    bf16_t b;
    b.u = (uint16_t)bf16x2_bits(f, 0.0f);
    return b;
  }
  __device__ __forceinline__ static float unpack(T v) {
    return __uint_as_float(((uint32_t)v.u) << 16);
  }
};

// Filtered-projection ("slab") layout: [k][t][y / YB][x][y % YB]. A voxel
// column needs, per tilt, the two taps (x, x+1) of YB consecutive detector
// rows: 2 * YB contiguous values.
__device__ __forceinline__ size_t slab_index(int k, int t, int y, int x,
                                             int T, int NYB, int N, int YB) {
  return (((size_t)k * T + t) * NYB + (size_t)(y / YB)) * (size_t)N * YB +
         (size_t)x * YB + (size_t)(y % YB);
}

// Workgroups are dispatched round-robin over the 8 XCDs of the chip. This
// maps the hardware block id to a logical id so that each XCD receives one
// contiguous run of logical blocks (neighbouring tiles then share that
// XCD's L2).
__device__ __forceinline__ unsigned xcd_remap(unsigned bid, unsigned nb) {
  constexpr unsigned kXcd = 8;
  const unsigned q = nb / kXcd, r = nb % kXcd;
  const unsigned x = bid % kXcd, i = bid / kXcd;
  return x < r ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

// Block-wide max of a non-negative float; result valid in all threads.
template <int NTHREADS>
__device__ __forceinline__ float block_max(float v, float *scratch) {
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) scratch[wave] = v;
  lds_barrier();
  float r = 0.0f;
  constexpr int NW = (NTHREADS + 63) / 64;
#pragma unroll
  for (int w = 0; w < NW; ++w) r = fmaxf(r, scratch[w]);
  lds_barrier();
  return r;
}

}  // namespace gx
