// Voxel-driven backprojection on gfx950.
//
// Reference semantics (src/libguanaco/reconstructor.cpp:14-90): for every
// voxel (x, z) of a slice and every tilt t, p = x c - z s + offset; with
// several defocus slabs the slab is floor(h * dscale + doffset + 0.5) clamped,
// h = -x s - z c; the value (1-t) row[i0] + t row[i0+1] with i0 = floor(p -
// 0.5) is added, times the tilt's scale, only when 0 <= i0 < N-1.
//
// The geometry of a voxel column does not depend on the detector row y, so
// each thread owns one (x, z) and YB consecutive rows: the coordinate work is
// paid once per (voxel column, tilt) and the taps of all YB rows are one
// contiguous 2*YB-value read from the [k][t][y/YB][x][y%YB] layout. A
// workgroup is a 32 (x) by 8 (z) tile; consecutive logical workgroups are
// z-neighbours and are kept on one XCD so their overlapping row segments hit
// the same L2. The volume is written once, [y][z][x] with x fastest.
//
// kF32 reproduces the reference's FP32 operation order exactly (compiled
// without FMA contraction): per-tilt constants computed on the host as the
// reference does, tilt-outer accumulation order, no reassociation. kF16 and
// kBF16 (whose error budget is set by the storage) compute p with two fused
// multiply-adds from a table holding offset - 1/2 and the scale times the
// normalisation the filter pass divided out, and accumulate with FP32 fused
// multiply-adds (BF16 taps rounded to FP16 first: bf16_tap). The per-tilt
// tables are read with scalar loads.
#include "gx_common.hpp"
#include "gx_kernels.hpp"

namespace gx {

constexpr int kTX = 32, kTZ = 8;

// Per-tilt tables are read through the constant address space: uniform
// indices become scalar loads, independent of the workgroup barriers.
typedef const __attribute__((address_space(4))) float kfloat;
typedef const __attribute__((address_space(4))) int kint;

__device__ __forceinline__ float4 kload(const float4 *p, size_t i) {
  kfloat *q = (kfloat *)p + 4 * i;
  return make_float4(q[0], q[1], q[2], q[3]);
}
__device__ __forceinline__ int2 kload(const int2 *p, size_t i) {
  kint *q = (kint *)p + 2 * i;
  return make_int2(q[0], q[1]);
}
__device__ __forceinline__ float kload(const float *p, size_t i) {
  return ((kfloat *)p)[i];
}

template <int YB, int STORE>
struct Taps;

// BF16 taps enter the arithmetic rounded to FP16, as the staged kernel keeps
// them in LDS (bf16_pair_to_f16): exact for magnitudes in [2^-14, 65504]
// (eight significant bits fit FP16's eleven), and within 2^-25 below, where
// FP16 is subnormal. The normalised slabs stay far below 65504, as they must
// for FP16 storage. The staged kernel then interpolates with
// v_fma_mix_f32, one instruction per tap, like FP16, where BF16 would need a
// shift or mask before each FMA; the direct kernels round the same way so
// that every kernel gives the same bits.
__device__ __forceinline__ float bf16_tap(float v) {
  return (float)(_Float16)v;
}
__device__ __forceinline__ uint32_t bf16_pair_to_f16(uint32_t w) {
  typedef float f32x2 __attribute__((ext_vector_type(2)));
  typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
  const f32x2 f = {__uint_as_float(w << 16), __uint_as_float(w & 0xffff0000u)};
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(f, f16x2));
}
template <int STORE>
__device__ __forceinline__ float tap_value(typename StoreT<STORE>::T v) {
  if constexpr (STORE == kBF16) {
    return bf16_tap(StoreT<STORE>::unpack(v));
  } else {
    return StoreT<STORE>::unpack(v);
  }
}

template <int YB>
struct Taps<YB, kF32> {
  float v0[YB], v1[YB];
  // 16-byte word i of each tap: rows 4i .. 4i + 3
  __device__ __forceinline__ void unpack(int i, uint4 a, uint4 b) {
    v0[4 * i] = __uint_as_float(a.x); v0[4 * i + 1] = __uint_as_float(a.y);
    v0[4 * i + 2] = __uint_as_float(a.z); v0[4 * i + 3] = __uint_as_float(a.w);
    v1[4 * i] = __uint_as_float(b.x); v1[4 * i + 1] = __uint_as_float(b.y);
    v1[4 * i + 2] = __uint_as_float(b.z); v1[4 * i + 3] = __uint_as_float(b.w);
  }
  // The YB values of each tap as YB / 4 16-byte words at q0 and at q1
  // (global memory).
  __device__ __forceinline__ void load16(const uint4 *q0, const uint4 *q1) {
    static_assert(YB % 4 == 0, "16-byte taps need YB % 4 == 0");
#pragma unroll
    for (int i = 0; i < YB / 4; ++i) unpack(i, q0[i], q1[i]);
  }
  // word-major runs of the staged kernel: word i of position p at q[i SW]
  template <int SW>
  __device__ __forceinline__ void load16_strided(const uint4 *q) {
#pragma unroll
    for (int i = 0; i < YB / 4; ++i) unpack(i, q[i * SW], q[i * SW + 1]);
  }
  __device__ __forceinline__ void load(const float *p) {
    if constexpr (YB % 4 == 0) {
      const uint4 *q = (const uint4 *)__builtin_assume_aligned(p, 16);
      load16(q, q + YB / 4);
    } else {
#pragma unroll
      for (int i = 0; i < YB; ++i) {
        v0[i] = p[i];
        v1[i] = p[YB + i];
      }
    }
  }
};

template <int YB, int STORE>
struct Taps {
  using T = typename StoreT<STORE>::T;
  static_assert(YB % 8 == 0, "reduced storage needs YB % 8 == 0");
  float v0[YB], v1[YB];
  __device__ __forceinline__ void load(const T *p) {
    const uint4 *q = (const uint4 *)__builtin_assume_aligned(p, 16);
    load16(q, q + YB / 8);
  }
  // 16-byte word i of each tap: rows 8i .. 8i + 7, two to a 32-bit word
  __device__ __forceinline__ void unpack(int i, uint4 a, uint4 b) {
    const uint32_t wa[4] = {a.x, a.y, a.z, a.w};
    const uint32_t wb[4] = {b.x, b.y, b.z, b.w};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      v0[8 * i + 2 * j] = lo(wa[j]);
      v0[8 * i + 2 * j + 1] = hi(wa[j]);
      v1[8 * i + 2 * j] = lo(wb[j]);
      v1[8 * i + 2 * j + 1] = hi(wb[j]);
    }
  }
  __device__ __forceinline__ void load16(const uint4 *q0, const uint4 *q1) {
#pragma unroll
    for (int i = 0; i < YB / 8; ++i) unpack(i, q0[i], q1[i]);
  }
  // The staged kernel's word-major runs: word i of position p at q[i SW]
  // (q at position p), position p + 1 one word on.
  template <int SW>
  __device__ __forceinline__ void load16_strided(const uint4 *q) {
#pragma unroll
    for (int i = 0; i < YB / 8; ++i) unpack(i, q[i * SW], q[i * SW + 1]);
  }
  __device__ __forceinline__ static float lo(uint32_t w) {
    if constexpr (STORE == kF16) {
      return __half2float(__ushort_as_half((unsigned short)(w & 0xffffu)));
    } else {
      return bf16_tap(__uint_as_float(w << 16));
    }
  }
  __device__ __forceinline__ static float hi(uint32_t w) {
    if constexpr (STORE == kF16) {
      return __half2float(__ushort_as_half((unsigned short)(w >> 16)));
    } else {
      return bf16_tap(__uint_as_float(w & 0xffff0000u));
    }
  }
};

// reconstructor.cpp:56-58: half-integer voxel centres (exact in float).
__device__ __forceinline__ float voxel_coord(int i, int n) {
  return (float)((double)i - 0.5 * (double)n + 0.5);
}

// reconstructor.cpp:59-60,76-79: detector position p = x c - z s + offset,
// i0 = floor(p - 0.5), t = (p - 0.5) - i0. EXACT: the reference's operation
// order; otherwise cf.z already holds offset - 1/2 and p is two fused
// multiply-adds.
template <bool EXACT>
__device__ __forceinline__ int tap_index(float x, float z, float4 cf,
                                         float &tt) {
  float p;
  if constexpr (EXACT) {
    p = x * cf.x - z * cf.y + cf.z;
    p = p - 0.5f;
  } else {
    p = fmaf(x, cf.x, fmaf(-z, cf.y, cf.z));
  }
  const float fi = floorf(p);
  tt = p - fi;
  return (int)fi;
}

// reconstructor.cpp:66-73: slab k = clamp(floor(h * dscale + doffset + 0.5)),
// h = -x s - z c. floor((double)d + 0.5) without FP64: round half to even,
// then move exact ties up.
__device__ __forceinline__ float slab_coord(float x, float z, float4 cf,
                                            float dscale, float doffset) {
  const float h = -x * cf.y - z * cf.x;
  return h * dscale + doffset;
}
__device__ __forceinline__ int slab_of(float x, float z, float4 cf,
                                       float dscale, float doffset, int ndf) {
  const float d = slab_coord(x, z, cf, dscale, doffset);
  float n = rintf(d);
  if (d - n == 0.5f) n += 1.0f;
  return min(max((int)n, 0), ndf - 1);
}

// ------------------------------------------------------- per-tilt tables --

// out[t] = (c, s, offset - 1/2, scale * normalisation): the normalisation
// float(tmax[t]) * nrm_mul is what the filter pass divided the stored values
// by (1 where it is not positive, or without tmax). Entries past T are zero.
__global__ void k_bp_coefficients(const float4 *__restrict__ coef,
                                  const uint32_t *__restrict__ tmax,
                                  float nrm_mul, int T, int Tpad,
                                  float4 *__restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= Tpad) return;
  float4 r = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (t < T) {
    const float4 cf = coef[t];
    float nrm = tmax != nullptr ? __uint_as_float(tmax[t]) * nrm_mul : 1.0f;
    if (!(nrm > 0.0f)) nrm = 1.0f;
    r = make_float4(cf.x, cf.y, cf.z - 0.5f, cf.w * nrm);
  }
  out[t] = r;
}

hipError_t launch_bp_coefficients(const float4 *coef, const uint32_t *tmax,
                                  float nrm_mul, int T, float4 *out,
                                  hipStream_t s) {
  const int Tpad = bp_tilt_pad(T);
  hipLaunchKernelGGL(k_bp_coefficients, dim3((Tpad + 63) / 64), dim3(64), 0, s,
                     coef, tmax, nrm_mul, T, Tpad, out);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// LDS-staged form.
//
// p is a chain of correctly rounded float operations, each monotone in x and
// in z, so over a 32 x 8 tile i0 is smallest at one of the four corners, and
// the exact range of p is at most 31|c| + 7|s| <= 31.79 sqrt(c^2 + s^2)
// (c^2 + s^2 = 1 / (cos^2 + sin^2) to rounding): the taps of one tilt lie in
// a window of 34 consecutive positions; kWPOS leaves a margin for rounding.
// Likewise h spans at most 31.79 units, so a tile touches at most
// ceil(31.79 |dscale|) + 1 consecutive slabs; the host picks KS, the slabs
// staged per tilt, no smaller than that. The window base and first slab of
// every (tile, tilt) come from a table computed once per plan with the same
// arithmetic as the kernel. For a batch of TB tilts the workgroup copies, per
// tilt and staged slab, kWPOS positions of its row block starting at the
// window base (one contiguous run of the slab, YB rows per position) into
// LDS, then interpolates from LDS with the same arithmetic per voxel as the
// direct kernel. LDS indices are clamped, so coefficients that break the
// bound (non-finite angles, which the host rejects) can only give wrong
// values, never an access outside the buffer.
//
// The runs are loaded straight into LDS (buffer_load ... lds), not through
// registers; BF16 runs are then converted to FP16 in place by the lanes that
// loaded them (bf16_tap). Two LDS buffers: the loads of batch b + 1 are
// issued before batch b is interpolated, one barrier per batch.
constexpr int kWPOS = 40;

template <int YB, int STORE, int KS>
struct Staging {
  using T = typename StoreT<STORE>::T;
  static constexpr int kPosBytes = YB * (int)sizeof(T);
  static constexpr int kPosWords = kPosBytes / 16;      // 16-byte words
  static constexpr int kTB = 256 / (kPosBytes * KS);    // tilts per batch
  static constexpr int kNSlot = kTB * KS;  // (tilt, slab) runs per batch
  static constexpr int kChunks = kWPOS * kPosWords;  // words per run
  static constexpr int kCPL = (kChunks + 63) / 64;   // words per lane per run
  static constexpr int kSPW = kNSlot / (kTX * kTZ / 64);  // runs per wave
  // buffer_load ... lds writes lane l of a load to slot l of a dense 1 KiB
  // block, so a run is kRunSlots 16-byte slots, slot s < kChunks holding
  // word s / kWPOS of position s % kWPOS. In this word-major order the reads
  // of lanes at consecutive positions are consecutive 16-byte words, which
  // gfx950 serves without bank conflicts (as it does a padded odd stride).
  static constexpr int kRunSlots = kCPL * 64;
  static constexpr int kBufWords = kNSlot * kRunSlots;
  static_assert(kPosBytes % 16 == 0, "a position is whole 16-byte words");
  static_assert(kTB >= 1 && kBpTiltPad % kTB == 0,
                "batches must tile the padded tilt tables");
  static_assert(kNSlot % (kTX * kTZ / 64) == 0, "runs split evenly on waves");
};

// One 16-byte raw buffer load per lane, written to LDS at dst + 16 lane
// instead of to registers (M0 holds dst). Out-of-range offsets write zeros.
__device__ __forceinline__ void load_to_lds16(__amdgpu_buffer_rsrc_t rs,
                                              uint4 *dst, int off) {
  __builtin_amdgcn_raw_ptr_buffer_load_lds(
      rs, (__attribute__((address_space(3))) void *)dst, 16, off, 0, 0, 0);
}

// Window of one tilt over the tile's active voxels [x0, x1] x [z0, z1]:
// smallest i0 and smallest k over the four corners, with kOneSlab set in .y
// when the largest k over the corners is the same: slab_of is monotone in x
// and in z like p, so every voxel of the tile then takes slab k.
constexpr int kOneSlab = 1 << 30;
constexpr int kSlabMask = kOneSlab - 1;

template <bool EXACT, bool MULTI>
__device__ __forceinline__ int2 tile_window(float4 cf, float x0, float x1,
                                            float z0, float z1, float dscale,
                                            float doffset, int ndf) {
  int w = 0x7fffffff, k = 0x7fffffff, kmax = -1;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const float x = (c & 1) ? x1 : x0;
    const float z = (c & 2) ? z1 : z0;
    float tt;
    w = min(w, tap_index<EXACT>(x, z, cf, tt));
    if constexpr (MULTI) {
      const int kc = slab_of(x, z, cf, dscale, doffset, ndf);
      k = min(k, kc);
      kmax = max(kmax, kc);
    }
  }
  return make_int2(w, MULTI ? (k | (kmax == k ? kOneSlab : 0)) : 0);
}

// win[tile * Tpad + t], tile = bx * nbz + bz (the kernels' order).
template <bool EXACT, bool MULTI>
__global__ void k_bp_windows(const float4 *__restrict__ coef, int T, int Tpad,
                             int W, int Z, int nbx, int nbz, float dscale,
                             float doffset, int ndf, int2 *__restrict__ win) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= (size_t)nbx * nbz * Tpad) return;
  const int t = (int)(i % Tpad);
  const int tile = (int)(i / Tpad);
  const int bz = tile % nbz, bx = tile / nbz;
  int2 r = make_int2(0, 0);
  if (t < T) {
    const float x0 = voxel_coord(bx * kTX, W);
    const float x1 = voxel_coord(min(bx * kTX + kTX, W) - 1, W);
    const float z0 = voxel_coord(bz * kTZ, Z);
    const float z1 = voxel_coord(min(bz * kTZ + kTZ, Z) - 1, Z);
    r = tile_window<EXACT, MULTI>(coef[t], x0, x1, z0, z1, dscale, doffset,
                                  ndf);
  }
  win[i] = r;
}

size_t bp_window_entries(int W, int Z, int T) {
  const size_t nbx = (W + kTX - 1) / kTX, nbz = (Z + kTZ - 1) / kTZ;
  return nbx * nbz * (size_t)bp_tilt_pad(T);
}

hipError_t launch_bp_windows(const float4 *coef, int exact, int T, int W,
                             int Z, float dscale, float doffset, int ndf,
                             int2 *win, hipStream_t s) {
  const int nbx = (W + kTX - 1) / kTX, nbz = (Z + kTZ - 1) / kTZ;
  const size_t n = bp_window_entries(W, Z, T);
  const dim3 g((unsigned)((n + 255) / 256)), b(256);
  const int Tpad = bp_tilt_pad(T);
  const bool multi = ndf > 1;
  if (exact && multi)
    hipLaunchKernelGGL((k_bp_windows<true, true>), g, b, 0, s, coef, T, Tpad,
                       W, Z, nbx, nbz, dscale, doffset, ndf, win);
  else if (exact)
    hipLaunchKernelGGL((k_bp_windows<true, false>), g, b, 0, s, coef, T, Tpad,
                       W, Z, nbx, nbz, dscale, doffset, ndf, win);
  else if (multi)
    hipLaunchKernelGGL((k_bp_windows<false, true>), g, b, 0, s, coef, T, Tpad,
                       W, Z, nbx, nbz, dscale, doffset, ndf, win);
  else
    hipLaunchKernelGGL((k_bp_windows<false, false>), g, b, 0, s, coef, T, Tpad,
                       W, Z, nbx, nbz, dscale, doffset, ndf, win);
  return hipGetLastError();
}

// ------------------------------------------------------------ kernels -----

// Accumulate one tilt's taps (reconstructor.cpp:82-84).
template <int YB, int STORE>
__device__ __forceinline__ void accumulate(float (&acc)[YB],
                                           const Taps<YB, STORE> &tp,
                                           float4 cf, float tt) {
  if constexpr (STORE == kF32) {
    const float w0 = 1.0f - tt;
#pragma unroll
    for (int r = 0; r < YB; ++r)
      acc[r] = acc[r] + cf.w * (w0 * tp.v0[r] + tt * tp.v1[r]);
  } else {
    const float w1 = cf.w * tt;  // cf.w: scale x normalisation
    const float w0 = cf.w - w1;
#pragma unroll
    for (int r = 0; r < YB; ++r) {
      acc[r] = fmaf(w0, tp.v0[r], acc[r]);
      acc[r] = fmaf(w1, tp.v1[r], acc[r]);
    }
  }
}

template <int YB, int STORE, bool MULTI>
__global__ __launch_bounds__(kTX *kTZ) void k_bp(
    const typename StoreT<STORE>::T *__restrict__ slab,
    const float4 *__restrict__ coef, float dscale, float doffset, int ndf,
    int N, int T, int NYB, int W, int Z, int Y, int yb0, int y_base,
    int slab_yb0,
    float *__restrict__ vol, int nbx, int nbz, unsigned nblocks) {
  constexpr bool kExact = STORE == kF32;
  const unsigned lb = xcd_remap(blockIdx.x, nblocks);
  const int bz = (int)(lb % (unsigned)nbz);
  const int bx = (int)((lb / (unsigned)nbz) % (unsigned)nbx);
  const int byb = yb0 + (int)(lb / ((unsigned)nbz * (unsigned)nbx));
  const int ix = bx * kTX + (int)(threadIdx.x % kTX);
  const int iz = bz * kTZ + (int)(threadIdx.x / kTX);
  const bool active = ix < W && iz < Z;
  const float x = voxel_coord(ix, W);
  const float z = voxel_coord(iz, Z);

  float acc[YB];
#pragma unroll
  for (int r = 0; r < YB; ++r) acc[r] = 0.0f;

  const size_t row_len = (size_t)N * YB;
  const size_t kstride = (size_t)T * NYB * row_len;
  const typename StoreT<STORE>::T *yblock = slab + (size_t)(byb - slab_yb0) * row_len;

  for (int t = 0; t < T; ++t) {
    const float4 cf = kload(coef, t);
    int k = 0;
    if constexpr (MULTI) k = slab_of(x, z, cf, dscale, doffset, ndf);
    float tt;
    const int i0 = tap_index<kExact>(x, z, cf, tt);
    if (active && (unsigned)i0 < (unsigned)(N - 1)) {
      Taps<YB, STORE> tp;
      tp.load(yblock + (size_t)t * NYB * row_len + (size_t)k * kstride +
              (size_t)i0 * YB);
      accumulate<YB, STORE>(acc, tp, cf, tt);
    }
  }
  if (active) {
#pragma unroll
    for (int r = 0; r < YB; ++r) {
      const int y = byb * YB + r;
      if (y < Y) vol[((size_t)(y - y_base) * Z + iz) * W + ix] = acc[r];
    }
  }
}

// Per-row rotation centres (one centre per slice, mp.py:144-154): the
// direct kernel with each row's own offset, so each row has its own taps.
// The slab depends on h only, which no centre changes. kF32 keeps the
// reference's order: p = (x c - z s) + offset_y, then p - 1/2.
template <int YB, int STORE, bool MULTI>
__global__ __launch_bounds__(kTX *kTZ) void k_bp_rows(
    const typename StoreT<STORE>::T *__restrict__ slab,
    const float4 *__restrict__ coef, const float *__restrict__ row_off,
    int Y_tab, float dscale, float doffset, int ndf, int N, int T, int NYB,
    int W, int Z, int Y, int yb0, int y_base, int slab_yb0,
    float *__restrict__ vol,
    int nbx, int nbz, unsigned nblocks) {
  constexpr bool kExact = STORE == kF32;
  using ST = StoreT<STORE>;
  const unsigned lb = xcd_remap(blockIdx.x, nblocks);
  const int bz = (int)(lb % (unsigned)nbz);
  const int bx = (int)((lb / (unsigned)nbz) % (unsigned)nbx);
  const int byb = yb0 + (int)(lb / ((unsigned)nbz * (unsigned)nbx));
  const int ix = bx * kTX + (int)(threadIdx.x % kTX);
  const int iz = bz * kTZ + (int)(threadIdx.x / kTX);
  const bool active = ix < W && iz < Z;
  const float x = voxel_coord(ix, W);
  const float z = voxel_coord(iz, Z);
  // rows of this block that exist (the same for the whole workgroup)
  const int nr = min(YB, Y - byb * YB);

  float acc[YB];
#pragma unroll
  for (int r = 0; r < YB; ++r) acc[r] = 0.0f;

  const size_t row_len = (size_t)N * YB;
  const size_t kstride = (size_t)T * NYB * row_len;
  const typename ST::T *yblock = slab + (size_t)(byb - slab_yb0) * row_len;

  for (int t = 0; t < T; ++t) {
    const float4 cf = kload(coef, t);
    int k = 0;
    if constexpr (MULTI) k = slab_of(x, z, cf, dscale, doffset, ndf);
    const typename ST::T *base =
        yblock + (size_t)t * NYB * row_len + (size_t)k * kstride;
    const float xc = x * cf.x - z * cf.y;  // kF32: the reference's first part
    const size_t o = (size_t)t * Y_tab + (size_t)byb * YB;
#pragma unroll
    for (int r = 0; r < YB; ++r) {
      if (r >= nr) break;
      const float off = kload(row_off, o + r);
      float p;
      if constexpr (kExact) {
        p = xc + off;
        p = p - 0.5f;
      } else {
        p = fmaf(x, cf.x, fmaf(-z, cf.y, off));
      }
      const float fi = floorf(p);
      const float tt = p - fi;
      const int i0 = (int)fi;
      if (active && (unsigned)i0 < (unsigned)(N - 1)) {
        const float v0 = tap_value<STORE>(base[(size_t)i0 * YB + r]);
        const float v1 = tap_value<STORE>(base[(size_t)(i0 + 1) * YB + r]);
        if constexpr (kExact) {
          acc[r] = acc[r] + cf.w * ((1.0f - tt) * v0 + tt * v1);
        } else {
          const float w1 = cf.w * tt;
          acc[r] = fmaf(cf.w - w1, v0, acc[r]);
          acc[r] = fmaf(w1, v1, acc[r]);
        }
      }
    }
  }
  if (active) {
#pragma unroll
    for (int r = 0; r < YB; ++r) {
      const int y = byb * YB + r;
      if (y < Y) vol[((size_t)(y - y_base) * Z + iz) * W + ix] = acc[r];
    }
  }
}

template <int YB, int STORE, bool MULTI, int KS>
__global__ __launch_bounds__(kTX *kTZ) void k_bp_lds(
    const typename StoreT<STORE>::T *__restrict__ slab,
    const float4 *__restrict__ coef, const int2 *__restrict__ win,
    float dscale, float doffset, int ndf, int N, int T, int NYB, int W, int Z,
    int Y, int yb0, int y_base, int slab_yb0, float *__restrict__ vol,
    int nbx, int nbz,
    unsigned nblocks) {
  using S = Staging<YB, STORE, KS>;
  using ST = typename StoreT<STORE>::T;
  constexpr bool kExact = STORE == kF32;
  // BF16 runs are staged as FP16 (bf16_tap)
  constexpr int kLds = STORE == kBF16 ? kF16 : STORE;
  static_assert(MULTI || KS == 1, "one slab without defocus slabs");
  // Two buffers as two arrays: with distinct objects the compiler can tell
  // that reads of one do not wait for the direct loads into the other.
  __shared__ uint4 lds_a[S::kBufWords];
  __shared__ uint4 lds_b[S::kBufWords];

  const unsigned lb = xcd_remap(blockIdx.x, nblocks);
  const int bz = (int)(lb % (unsigned)nbz);
  const int bx = (int)((lb / (unsigned)nbz) % (unsigned)nbx);
  const int byb = yb0 + (int)(lb / ((unsigned)nbz * (unsigned)nbx));
  const int lane = (int)(threadIdx.x & 63), wave = (int)(threadIdx.x >> 6);
  const int ix = bx * kTX + (int)(threadIdx.x % kTX);
  const int iz = bz * kTZ + (int)(threadIdx.x / kTX);
  const bool active = ix < W && iz < Z;
  const float x = voxel_coord(ix, W);
  const float z = voxel_coord(iz, Z);
  const int2 *tw = win + (size_t)(bx * nbz + bz) * bp_tilt_pad(T);

  const size_t row_len = (size_t)N * YB;
  const size_t tstride = (size_t)NYB * row_len;
  const size_t kstride = (size_t)T * tstride;
  const ST *yblock = slab + (size_t)(byb - slab_yb0) * row_len;

  // The staging runs of a wave are wave-uniform: with the wave index taken
  // from the first lane, each run's window entry is a scalar load and its
  // buffer descriptor lives in scalar registers.
  const int wave_u = __builtin_amdgcn_readfirstlane(wave);
  // each lane's byte offset within a run past its window base (-1: the
  // slots past the run), the same for every batch
  int dslot[S::kCPL];
#pragma unroll
  for (int c = 0; c < S::kCPL; ++c) {
    const int s = lane + 64 * c;
    const int wi = s / kWPOS, p = s - wi * kWPOS;
    dslot[c] = s < S::kChunks
                   ? (p * YB + wi * (16 / (int)sizeof(ST))) * (int)sizeof(ST)
                   : -1;
  }
  auto load_batch = [&](int tb, uint4 *dst) {
#pragma unroll
    for (int i = 0; i < S::kSPW; ++i) {
      const int q = wave_u + (kTX * kTZ / 64) * i;
      const int j = q / KS, sl = q % KS;
      const int t = tb + j;
      const int2 w = kload(tw, t);
      const int k = MULTI ? (w.y & kSlabMask) + sl : 0;
      // a tile within one slab at this tilt reads only its first run
      const bool ok = t < T && k < ndf && (sl == 0 || !(w.y & kOneSlab));
      const ST *src = yblock + (size_t)t * tstride + (size_t)k * kstride;
      // One range-checked raw buffer per run: its N positions of YB values.
      // A load at an offset outside [0, N * YB * sizeof(T)) returns zeros
      // and touches no memory, so positions outside [0, N), the slots past
      // the window (offset -1) and runs past the last tilt or slab (no
      // records) all read as zero without branches.
      const auto rs = __builtin_amdgcn_make_buffer_rsrc(
          (void *)(ok ? src : yblock), (short)0,
          ok ? (int)(row_len * sizeof(ST)) : 0, 0x00020000);
      // 64 slots of the run per load, lane l to slot l of its 1 KiB
      const int base = w.x * YB * (int)sizeof(ST);
#pragma unroll
      for (int c = 0; c < S::kCPL; ++c)
        load_to_lds16(rs, dst + q * S::kRunSlots + 64 * c,
                      dslot[c] >= 0 ? base + dslot[c] : -1);
    }
  };
  // After this wave's loads have landed: BF16 slots to FP16, each by the
  // lane that loaded it (no other wave touches them before the barrier).
  auto convert_batch = [&](uint4 *dst) {
    if constexpr (STORE == kBF16) {
#pragma unroll
      for (int i = 0; i < S::kSPW; ++i) {
        const int q = wave + (kTX * kTZ / 64) * i;
#pragma unroll
        for (int c = 0; c < S::kCPL; ++c) {
          uint4 *d = dst + q * S::kRunSlots + lane + 64 * c;
          const uint4 v = *d;
          *d = make_uint4(bf16_pair_to_f16(v.x), bf16_pair_to_f16(v.y),
                          bf16_pair_to_f16(v.z), bf16_pair_to_f16(v.w));
        }
      }
    }
  };

  float acc[YB];
#pragma unroll
  for (int r = 0; r < YB; ++r) acc[r] = 0.0f;

  // One batch: prefetch the next into `nxt`, interpolate from `cur`.
  auto batch = [&](int b, const uint4 *cur, uint4 *nxt) {
    const int tb = b * S::kTB;
    const bool more = tb + S::kTB < T;
    if (more) load_batch(tb + S::kTB, nxt);
#pragma unroll
    for (int j = 0; j < S::kTB; ++j) {
      const int t = tb + j;
      if (t < T) {
        const float4 cf = kload(coef, t);
        const int2 w = kload(tw, t);
        int ks = 0;
        if constexpr (MULTI) {
          // A tile within one slab at this tilt needs no per-voxel slab
          // choice (w is the workgroup's own entry: the branch is uniform).
          if (!(w.y & kOneSlab)) {
            const int kf = w.y & kSlabMask;
            if constexpr (KS == 2) {
              // The tile's voxels take slab kf or kf + 1 (<= ndf - 1, as the
              // tile spans two), and floor(d + 1/2) >= kf + 1 exactly when
              // d >= kf + 1/2: one comparison of the float d with a
              // representable bound decides as slab_of does.
              const float d = slab_coord(x, z, cf, dscale, doffset);
              ks = d >= (float)kf + 0.5f ? 1 : 0;
            } else {
              ks = slab_of(x, z, cf, dscale, doffset, ndf) - kf;
              ks = min(max(ks, 0), KS - 1);
            }
          }
        }
        float tt;
        const int i0 = tap_index<kExact>(x, z, cf, tt);
        if (active && (unsigned)i0 < (unsigned)(N - 1)) {
          const int rel = min(max(i0 - w.x, 0), kWPOS - 2);
          // word-major run: word wi of position p in slot wi kWPOS + p
          Taps<YB, kLds> tp;
          tp.template load16_strided<kWPOS>(cur + (j * KS + ks) * S::kRunSlots +
                                            rel);
          accumulate<YB, kLds>(acc, tp, cf, tt);
        }
      }
    }
    if (more) {
      // this wave's loads have landed; the barrier then covers the other
      // waves' too
      asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
      convert_batch(nxt);
    }
    lds_barrier();
  };

  load_batch(0, lds_a);
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  convert_batch(lds_a);
  lds_barrier();
  const int nb = (T + S::kTB - 1) / S::kTB;
  for (int b = 0; b < nb; b += 2) {
    batch(b, lds_a, lds_b);
    if (b + 1 < nb) batch(b + 1, lds_b, lds_a);
  }
  if (active) {
#pragma unroll
    for (int r = 0; r < YB; ++r) {
      const int y = byb * YB + r;
      if (y < Y) vol[((size_t)(y - y_base) * Z + iz) * W + ix] = acc[r];
    }
  }
}

// ------------------------------------------------------------ launchers ---
template <int YB, int STORE, bool MULTI, int KS>
static hipError_t launch_bp_lds_k(const BpArgs &a, hipStream_t s) {
  using ST = typename StoreT<STORE>::T;
  const int nbx = (a.W + kTX - 1) / kTX;
  const int nbz = (a.Z + kTZ - 1) / kTZ;
  if (a.yb1 <= a.yb0) return hipSuccess;
  if (a.win == nullptr) return hipErrorInvalidValue;
  const unsigned nb =
      (unsigned)nbx * (unsigned)nbz * (unsigned)(a.yb1 - a.yb0);
  hipLaunchKernelGGL((k_bp_lds<YB, STORE, MULTI, KS>), dim3(nb),
                     dim3(kTX * kTZ), 0, s, (const ST *)a.slab, a.coef, a.win,
                     a.dscale, a.doffset, a.ndf, a.N, a.T, a.NYB, a.W, a.Z,
                     a.Y, a.yb0, a.y_base, a.slab_yb0, a.vol, nbx, nbz, nb);
  return hipGetLastError();
}

template <int YB, int STORE>
static hipError_t launch_bp_lds_t(const BpArgs &a, hipStream_t s) {
  if (a.ndf == 1) return launch_bp_lds_k<YB, STORE, false, 1>(a, s);
  switch (a.staged) {
    case 2: return launch_bp_lds_k<YB, STORE, true, 2>(a, s);
    case 4: return launch_bp_lds_k<YB, STORE, true, 4>(a, s);
    default: return hipErrorInvalidValue;
  }
}

template <int YB, int STORE>
static hipError_t launch_bp_t(const BpArgs &a, hipStream_t s) {
  using ST = typename StoreT<STORE>::T;
  const int nbx = (a.W + kTX - 1) / kTX;
  const int nbz = (a.Z + kTZ - 1) / kTZ;
  if (a.yb1 <= a.yb0) return hipSuccess;
  const unsigned nb =
      (unsigned)nbx * (unsigned)nbz * (unsigned)(a.yb1 - a.yb0);
  if (a.ndf > 1) {
    hipLaunchKernelGGL((k_bp<YB, STORE, true>), dim3(nb), dim3(kTX * kTZ), 0,
                       s, (const ST *)a.slab, a.coef, a.dscale, a.doffset,
                       a.ndf, a.N, a.T, a.NYB, a.W, a.Z, a.Y, a.yb0, a.y_base, a.slab_yb0,
                       a.vol, nbx, nbz, nb);
  } else {
    hipLaunchKernelGGL((k_bp<YB, STORE, false>), dim3(nb), dim3(kTX * kTZ), 0,
                       s, (const ST *)a.slab, a.coef, a.dscale, a.doffset,
                       a.ndf, a.N, a.T, a.NYB, a.W, a.Z, a.Y, a.yb0, a.y_base, a.slab_yb0,
                       a.vol, nbx, nbz, nb);
  }
  return hipGetLastError();
}

template <int YB, int STORE>
static hipError_t launch_bp_rows_t(const BpArgs &a, hipStream_t s) {
  using ST = typename StoreT<STORE>::T;
  const int nbx = (a.W + kTX - 1) / kTX;
  const int nbz = (a.Z + kTZ - 1) / kTZ;
  if (a.yb1 <= a.yb0) return hipSuccess;
  const unsigned nb =
      (unsigned)nbx * (unsigned)nbz * (unsigned)(a.yb1 - a.yb0);
  if (a.ndf > 1) {
    hipLaunchKernelGGL((k_bp_rows<YB, STORE, true>), dim3(nb),
                       dim3(kTX * kTZ), 0, s, (const ST *)a.slab, a.coef,
                       a.row_off, a.Y_tab, a.dscale, a.doffset, a.ndf, a.N,
                       a.T, a.NYB, a.W, a.Z, a.Y, a.yb0, a.y_base, a.slab_yb0, a.vol,
                       nbx, nbz, nb);
  } else {
    hipLaunchKernelGGL((k_bp_rows<YB, STORE, false>), dim3(nb),
                       dim3(kTX * kTZ), 0, s, (const ST *)a.slab, a.coef,
                       a.row_off, a.Y_tab, a.dscale, a.doffset, a.ndf, a.N,
                       a.T, a.NYB, a.W, a.Z, a.Y, a.yb0, a.y_base, a.slab_yb0, a.vol,
                       nbx, nbz, nb);
  }
  return hipGetLastError();
}

hipError_t launch_bp(const BpArgs &a, int YB, int store, hipStream_t s) {
  if (a.row_off != nullptr) {
    if (store == kF32) {
      if (YB == 8) return launch_bp_rows_t<8, kF32>(a, s);
      if (YB == 16) return launch_bp_rows_t<16, kF32>(a, s);
    } else if (store == kF16) {
      if (YB == 16) return launch_bp_rows_t<16, kF16>(a, s);
      if (YB == 32) return launch_bp_rows_t<32, kF16>(a, s);
    } else if (store == kBF16) {
      if (YB == 16) return launch_bp_rows_t<16, kBF16>(a, s);
      if (YB == 32) return launch_bp_rows_t<32, kBF16>(a, s);
    }
    return hipErrorInvalidValue;
  }
  if (a.staged) {
    if (store == kF32) {
      if (YB == 8) return launch_bp_lds_t<8, kF32>(a, s);
      if (YB == 16) return launch_bp_lds_t<16, kF32>(a, s);
    } else if (store == kF16) {
      if (YB == 16) return launch_bp_lds_t<16, kF16>(a, s);
      if (YB == 32) return launch_bp_lds_t<32, kF16>(a, s);
    } else if (store == kBF16) {
      if (YB == 16) return launch_bp_lds_t<16, kBF16>(a, s);
      if (YB == 32) return launch_bp_lds_t<32, kBF16>(a, s);
    }
    return hipErrorInvalidValue;
  }
  if (store == kF32) {
    if (YB == 1) return launch_bp_t<1, kF32>(a, s);
    if (YB == 8) return launch_bp_t<8, kF32>(a, s);
    if (YB == 16) return launch_bp_t<16, kF32>(a, s);
  } else if (store == kF16) {
    if (YB == 16) return launch_bp_t<16, kF16>(a, s);
    if (YB == 32) return launch_bp_t<32, kF16>(a, s);
  } else if (store == kBF16) {
    if (YB == 16) return launch_bp_t<16, kBF16>(a, s);
    if (YB == 32) return launch_bp_t<32, kBF16>(a, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace gx
