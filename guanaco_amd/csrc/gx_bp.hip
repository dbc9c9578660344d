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
// workgroup is a 64 (x) by 4 (z) tile; consecutive logical workgroups are
// z-neighbours and are kept on one XCD so their overlapping row segments hit
// the same L2. The volume is written once, [y][z][x] with x fastest.
//
// kF32 reproduces the reference's FP32 operation order exactly (compiled
// without FMA contraction): per-tilt constants computed on the host as the
// reference does, tilt-outer accumulation order, no reassociation. kF16 and
// kBF16 keep the same coordinate decisions, fold the normalisation into the
// per-tilt scale and accumulate with FP32 fused multiply-adds.
#include "gx_common.hpp"
#include "gx_kernels.hpp"

namespace gx {

constexpr int kTX = 64, kTZ = 4;

template <int YB, int STORE>
struct Taps;

template <int YB>
struct Taps<YB, kF32> {
  float v0[YB], v1[YB];
  __device__ __forceinline__ void load(const float *p) {
    if constexpr (YB % 4 == 0) {
      const float4 *q = (const float4 *)__builtin_assume_aligned(p, 16);
#pragma unroll
      for (int i = 0; i < YB / 4; ++i) {
        const float4 a = q[i], b = q[YB / 4 + i];
        v0[4 * i] = a.x; v0[4 * i + 1] = a.y; v0[4 * i + 2] = a.z; v0[4 * i + 3] = a.w;
        v1[4 * i] = b.x; v1[4 * i + 1] = b.y; v1[4 * i + 2] = b.z; v1[4 * i + 3] = b.w;
      }
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
#pragma unroll
    for (int i = 0; i < YB / 8; ++i) {
      const uint4 a = q[i], b = q[YB / 8 + i];
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
  }
  __device__ __forceinline__ static float lo(uint32_t w) {
    if constexpr (STORE == kF16) {
      return __half2float(__ushort_as_half((unsigned short)(w & 0xffffu)));
    } else {
      return __uint_as_float(w << 16);
    }
  }
  __device__ __forceinline__ static float hi(uint32_t w) {
    if constexpr (STORE == kF16) {
      return __half2float(__ushort_as_half((unsigned short)(w >> 16)));
    } else {
      return __uint_as_float(w & 0xffff0000u);
    }
  }
};

template <int YB, int STORE, bool MULTI>
__global__ __launch_bounds__(kTX *kTZ) void k_bp(
    const typename StoreT<STORE>::T *__restrict__ slab,
    const float4 *__restrict__ coef, const uint32_t *__restrict__ tmax,
    float nrm_mul, float dscale, float doffset, int ndf, int N, int T,
    int NYB, int W, int Z, int Y, int yb0, int y_base,
    float *__restrict__ vol, int nbx, int nbz, unsigned nblocks) {
  const unsigned lb = xcd_remap(blockIdx.x, nblocks);
  const int bz = (int)(lb % (unsigned)nbz);
  const int bx = (int)((lb / (unsigned)nbz) % (unsigned)nbx);
  const int byb = yb0 + (int)(lb / ((unsigned)nbz * (unsigned)nbx));
  const int ix = bx * kTX + (int)(threadIdx.x % kTX);
  const int iz = bz * kTZ + (int)(threadIdx.x / kTX);
  const bool active = ix < W && iz < Z;

  // reconstructor.cpp:56-58: half-integer voxel centres (exact in float)
  const float x = (float)((double)ix - 0.5 * (double)W + 0.5);
  const float z = (float)((double)iz - 0.5 * (double)Z + 0.5);

  float acc[YB];
#pragma unroll
  for (int r = 0; r < YB; ++r) acc[r] = 0.0f;

  const size_t row_len = (size_t)N * YB;
  const size_t kstride = (size_t)T * NYB * row_len;
  const typename StoreT<STORE>::T *yblock = slab + (size_t)byb * row_len;

  for (int t = 0; t < T; ++t) {
    const float4 cf = coef[t];
    float p = x * cf.x - z * cf.y + cf.z;
    int k = 0;
    if constexpr (MULTI) {
      const float h = -x * cf.y - z * cf.x;
      const float d = h * dscale + doffset;
      // floor((double)d + 0.5) without FP64: round half to even, then move
      // exact ties up.
      float n = rintf(d);
      if (d - n == 0.5f) n += 1.0f;
      k = min(max((int)n, 0), ndf - 1);
    }
    p = p - 0.5f;
    const float fi = floorf(p);
    const float tt = p - fi;
    const int i0 = (int)fi;
    if (active && (unsigned)i0 < (unsigned)(N - 1)) {
      const auto *src = yblock + (size_t)t * NYB * row_len +
                        (size_t)k * kstride + (size_t)i0 * YB;
      Taps<YB, STORE> tp;
      tp.load(src);
      if constexpr (STORE == kF32) {
        const float w0 = 1.0f - tt;
#pragma unroll
        for (int r = 0; r < YB; ++r)
          acc[r] = acc[r] + cf.w * (w0 * tp.v0[r] + tt * tp.v1[r]);
      } else {
        float nrm = __uint_as_float(tmax[t]) * nrm_mul;
        if (!(nrm > 0.0f)) nrm = 1.0f;
        const float sc = cf.w * nrm;
        const float w1 = sc * tt;
        const float w0 = sc - w1;
#pragma unroll
        for (int r = 0; r < YB; ++r) {
          acc[r] = fmaf(w0, tp.v0[r], acc[r]);
          acc[r] = fmaf(w1, tp.v1[r], acc[r]);
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
                       s, (const ST *)a.slab, a.coef, a.tmax, a.nrm_mul,
                       a.dscale, a.doffset, a.ndf, a.N, a.T, a.NYB, a.W, a.Z,
                       a.Y, a.yb0, a.y_base, a.vol, nbx, nbz, nb);
  } else {
    hipLaunchKernelGGL((k_bp<YB, STORE, false>), dim3(nb), dim3(kTX * kTZ), 0,
                       s, (const ST *)a.slab, a.coef, a.tmax, a.nrm_mul,
                       a.dscale, a.doffset, a.ndf, a.N, a.T, a.NYB, a.W, a.Z,
                       a.Y, a.yb0, a.y_base, a.vol, nbx, nbz, nb);
  }
  return hipGetLastError();
}

hipError_t launch_bp(const BpArgs &a, int YB, int store, hipStream_t s) {
  if (store == kF32) {
    if (YB == 1) return launch_bp_t<1, kF32>(a, s);
    if (YB == 8) return launch_bp_t<8, kF32>(a, s);
  } else if (store == kF16) {
    if (YB == 16) return launch_bp_t<16, kF16>(a, s);
  } else if (store == kBF16) {
    if (YB == 16) return launch_bp_t<16, kBF16>(a, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace gx
