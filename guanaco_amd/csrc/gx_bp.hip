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
  // The 2 * YB values as YB / 2 16-byte words (global memory or LDS).
  __device__ __forceinline__ void load16(const uint4 *q) {
    static_assert(YB % 4 == 0, "16-byte taps need YB % 4 == 0");
#pragma unroll
    for (int i = 0; i < YB / 4; ++i) {
      const uint4 a = q[i], b = q[YB / 4 + i];
      v0[4 * i] = __uint_as_float(a.x); v0[4 * i + 1] = __uint_as_float(a.y);
      v0[4 * i + 2] = __uint_as_float(a.z); v0[4 * i + 3] = __uint_as_float(a.w);
      v1[4 * i] = __uint_as_float(b.x); v1[4 * i + 1] = __uint_as_float(b.y);
      v1[4 * i + 2] = __uint_as_float(b.z); v1[4 * i + 3] = __uint_as_float(b.w);
    }
  }
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
    load16((const uint4 *)__builtin_assume_aligned(p, 16));
  }
  __device__ __forceinline__ void load16(const uint4 *q) {
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

// reconstructor.cpp:56-58: half-integer voxel centres (exact in float).
__device__ __forceinline__ float voxel_coord(int i, int n) {
  return (float)((double)i - 0.5 * (double)n + 0.5);
}

// reconstructor.cpp:59-60,76-79: detector position p = x c - z s + offset,
// i0 = floor(p - 0.5), t = (p - 0.5) - i0.
__device__ __forceinline__ int tap_index(float x, float z, float4 cf,
                                         float &tt) {
  float p = x * cf.x - z * cf.y + cf.z;
  p = p - 0.5f;
  const float fi = floorf(p);
  tt = p - fi;
  return (int)fi;
}

// reconstructor.cpp:66-73: slab k = clamp(floor(h * dscale + doffset + 0.5)),
// h = -x s - z c. floor((double)d + 0.5) without FP64: round half to even,
// then move exact ties up.
__device__ __forceinline__ int slab_of(float x, float z, float4 cf,
                                       float dscale, float doffset, int ndf) {
  const float h = -x * cf.y - z * cf.x;
  const float d = h * dscale + doffset;
  float n = rintf(d);
  if (d - n == 0.5f) n += 1.0f;
  return min(max((int)n, 0), ndf - 1);
}

// Per-tilt multiplier of the reduced-storage forms: the reference's scale
// times the normalisation the filter pass divided out.
__device__ __forceinline__ float reduced_scale(float w, const uint32_t *tmax,
                                               int t, float nrm_mul) {
  float nrm = __uint_as_float(tmax[t]) * nrm_mul;
  if (!(nrm > 0.0f)) nrm = 1.0f;
  return w * nrm;
}

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
  const float x = voxel_coord(ix, W);
  const float z = voxel_coord(iz, Z);

  float acc[YB];
#pragma unroll
  for (int r = 0; r < YB; ++r) acc[r] = 0.0f;

  const size_t row_len = (size_t)N * YB;
  const size_t kstride = (size_t)T * NYB * row_len;
  const typename StoreT<STORE>::T *yblock = slab + (size_t)byb * row_len;

  for (int t = 0; t < T; ++t) {
    const float4 cf = coef[t];
    int k = 0;
    if constexpr (MULTI) k = slab_of(x, z, cf, dscale, doffset, ndf);
    float tt;
    const int i0 = tap_index(x, z, cf, tt);
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
        const float sc = reduced_scale(cf.w, tmax, t, nrm_mul);
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

// ---------------------------------------------------------------------------
// LDS-staged form.
//
// p is a chain of correctly rounded float operations, each monotone in x and
// in z, so over a 64 x 4 tile i0 is smallest at one of the four corners, and
// the exact range of p is at most 63|c| + 3|s| <= 63.08 sqrt(c^2 + s^2)
// (c^2 + s^2 = 1 / (cos^2 + sin^2) to rounding): the taps of one tilt lie in
// a window of 66 consecutive positions; kWPOS leaves a margin for rounding.
// Likewise h spans at most 63.08 units, so a tile touches at most
// ceil(63.08 |dscale|) + 1 consecutive slabs; the host picks KS, the slabs
// staged per tilt, no smaller than that. For a batch of TB tilts the
// workgroup copies, per tilt and staged slab, kWPOS positions of its row
// block starting at the window base (one contiguous run of the slab, YB rows
// per position) into LDS, then interpolates from LDS with the same
// arithmetic per voxel as k_bp. LDS indices are clamped, so coefficients
// that break the bound (non-finite angles, which the host rejects) can only
// give wrong values, never an access outside the buffer.
//
// The coefficients, window bases and first staged slabs of 64 consecutive
// tilts are computed by the 64 lanes of every wave (lane j for tilt j) and
// read back with readlane. Two LDS buffers: the global loads of batch b + 1
// are issued before batch b is interpolated and written to LDS after it, one
// barrier per batch.
constexpr int kWPOS = 72;

template <int YB, int STORE, int KS>
struct Staging {
  using T = typename StoreT<STORE>::T;
  static constexpr int kPosBytes = YB * (int)sizeof(T);
  static constexpr int kTB = 256 / (kPosBytes * KS);  // tilts per batch
  static constexpr int kNSlot = kTB * KS;  // (tilt, slab) runs per batch
  static constexpr int kSlotBytes = kWPOS * kPosBytes;
  static constexpr int kBufBytes = kNSlot * kSlotBytes;
  static constexpr int kChunks = kSlotBytes / 16;  // 16-byte words per run
  static constexpr int kCPL = (kChunks + 63) / 64; // words per lane per run
  static constexpr int kSPW = kNSlot / (kTX * kTZ / 64);  // runs per wave
  static_assert(kPosBytes % 16 == 0, "a position is whole 16-byte words");
  static_assert(kTB >= 1 && 64 % kTB == 0, "batches must tile 64 tilts");
  static_assert(kNSlot % (kTX * kTZ / 64) == 0, "runs split evenly on waves");
};

// Window of one tilt over the tile's active voxels [x0, x1] x [z0, z1]:
// smallest i0 and smallest k over the four corners.
template <bool MULTI>
__device__ __forceinline__ void tile_window(float4 cf, float x0, float x1,
                                            float z0, float z1, float dscale,
                                            float doffset, int ndf, int &wb,
                                            int &kb) {
  int w = 0x7fffffff, k = 0x7fffffff;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const float x = (c & 1) ? x1 : x0;
    const float z = (c & 2) ? z1 : z0;
    float tt;
    w = min(w, tap_index(x, z, cf, tt));
    if constexpr (MULTI) k = min(k, slab_of(x, z, cf, dscale, doffset, ndf));
  }
  wb = w;
  kb = MULTI ? k : 0;
}

__device__ __forceinline__ float readlane_f(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

template <int YB, int STORE, bool MULTI, int KS>
__global__ __launch_bounds__(kTX *kTZ) void k_bp_lds(
    const typename StoreT<STORE>::T *__restrict__ slab,
    const float4 *__restrict__ coef, const uint32_t *__restrict__ tmax,
    float nrm_mul, float dscale, float doffset, int ndf, int N, int T,
    int NYB, int W, int Z, int Y, int yb0, int y_base,
    float *__restrict__ vol, int nbx, int nbz, unsigned nblocks) {
  using S = Staging<YB, STORE, KS>;
  using ST = typename StoreT<STORE>::T;
  constexpr bool kReduced = STORE != kF32;
  static_assert(MULTI || KS == 1, "one slab without defocus slabs");
  __shared__ uint4 lds[2 * S::kBufBytes / 16];

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
  const float tx0 = voxel_coord(bx * kTX, W);
  const float tx1 = voxel_coord(min(bx * kTX + kTX, W) - 1, W);
  const float tz0 = voxel_coord(bz * kTZ, Z);
  const float tz1 = voxel_coord(min(bz * kTZ + kTZ, Z) - 1, Z);

  const size_t row_len = (size_t)N * YB;
  const size_t tstride = (size_t)NYB * row_len;
  const size_t kstride = (size_t)T * tstride;
  const ST *yblock = slab + (size_t)byb * row_len;

  // Per-lane tables of a 64-tilt chunk (lane j: tilt tc + j): coefficients
  // (the reduced forms carry their per-tilt scale in .w), window base and
  // first staged slab. pf: chunk of the batch being loaded; cp: chunk of the
  // batch being interpolated.
  float4 cpf;
  int wpf, kpf;
  auto chunk_tables = [&](int tc) {
    const int t = tc + lane;
    cpf = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    wpf = kpf = 0;
    if (t < T) {
      cpf = coef[t];
      tile_window<MULTI>(cpf, tx0, tx1, tz0, tz1, dscale, doffset, ndf, wpf,
                         kpf);
      if constexpr (kReduced) cpf.w = reduced_scale(cpf.w, tmax, t, nrm_mul);
    }
  };

  uint4 pre[S::kSPW][S::kCPL];
  auto load_batch = [&](int tb) {
#pragma unroll
    for (int i = 0; i < S::kSPW; ++i) {
      const int q = wave + (kTX * kTZ / 64) * i;
      const int j = q / KS, sl = q % KS;
      const int t = tb + j;
      const int tl = t & 63;
      const int wb = __builtin_amdgcn_readlane(wpf, tl);
      const int k = MULTI ? __builtin_amdgcn_readlane(kpf, tl) + sl : 0;
      const bool ok = t < T && k < ndf;
      const ST *src = yblock + (size_t)t * tstride + (size_t)k * kstride;
#pragma unroll
      for (int c = 0; c < S::kCPL; ++c) {
        const int ch = lane + 64 * c;
        const int pos = wb + ch / (S::kPosBytes / 16);
        uint4 v = make_uint4(0u, 0u, 0u, 0u);
        if (ok && ch < S::kChunks && (unsigned)pos < (unsigned)N)
          v = *(const uint4 *)(src + ((long)wb * YB +
                                      (long)ch * (16 / (int)sizeof(ST))));
        pre[i][c] = v;
      }
    }
  };
  auto store_batch = [&](int buf) {
#pragma unroll
    for (int i = 0; i < S::kSPW; ++i) {
      const int q = wave + (kTX * kTZ / 64) * i;
#pragma unroll
      for (int c = 0; c < S::kCPL; ++c) {
        const int ch = lane + 64 * c;
        if (ch < S::kChunks)
          lds[(buf * S::kBufBytes + q * S::kSlotBytes) / 16 + ch] = pre[i][c];
      }
    }
  };

  float acc[YB];
#pragma unroll
  for (int r = 0; r < YB; ++r) acc[r] = 0.0f;

  chunk_tables(0);
  load_batch(0);
  store_batch(0);
  lds_barrier();
  float4 ccp = cpf;
  int wcp = wpf, kcp = kpf;
  const int nb = (T + S::kTB - 1) / S::kTB;
  for (int b = 0; b < nb; ++b) {
    const int tb = b * S::kTB;
    if ((tb & 63) == 0) {
      ccp = cpf;
      wcp = wpf;
      kcp = kpf;
    }
    const bool more = b + 1 < nb;
    if (more) {
      if (((tb + S::kTB) & 63) == 0) chunk_tables(tb + S::kTB);
      load_batch(tb + S::kTB);
    }
    const int boff = (b & 1) * (S::kBufBytes / 16);
    for (int j = 0; j < S::kTB; ++j) {
      const int t = tb + j;
      if (t >= T) break;
      const int tl = t & 63;
      const float4 cf =
          make_float4(readlane_f(ccp.x, tl), readlane_f(ccp.y, tl),
                      readlane_f(ccp.z, tl), readlane_f(ccp.w, tl));
      int ks = 0;
      if constexpr (MULTI) {
        ks = slab_of(x, z, cf, dscale, doffset, ndf) -
             __builtin_amdgcn_readlane(kcp, tl);
        ks = min(max(ks, 0), KS - 1);
      }
      float tt;
      const int i0 = tap_index(x, z, cf, tt);
      if (active && (unsigned)i0 < (unsigned)(N - 1)) {
        int rel = i0 - __builtin_amdgcn_readlane(wcp, tl);
        rel = min(max(rel, 0), kWPOS - 2);
        Taps<YB, STORE> tp;
        tp.load16(&lds[boff + ((j * KS + ks) * S::kSlotBytes +
                               rel * S::kPosBytes) / 16]);
        if constexpr (!kReduced) {
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
    }
    if (more) store_batch((b + 1) & 1);
    lds_barrier();
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
static hipError_t launch_bp_lds_k(const BpArgs &a, hipStream_t s) {
  using ST = typename StoreT<STORE>::T;
  const int nbx = (a.W + kTX - 1) / kTX;
  const int nbz = (a.Z + kTZ - 1) / kTZ;
  if (a.yb1 <= a.yb0) return hipSuccess;
  const unsigned nb =
      (unsigned)nbx * (unsigned)nbz * (unsigned)(a.yb1 - a.yb0);
  hipLaunchKernelGGL((k_bp_lds<YB, STORE, MULTI, KS>), dim3(nb),
                     dim3(kTX * kTZ), 0, s, (const ST *)a.slab, a.coef, a.tmax,
                     a.nrm_mul, a.dscale, a.doffset, a.ndf, a.N, a.T, a.NYB,
                     a.W, a.Z, a.Y, a.yb0, a.y_base, a.vol, nbx, nbz, nb);
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
