// LDS-resident Stockham FFT building blocks for gfx950.
//
// One transform of N points (N a power of two) is computed by N/E threads of
// a workgroup, each holding E complex values in registers. The radix plan is
// E, <middle radices>, E: the first stage reads the thread's registers
// (x[tid + m*N/E]) and the last stage leaves X[tid + m*N/E] in registers, so
// producers and consumers (global loads, the CTF flip, the ramp multiply,
// stores) fuse directly into the first and last stages. Middle stages
// exchange through LDS. Several transforms can share a workgroup; each uses
// its own LDS region, addressed as lds[i * LS] so that columns can be
// interleaved. Every thread of the workgroup must call run() (it contains
// workgroup barriers), unless WAVE is set: then one transform is computed by
// exactly one wave (N / E == 64) in its own LDS region and the exchanges are
// ordered with wave-scope syncs only (a wave's LDS operations complete in
// issue order), so the waves of a workgroup run their transforms
// independently.
//
// Conventions follow FFTW (the reference's host FFT, src/libguanaco/
// fft.cpp:38-71): DIR = -1 is exp(-2 pi i jk/N), DIR = +1 is exp(+...), both
// unnormalised. Twiddles come from a table W[j] = exp(-2 pi i j / kTwN)
// computed in double on the host and rounded to float.
#pragma once

#include <hip/hip_runtime.h>

#include "gx_common.hpp"

namespace gx {

constexpr int kTwN = 16384;  // twiddle table length: largest supported N

// LDS slot of element i of a transform computed with E values per thread:
// one float2 of padding after every E, so that the first stage's stores
// (lane tid writes elements tid * E .. tid * E + E - 1) and the strided
// accesses of the later stages fall in different banks. Every stride the
// stages use (E, N/E, N/R, and the products of earlier radices) is a multiple
// of E, so slot(i + c) = slot(i) + c + c / E: a per-thread base plus constant
// offsets. A transform's region holds fft_lds_len(N, E) * LS float2.
__host__ __device__ constexpr int fft_lds_len(int n, int e) {
  return n + n / e;
}

// The transforms may fuse multiply-adds (FP32 rounding of order 1e-7): only
// the backprojector reproduces the reference's operation order exactly, and
// it does not use these helpers.
__host__ __device__ __forceinline__ float2 cadd(float2 a, float2 b) {
  return make_float2(a.x + b.x, a.y + b.y);
}
__host__ __device__ __forceinline__ float2 csub(float2 a, float2 b) {
  return make_float2(a.x - b.x, a.y - b.y);
}
__host__ __device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  #pragma clang fp contract(fast)
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// multiply by DIR * i
template <int DIR>
__host__ __device__ __forceinline__ float2 mul_i(float2 a) {
  return DIR < 0 ? make_float2(a.y, -a.x) : make_float2(-a.y, a.x);
}
template <int DIR>
__device__ __forceinline__ float2 twiddle(const float2 *__restrict__ tw,
                                          int idx) {
  float2 w = tw[idx];
  return DIR < 0 ? w : make_float2(w.x, -w.y);
}

// ---- radix kernels: v[k] <- sum_n v[n] exp(DIR 2 pi i nk / R) -------------

template <int DIR>
__host__ __device__ __forceinline__ void dft4(float2 &a0, float2 &a1, float2 &a2,
                                     float2 &a3) {
  float2 t0 = cadd(a0, a2), t1 = csub(a0, a2);
  float2 t2 = cadd(a1, a3), t3 = mul_i<DIR>(csub(a1, a3));
  a0 = cadd(t0, t2);
  a2 = csub(t0, t2);
  a1 = cadd(t1, t3);
  a3 = csub(t1, t3);
}

// a * exp(DIR i pi j / 8) for the constant j (0 <= j < 16)
template <int DIR, int J>
__host__ __device__ __forceinline__ float2 rot16(float2 a) {
  constexpr int j = J & 15;
  if constexpr (j == 0) {
    return a;
  } else if constexpr (j == 4) {
    return mul_i<DIR>(a);
  } else if constexpr (j == 8) {
    return make_float2(-a.x, -a.y);
  } else if constexpr (j == 12) {
    return mul_i<-DIR>(a);
  } else {
    constexpr float C1 = 0.92387953251128674f;  // cos(pi/8)
    constexpr float S1 = 0.38268343236508978f;  // sin(pi/8)
    constexpr float R2 = 0.70710678118654752f;  // cos(pi/4)
    constexpr float ct[16] = {1, C1, R2, S1, 0, -S1, -R2, -C1,
                              -1, -C1, -R2, -S1, 0, S1, R2, C1};
    constexpr float st[16] = {0, S1, R2, C1, 1, C1, R2, S1,
                              0, -S1, -R2, -C1, -1, -C1, -R2, -S1};
    return cmul(a, make_float2(ct[j], DIR * st[j]));
  }
}

template <int R, int DIR>
__host__ __device__ __forceinline__ void dft(float2 *v) {
  #pragma clang fp contract(fast)
  if constexpr (R == 1) {
    return;
  } else if constexpr (R == 2) {
    float2 t = v[0];
    v[0] = cadd(t, v[1]);
    v[1] = csub(t, v[1]);
  } else if constexpr (R == 4) {
    dft4<DIR>(v[0], v[1], v[2], v[3]);
  } else if constexpr (R == 8) {
    float2 e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
    float2 o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4<DIR>(e0, e1, e2, e3);
    dft4<DIR>(o0, o1, o2, o3);
    o1 = rot16<DIR, 2>(o1);
    o2 = rot16<DIR, 4>(o2);
    o3 = rot16<DIR, 6>(o3);
    v[0] = cadd(e0, o0);
    v[4] = csub(e0, o0);
    v[1] = cadd(e1, o1);
    v[5] = csub(e1, o1);
    v[2] = cadd(e2, o2);
    v[6] = csub(e2, o2);
    v[3] = cadd(e3, o3);
    v[7] = csub(e3, o3);
  } else if constexpr (R == 16) {
    // 4 x 4: X[k1 + 4 k2] = sum_n2 W4^{n2 k2} W16^{n2 k1} DFT4_n1(v[n2+4n1])
    float2 b[16];
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
      float2 a0 = v[n2], a1 = v[n2 + 4], a2 = v[n2 + 8], a3 = v[n2 + 12];
      dft4<DIR>(a0, a1, a2, a3);
      b[n2 * 4 + 0] = a0;
      b[n2 * 4 + 1] = a1;
      b[n2 * 4 + 2] = a2;
      b[n2 * 4 + 3] = a3;
    }
    // twiddle W16^{n2 k1}
    b[5] = rot16<DIR, 1>(b[5]);
    b[6] = rot16<DIR, 2>(b[6]);
    b[7] = rot16<DIR, 3>(b[7]);
    b[9] = rot16<DIR, 2>(b[9]);
    b[10] = rot16<DIR, 4>(b[10]);
    b[11] = rot16<DIR, 6>(b[11]);
    b[13] = rot16<DIR, 3>(b[13]);
    b[14] = rot16<DIR, 6>(b[14]);
    b[15] = rot16<DIR, 9>(b[15]);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
      float2 a0 = b[k1], a1 = b[4 + k1], a2 = b[8 + k1], a3 = b[12 + k1];
      dft4<DIR>(a0, a1, a2, a3);
      v[k1] = a0;
      v[k1 + 4] = a1;
      v[k1 + 8] = a2;
      v[k1 + 12] = a3;
    }
  }
}

// ---- radix plan -----------------------------------------------------------

__host__ __device__ constexpr int ilog2c(int n) {
  int l = 0;
  while ((1 << l) < n) ++l;
  return l;
}

// Plan for N with end radix E: E, middle..., E where the middle product
// N / E^2 is split into radices <= E, largest first. N == E is one stage.
__host__ __device__ constexpr int plan_stages(int N, int E) {
  if (N <= E) return 1;
  int mid = N / (E * E), s = 2;
  while (mid > 1) {
    int r = mid >= E ? E : mid;
    mid /= r;
    ++s;
  }
  return s;
}
__host__ __device__ constexpr int plan_radix(int N, int E, int stage) {
  int ns = plan_stages(N, E);
  if (ns == 1) return N;
  if (stage == 0 || stage == ns - 1) return E;
  int mid = N / (E * E);
  for (int s = 1; s < ns - 1; ++s) {
    int r = mid >= E ? E : mid;
    if (s == stage) return r;
    mid /= r;
  }
  return 1;
}
// product of the radices of the stages before `stage`
__host__ __device__ constexpr int plan_ns(int N, int E, int stage) {
  int p = 1;
  for (int s = 0; s < stage; ++s) p *= plan_radix(N, E, s);
  return p;
}
// Default end radix: 16 for N >= 256, else the largest E with E^2 <= N
// (single stage up to 16 points when N is itself a valid radix).
__host__ __device__ constexpr int default_e(int N) {
  return N >= 256 ? 16 : (N >= 64 ? 8 : (N >= 16 ? 4 : (N >= 4 ? 2 : N)));
}

// Orders the LDS accesses of one wave before its later ones: no instruction
// is needed in hardware; the fences and the empty asm keep the compiler from
// moving LDS accesses across it.
__device__ __forceinline__ void wave_lds_sync() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  asm volatile("" ::: "memory");
}

template <bool WAVE>
__device__ __forceinline__ void fft_sync() {
  if constexpr (WAVE) {
    wave_lds_sync();
  } else {
    lds_barrier();
  }
}

// Per-thread twiddle factors of one N-point transform with E values per
// thread (raw values W^j = exp(-2 pi i j / kTwN); a transform with DIR = +1
// conjugates them). Loaded once, before the first stage, so that their
// latency overlaps the first radix pass, and reusable by every transform of
// that size the thread computes.
//
// Middle stage S multiplies element m of each group by W^(m k kTwN/(Ns R)),
// k = b mod Ns. Every b = tid + q N/E of a thread has the same k (Ns divides
// N/E), so a stage needs R - 1 values; a radix-8 or -16 stage keeps W^(j ...)
// for j = 1..3 and W^(4j ...) for j < R/4, and forms the others as their
// products (one rounding), which saves 6 registers (R = 8) or 18 (R = 16). The last stage needs
// W^(m tid kTwN/N) for m < E: formed the same way as lo[m mod 4] * hi[m / 4]
// from at most three values of each kind.
//
// The device table holds W^j (kTwN entries) followed, for every (N, E), by
// these per-thread values laid out [value][tid], so that each load of a
// wave is one contiguous run.
// values kept for a middle stage of radix R: W^j (j = 1..3) and W^(4j)
// (j = 1..R/4 - 1) for R = 8 and 16
__host__ __device__ constexpr int tw_mid_count(int R) {
  return R >= 8 ? 3 + R / 4 - 1 : R - 1;
}
__host__ __device__ constexpr int tw_mid_before(int N, int E, int S) {
  int c = 0;
  const int ns = plan_stages(N, E);
  for (int s = 1; s < S && s < ns - 1; ++s)
    c += tw_mid_count(plan_radix(N, E, s));
  return c;
}
__host__ __device__ constexpr int tw_lo_count(int E) { return E - 1 < 3 ? E - 1 : 3; }
__host__ __device__ constexpr int tw_hi_count(int E) { return E / 4 > 1 ? E / 4 - 1 : 0; }
// per-thread values of (N, E); 0 for single-stage or invalid shapes
__host__ __device__ constexpr int tw_count(int N, int E) {
  if (N <= E || N < E * E) return 0;
  return tw_mid_before(N, E, plan_stages(N, E)) + tw_lo_count(E) +
         tw_hi_count(E);
}
// offset of the (N, E) block after the kTwN-entry table; blocks are ordered
// by log2 N (1..14), then E = 2, 4, 8, 16
__host__ __device__ constexpr size_t tw_block_offset(int N, int E) {
  size_t off = 0;
  for (int ln = 1; ln <= 14; ++ln) {
    for (int e = 2; e <= 16; e *= 2) {
      if ((1 << ln) == N && e == E) return off;
      off += (size_t)tw_count(1 << ln, e) * (size_t)((1 << ln) / e);
    }
  }
  return off;
}
__host__ __device__ constexpr size_t tw_table_len() {
  return (size_t)kTwN + tw_block_offset(1 << 15, 2);
}

template <int N, int E>
struct FftTwiddles {
  static constexpr int NS = plan_stages(N, E);
  static constexpr int NT = N / E;
  static constexpr int kMid = NS > 1 ? tw_mid_before(N, E, NS) : 0;
  static constexpr int kLo = tw_lo_count(E);
  static constexpr int kHi = tw_hi_count(E);
  static constexpr int mid_before(int S) { return tw_mid_before(N, E, S); }
  static constexpr size_t kOffset = kTwN + tw_block_offset(N, E);
  float2 mid[kMid > 0 ? kMid : 1];
  float2 lo[kLo > 0 ? kLo : 1];
  float2 hi[kHi > 0 ? kHi : 1];

  // tw: the device table (twiddles())
  __device__ __forceinline__ void load(int tid, const float2 *__restrict__ tw) {
    if constexpr (NS > 1) {
      static_assert(tw_count(N, E) == kMid + kLo + kHi, "table layout");
      const float2 *t = tw + kOffset + tid;
#pragma unroll
      for (int i = 0; i < kMid; ++i) mid[i] = t[i * NT];
#pragma unroll
      for (int i = 0; i < kLo; ++i) lo[i] = t[(kMid + i) * NT];
#pragma unroll
      for (int i = 0; i < kHi; ++i) hi[i] = t[(kMid + kLo + i) * NT];
    }
  }
  // middle-stage factor m (0 < m < R) of the stage whose values start at
  // mid[m0]: kept as is, or for R = 8, 16 the product lo(m mod 4) hi(m / 4)
  template <int R>
  __device__ __forceinline__ static float2 mid_at(const float2 *mid, int m0,
                                                  int m) {
    #pragma clang fp contract(fast)
    if constexpr (R >= 8) {
      const int m1 = m & 3, m2 = m >> 2;
      if (m2 == 0) return mid[m0 + m1 - 1];
      if (m1 == 0) return mid[m0 + 3 + m2 - 1];
      return cmul(mid[m0 + m1 - 1], mid[m0 + 3 + m2 - 1]);
    } else {
      return mid[m0 + m - 1];
    }
  }
  // last-stage factor for element m (0 < m < E)
  __device__ __forceinline__ float2 last(int m) const {
    #pragma clang fp contract(fast)
    const int m1 = m & 3, m2 = m >> 2;
    if (m2 == 0) return lo[m1 - 1];
    if (m1 == 0) return hi[m2 - 1];
    return cmul(lo[m1 - 1], hi[m2 - 1]);
  }
};

// Host: the per-thread block of (N, E) from the W^j table w (kTwN entries),
// the same values FftTwiddles would take from it.
inline void tw_fill_block(int N, int E, const float2 *w, float2 *out) {
  const int NT = N / E, ns = plan_stages(N, E);
  const int nmid = tw_mid_before(N, E, ns), nlo = tw_lo_count(E);
  for (int tid = 0; tid < NT; ++tid) {
    for (int s = 1; s < ns - 1; ++s) {
      const int R = plan_radix(N, E, s), Ns = plan_ns(N, E, s);
      const int TWS = kTwN / (Ns * R), k = tid & (Ns - 1);
      const int base = tw_mid_before(N, E, s);
      if (R >= 8) {
        for (int j = 1; j <= 3; ++j) out[(base + j - 1) * NT + tid] = w[j * k * TWS];
        for (int j = 1; j < R / 4; ++j)
          out[(base + 3 + j - 1) * NT + tid] = w[4 * j * k * TWS];
      } else {
        for (int m = 1; m < R; ++m)
          out[(base + m - 1) * NT + tid] = w[m * k * TWS];
      }
    }
    const int TWS = kTwN / N;
    for (int m = 1; m <= nlo; ++m)
      out[(nmid + m - 1) * NT + tid] = w[m * tid * TWS];
    for (int m = 1; m <= tw_hi_count(E); ++m)
      out[(nmid + nlo + m - 1) * NT + tid] = w[4 * m * tid * TWS];
  }
}

template <int DIR>
__device__ __forceinline__ float2 dir_conj(float2 w) {
  return DIR < 0 ? w : make_float2(w.x, -w.y);
}

template <int N, int E, int DIR, int LS, bool WAVE = false>
struct Fft {
  static_assert((N & (N - 1)) == 0, "N must be a power of two");
  static_assert(N <= kTwN, "N exceeds the twiddle table");
  static constexpr int NT = N / E;  // threads per transform
  static_assert(!WAVE || (NT <= 64 && 64 % NT == 0),
                "a wave-scope transform lies within one wave");
  static constexpr int NS = plan_stages(N, E);
  static_assert(NS == 1 || N >= E * E, "invalid end radix");
  static constexpr int kLds = fft_lds_len(N, E);
  using Twiddles = FftTwiddles<N, E>;
  // padded slot of element i, and of a constant offset c (a multiple of E)
  __device__ __forceinline__ static int slot(int i) { return i + i / E; }
  static constexpr int off(int c) { return c + c / E; }

  template <int S>
  __device__ __forceinline__ static void middle(float2 *lds, int tid,
                                                const Twiddles &tw) {
    #pragma clang fp contract(fast)
    if constexpr (S < NS - 1) {
      constexpr int R = plan_radix(N, E, S);
      constexpr int Ns = plan_ns(N, E, S);
      constexpr int NB = N / R;
      constexpr int Q = E / R;
      constexpr int M0 = Twiddles::mid_before(S);
      float2 w[E];
      const int sb = slot(tid);
#pragma unroll
      for (int q = 0; q < Q; ++q) {
#pragma unroll
        for (int m = 0; m < R; ++m)
          w[q * R + m] = lds[(sb + off(q * NT + m * NB)) * LS];
      }
      fft_sync<WAVE>();
      const int k = tid & (Ns - 1);  // the same for every b = tid + q NT
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const int b = tid + q * NT;
#pragma unroll
        for (int m = 1; m < R; ++m)
          w[q * R + m] =
              cmul(w[q * R + m], dir_conj<DIR>(Twiddles::template mid_at<R>(
                                     tw.mid, M0, m)));
        dft<R, DIR>(&w[q * R]);
        const int sbase = slot((b - k) * R + k);
#pragma unroll
        for (int m = 0; m < R; ++m) lds[(sbase + off(m * Ns)) * LS] = w[q * R + m];
      }
      fft_sync<WAVE>();
      middle<S + 1>(lds, tid, tw);
    }
  }

  // In: v[m] = x[tid + m*NT]. Out: v[m] = X[tid + m*NT]. lds holds kLds * LS
  // float2 (only every LS-th used). Returns with lds free for reuse.
  __device__ __forceinline__ static void run(float2 (&v)[E], float2 *lds,
                                             int tid, const Twiddles &tw) {
    #pragma clang fp contract(fast)
    if constexpr (NS == 1) {
      dft<E, DIR>(v);
    } else {
      dft<E, DIR>(v);
#pragma unroll
      for (int m = 0; m < E; ++m) lds[(tid * (E + 1) + m) * LS] = v[m];
      fft_sync<WAVE>();
      middle<1>(lds, tid, tw);
#pragma unroll
      for (int m = 0; m < E; ++m) v[m] = lds[(slot(tid) + off(m * NT)) * LS];
      fft_sync<WAVE>();
#pragma unroll
      for (int m = 1; m < E; ++m) v[m] = cmul(v[m], dir_conj<DIR>(tw.last(m)));
      dft<E, DIR>(v);
    }
  }
  // Same, loading this thread's twiddles first.
  __device__ __forceinline__ static void run(float2 (&v)[E], float2 *lds,
                                             int tid,
                                             const float2 *__restrict__ tw) {
    Twiddles t;
    t.load(tid, tw);
    run(v, lds, tid, t);
  }
};

}  // namespace gx
