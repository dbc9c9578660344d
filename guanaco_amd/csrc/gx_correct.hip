// CTF phase-flip correction and ramp filtering on gfx950.
//
// Reference semantics (src/libguanaco/correct.cpp:30-90, filter.cpp:30-97):
//   corrected_k = Re(IFFT2_unnorm(FFT2(image) * -sign(Im ctf_k)))
//   filtered    = first N of Re(IFFT_2N(ramp * FFT_2N(pad(row)))) / 2N
// Here the 2D transforms are real-to-complex: a Hermitian half spectrum
// (X/2 + 1 columns, padded to a multiple of CB) is computed once per tilt,
// flipped with a per-defocus mask, inverse transformed along columns, and
// the final row pass fuses the complex-to-real row transform with the ramp
// filter, the per-tilt output scale and the store into the backprojection
// layout. Two real rows travel as one complex row (a + i b) through every
// row transform, and the ramp (real and even) keeps them separable.
#include "gx_common.hpp"
#include "gx_fft.hpp"
#include "gx_kernels.hpp"

namespace gx {

// ---------------------------------------------------------------- helpers --

__global__ void k_fill_u32(uint32_t *__restrict__ p, uint32_t v, size_t n) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

hipError_t launch_fill_u32(uint32_t *p, uint32_t v, size_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  size_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)blocks), dim3(256), 0, s, p,
                     v, n);
  return hipGetLastError();
}

// Plain copy, the calibration kernel for HBM byte counters.
__global__ void k_copy_u32(uint32_t *__restrict__ dst,
                           const uint32_t *__restrict__ src, size_t n) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) dst[i] = src[i];
}

hipError_t launch_copy_u32(uint32_t *dst, const uint32_t *src, size_t n,
                           hipStream_t s) {
  if (n == 0) return hipSuccess;
  size_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_copy_u32, dim3((unsigned)blocks), dim3(256), 0, s, dst,
                     src, n);
  return hipGetLastError();
}

// ------------------------------------------------------------- CTF masks --

// The reference stores H = exp(-i(chi - phi)) as complex64 (correct.py:192)
// and flips by -sign(Im H) (correct.h:30-39). Im H = sin(-(chi - phi)) is
// evaluated in FP64; its float cast is +-0 exactly when |s| <= 2^-150.
__device__ __forceinline__ int flip_sign(double s) {
  const double tiny = 0x1p-150;
  return -((s > tiny) - (s < -tiny));
}

// ctf.h:88-101 (get_r_and_theta) and ctf.h:233-238 (get_chi), FP64.
__device__ double chi_at(int x, int y, int w, int h, double inv_ps, double l,
                         double df, double Cs, double Ca, double Pa) {
  const size_t w2 = (size_t)w / 2, h2 = (size_t)h / 2;
  const double dx =
      ((double)(((size_t)x + w2) % (size_t)w) - (double)w2) / (double)w;
  const double dy =
      ((double)(((size_t)y + h2) % (size_t)h) - (double)h2) / (double)h;
  const double q = sqrt((dx * dx + dy * dy) * inv_ps * inv_ps);
  double d = df;
  if (Ca != 0.0) d = df + Ca * cos(2 * (atan2(dy, dx) - Pa));
  return M_PI * (Cs * l * l * l * q * q * q * q / 2 - l * d * q * q);
}

// Where the flip-mask value of (defocus k, column block cb, row ky, column
// ci) lies. The power-of-two column pass (k_cols) reads it in thread order,
// [k][cb][tid][ci][m] with ky = tid + m NT (NT = Y / E threads per column):
// the E values one thread multiplies for one defocus are one E-byte load.
// The Bluestein column pass reads [k][cb][ky][ci].
__host__ __device__ __forceinline__ size_t mask_offset(int k, int cb, int ky,
                                                       int ci, int XB, int Y,
                                                       bool thread_order) {
  if (!thread_order) return (((size_t)k * XB + cb) * Y + ky) * CB + ci;
  const int E = default_e(Y), NT = Y / E;
  return ((((size_t)k * XB + cb) * NT + ky % NT) * CB + ci) * E + ky / NT;
}

// mask[k][cb][ky][ci] = M(k) + M(-k) in {-2..2}: twice the Hermitian part of
// the flip, which is what Re(IFFT2(...)) keeps. Columns past X/2 are 0.
__global__ void k_ctf_mask(int8_t *__restrict__ mask, int X, int Y, int XB,
                           int ndf, const double *__restrict__ defocus,
                           double l, double Cs, double Ca, double Pa,
                           double phi, double inv_ps, bool thread_order) {
  const size_t total = (size_t)ndf * XB * Y * CB;
  const size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (e >= total) return;
  const int ci = (int)(e % CB);
  const int ky = (int)((e / CB) % Y);
  const int cb = (int)((e / ((size_t)CB * Y)) % XB);
  const int k = (int)(e / ((size_t)CB * Y * XB));
  const int kx = cb * CB + ci;
  int v = 0;
  if (kx <= X / 2) {
    const double df = defocus[k];
    const double c1 = chi_at(kx, ky, X, Y, inv_ps, l, df, Cs, Ca, Pa);
    const int m1 = flip_sign(sin(-(c1 - phi)));
    int m2 = m1;
    if (Ca != 0.0) {
      const int mx = (X - kx) % X, my = (Y - ky) % Y;
      const double c2 = chi_at(mx, my, X, Y, inv_ps, l, df, Cs, Ca, Pa);
      m2 = flip_sign(sin(-(c2 - phi)));
    }
    v = m1 + m2;
  }
  mask[mask_offset(k, cb, ky, ci, XB, Y, thread_order)] = (int8_t)v;
}

hipError_t launch_ctf_mask(int8_t *mask, int X, int Y, int XB, int ndf,
                           const double *d_defocus, double l, double Cs,
                           double Ca, double Pa, double phi, double ps,
                           bool thread_order, hipStream_t s) {
  const size_t total = (size_t)ndf * XB * Y * CB;
  hipLaunchKernelGGL(k_ctf_mask, dim3((unsigned)((total + 255) / 256)),
                     dim3(256), 0, s, mask, X, Y, XB, ndf, d_defocus, l, Cs,
                     Ca, Pa, phi, 1.0 / ps, thread_order);
  return hipGetLastError();
}

// Same mask from a caller-supplied complex64 CTF stack (ndf, Y, X), as the
// reference's corr() receives it (ext/correct.h:32-76).
__global__ void k_mask_from_ctf(int8_t *__restrict__ mask,
                                const float2 *__restrict__ ctf, int X, int Y,
                                int XB, int ndf, bool thread_order) {
  const size_t total = (size_t)ndf * XB * Y * CB;
  const size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (e >= total) return;
  const int ci = (int)(e % CB);
  const int ky = (int)((e / CB) % Y);
  const int cb = (int)((e / ((size_t)CB * Y)) % XB);
  const int k = (int)(e / ((size_t)CB * Y * XB));
  const int kx = cb * CB + ci;
  int v = 0;
  if (kx <= X / 2) {
    const float2 *c = ctf + (size_t)k * X * Y;
    const float a = c[(size_t)ky * X + kx].y;
    const int mx = (X - kx) % X, my = (Y - ky) % Y;
    const float b = c[(size_t)my * X + mx].y;
    v = -((0.0f < a) - (a < 0.0f)) - ((0.0f < b) - (b < 0.0f));
  }
  mask[mask_offset(k, cb, ky, ci, XB, Y, thread_order)] = (int8_t)v;
}

hipError_t launch_mask_from_ctf(int8_t *mask, const float2 *ctf, int X, int Y,
                                int XB, int ndf, bool thread_order,
                                hipStream_t s) {
  const size_t total = (size_t)ndf * XB * Y * CB;
  hipLaunchKernelGGL(k_mask_from_ctf, dim3((unsigned)((total + 255) / 256)),
                     dim3(256), 0, s, mask, ctf, X, Y, XB, ndf, thread_order);
  return hipGetLastError();
}

// ------------------------------------------------------ CTF generation ---

// The CTF on the FFT-ordered grid (ctf.h:466-478 get_ctf_n, ctf.h:489-503
// get_ctf_n_simple), FP64 throughout, stored as complex64: with the
// envelopes (ctf.h:188-223, 250-258) A exp(log Es + log Et) exp(-i (chi -
// phi)), without them exp(-i (chi - phi)). recentre puts the zero frequency
// at (h/2, w/2), as numpy.fft.fftshift does. blockIdx.y = defocus.
template <bool ENVELOPE>
__global__ void k_ctf_grid(float2 *__restrict__ out, int w, int h,
                           double inv_ps, const double *__restrict__ defocus,
                           CtfParams c, int recentre) {
  const size_t n = (size_t)w * h;
  const size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int x = (int)(e % w), y = (int)(e / w);
  const size_t w2 = (size_t)w / 2, h2 = (size_t)h / 2;
  const double dx =
      ((double)(((size_t)x + w2) % (size_t)w) - (double)w2) / (double)w;
  const double dy =
      ((double)(((size_t)y + h2) % (size_t)h) - (double)h2) / (double)h;
  const double q = sqrt((dx * dx + dy * dy) * inv_ps * inv_ps);
  const double theta = atan2(dy, dx);
  const double df = defocus[blockIdx.y] + c.Ca * cos(2 * (theta - c.Pa));
  // the reference's operation order (ctf.h:234-238, 189-223)
  const double chi =
      M_PI * (c.Cs * c.l * c.l * c.l * q * q * q * q / 2 - c.l * df * q * q);
  double mag = 1.0;
  if constexpr (ENVELOPE) {
    const double u =
        1 + 2 * M_PI * M_PI * c.theta_c * c.theta_c * c.dd * c.dd * q * q;
    const double v = c.Cs * c.l * c.l * c.l * q * q * q - df * c.l * q;
    const double log_es =
        -(M_PI * M_PI * c.theta_c * c.theta_c * v * v / (c.l * c.l * u));
    const double log_et =
        -(M_PI * M_PI * c.l * c.l * c.dd * c.dd * q * q * q * q / (2 * u));
    mag = (1 / sqrt(u)) * exp(log_es + log_et);
  }
  double sn, cs;
  sincos(-(chi - c.phi), &sn, &cs);
  const int ox = recentre ? (int)(((size_t)x + w2) % (size_t)w) : x;
  const int oy = recentre ? (int)(((size_t)y + h2) % (size_t)h) : y;
  out[(size_t)blockIdx.y * n + (size_t)oy * w + ox] =
      make_float2((float)(mag * cs), (float)(mag * sn));
}

hipError_t launch_ctf_grid(float2 *out, int w, int h, double ps,
                           const double *d_defocus, int ndf,
                           const CtfParams &c, bool envelope, bool recentre,
                           hipStream_t s) {
  const size_t n = (size_t)w * h;
  dim3 grid((unsigned)((n + 255) / 256), (unsigned)ndf);
  if (envelope) {
    hipLaunchKernelGGL(k_ctf_grid<true>, grid, dim3(256), 0, s, out, w, h,
                       1.0 / ps, d_defocus, c, recentre ? 1 : 0);
  } else {
    hipLaunchKernelGGL(k_ctf_grid<false>, grid, dim3(256), 0, s, out, w, h,
                       1.0 / ps, d_defocus, c, recentre ? 1 : 0);
  }
  return hipGetLastError();
}

// ------------------------------------------------- pass 1: rows, R2C -------

// 2 * RP image rows per workgroup (RP complex pairs). Writes the half
// spectrum F[t][cb][y][ci]; for each cb the 2 * RP rows form one contiguous
// run of 2 * RP * 32 bytes. Optionally accumulates max |image| per tilt (for
// reduced-precision storage) with one atomic per workgroup.
template <int X, int RP>
__global__ __launch_bounds__(RP * (X / default_e(X))) void k_rows_r2c(
    const float *__restrict__ img, size_t tstride, int Y, int XB,
    float2 *__restrict__ F, uint32_t *__restrict__ tmax, int t0,
    const float2 *__restrict__ tw) {
  constexpr int E = default_e(X);
  constexpr int NT = X / E;
  __shared__ float2 s[RP][fft_lds_len(X, E)];
  __shared__ float red[16];
  const int p = threadIdx.x / NT, tid = threadIdx.x % NT;
  const int t = blockIdx.y;
  const int ya = blockIdx.x * (2 * RP) + 2 * p, yb = ya + 1;
  const float *im = img + (size_t)t * tstride;
  float2 v[E];
  float amax = 0.0f;
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int n = tid + m * NT;
    const float a = ya < Y ? im[(size_t)ya * X + n] : 0.0f;
    const float b = yb < Y ? im[(size_t)yb * X + n] : 0.0f;
    v[m] = make_float2(a, b);
    amax = fmaxf(amax, fmaxf(fabsf(a), fabsf(b)));
  }
  // a pair of rows is one wave (or part of one) up to 1024 points: its
  // transform synchronises only that wave
  Fft<X, E, -1, 1, (NT <= 64)>::run(v, s[p], tid, tw);
#pragma unroll
  for (int m = 0; m < E; ++m) s[p][tid + m * NT] = v[m];
  lds_barrier();
  if (tmax != nullptr) {
    amax = block_max<RP * NT>(amax, red);
    if (threadIdx.x == 0) atomicMax(&tmax[t0 + t], __float_as_uint(amax));
  }
  const int ncol = XB * CB;
  for (int e = threadIdx.x; e < 2 * RP * ncol; e += RP * NT) {
    const int ci = e % CB, r = (e / CB) % (2 * RP), cb = e / (2 * RP * CB);
    const int kx = cb * CB + ci, y = blockIdx.x * (2 * RP) + r;
    if (y >= Y) continue;
    float2 val = make_float2(0.0f, 0.0f);
    if (kx <= X / 2) {
      const float2 z = s[r / 2][kx];
      const float2 zm = s[r / 2][(X - kx) & (X - 1)];
      val = (r & 1) == 0 ? make_float2((z.x + zm.x) * 0.5f, (z.y - zm.y) * 0.5f)
                         : make_float2((z.y + zm.y) * 0.5f, (zm.x - z.x) * 0.5f);
    }
    F[(((size_t)t * XB + cb) * Y + y) * CB + ci] = val;
  }
}

// ---------------------------- pass 2: columns, flip per defocus, inverse ---

// One block of CB columns (CB * Y / E threads). The forward column transform
// stays in registers; every defocus k reuses it: multiply by the mask,
// inverse transform, store G[t][k][cb][y][ci]. M = int8_t: the flip mask
// (M(k) + M(-k), scaled by 1/2); M = float: a real transfer function.
template <typename M>
__device__ __forceinline__ float mask_value(M m) {
  if constexpr (sizeof(M) == 1) {
    return 0.5f * (float)m;
  } else {
    return (float)m;
  }
}

// CW columns per workgroup (WPE waves per SIMD bound the registers). The
// grid is 1D over (column part, tilt, column block), remapped so that each
// XCD takes consecutive ones: the parts of a column block, which share its
// 32-byte row segments, meet in one L2, and so do the launch's tilts of one
// column block, which read the same flip masks (every defocus of the block:
// ndf Y CB bytes), so that most of them find the masks in L2 (fetched bytes
// per 7-tilt launch of 4096-point columns with 32 defoci: 1.86 -> 0.91 GB,
// 3.25 -> 3.10 ms). The launches use CW = CB
// up to 2048 points: a wave then stores 16 rows x 4 columns, whole 128-byte
// lines of G. With one column per workgroup at 4096 points the partial-line
// stores doubled the bytes written to HBM (4.5 GB per tilt at 32 defoci,
// 2.15 GB compulsory); at 4096 and 8192 points two columns per workgroup
// (two 512-thread workgroups per CU at 4096, so one's barriers overlap the
// other's work) measured faster than four (C4 column pass 58.3 -> 52.9 ms).
//
// Registers: the forward spectrum (2E) stays live across the defocus loop;
// the per-row store addresses and the twiddles (about 50 registers) would
// too if the compiler hoisted them, and at the 128 registers of four waves
// per SIMD they spilled (22-68 registers, up to 276 bytes per lane of
// scratch, which made most of the pass's memory traffic). The thread
// indices therefore pass through an empty asm each defocus, and the
// twiddles are reloaded (from the cache) per inverse transform: 106-108
// registers, no scratch.
template <int Y, typename M, int CW, int WPE = 4>
__global__ __launch_bounds__(CW *(Y / default_e(Y)))
__attribute__((amdgpu_waves_per_eu(WPE))) void k_cols(
    const float2 *__restrict__ F, const M *__restrict__ mask,
    float2 *__restrict__ G, int XB, int ndf, const float2 *__restrict__ tw,
    int ylo, int yhi, const int2 *__restrict__ krange) {
  constexpr int E = default_e(Y);
  constexpr int NT = Y / E;
  constexpr int H = CB / CW;  // workgroups per column block
  static_assert(CB % CW == 0, "whole column blocks");
  __shared__ float2 s[fft_lds_len(Y, E) * CW];
  const int cl = threadIdx.x % CW, tid = threadIdx.x / CW;
  // (column part, tilt, column block), column part fastest
  const unsigned lb = xcd_remap(blockIdx.x, gridDim.x);
  const unsigned nt = gridDim.x / (H * (unsigned)XB);  // tilts of the launch
  const int ci = (int)(lb % H) * CW + cl;
  const int t = (int)((lb / H) % nt);
  const int cb = (int)(lb / (H * nt));
  const float2 *src = F + ((size_t)t * XB + cb) * Y * CB;
  typename Fft<Y, E, -1, CW>::Twiddles twy;  // shared by every transform here
  twy.load(tid, tw);
  float2 v[E];
#pragma unroll
  for (int m = 0; m < E; ++m) v[m] = src[(size_t)(tid + m * NT) * CB + ci];
  Fft<Y, E, -1, CW>::run(v, s + cl, tid, twy);
  // int8 flip masks in thread order (mask_offset): this thread's E values of
  // one defocus are one E-byte load (loading a defocus ahead measured slower:
  // the extra registers spill)
  using MaskRun = typename std::conditional<
      E == 16, uint4,
      typename std::conditional<
          E == 8, uint2,
          typename std::conditional<E == 4, uint32_t, uint16_t>::type>::type>::
      type;
  constexpr bool kRun = sizeof(M) == 1 && sizeof(MaskRun) == E;
  const size_t run0 = ((size_t)cb * NT + tid) * CB + ci;  // in E-byte runs
  // only the defoci the volume reads at this tilt (all without krange)
  const int2 kr = krange != nullptr ? krange[t] : make_int2(0, ndf - 1);
  for (int k = kr.x; k <= kr.y; ++k) {
    // the thread's row and column indices pass through an empty asm each
    // defocus, so the per-row store addresses are formed in the iteration
    // instead of being held across the loop
    int tk = tid, ck = ci;
    asm volatile("" : "+v"(tk), "+v"(ck));
    float2 w[E];
    if constexpr (kRun) {
      int8_t mv[E];
      const MaskRun cur = reinterpret_cast<const MaskRun *>(
            mask)[(size_t)k * XB * NT * CB + run0];
      __builtin_memcpy(mv, &cur, E);
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const float f = mask_value(mv[m]);
        w[m] = make_float2(v[m].x * f, v[m].y * f);
      }
    } else {
      const M *mk = mask + ((size_t)k * XB + cb) * Y * CB;
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const float f = mask_value(mk[(size_t)(tid + m * NT) * CB + ci]);
        w[m] = make_float2(v[m].x * f, v[m].y * f);
      }
    }
    // the twiddles are loaded again for each defocus (from the cache) rather
    // than held across the loop
    typename Fft<Y, E, 1, CW>::Twiddles twk;
    twk.load(tk, tw);
    Fft<Y, E, 1, CW>::run(w, s + cl, tk, twk);
    // only the rows [ylo, yhi) that the row pass reads (a band of a
    // banded plan; every row otherwise)
    float2 *dst = G + (((size_t)t * ndf + k) * XB + cb) * Y * CB +
                  ((size_t)tk * CB + ck);
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const int y = tk + m * NT;
      if (y >= ylo && y < yhi) dst[(size_t)(m * NT) * CB] = w[m];
    }
  }
}

// ------------------------------------------- pass 3: rows, C2R + filter ----

// Load Z[kx] = A[kx] + i B[kx] (Hermitian-extended) for rows ya, yb of one
// (tilt, defocus) half spectrum g[cb][y][ci].
__device__ __forceinline__ float2 load_pair_spectrum(const float2 *g, int X,
                                                     int Y, int ya, int yb,
                                                     int kx) {
  const int kk = kx <= X / 2 ? kx : X - kx;
  const int cb = kk / CB, ci = kk % CB;
  // Rows past the image load row Y - 1 and are zeroed afterwards: selecting
  // values rather than addresses keeps these plain global loads (a select
  // between the row and a local zero becomes a generic-address load of a
  // selected pointer). One (tilt, defocus) spectrum holds fewer than 2^32
  // elements, so the offsets are 32-bit.
  const unsigned base = (unsigned)cb * (unsigned)Y;
  const float2 A0 = g[(base + (unsigned)min(ya, Y - 1)) * CB + ci];
  const float2 B0 = g[(base + (unsigned)min(yb, Y - 1)) * CB + ci];
  const float2 zero = make_float2(0.0f, 0.0f);
  const float2 A = ya < Y ? A0 : zero;
  const float2 B = yb < Y ? B0 : zero;
  return kx <= X / 2 ? make_float2(A.x - B.y, A.y + B.x)
                     : make_float2(A.x + B.y, B.x - A.y);
}

// load_pair_spectrum for the E columns kx = tid + m NT of one thread, both
// rows (ya, ya + 1) inside the image. In the [cb][y][ci] layout (CB = 4)
// column kk of row y is element (kk / 4 * Y + y) * 4 + kk % 4, and row ya + 1
// lies 32 bytes after row ya. Columns kx <= X / 2 (m < E / 2, and m = E / 2
// at tid = 0) are read as they are, the others as the mirrored column
// X - kx, conjugated; the element offsets are then
//   kk = kx:      lo + m NT Y,         lo = ((tid / 4) Y + ya) 4 + tid % 4
//   kk = X - kx:  hi + (X - m NT) Y,   hi = (ya - ceil(tid / 4) Y) 4 + (-tid) % 4
// (the two agree at m = E / 2, tid = 0), i.e. one per-thread value plus a
// uniform one. Byte offsets are 32-bit: one (tilt, defocus) spectrum holds
// fewer than 2^28 elements.
template <int X, int E>
__device__ __forceinline__ void load_pair_rows(const float2 *__restrict__ g,
                                               int Y, int ya, int tid,
                                               float2 (&z)[E]) {
  constexpr int NT = X / E;
  const char *gb = reinterpret_cast<const char *>(g);
  const unsigned lo =
      ((((unsigned)(tid >> 2) * (unsigned)Y + (unsigned)ya) << 2) +
       (unsigned)(tid & 3)) << 3;
  const unsigned hi =
      ((((unsigned)ya - (unsigned)((tid + 3) >> 2) * (unsigned)Y) << 2) +
       (unsigned)((-tid) & 3)) << 3;
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const unsigned ob =
        m < E / 2 ? lo + (unsigned)(m * NT) * (unsigned)Y * 8u
                  : hi + (unsigned)(X - m * NT) * (unsigned)Y * 8u;
    const float2 A = *reinterpret_cast<const float2 *>(gb + ob);
    const float2 B = *reinterpret_cast<const float2 *>(gb + ob + 32);
    const float2 u = make_float2(A.x - B.y, A.y + B.x);
    const float2 v = make_float2(A.x + B.y, B.x - A.y);
    z[m] = m < E / 2 ? u : (m > E / 2 ? v : (tid == 0 ? u : v));
  }
}

__device__ __forceinline__ float inv_norm(const uint32_t *tmax, int t,
                                          float mul) {
  if (tmax == nullptr) return 1.0f;
  const float n = __uint_as_float(tmax[t]) * mul;
  return n > 0.0f ? 1.0f / n : 1.0f;
}

// Two values of one detector column from rows (y, y + 1), packed as they lie
// in the slab (row y first).
template <int STORE>
struct RowPair;
template <>
struct RowPair<kF32> {
  using T = float2;
  __device__ __forceinline__ static T pack(float a, float b) {
    return make_float2(a, b);
  }
};
template <>
struct RowPair<kF16> {
  using T = uint32_t;
  __device__ __forceinline__ static T pack(float a, float b) {
    return (uint32_t)__half_as_ushort(__float2half(a)) |
           ((uint32_t)__half_as_ushort(__float2half(b)) << 16);
  }
};
template <>
struct RowPair<kBF16> {
  using T = uint32_t;
  __device__ __forceinline__ static T pack(float a, float b) {
    return bf16x2_bits(a, b);
  }
};

// Rows of fewer than 256 points (k_rows_c2r_ramp below takes the others).
// One workgroup filters RB = 4 * IT consecutive rows of one (tilt, defocus)
// spectrum, RB dividing the slab's row blocking: two pair slots (L = 2X
// threads-per-pair layout shared by the X-point inverse, end radix EL/2, and
// the L-point filter, end radix EL) each take IT row pairs in turn and keep
// their outputs in registers, so that each thread finally writes the 2 * IT
// consecutive rows of its columns as whole 16-byte stores into the
// [y / YB][x][y % YB] slab layout.
template <int X, int STORE, int IT>
__global__ __launch_bounds__(2 * (2 * X / default_e(2 * X))) void
    k_rows_c2r_filter(const float2 *__restrict__ G, int Y, int XB, int ndf,
                      int t0, int T, const float *__restrict__ wscale,
                      const uint32_t *__restrict__ tmax, float nrm_mul,
                      typename StoreT<STORE>::T *__restrict__ slab, int NYB,
                      int YB, const float2 *__restrict__ tw, int y0, int ny,
                      const int2 *__restrict__ krange) {
  constexpr int L = 2 * X;
  constexpr int EL = default_e(L);
  constexpr int EX = EL / 2;
  constexpr int NT = L / EL;
  static_assert(X / EX == NT, "thread layouts must agree");
  using PT = typename RowPair<STORE>::T;
  constexpr bool kWide = (IT * sizeof(PT)) % 16 == 0;  // whole 16-byte stores
  __shared__ float2 s[2][fft_lds_len(L, EL)];
  const int p = threadIdx.x / NT, tid = threadIdx.x % NT;
  const int k = blockIdx.y, tl = blockIdx.z, t = t0 + tl;
  if (krange != nullptr && (k < krange[t].x || k > krange[t].y)) return;
  // first row of the slot within the band [y0, y0 + ny) the slab holds
  const int r0 = blockIdx.x * (4 * IT) + p * (2 * IT);
  const float2 *g = G + ((size_t)tl * ndf + k) * XB * Y * CB;

  // the L-point twiddles serve both L-point transforms of every row pair
  typename Fft<L, EL, -1, 1>::Twiddles twl;
  twl.load(tid, tw);
  const float sc =
      wscale[t] * (STORE == kF32 ? 1.0f : inv_norm(tmax, t, nrm_mul));
  // filter.cpp:30-45: |j| / X on the L = 2X grid, times the 1/L of the
  // unnormalised inverse; X and L are powers of two, so j / (X L) is exact
  // and equals the reference's (j / X) * (1 / L)
  constexpr float kRamp = 1.0f / ((float)X * (float)L);
  PT out[EL / 2][IT];
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int ya = y0 + r0 + 2 * it, yb = ya + 1;
    float2 v[EX];
#pragma unroll
    for (int m = 0; m < EX; ++m)
      v[m] = load_pair_spectrum(g, X, Y, ya, yb, tid + m * NT);
    Fft<X, EX, 1, 1>::run(v, s[p], tid, tw);  // v = a + i b (unnormalised)
#pragma unroll
    for (int m = 0; m < EX; ++m) s[p][tid + m * NT] = v[m];
    lds_barrier();
    float2 u[EL];
#pragma unroll
    for (int m = 0; m < EL; ++m)
      u[m] = m < EL / 2 ? s[p][tid + m * NT] : make_float2(0.0f, 0.0f);
    lds_barrier();
    Fft<L, EL, -1, 1>::run(u, s[p], tid, twl);
#pragma unroll
    for (int m = 0; m < EL; ++m) {
      const int j = tid + m * NT;
      const float r = (float)(j <= X ? j : L - j) * kRamp;
      u[m] = make_float2(u[m].x * r, u[m].y * r);
    }
    Fft<L, EL, 1, 1>::run(u, s[p], tid, twl);
#pragma unroll
    for (int m = 0; m < EL / 2; ++m)
      out[m][it] = RowPair<STORE>::pack(u[m].x * sc, u[m].y * sc);
  }
  using ST = typename StoreT<STORE>::T;
  if (r0 >= ny) return;
  const bool whole = r0 + 2 * IT <= ny;
#pragma unroll
  for (int m = 0; m < EL / 2; ++m) {
    const int n = tid + m * NT;
    ST *dst = slab + slab_index(k, t, r0, n, T, NYB, X, YB);
    if (whole && kWide) {
      uint4 *d4 = reinterpret_cast<uint4 *>(dst);
#pragma unroll
      for (int q = 0; q < (int)(IT * sizeof(PT) / 16); ++q) {
        uint4 w;
        __builtin_memcpy(&w, reinterpret_cast<const char *>(out[m]) + 16 * q,
                         16);
        d4[q] = w;
      }
    } else if (whole) {
      PT *dp = reinterpret_cast<PT *>(dst);
#pragma unroll
      for (int it = 0; it < IT; ++it) dp[it] = out[m][it];
    } else {
      const ST *src = reinterpret_cast<const ST *>(out[m]);
#pragma unroll
      for (int r = 0; r < 2 * IT; ++r)  // constant indices keep out[] in VGPRs
        if (r0 + r < ny) dst[r] = src[r];
    }
  }
}

// Writes a row group of the slab: P slots, slot p holding rows (R0 + 2p,
// R0 + 2p + 1) of column tid + m NT in out[m] (E = X / NT per thread), of
// which the first ncols columns are stored ([y / YB][x][y % YB], rows of
// ncols). s: the workgroup's P slot regions of fft_lds_len(X, E) float2.
// One slot, and the last, partial row group, store from registers; otherwise
// the slots stage [column][slot] in LDS (stride P + 1 PT: odd in 4-byte
// words for 16-bit storage, so a wave's column writes fall in distinct
// banks), in NCH passes of XC columns when the regions do not hold all X
// columns at once, and the workgroup writes each column's RB = 2P rows as
// whole 4-, 8- or 16-byte runs.
template <int X, int STORE, int P>
__device__ __forceinline__ void store_row_group(
    const typename RowPair<STORE>::T (&out)[default_e(X)], float2 *s, int p,
    int tid, typename StoreT<STORE>::T *__restrict__ slab, int k, int t,
    int R0, int Y, int ncols, int T, int NYB, int YB) {
  constexpr int E = default_e(X);
  constexpr int NT = X / E;
  using PT = typename RowPair<STORE>::T;
  using ST = typename StoreT<STORE>::T;
  const int ya = R0 + 2 * p, yb = ya + 1;
  if (P == 1 || R0 + 2 * P > Y) {
    if (ya >= Y) return;
    // ya is even: with an even row blocking the pair is one aligned PT;
    // with YB = 1 (one row per block) row yb lies a whole row further
    const bool pair = yb < Y && YB % 2 == 0;
    // column n: base + n YB elements (32-bit offsets within one slab row
    // block of ncols * YB elements)
    ST *const base = slab + slab_index(k, t, ya, 0, T, NYB, ncols, YB);
    ST *const base_b =
        yb < Y ? slab + slab_index(k, t, yb, 0, T, NYB, ncols, YB) : base;
    const unsigned xo = (unsigned)tid * (unsigned)YB;
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const int n = tid + m * NT;
      if (n >= ncols) continue;
      const unsigned o = xo + (unsigned)(m * NT) * (unsigned)YB;
      ST *dst = base + o;
      if (pair) {
        *reinterpret_cast<PT *>(dst) = out[m];
      } else {
        const ST *v = reinterpret_cast<const ST *>(&out[m]);
        dst[0] = v[0];
        if (yb < Y) base_b[o] = v[1];
      }
    }
    return;
  }
  constexpr int NCH = sizeof(PT) / 4;
  constexpr int XC = X / NCH;
  constexpr int XS = P + 1;
  static_assert((size_t)XC * XS * sizeof(PT) <=
                    (size_t)P * fft_lds_len(X, E) * sizeof(float2),
                "staging fits");
  constexpr int QW = P * sizeof(PT) < 16 ? P : 16 / (int)sizeof(PT);
  constexpr int NQ = P / QW;  // runs per column
  using RunT = typename std::conditional<
      QW * sizeof(PT) == 16, uint4,
      typename std::conditional<QW * sizeof(PT) == 8, uint2,
                                uint32_t>::type>::type;
  static_assert(QW * sizeof(PT) == sizeof(RunT), "whole runs");
  PT *st = reinterpret_cast<PT *>(s);
  ST *base = slab + slab_index(k, t, R0, 0, T, NYB, ncols, YB);
#pragma unroll
  for (int c = 0; c < NCH; ++c) {
    lds_barrier();  // the slots' transforms, or the previous pass, are done
#pragma unroll
    for (int m = c * (E / NCH); m < (c + 1) * (E / NCH); ++m)
      st[(tid + m * NT - c * XC) * XS + p] = out[m];
    lds_barrier();
    for (int q = threadIdx.x; q < XC * NQ; q += P * NT) {
      const int x = q / NQ, j = q % NQ;
      if (c * XC + x >= ncols) break;  // q grows with x
      PT v[QW];
#pragma unroll
      for (int i = 0; i < QW; ++i) v[i] = st[x * XS + j * QW + i];
      RunT w;
      __builtin_memcpy(&w, v, sizeof(RunT));
      *reinterpret_cast<RunT *>(base + (size_t)(c * XC + x) * YB +
                                2 * QW * j) = w;
    }
  }
}

// The same ramp-filtered inverse as k_rows_c2r_filter, as four X-point
// transforms per row pair instead of one X-point and two 2X-point ones. With
// z = IFFT_X(Z) (the row pair, unnormalised) zero-padded to L = 2X:
//   FFT_L(z)[2i]     = X Z[i]
//   FFT_L(z)[2i + 1] = FFT_X(z[n] c(n))[i],          c(n) = exp(-i pi n / X)
//   IFFT_L(U)[n]     = IFFT_X(U[2i])[n] + conj(c(n)) IFFT_X(U[2i + 1])[n]
// for n < X, so the filtered rows are IFFT_X(X r[2i] Z) + conj(c) IFFT_X(
// r[2i + 1] FFT_X(c z)), r the ramp of the 2X grid. Every transform is X
// points: a slot (one row pair) needs X / 16 threads and an X-point LDS
// region, and at X <= 1024 a slot lies within one wave and synchronises
// only itself (the workgroups and their slab stores: k_rows_c2r_ramp).
//
// ramp_pair: one row pair (ya, yb) of one (tilt, defocus) spectrum g through
// the four transforms: out[m] = rows (ya, yb) of column tid + m NT, times sc,
// packed. ls: the pair's X-point LDS region; twx: this thread's X-point
// twiddles; tw: the device table (c(n) = c(tid) exp(-i pi m / E) for
// n = tid + m NT, both factors from its W^j entries).
template <int X, int STORE, bool WAVE>
__device__ __forceinline__ void ramp_pair(
    const float2 *g, int Y, int ya, int yb, float2 *ls, int tid,
    const typename Fft<X, default_e(X), 1, 1, WAVE>::Twiddles &twx,
    const float2 *__restrict__ tw, float sc,
    typename RowPair<STORE>::T (&out)[default_e(X)]) {
  constexpr int E = default_e(X);
  constexpr int NT = X / E;
  constexpr int L = 2 * X;
  // filter.cpp:30-45: r[j] = |j| / X on the 2X grid, times the 1/L of the
  // unnormalised inverse (powers of two: exact as the reference's product)
  constexpr float kRamp = 1.0f / ((float)X * (float)L);
  float2 z[E], e[E];
  if (yb < Y) {
    load_pair_rows<X, E>(g, Y, ya, tid, z);
  } else {
#pragma unroll
    for (int m = 0; m < E; ++m)
      z[m] = load_pair_spectrum(g, X, Y, ya, yb, tid + m * NT);
  }
  // even bins: X r[2n] Z[n] = (fold(2n) / L) Z[n]
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int n = tid + m * NT;
    const float r = (float)(2 * n <= X ? 2 * n : L - 2 * n) * (1.0f / L);
    e[m] = make_float2(z[m].x * r, z[m].y * r);
  }
  Fft<X, E, 1, 1, WAVE>::run(e, ls, tid, twx);
  // odd bins: the rows, twiddled by c(n), forward, times r[2n + 1]
  Fft<X, E, 1, 1, WAVE>::run(z, ls, tid, twx);
  const float2 ct = tw[tid * (kTwN / L)];  // c(tid) = W^(tid kTwN / L)
#pragma unroll
  for (int m = 0; m < E; ++m)
    z[m] = cmul(z[m], cmul(ct, tw[m * (kTwN / (2 * E))]));
  Fft<X, E, -1, 1, WAVE>::run(z, ls, tid, twx);
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int n = tid + m * NT;
    const float r = (float)(2 * n + 1 <= X ? 2 * n + 1 : L - 2 * n - 1) * kRamp;
    z[m] = make_float2(z[m].x * r, z[m].y * r);
  }
  Fft<X, E, 1, 1, WAVE>::run(z, ls, tid, twx);
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const float2 c = cmul(ct, tw[m * (kTwN / (2 * E))]);  // e + conj(c) z
    const float ox = fmaf(c.x, z[m].x, fmaf(c.y, z[m].y, e[m].x));
    const float oy = fmaf(c.x, z[m].y, fmaf(-c.y, z[m].x, e[m].y));
    out[m] = RowPair<STORE>::pack(ox * sc, oy * sc);
  }
}

// The grid is 1D over (row group, defocus, tilt), or over (row group, entry
// of `pairs`) when the plan lists the (tilt, defocus) pairs its volume reads
// (absolute tilts, those of this launch), so that no workgroup is idle.
// The workgroup's P slots each take one row pair (IT = 1) and meet once, to
// stage their outputs in LDS column by column (store_row_group), so that the
// workgroup writes each column's RB = 2P rows (RB dividing the slab's row
// blocking) as whole 8- or 16-byte runs. Or each slot takes two pairs in
// turn (IT = 2, 16-bit storage at 4096 points, where the one slot is the
// whole 256-thread workgroup; at 1024 and 2048 points it measured slower):
// slot p then takes rows R0 + 4p .. R0 + 4p + 3, the first pair's packed
// rows wait in an LDS stash (X words per slot, read back by the thread that
// wrote them) rather than in registers, and each column's four rows leave
// as one 8-byte write.
template <int X, int STORE, int P, int IT>
__global__ __launch_bounds__(P *(X / default_e(X)))
__attribute__((amdgpu_waves_per_eu(X >= 8192 ? 2 : 3))) void k_rows_c2r_ramp(
    const float2 *__restrict__ G, int Y, int XB, int ndf, int t0, int T,
    const float *__restrict__ wscale, const uint32_t *__restrict__ tmax,
    float nrm_mul, typename StoreT<STORE>::T *__restrict__ slab, int NYB,
    int YB, const float2 *__restrict__ tw, int y0, int ny,
    const int2 *__restrict__ pairs) {
  constexpr int E = default_e(X);
  constexpr int NT = X / E;
  constexpr bool WAVE = NT <= 64;
  constexpr int L = 2 * X;
  constexpr int SL = fft_lds_len(X, E);
  constexpr int RB = 2 * P * IT;  // rows per workgroup
  using PT = typename RowPair<STORE>::T;
  using ST = typename StoreT<STORE>::T;
  static_assert(E == 16, "sizes of 256 points and more");
  static_assert(IT == 1 || (IT == 2 && sizeof(PT) == 4),
                "two pairs per slot: 16-bit storage");
  __shared__ float2 s[P * SL];
  __shared__ PT stash[IT > 1 ? P * X : 1];
  const int p = threadIdx.x / NT, tid = threadIdx.x % NT;
  // A 1D grid over (row group, defocus, tilt), remapped so that each XCD
  // takes consecutive row groups: the RB-row runs of one slab row block, and
  // the halves of the 128-byte spectrum lines that row groups share, then
  // meet in that XCD's L2 instead of reaching HBM as partial lines.
  // Row groups cover the band [y0, y0 + ny) the slab holds; R0 counts from
  // y0, and the spectrum rows read are y0 + R0 onwards.
  const unsigned nry = (unsigned)(ny + RB - 1) / RB;
  const unsigned lb = xcd_remap(blockIdx.x, gridDim.x);
  // (tilt, defocus): from the list of those the volume reads, or every one
  int k, tl;
  if (pairs != nullptr) {
    const int2 pk = pairs[lb / nry];
    tl = pk.x - t0;
    k = pk.y;
  } else {
    k = (int)((lb / nry) % (unsigned)ndf);
    tl = (int)(lb / (nry * (unsigned)ndf));
  }
  const int t = t0 + tl;
  const int R0 = (int)(lb % nry) * RB;
  const int ya0 = y0 + R0;
  const float2 *g = G + ((size_t)tl * ndf + k) * XB * Y * CB;
  float2 *ls = s + p * SL;

  typename Fft<X, E, 1, 1, WAVE>::Twiddles twx;  // both directions
  twx.load(tid, tw);
  const float sc =
      wscale[t] * (STORE == kF32 ? 1.0f : inv_norm(tmax, t, nrm_mul));

  PT out[E];
  if constexpr (IT == 1) {
    ramp_pair<X, STORE, WAVE>(g, Y, ya0 + 2 * p, ya0 + 2 * p + 1, ls, tid,
                              twx, tw, sc, out);
    store_row_group<X, STORE, P>(out, s, p, tid, slab, k, t, R0, ny, X, T,
                                 NYB, YB);
  } else {
    const int Rs = R0 + 4 * p;  // this slot's four rows (band-relative)
    const int ys = y0 + Rs;
    PT *sp = stash + p * X;
    ramp_pair<X, STORE, WAVE>(g, Y, ys, ys + 1, ls, tid, twx, tw, sc, out);
#pragma unroll
    for (int m = 0; m < E; ++m) sp[tid + m * NT] = out[m];
    // The second pair recomputes its per-column values (offsets, ramp)
    // instead of keeping the first pair's: held across its transforms they
    // would take more registers than the cap leaves.
    int tid2 = tid;
    asm volatile("" : "+v"(tid2)::"memory");
    ramp_pair<X, STORE, WAVE>(g, Y, ys + 2, ys + 3, ls, tid2, twx, tw, sc,
                              out);
    if (Rs >= ny) return;
    const bool whole = Rs + 4 <= ny;  // Rs % 4 == 0: the run is 8-byte aligned
    // column n of the four rows: base + n YB elements (32-bit byte offsets
    // within one slab row block of X * YB elements)
    char *base = reinterpret_cast<char *>(
        slab + slab_index(k, t, Rs, 0, T, NYB, X, YB));
    const unsigned xb = (unsigned)tid * (unsigned)YB * (unsigned)sizeof(ST);
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const int n = tid + m * NT;
      ST *dst = reinterpret_cast<ST *>(
          base + xb + (unsigned)(m * NT) * (unsigned)YB * (unsigned)sizeof(ST));
      const uint2 w = make_uint2(sp[n], out[m]);
      if (whole) {
        *reinterpret_cast<uint2 *>(dst) = w;
      } else {
        const ST *v = reinterpret_cast<const ST *>(&w);
#pragma unroll
        for (int r = 0; r < 4; ++r)
          if (Rs + r < ny) dst[r] = v[r];
      }
    }
  }
}

// 8192-point halves: a 512-thread workgroup with 68 KB of LDS, so at most
// two waves per SIMD share a CU anyway and the kernel may take 256 VGPRs.
template <int X, int STORE, int P>
__global__ __launch_bounds__(P *(X / default_e(X)))
__attribute__((amdgpu_waves_per_eu(X >= 8192 ? 2 : 3))) void k_filter_ramp(
    const float *__restrict__ src, size_t st, size_t sk, size_t sy, int N,
    int Y, int T, int ndf, const float *__restrict__ ramp,
    const float *__restrict__ wscale, const uint32_t *__restrict__ tmax,
    float nrm_mul, typename StoreT<STORE>::T *__restrict__ slab, int NYB,
    int YB, const float2 *__restrict__ tw, const int2 *__restrict__ krange) {
  constexpr int E = default_e(X);
  constexpr int NT = X / E;
  constexpr bool WAVE = NT <= 64;
  constexpr int L = 2 * X;
  constexpr int SL = fft_lds_len(X, E);
  using PT = typename RowPair<STORE>::T;
  static_assert(E == 16, "sizes of 256 points and more");
  __shared__ float2 s[P * SL];
  const int p = threadIdx.x / NT, tid = threadIdx.x % NT;
  // 1D grid over (row group, tilt, slab), XCD-aware as in k_rows_c2r_ramp
  const unsigned nry = (unsigned)(Y + 2 * P - 1) / (2 * P);
  const unsigned lb = xcd_remap(blockIdx.x, gridDim.x);
  const int t = (int)((lb / nry) % (unsigned)T);
  const int k = (int)(lb / (nry * (unsigned)T));
  if (krange != nullptr && (k < krange[t].x || k > krange[t].y)) return;
  const int R0 = (int)(lb % nry) * (2 * P);
  const int ya = R0 + 2 * p, yb = ya + 1;
  const float *base = src + t * st + k * sk;
  const float *a = base + (size_t)min(ya, Y - 1) * sy;
  const float *b = base + (size_t)min(yb, Y - 1) * sy;
  float2 *ls = s + p * SL;

  typename Fft<X, E, -1, 1, WAVE>::Twiddles twx;  // both directions
  twx.load(tid, tw);
  constexpr int kHalf = kTwN / L;  // W^(kHalf n) = c(n)
  const float2 ct = tw[tid * kHalf];
  const float sc =
      wscale[t] * (STORE == kF32 ? 1.0f : inv_norm(tmax, t, nrm_mul));

  float2 z[E], w[E];
#pragma unroll
  for (int m = 0; m < E; ++m) {
    // clamped loads, then zeroed values (no select between addresses)
    const int n = tid + m * NT;
    const int nc = min(n, N - 1);
    const float va = a[nc], vb = b[nc];
    z[m] = make_float2(ya < Y && n < N ? va : 0.0f,
                       yb < Y && n < N ? vb : 0.0f);
    w[m] = cmul(z[m], cmul(ct, tw[m * (kTwN / (2 * E))]));
  }
  // even bins
  Fft<X, E, -1, 1, WAVE>::run(z, ls, tid, twx);
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int j = 2 * (tid + m * NT);
    const float r = ramp[j <= X ? j : L - j];
    z[m] = make_float2(z[m].x * r, z[m].y * r);
  }
  Fft<X, E, 1, 1, WAVE>::run(z, ls, tid, twx);
  // odd bins
  Fft<X, E, -1, 1, WAVE>::run(w, ls, tid, twx);
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int j = 2 * (tid + m * NT) + 1;
    const float r = ramp[j <= X ? j : L - j];
    w[m] = make_float2(w[m].x * r, w[m].y * r);
  }
  Fft<X, E, 1, 1, WAVE>::run(w, ls, tid, twx);
  PT out[E];
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const float2 c = cmul(ct, tw[m * (kTwN / (2 * E))]);
    const float2 o = cmul(w[m], make_float2(c.x, -c.y));
    out[m] = RowPair<STORE>::pack((z[m].x + o.x) * sc, (z[m].y + o.y) * sc);
  }
  store_row_group<X, STORE, P>(out, s, p, tid, slab, k, t, R0, Y, N, T, NYB,
                               YB);
}

// Unfiltered complex-to-real rows for corr(): out[t][k][y][x] (float).
template <int X>
__global__ __launch_bounds__(2 * (X / default_e(X))) void k_rows_c2r(
    const float2 *__restrict__ G, int Y, int XB, int ndf,
    float *__restrict__ out, const float2 *__restrict__ tw) {
  constexpr int E = default_e(X);
  constexpr int NT = X / E;
  __shared__ float2 s[2][fft_lds_len(X, E)];
  const int p = threadIdx.x / NT, tid = threadIdx.x % NT;
  const int k = blockIdx.y, tl = blockIdx.z;
  const int ya = blockIdx.x * 4 + 2 * p, yb = ya + 1;
  const float2 *g = G + ((size_t)tl * ndf + k) * XB * Y * CB;
  float2 v[E];
#pragma unroll
  for (int m = 0; m < E; ++m)
    v[m] = load_pair_spectrum(g, X, Y, ya, yb, tid + m * NT);
  Fft<X, E, 1, 1>::run(v, s[p], tid, tw);
  float *o = out + ((size_t)tl * ndf + k) * Y * X;
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int n = tid + m * NT;
    if (ya < Y) o[(size_t)ya * X + n] = v[m].x;
    if (yb < Y) o[(size_t)yb * X + n] = v[m].y;
  }
}

// ----------------------------------------------- filter from real rows ----

// Rows r = (k, t, y) with y fastest, read at src + t*st + k*sk + y*sy, each
// N samples, zero-padded to L >= 2N - 1 (the ramp table accounts for L).
// Two row pairs (four rows) per workgroup.
template <int L, int STORE>
__global__ __launch_bounds__(2 * (L / default_e(L))) void k_filter_rows(
    const float *__restrict__ src, size_t st, size_t sk, size_t sy, int N,
    int Y, int T, int ndf, const float *__restrict__ ramp,
    const float *__restrict__ wscale, const uint32_t *__restrict__ tmax,
    float nrm_mul, typename StoreT<STORE>::T *__restrict__ slab, int NYB,
    int YB, const float2 *__restrict__ tw) {
  constexpr int EL = default_e(L);
  constexpr int NT = L / EL;
  __shared__ float2 s[2][fft_lds_len(L, EL)];
  const int p = threadIdx.x / NT, tid = threadIdx.x % NT;
  const long nrows = (long)ndf * T * Y;
  const long r0 = ((long)blockIdx.x * 2 + p) * 2, r1 = r0 + 1;
  int k0 = 0, t0 = 0, y0 = 0, k1 = 0, t1 = 0, y1 = 0;
  if (r0 < nrows) {
    y0 = (int)(r0 % Y);
    t0 = (int)((r0 / Y) % T);
    k0 = (int)(r0 / ((long)Y * T));
  }
  if (r1 < nrows) {
    y1 = (int)(r1 % Y);
    t1 = (int)((r1 / Y) % T);
    k1 = (int)(r1 / ((long)Y * T));
  }
  const float *a = src + t0 * st + k0 * sk + y0 * sy;
  const float *b = src + t1 * st + k1 * sk + y1 * sy;
  float2 u[EL];
#pragma unroll
  for (int m = 0; m < EL; ++m) {
    const int n = tid + m * NT;
    const float va = (r0 < nrows && n < N) ? a[n] : 0.0f;
    const float vb = (r1 < nrows && n < N) ? b[n] : 0.0f;
    u[m] = make_float2(va, vb);
  }
  Fft<L, EL, -1, 1>::run(u, s[p], tid, tw);
#pragma unroll
  for (int m = 0; m < EL; ++m) {
    const int j = tid + m * NT;
    const float r = ramp[j <= L / 2 ? j : L - j];
    u[m] = make_float2(u[m].x * r, u[m].y * r);
  }
  Fft<L, EL, 1, 1>::run(u, s[p], tid, tw);
  using ST = StoreT<STORE>;
  const float sa = r0 < nrows ? wscale[t0] * (STORE == kF32 ? 1.0f : inv_norm(tmax, t0, nrm_mul)) : 0.0f;
  const float sb = r1 < nrows ? wscale[t1] * (STORE == kF32 ? 1.0f : inv_norm(tmax, t1, nrm_mul)) : 0.0f;
#pragma unroll
  for (int m = 0; m < EL; ++m) {
    const int n = tid + m * NT;
    if (n < N) {
      if (r0 < nrows) slab[slab_index(k0, t0, y0, n, T, NYB, N, YB)] = ST::pack(u[m].x * sa);
      if (r1 < nrows) slab[slab_index(k1, t1, y1, n, T, NYB, N, YB)] = ST::pack(u[m].y * sb);
    }
  }
}

// The volume path's filter: one workgroup filters RB = 4 * IT consecutive
// rows y of one (tilt, slab), RB dividing the slab's row blocking, with the
// outputs held in registers and written as whole 16-byte row runs (as
// k_rows_c2r_filter does after the correction).
template <int L, int STORE, int IT>
__global__ __launch_bounds__(2 * (L / default_e(L))) void k_filter_rows_blk(
    const float *__restrict__ src, size_t st, size_t sk, size_t sy, int N,
    int Y, int T, const float *__restrict__ ramp,
    const float *__restrict__ wscale, const uint32_t *__restrict__ tmax,
    float nrm_mul, typename StoreT<STORE>::T *__restrict__ slab, int NYB,
    int YB, const float2 *__restrict__ tw) {
  constexpr int EL = default_e(L);
  constexpr int NT = L / EL;
  using PT = typename RowPair<STORE>::T;
  constexpr bool kWide = (IT * sizeof(PT)) % 16 == 0;
  __shared__ float2 s[2][fft_lds_len(L, EL)];
  const int p = threadIdx.x / NT, tid = threadIdx.x % NT;
  const int t = blockIdx.y, k = blockIdx.z;
  const int r0 = blockIdx.x * (4 * IT) + p * (2 * IT);
  const float *base = src + t * st + k * sk;
  typename Fft<L, EL, -1, 1>::Twiddles twl;
  twl.load(tid, tw);
  const float sc =
      wscale[t] * (STORE == kF32 ? 1.0f : inv_norm(tmax, t, nrm_mul));
  PT out[EL / 2][IT];
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int ya = r0 + 2 * it, yb = ya + 1;
    const float *a = base + (size_t)ya * sy;
    const float *b = base + (size_t)yb * sy;
    float2 u[EL];
#pragma unroll
    for (int m = 0; m < EL; ++m) {
      const int n = tid + m * NT;
      const float va = (ya < Y && n < N) ? a[n] : 0.0f;
      const float vb = (yb < Y && n < N) ? b[n] : 0.0f;
      u[m] = make_float2(va, vb);
    }
    Fft<L, EL, -1, 1>::run(u, s[p], tid, twl);
#pragma unroll
    for (int m = 0; m < EL; ++m) {
      const int j = tid + m * NT;
      const float r = ramp[j <= L / 2 ? j : L - j];
      u[m] = make_float2(u[m].x * r, u[m].y * r);
    }
    Fft<L, EL, 1, 1>::run(u, s[p], tid, twl);
#pragma unroll
    for (int m = 0; m < EL / 2; ++m)
      out[m][it] = RowPair<STORE>::pack(u[m].x * sc, u[m].y * sc);
  }
  using ST = typename StoreT<STORE>::T;
  if (r0 >= Y) return;
  const bool whole = r0 + 2 * IT <= Y;
#pragma unroll
  for (int m = 0; m < EL / 2; ++m) {
    const int n = tid + m * NT;
    if (n >= N) continue;
    ST *dst = slab + slab_index(k, t, r0, n, T, NYB, N, YB);
    if (whole && kWide) {
      uint4 *d4 = reinterpret_cast<uint4 *>(dst);
#pragma unroll
      for (int q = 0; q < (int)(IT * sizeof(PT) / 16); ++q) {
        uint4 w;
        __builtin_memcpy(&w, reinterpret_cast<const char *>(out[m]) + 16 * q,
                         16);
        d4[q] = w;
      }
    } else if (whole) {
      PT *dp = reinterpret_cast<PT *>(dst);
#pragma unroll
      for (int it = 0; it < IT; ++it) dp[it] = out[m][it];
    } else {
      const ST *src2 = reinterpret_cast<const ST *>(out[m]);
#pragma unroll
      for (int r = 0; r < 2 * IT; ++r)  // constant indices keep out[] in VGPRs
        if (r0 + r < Y) dst[r] = src2[r];
    }
  }
}

// max |row| per tilt over rows (k, t, y), for reduced-precision scaling:
// one workgroup per (slab, tilt, group of kAbsRows rows), one atomic each.
constexpr int kAbsRows = 16;
__global__ void k_absmax_rows(const float *__restrict__ src, size_t st,
                              size_t sk, size_t sy, int N, int Y, int T,
                              int ndf, uint32_t *__restrict__ tmax) {
  __shared__ float red[4];
  const int ng = (Y + kAbsRows - 1) / kAbsRows;
  const long b = blockIdx.x;
  const int g = (int)(b % ng), t = (int)((b / ng) % T);
  const int k = (int)(b / ((long)ng * T));
  const float *base = src + t * st + k * sk;
  const int y0 = g * kAbsRows, y1 = min(Y, y0 + kAbsRows);
  float m = 0.0f;
  for (int y = y0; y < y1; ++y) {
    const float *a = base + (size_t)y * sy;
    for (int n = threadIdx.x; n < N; n += blockDim.x)
      m = fmaxf(m, fabsf(a[n]));
  }
  m = block_max<256>(m, red);
  if (threadIdx.x == 0) atomicMax(&tmax[t], __float_as_uint(m));
}

// ------------------------------------------ any size: Bluestein transforms -
//
// An n-point DFT for n not a power of two (or below the power-of-two kernels'
// minimum) as a chirp-z transform on the M-point power-of-two FFTs, M >=
// 2n - 1: with w_m = exp(-i pi m^2 / n), nk = (n^2 + k^2 - (n-k)^2) / 2 gives
//   forward  X_k = w_k sum_j (x_j w_j) conj(w_(k-j)),
//   inverse  x_j = conj(w_j) sum_k (X_k conj(w_k)) w_(j-k),
// a circular convolution of length M computed as IFFT(FFT(u) FFT(b)) with
// FFT(b) / M tabulated per plan (BluesteinTable). The chirp and the kernel
// spectra are computed in FP64 on the host (m^2 mod 2n in integers) and
// rounded once.
// v[m] holds x[tid + m NT] for tid + m NT < n (anything beyond is ignored);
// on return v[m] = X[tid + m NT] for indices < n and 0 beyond.
template <int M, int E, int DIR, int LS>
__device__ __forceinline__ void bluestein(
    float2 (&v)[E], float2 *lds, int tid,
    const typename Fft<M, E, -1, LS>::Twiddles &tw, const BluesteinTable &b) {
  constexpr int NT = M / E;
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int j = tid + m * NT;
    if (j < b.n) {
      const float2 c = b.chirp[j];
      v[m] = cmul(v[m], DIR < 0 ? c : make_float2(c.x, -c.y));
    } else {
      v[m] = make_float2(0.0f, 0.0f);
    }
  }
  Fft<M, E, -1, LS>::run(v, lds, tid, tw);
  const float2 *B = DIR < 0 ? b.bf : b.bi;
#pragma unroll
  for (int m = 0; m < E; ++m) v[m] = cmul(v[m], B[tid + m * NT]);
  Fft<M, E, 1, LS>::run(v, lds, tid, tw);
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int j = tid + m * NT;
    if (j < b.n) {
      const float2 c = b.chirp[j];
      v[m] = cmul(v[m], DIR < 0 ? c : make_float2(c.x, -c.y));
    } else {
      v[m] = make_float2(0.0f, 0.0f);
    }
  }
}

// Bluestein row passes: P row pairs per workgroup, one M-point slot each
// (P = 2; P = 1 at M = 16384, whose slot is a whole 1024-thread workgroup
// with 139 KB of LDS).
constexpr int bs_row_slots(int M) { return M >= 16384 ? 1 : 2; }

// Pass 1 for any X: rows (ya, yb) as one complex row, forward X-point DFT,
// then the Hermitian split into the two half spectra (as k_rows_r2c).
template <int M, int P = bs_row_slots(M)>
__global__ __launch_bounds__(P * (M / default_e(M))) void k_rows_r2c_bs(
    const float *__restrict__ img, size_t tstride, int X, int Y, int XB,
    float2 *__restrict__ F, uint32_t *__restrict__ tmax, int t0,
    const float2 *__restrict__ tw, BluesteinTable bx) {
  constexpr int E = default_e(M);
  constexpr int NT = M / E;
  constexpr int RB = 2 * P;  // rows per workgroup
  __shared__ float2 s[P][fft_lds_len(M, E)];
  __shared__ float red[16];
  const int p = threadIdx.x / NT, tid = threadIdx.x % NT;
  const int t = blockIdx.y;
  const int ya = blockIdx.x * RB + 2 * p, yb = ya + 1;
  const float *im = img + (size_t)t * tstride;
  typename Fft<M, E, -1, 1>::Twiddles twm;
  twm.load(tid, tw);
  float2 v[E];
  float amax = 0.0f;
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int n = tid + m * NT;
    const float a = (ya < Y && n < X) ? im[(size_t)ya * X + n] : 0.0f;
    const float b = (yb < Y && n < X) ? im[(size_t)yb * X + n] : 0.0f;
    v[m] = make_float2(a, b);
    amax = fmaxf(amax, fmaxf(fabsf(a), fabsf(b)));
  }
  bluestein<M, E, -1, 1>(v, s[p], tid, twm, bx);
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int n = tid + m * NT;
    if (n < X) s[p][n] = v[m];
  }
  lds_barrier();
  if (tmax != nullptr) {
    amax = block_max<P * NT>(amax, red);
    if (threadIdx.x == 0) atomicMax(&tmax[t0 + t], __float_as_uint(amax));
  }
  const int ncol = XB * CB;
  for (int e = threadIdx.x; e < RB * ncol; e += P * NT) {
    const int ci = e % CB, r = (e / CB) % RB, cb = e / (RB * CB);
    const int kx = cb * CB + ci, y = blockIdx.x * RB + r;
    if (y >= Y) continue;
    float2 val = make_float2(0.0f, 0.0f);
    if (kx <= X / 2) {
      const float2 z = s[r / 2][kx];
      const float2 zm = s[r / 2][(X - kx) % X];
      val = (r & 1) == 0 ? make_float2((z.x + zm.x) * 0.5f, (z.y - zm.y) * 0.5f)
                         : make_float2((z.y + zm.y) * 0.5f, (zm.x - z.x) * 0.5f);
    }
    F[(((size_t)t * XB + cb) * Y + y) * CB + ci] = val;
  }
}

// Pass 2 for any Y: CW of the CB columns of a block per workgroup (LDS
// holds CW M-point transforms), forward Y-point DFT, then per defocus the
// flip and the inverse Y-point DFT.
template <int M, int CW>
__global__ __launch_bounds__(CW *(M / default_e(M))) void k_cols_bs(
    const float2 *__restrict__ F, const int8_t *__restrict__ mask,
    float2 *__restrict__ G, int Y, int XB, int ndf,
    const float2 *__restrict__ tw, BluesteinTable by, int ylo, int yhi,
    const int2 *__restrict__ krange) {
  constexpr int E = default_e(M);
  constexpr int NT = M / E;
  static_assert(CB % CW == 0, "column groups split a block");
  __shared__ float2 s[fft_lds_len(M, E) * CW];
  const int c = threadIdx.x % CW, tid = threadIdx.x / CW;
  const int cb = blockIdx.x / (CB / CW);
  const int ci = (blockIdx.x % (CB / CW)) * CW + c;
  const int t = blockIdx.y;
  const float2 *src = F + ((size_t)t * XB + cb) * Y * CB;
  typename Fft<M, E, -1, CW>::Twiddles twm;
  twm.load(tid, tw);
  float2 v[E];
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int y = tid + m * NT;
    v[m] = y < Y ? src[(size_t)y * CB + ci] : make_float2(0.0f, 0.0f);
  }
  bluestein<M, E, -1, CW>(v, s + c, tid, twm, by);
  const int2 kr = krange != nullptr ? krange[t] : make_int2(0, ndf - 1);
  for (int k = kr.x; k <= kr.y; ++k) {
    // as in k_cols: indices laundered and twiddles reloaded per defocus, so
    // that addresses and twiddles are not held across the loop
    int tk = tid, ck = ci;
    asm volatile("" : "+v"(tk), "+v"(ck));
    const int8_t *mk = mask + ((size_t)k * XB + cb) * Y * CB;
    float2 w[E];
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const int y = tk + m * NT;
      const float f = y < Y ? mask_value(mk[(size_t)y * CB + ck]) : 0.0f;
      w[m] = make_float2(v[m].x * f, v[m].y * f);
    }
    typename Fft<M, E, -1, CW>::Twiddles twk;
    twk.load(tk, tw);
    bluestein<M, E, 1, CW>(w, s + c, tk, twk, by);
    float2 *dst = G + (((size_t)t * ndf + k) * XB + cb) * Y * CB + ck;
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const int y = tk + m * NT;
      if (y >= ylo && y < yhi) dst[(size_t)y * CB] = w[m];
    }
  }
}

// Pass 3 for any X, unfiltered (corr()): out[t][k][y][x].
template <int M, int P = bs_row_slots(M)>
__global__ __launch_bounds__(P * (M / default_e(M))) void k_rows_c2r_bs(
    const float2 *__restrict__ G, int X, int Y, int XB, int ndf,
    float *__restrict__ out, const float2 *__restrict__ tw, BluesteinTable bx) {
  constexpr int E = default_e(M);
  constexpr int NT = M / E;
  __shared__ float2 s[P][fft_lds_len(M, E)];
  const int p = threadIdx.x / NT, tid = threadIdx.x % NT;
  const int k = blockIdx.y, tl = blockIdx.z;
  const int ya = blockIdx.x * (2 * P) + 2 * p, yb = ya + 1;
  const float2 *g = G + ((size_t)tl * ndf + k) * XB * Y * CB;
  typename Fft<M, E, -1, 1>::Twiddles twm;
  twm.load(tid, tw);
  float2 v[E];
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int n = tid + m * NT;
    v[m] = n < X ? load_pair_spectrum(g, X, Y, ya, yb, n)
                 : make_float2(0.0f, 0.0f);
  }
  bluestein<M, E, 1, 1>(v, s[p], tid, twm, bx);
  float *o = out + ((size_t)tl * ndf + k) * Y * X;
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int n = tid + m * NT;
    if (n < X) {
      if (ya < Y) o[(size_t)ya * X + n] = v[m].x;
      if (yb < Y) o[(size_t)yb * X + n] = v[m].y;
    }
  }
}

// Pass 3 for any X, fused with the ramp filter: the corrected rows (X
// samples, unnormalised inverse) are zero-padded to M >= 2X - 1 and filtered
// with the truncated-kernel ramp table, as k_filter_rows does for real rows.
template <int M, int STORE, int P = bs_row_slots(M)>
__global__ __launch_bounds__(P * (M / default_e(M))) void
    k_rows_c2r_filter_bs(const float2 *__restrict__ G, int X, int Y, int XB,
                         int ndf, int t0, int T,
                         const float *__restrict__ ramp,
                         const float *__restrict__ wscale,
                         const uint32_t *__restrict__ tmax, float nrm_mul,
                         typename StoreT<STORE>::T *__restrict__ slab,
                         int NYB, int YB, const float2 *__restrict__ tw,
                         BluesteinTable bx, int y0, int ny,
                         const int2 *__restrict__ krange) {
  constexpr int E = default_e(M);
  constexpr int NT = M / E;
  __shared__ float2 s[P][fft_lds_len(M, E)];
  const int p = threadIdx.x / NT, tid = threadIdx.x % NT;
  const int k = blockIdx.y, tl = blockIdx.z, t = t0 + tl;
  if (krange != nullptr && (k < krange[t].x || k > krange[t].y)) return;
  // band-relative rows ra, rb of the slab; spectrum rows y0 + ra, y0 + rb
  const int ra = blockIdx.x * (2 * P) + 2 * p, rb = ra + 1;
  const int ya = y0 + ra, yb = ya + 1;
  const float2 *g = G + ((size_t)tl * ndf + k) * XB * Y * CB;
  typename Fft<M, E, -1, 1>::Twiddles twm;
  twm.load(tid, tw);
  float2 v[E];
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int n = tid + m * NT;
    v[m] = n < X ? load_pair_spectrum(g, X, Y, ya, yb, n)
                 : make_float2(0.0f, 0.0f);
  }
  bluestein<M, E, 1, 1>(v, s[p], tid, twm, bx);  // zero beyond X
  Fft<M, E, -1, 1>::run(v, s[p], tid, twm);
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int j = tid + m * NT;
    const float r = ramp[j <= M / 2 ? j : M - j];
    v[m] = make_float2(v[m].x * r, v[m].y * r);
  }
  Fft<M, E, 1, 1>::run(v, s[p], tid, twm);
  const float sc =
      wscale[t] * (STORE == kF32 ? 1.0f : inv_norm(tmax, t, nrm_mul));
  using ST = StoreT<STORE>;
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int n = tid + m * NT;
    if (n < X) {
      if (ra < ny) slab[slab_index(k, t, ra, n, T, NYB, X, YB)] = ST::pack(v[m].x * sc);
      if (rb < ny) slab[slab_index(k, t, rb, n, T, NYB, X, YB)] = ST::pack(v[m].y * sc);
    }
  }
}

// ------------------------------------------------------------ launchers ---

#define GX_POW2_CASES(M) \
  M(4) M(8) M(16) M(32) M(64) M(128) M(256) M(512) M(1024) M(2048) M(4096) \
  M(8192)

hipError_t launch_rows_r2c(const float *img, size_t tstride, int X, int Y,
                           int XB, int ntilts, float2 *F, uint32_t *tmax,
                           int t0, hipStream_t s) {
  hipError_t err;
  const float2 *tw = twiddles(&err);
  if (err != hipSuccess) return err;
  // more row pairs per workgroup at 512 to 2048 points (each column block
  // then written as 256 or 512 contiguous bytes), two elsewhere
  switch (X) {
#define C_(n)                                                                  \
  case n: {                                                                    \
    constexpr int rp =                                                         \
        n == 1024 ? 8 : (n == 512 || n == 2048) ? 4 : n == 8192 ? 1 : 2;       \
    const dim3 grid((Y + 2 * rp - 1) / (2 * rp), ntilts);                      \
    hipLaunchKernelGGL((k_rows_r2c<n, rp>), grid,                              \
                       dim3(rp * (n / default_e(n))), 0, s, img, tstride, Y,   \
                       XB, F, tmax, t0, tw);                                   \
    break;                                                                     \
  }
    GX_POW2_CASES(C_)
#undef C_
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <typename M>
static hipError_t launch_cols_t(const float2 *F, const M *mask, float2 *G,
                                int Y, int XB, int ndf, int ntilts,
                                hipStream_t s, int ylo, int yhi,
                                const int2 *krange) {
  if (yhi < 0) yhi = Y;
  hipError_t err;
  const float2 *tw = twiddles(&err);
  if (err != hipSuccess) return err;
  const dim3 grid(XB * ntilts);  // one column part per column block
  switch (Y) {
#define C_(n)                                                              \
  case n:                                                                  \
    hipLaunchKernelGGL((k_cols<n, M, CB>), grid,                           \
                       dim3(CB * (n / default_e(n))), 0, s, F, mask, G, XB, \
                       ndf, tw, ylo, yhi, krange);                         \
    break;
    C_(4) C_(8) C_(16) C_(32) C_(64) C_(128) C_(256) C_(512) C_(1024)
    C_(2048)
#undef C_
    // two columns per workgroup (512 threads and 70 KB of LDS at 4096, 1024
    // and 139 KB at 8192), the two halves of a column block on one XCD
#define C_(n)                                                              \
  case n:                                                                  \
    hipLaunchKernelGGL((k_cols<n, M, 2>), dim3(2 * XB * ntilts),           \
                       dim3(2 * (n / default_e(n))), 0, s, F, mask, G, XB, \
                       ndf, tw, ylo, yhi, krange);                         \
    break;
    C_(4096) C_(8192)
#undef C_
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_cols(const float2 *F, const int8_t *mask, float2 *G, int Y,
                       int XB, int ndf, int ntilts, hipStream_t s, int ylo,
                       int yhi, const int2 *krange) {
  return launch_cols_t<int8_t>(F, mask, G, Y, XB, ndf, ntilts, s, ylo, yhi,
                               krange);
}

hipError_t launch_cols_transfer(const float2 *F, const float *transfer,
                                float2 *G, int Y, int XB, int ntilts,
                                hipStream_t s) {
  return launch_cols_t<float>(F, transfer, G, Y, XB, 1, ntilts, s, 0, Y,
                              nullptr);
}

// Rows per workgroup: the slab's row blocking YB, at most 16 (float32) or
// 32 (16-bit) rows; 4 for 4096-point rows, whose 1024-thread workgroups
// leave no registers for held outputs.
inline int c2r_filter_rows(int X, int store, int YB) {
  if (X >= 4096) return 4;
  const int cap = store == kF32 ? 16 : 32;
  return YB < cap ? YB : cap;
}

template <int STORE, int IT>
static hipError_t launch_c2r_filter_it(const float2 *G, int X, int Y, int XB,
                                       int ndf, int t0, int ntilts, int T,
                                       const float *wscale,
                                       const uint32_t *tmax, float nrm_mul,
                                       void *slab, int NYB, int YB,
                                       hipStream_t s, const float2 *tw,
                                       int y0, int ny, const int2 *krange) {
  using ST = typename StoreT<STORE>::T;
  dim3 grid((ny + 4 * IT - 1) / (4 * IT), ndf, ntilts);
  switch (X) {
#define C_(n)                                                                  \
  case n:                                                                      \
    hipLaunchKernelGGL((k_rows_c2r_filter<n, STORE, IT>), grid,                \
                       dim3(2 * (2 * n / default_e(2 * n))), 0, s, G, Y, XB,   \
                       ndf, t0, T, wscale, tmax, nrm_mul, (ST *)slab, NYB, YB, \
                       tw, y0, ny, krange);                                    \
    break;
    C_(8) C_(16) C_(32) C_(64) C_(128)
#undef C_
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// Slots (row pairs) per workgroup of k_rows_c2r_ramp: RB = 2P rows dividing
// the slab's row blocking. One-wave slots (X <= 1024) run 4 to a workgroup;
// the 2- and 4-wave slots of 2048 and 4096 points 2 and 1, so that
// workgroups that wait at each other's barriers stay small.
inline int c2r_ramp_slots(int X, int YB) {
  const int cap = X <= 1024 ? 4 : (X == 2048 ? 2 : 1);
  const int p = YB / 2 < cap ? YB / 2 : cap;
  return p < 1 ? 1 : p;  // YB = 1: one slot, rows stored one by one
}

template <int X, int STORE, int P, int IT = 1>
static void launch_c2r_ramp_x(dim3 grid, const float2 *G, int Y, int XB,
                              int ndf, int t0, int T, const float *wscale,
                              const uint32_t *tmax, float nrm_mul, void *slab,
                              int NYB, int YB, hipStream_t s,
                              const float2 *tw, int y0, int ny,
                              const int2 *pairs) {
  using ST = typename StoreT<STORE>::T;
  hipLaunchKernelGGL((k_rows_c2r_ramp<X, STORE, P, IT>), grid,
                     dim3(P * (X / default_e(X))), 0, s, G, Y, XB, ndf, t0, T,
                     wscale, tmax, nrm_mul, (ST *)slab, NYB, YB, tw, y0, ny,
                     pairs);
}

template <int STORE, int P>
static hipError_t launch_c2r_ramp_p(const float2 *G, int X, int Y, int XB,
                                    int ndf, int t0, int ntilts, int T,
                                    const float *wscale, const uint32_t *tmax,
                                    float nrm_mul, void *slab, int NYB, int YB,
                                    hipStream_t s, const float2 *tw, int y0,
                                    int ny, const int2 *pairs, int npairs) {
  // two row pairs per slot at 4096 and 8192 points with 16-bit storage
  const int rb = X >= 4096 && STORE != kF32 ? 4 * P : 2 * P;
  const size_t nb = (size_t)((ny + rb - 1) / rb) *
                    (pairs != nullptr ? (size_t)npairs : (size_t)ndf * ntilts);
  if (nb == 0) return hipSuccess;
  if (nb > 0xffffffffu) return hipErrorInvalidValue;
  dim3 grid((unsigned)nb);
#define A_ grid, G, Y, XB, ndf, t0, T, wscale, tmax, nrm_mul, slab, NYB, YB, s, tw, y0, ny, pairs
  switch (X) {
    case 256: launch_c2r_ramp_x<256, STORE, P>(A_); break;
    case 512: launch_c2r_ramp_x<512, STORE, P>(A_); break;
    case 1024: launch_c2r_ramp_x<1024, STORE, P>(A_); break;
    case 2048:
      if constexpr (P <= 2) {
        launch_c2r_ramp_x<2048, STORE, P>(A_);
        break;
      }
      return hipErrorInvalidValue;
    case 4096:
      if constexpr (P == 1 && STORE != kF32) {
        launch_c2r_ramp_x<4096, STORE, 1, 2>(A_);
        break;
      } else if constexpr (P == 1) {
        launch_c2r_ramp_x<4096, STORE, 1>(A_);
        break;
      }
      return hipErrorInvalidValue;
    case 8192:
      if constexpr (P == 1 && STORE != kF32) {
        launch_c2r_ramp_x<8192, STORE, 1, 2>(A_);
        break;
      } else if constexpr (P == 1) {
        launch_c2r_ramp_x<8192, STORE, 1>(A_);
        break;
      }
      return hipErrorInvalidValue;
    default:
      return hipErrorInvalidValue;
  }
#undef A_
  return hipGetLastError();
}

template <int STORE>
static hipError_t launch_c2r_filter_t(const float2 *G, int X, int Y, int XB,
                                      int ndf, int t0, int ntilts, int T,
                                      const float *wscale,
                                      const uint32_t *tmax, float nrm_mul,
                                      void *slab, int NYB, int YB,
                                      hipStream_t s, const float2 *tw, int y0,
                                      int ny, const int2 *krange,
                                      const int2 *pairs, int npairs) {
#define A_ G, X, Y, XB, ndf, t0, ntilts, T, wscale, tmax, nrm_mul, slab, NYB, YB, s, tw, y0, ny, pairs, npairs
  if (X >= 256 && YB >= 8) {
    switch (c2r_ramp_slots(X, YB)) {
      case 1: return launch_c2r_ramp_p<STORE, 1>(A_);
      case 2: return launch_c2r_ramp_p<STORE, 2>(A_);
      case 4: return launch_c2r_ramp_p<STORE, 4>(A_);
    }
    return hipErrorInvalidValue;
  }
#undef A_
#define A_ G, X, Y, XB, ndf, t0, ntilts, T, wscale, tmax, nrm_mul, slab, NYB, YB, s, tw, y0, ny, krange
  const int rb = c2r_filter_rows(X, STORE, YB);
  if (YB % rb != 0) return hipErrorInvalidValue;
  if (rb == 4) return launch_c2r_filter_it<STORE, 1>(A_);
  if constexpr (STORE == kF32) {
    if (rb == 8) return launch_c2r_filter_it<STORE, 2>(A_);
    if (rb == 16) return launch_c2r_filter_it<STORE, 4>(A_);
  } else {
    if (rb == 16) return launch_c2r_filter_it<STORE, 4>(A_);
    if (rb == 32) return launch_c2r_filter_it<STORE, 8>(A_);
  }
#undef A_
  return hipErrorInvalidValue;
}

hipError_t launch_rows_c2r_filter(const float2 *G, int X, int Y, int XB,
                                  int ndf, int t0, int ntilts, int T,
                                  const float *wscale,
                                  const uint32_t *tmax, float nrm_mul,
                                  int store, void *slab, int NYB, int YB,
                                  hipStream_t s, int y0, int ny,
                                  const int2 *krange, const int2 *pairs,
                                  int npairs) {
  hipError_t err;
  const float2 *tw = twiddles(&err);
  if (err != hipSuccess) return err;
  if (ny < 0) ny = Y - y0;
  if (y0 < 0 || y0 % YB != 0 || ny <= 0 || y0 + ny > Y)
    return hipErrorInvalidValue;
  switch (store) {
    case kF32:
      return launch_c2r_filter_t<kF32>(G, X, Y, XB, ndf, t0, ntilts, T,
                                       wscale, tmax, nrm_mul, slab, NYB, YB,
                                       s, tw, y0, ny, krange, pairs, npairs);
    case kF16:
      return launch_c2r_filter_t<kF16>(G, X, Y, XB, ndf, t0, ntilts, T,
                                       wscale, tmax, nrm_mul, slab, NYB, YB,
                                       s, tw, y0, ny, krange, pairs, npairs);
    case kBF16:
      return launch_c2r_filter_t<kBF16>(G, X, Y, XB, ndf, t0, ntilts, T,
                                        wscale, tmax, nrm_mul, slab,
                                        NYB, YB, s, tw, y0, ny, krange, pairs,
                                        npairs);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_rows_c2r(const float2 *G, int X, int Y, int XB, int ndf,
                           int ntilts, float *out, hipStream_t s) {
  hipError_t err;
  const float2 *tw = twiddles(&err);
  if (err != hipSuccess) return err;
  dim3 grid((Y + 3) / 4, ndf, ntilts);
  switch (X) {
#define C_(n)                                                                 \
  case n:                                                                     \
    hipLaunchKernelGGL(k_rows_c2r<n>, grid, dim3(2 * (n / default_e(n))), 0,  \
                       s, G, Y, XB, ndf, out, tw);                            \
    break;
    GX_POW2_CASES(C_)
#undef C_
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int STORE>
static hipError_t launch_filter_rows_t(const float *src, size_t st, size_t sk,
                                       size_t sy, int N, int L, int Y, int T,
                                       int ndf, const float *ramp,
                                       const float *wscale,
                                       const uint32_t *tmax, float nrm_mul,
                                       void *slab, int NYB, int YB,
                                       hipStream_t s, const float2 *tw) {
  using ST = typename StoreT<STORE>::T;
  const long nrows = (long)ndf * T * Y;
  dim3 grid((unsigned)((nrows + 3) / 4));
  switch (L) {
#define C_(n)                                                                 \
  case n:                                                                     \
    hipLaunchKernelGGL((k_filter_rows<n, STORE>), grid,                       \
                       dim3(2 * (n / default_e(n))), 0, s, src, st, sk, sy,   \
                       N, Y, T, ndf, ramp, wscale, tmax, nrm_mul, (ST *)slab, \
                       NYB, YB, tw);                                          \
    break;
    C_(8) C_(16) C_(32) C_(64) C_(128) C_(256) C_(512) C_(1024) C_(2048)
    C_(4096) C_(8192)
#undef C_
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int STORE, int IT>
static hipError_t launch_filter_blk_it(const float *src, size_t st, size_t sk,
                                       size_t sy, int N, int L, int Y, int T,
                                       int ndf, const float *ramp,
                                       const float *wscale,
                                       const uint32_t *tmax, float nrm_mul,
                                       void *slab, int NYB, int YB,
                                       hipStream_t s, const float2 *tw) {
  using ST = typename StoreT<STORE>::T;
  dim3 grid((Y + 4 * IT - 1) / (4 * IT), T, ndf);
  switch (L) {
#define C_(n)                                                                  \
  case n:                                                                      \
    hipLaunchKernelGGL((k_filter_rows_blk<n, STORE, IT>), grid,                \
                       dim3(2 * (n / default_e(n))), 0, s, src, st, sk, sy, N, \
                       Y, T, ramp, wscale, tmax, nrm_mul, (ST *)slab, NYB, YB, \
                       tw);                                                    \
    break;
    C_(8) C_(16) C_(32) C_(64) C_(128) C_(256)
#undef C_
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int X, int STORE, int P>
static void launch_filter_ramp_x(dim3 grid, const float *src, size_t st,
                                 size_t sk, size_t sy, int N, int Y, int T,
                                 int ndf, const float *ramp,
                                 const float *wscale, const uint32_t *tmax,
                                 float nrm_mul, void *slab, int NYB, int YB,
                                 hipStream_t s, const float2 *tw,
                                 const int2 *krange) {
  using ST = typename StoreT<STORE>::T;
  hipLaunchKernelGGL((k_filter_ramp<X, STORE, P>), grid,
                     dim3(P * (X / default_e(X))), 0, s, src, st, sk, sy, N,
                     Y, T, ndf, ramp, wscale, tmax, nrm_mul, (ST *)slab, NYB,
                     YB, tw, krange);
}

template <int STORE, int P>
static hipError_t launch_filter_ramp_p(const float *src, size_t st, size_t sk,
                                       size_t sy, int N, int L, int Y, int T,
                                       int ndf, const float *ramp,
                                       const float *wscale,
                                       const uint32_t *tmax, float nrm_mul,
                                       void *slab, int NYB, int YB,
                                       hipStream_t s, const float2 *tw,
                                       const int2 *krange) {
  const size_t nb = (size_t)((Y + 2 * P - 1) / (2 * P)) * T * ndf;
  if (nb > 0xffffffffu) return hipErrorInvalidValue;
  dim3 grid((unsigned)nb);
#define A_ grid, src, st, sk, sy, N, Y, T, ndf, ramp, wscale, tmax, nrm_mul, slab, NYB, YB, s, tw, krange
  switch (L / 2) {
    case 256: launch_filter_ramp_x<256, STORE, P>(A_); break;
    case 512: launch_filter_ramp_x<512, STORE, P>(A_); break;
    case 1024: launch_filter_ramp_x<1024, STORE, P>(A_); break;
    case 2048:
      if constexpr (P <= 2) {
        launch_filter_ramp_x<2048, STORE, P>(A_);
        break;
      }
      return hipErrorInvalidValue;
    case 4096:
      if constexpr (P <= 1) {
        launch_filter_ramp_x<4096, STORE, P>(A_);
        break;
      }
      return hipErrorInvalidValue;
    case 8192:
      if constexpr (P <= 1) {
        launch_filter_ramp_x<8192, STORE, P>(A_);
        break;
      }
      return hipErrorInvalidValue;
    default:
      return hipErrorInvalidValue;
  }
#undef A_
  return hipGetLastError();
}

template <int STORE>
static hipError_t launch_filter_ramp_t(const float *src, size_t st, size_t sk,
                                       size_t sy, int N, int L, int Y, int T,
                                       int ndf, const float *ramp,
                                       const float *wscale,
                                       const uint32_t *tmax, float nrm_mul,
                                       void *slab, int NYB, int YB,
                                       hipStream_t s, const float2 *tw,
                                       const int2 *krange) {
#define A_ src, st, sk, sy, N, L, Y, T, ndf, ramp, wscale, tmax, nrm_mul, slab, NYB, YB, s, tw, krange
  switch (c2r_ramp_slots(L / 2, YB)) {
    case 1: return launch_filter_ramp_p<STORE, 1>(A_);
    case 2: return launch_filter_ramp_p<STORE, 2>(A_);
    case 4: return launch_filter_ramp_p<STORE, 4>(A_);
  }
#undef A_
  return hipErrorInvalidValue;
}

hipError_t launch_filter_rows(const float *src, size_t st, size_t sk,
                              size_t sy, int N, int L, int Y, int T, int ndf,
                              const float *ramp, const float *wscale,
                              const uint32_t *tmax, float nrm_mul, int store,
                              void *slab, int NYB, int YB, hipStream_t s,
                              const float2 *tw, const int2 *krange) {
  if (tw == nullptr) {
    hipError_t err;
    tw = twiddles(&err);
    if (err != hipSuccess) return err;
  }
  // 512- to 16384-point filters: four half-length transforms per row pair
  // (16384 points, rows of 4097 to 8192: at any row blocking, one slot)
  if (L >= 512 && (L == 16384 || (L <= 8192 && YB >= 8))) {
#define A_ src, st, sk, sy, N, L, Y, T, ndf, ramp, wscale, tmax, nrm_mul, slab, NYB, YB, s, tw, krange
    switch (store) {
      case kF32: return launch_filter_ramp_t<kF32>(A_);
      case kF16: return launch_filter_ramp_t<kF16>(A_);
      case kBF16: return launch_filter_ramp_t<kBF16>(A_);
    }
#undef A_
    return hipErrorInvalidValue;
  }
  // shorter filters: row blocks of whole 16-byte runs where the slab layout
  // allows them
  const int rb = c2r_filter_rows(L / 2, store, YB);
  if (L <= 256 && YB >= 8 && YB % rb == 0) {
#define A_ src, st, sk, sy, N, L, Y, T, ndf, ramp, wscale, tmax, nrm_mul, slab, NYB, YB, s, tw
    if (store == kF32) {
      if (rb == 8) return launch_filter_blk_it<kF32, 2>(A_);
      if (rb == 16) return launch_filter_blk_it<kF32, 4>(A_);
    } else if (store == kF16) {
      if (rb == 16) return launch_filter_blk_it<kF16, 4>(A_);
      if (rb == 32) return launch_filter_blk_it<kF16, 8>(A_);
    } else if (store == kBF16) {
      if (rb == 16) return launch_filter_blk_it<kBF16, 4>(A_);
      if (rb == 32) return launch_filter_blk_it<kBF16, 8>(A_);
    }
#undef A_
  }
  switch (store) {
    case kF32:
      return launch_filter_rows_t<kF32>(src, st, sk, sy, N, L, Y, T, ndf,
                                        ramp, wscale, tmax, nrm_mul, slab,
                                        NYB, YB, s, tw);
    case kF16:
      return launch_filter_rows_t<kF16>(src, st, sk, sy, N, L, Y, T, ndf,
                                        ramp, wscale, tmax, nrm_mul, slab,
                                        NYB, YB, s, tw);
    case kBF16:
      return launch_filter_rows_t<kBF16>(src, st, sk, sy, N, L, Y, T, ndf,
                                         ramp, wscale, tmax, nrm_mul, slab,
                                         NYB, YB, s, tw);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_absmax_rows(const float *src, size_t st, size_t sk,
                              size_t sy, int N, int Y, int T, int ndf,
                              uint32_t *tmax, hipStream_t s) {
  const long nb = (long)ndf * T * ((Y + kAbsRows - 1) / kAbsRows);
  if (nb > 0xffffffffL) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_absmax_rows, dim3((unsigned)nb), dim3(256), 0, s,
                     src, st, sk, sy, N, Y, T, ndf, tmax);
  return hipGetLastError();
}

}  // namespace gx

namespace gx {

#define GX_BS_CASES(M) \
  M(8) M(16) M(32) M(64) M(128) M(256) M(512) M(1024) M(2048) M(4096) M(8192) \
  M(16384)

hipError_t launch_rows_r2c_bs(const float *img, size_t tstride, int X, int Y,
                              int XB, int ntilts, float2 *F, uint32_t *tmax,
                              int t0, int M, const BluesteinTable &bx,
                              hipStream_t s) {
  hipError_t err;
  const float2 *tw = twiddles(&err);
  if (err != hipSuccess) return err;
  switch (M) {
#define C_(n)                                                                \
  case n: {                                                                  \
    constexpr int P = bs_row_slots(n);                                       \
    const dim3 grid((Y + 2 * P - 1) / (2 * P), ntilts);                      \
    hipLaunchKernelGGL(k_rows_r2c_bs<n>, grid, dim3(P * (n / default_e(n))), \
                       0, s, img, tstride, X, Y, XB, F, tmax, t0, tw, bx);   \
    break;                                                                   \
  }
    GX_BS_CASES(C_)
#undef C_
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_cols_bs(const float2 *F, const int8_t *mask, float2 *G,
                          int Y, int XB, int ndf, int ntilts, int M,
                          const BluesteinTable &by, hipStream_t s, int ylo,
                          int yhi, const int2 *krange) {
  if (yhi < 0 || yhi > Y) yhi = Y;
  hipError_t err;
  const float2 *tw = twiddles(&err);
  if (err != hipSuccess) return err;
  // CB columns of a block per workgroup up to 2048-point transforms, one
  // column at 4096 and 8192 points (1024-thread workgroups would cap the
  // registers at 128 and spill) and at 16384 (a whole 1024-thread workgroup)
  switch (M) {
#define C_(n)                                                                 \
  case n:                                                                     \
    hipLaunchKernelGGL((k_cols_bs<n, CB>), dim3(XB, ntilts),                  \
                       dim3(CB * (n / default_e(n))), 0, s, F, mask, G, Y,    \
                       XB, ndf, tw, by, ylo, yhi, krange);                    \
    break;
    C_(8) C_(16) C_(32) C_(64) C_(128) C_(256) C_(512) C_(1024) C_(2048)
#undef C_
#define C_(n)                                                                 \
  case n:                                                                     \
    hipLaunchKernelGGL((k_cols_bs<n, 1>), dim3(XB * CB, ntilts),              \
                       dim3(n / default_e(n)), 0, s, F, mask, G, Y, XB, ndf,  \
                       tw, by, ylo, yhi, krange);                             \
    break;
    C_(4096) C_(8192) C_(16384)
#undef C_
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_rows_c2r_bs(const float2 *G, int X, int Y, int XB, int ndf,
                              int ntilts, float *out, int M,
                              const BluesteinTable &bx, hipStream_t s) {
  hipError_t err;
  const float2 *tw = twiddles(&err);
  if (err != hipSuccess) return err;
  switch (M) {
#define C_(n)                                                                \
  case n: {                                                                  \
    constexpr int P = bs_row_slots(n);                                       \
    const dim3 grid((Y + 2 * P - 1) / (2 * P), ndf, ntilts);                 \
    hipLaunchKernelGGL(k_rows_c2r_bs<n>, grid, dim3(P * (n / default_e(n))), \
                       0, s, G, X, Y, XB, ndf, out, tw, bx);                 \
    break;                                                                   \
  }
    GX_BS_CASES(C_)
#undef C_
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int STORE>
static hipError_t launch_c2r_filter_bs_t(const float2 *G, int X, int Y,
                                         int XB, int ndf, int t0, int ntilts,
                                         int T, const float *ramp,
                                         const float *wscale,
                                         const uint32_t *tmax, float nrm_mul,
                                         void *slab, int NYB, int YB, int M,
                                         const BluesteinTable &bx,
                                         hipStream_t s, const float2 *tw,
                                         int y0, int ny, const int2 *krange) {
  using ST = typename StoreT<STORE>::T;
  switch (M) {
#define C_(n)                                                                  \
  case n: {                                                                    \
    constexpr int P = bs_row_slots(n);                                         \
    const dim3 grid((ny + 2 * P - 1) / (2 * P), ndf, ntilts);                  \
    hipLaunchKernelGGL((k_rows_c2r_filter_bs<n, STORE>), grid,                 \
                       dim3(P * (n / default_e(n))), 0, s, G, X, Y, XB, ndf,   \
                       t0, T, ramp, wscale, tmax, nrm_mul, (ST *)slab, NYB, YB, \
                       tw, bx, y0, ny, krange);                                \
    break;                                                                     \
  }
    GX_BS_CASES(C_)
#undef C_
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_rows_c2r_filter_bs(const float2 *G, int X, int Y, int XB,
                                     int ndf, int t0, int ntilts, int T,
                                     const float *ramp, const float *wscale,
                                     const uint32_t *tmax, float nrm_mul,
                                     int store, void *slab, int NYB, int YB,
                                     int M, const BluesteinTable &bx,
                                     hipStream_t s, int y0, int ny,
                                     const int2 *krange) {
  hipError_t err;
  const float2 *tw = twiddles(&err);
  if (err != hipSuccess) return err;
  if (ny < 0) ny = Y - y0;
  if (y0 < 0 || y0 % YB != 0 || ny <= 0 || y0 + ny > Y)
    return hipErrorInvalidValue;
#define A_ G, X, Y, XB, ndf, t0, ntilts, T, ramp, wscale, tmax, nrm_mul, slab, NYB, YB, M, bx, s, tw, y0, ny, krange
  switch (store) {
    case kF32:
      return launch_c2r_filter_bs_t<kF32>(A_);
    case kF16:
      return launch_c2r_filter_bs_t<kF16>(A_);
    case kBF16:
      return launch_c2r_filter_bs_t<kBF16>(A_);
  }
#undef A_
  return hipErrorInvalidValue;
}

}  // namespace gx
