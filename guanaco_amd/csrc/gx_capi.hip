// C ABI of libguanaco_hip.so (declared in include/guanaco_hip.h).
//
// Host-side orchestration of the gfx950 kernels: argument checks with the
// reference's error behaviour (a non-zero status plus a message in place of
// guanaco::Error), the per-tilt constants computed in the reference's own
// float/double arithmetic, the ramp-filter table, and device-resident plans
// for whole volumes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "gx_common.hpp"
#include "gx_fft.hpp"
#include "gx_kernels.hpp"
#include "guanaco_hip.h"

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define GX_HIP(call)                                                         \
  do {                                                                       \
    hipError_t e_ = (call);                                                  \
    if (e_ != hipSuccess)                                                    \
      return fail(e_ == hipErrorOutOfMemory ? GX_ENOMEM : GX_EDEVICE,        \
                  "%s:%d %s: %s", __FILE__, __LINE__, #call,                 \
                  hipGetErrorString(e_));                                    \
  } while (0)

#define GX_CHECK(cond)                                                         \
  do {                                                                         \
    if (!(cond))                                                               \
      return fail(GX_EINVAL, "%s:%d ASSERT (%s) failed", __FILE__, __LINE__,   \
                  #cond);                                                      \
  } while (0)

// Device buffer released on scope exit.
struct DevBuf {
  void *p = nullptr;
  bool owned = true;  // false: a slice of a caller's workspace
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() {
    if (p && owned) (void)hipFree(p);
  }
  hipError_t alloc(size_t bytes) { return hipMalloc(&p, bytes ? bytes : 16); }
  void borrow(void *q) {
    p = q;
    owned = false;
  }
  template <typename T>
  T *as() const {
    return (T *)p;
  }
};

bool is_pow2(size_t n) { return n && !(n & (n - 1)); }

size_t next_pow2(size_t n) {
  size_t p = 1;
  while (p < n) p <<= 1;
  return p;
}

int select_device(int gpu_index) {
  if (gpu_index >= 0) GX_HIP(hipSetDevice(gpu_index));
  return GX_OK;
}

}  // namespace

// Host tables shared with the library's other translation units.
namespace gx {

// reconstructor.cpp:29-47 in the reference's float/double arithmetic.
void tilt_geometry(size_t T, const float *angles, float centre, float *coef) {
  const double output_scale = M_PI / (double)(2 * T);
  for (size_t t = 0; t < T; ++t) {
    const float theta = angles[t];
    float c = std::cos(theta);
    float s = std::sin(theta);
    const float det_x0 = -centre * c;
    const float det_y0 = -centre * s;
    const float ray_length = std::sqrt(c * c + s * s);
    const float d = c * c + s * s;
    const float offset = (det_y0 * (-s) - det_x0 * c) / d;
    const float scale =
        (float)(output_scale * (double)ray_length / (double)std::fabs(d));
    c = c / d;
    s = s / d;
    coef[4 * t + 0] = c;
    coef[4 * t + 1] = s;
    coef[4 * t + 2] = offset;
    coef[4 * t + 3] = scale;
  }
}

void defocus_coefficients(size_t ndf, float ps, float mn, float mx,
                          float *dscale, float *doffset) {
  const float ds = ndf > 1 ? (float)ndf * ps / (mx - mn) : 0.0f;
  *dscale = ds;
  *doffset = -ds * (mn / ps);
}

// The first and last defocus slab any voxel of the W x Z grid reads at each
// tilt: the backprojectors' slab choice (gx_bp.hip: slab_of, the voxel
// coordinates of voxel_coord) evaluated here in the same float operations at
// the four corners of the grid; the choice is monotone in x and in z, so the
// corners bound every voxel. coef: the per-tilt table of tilt_geometry.
void slab_ranges(size_t T, const float *coef, int W, int Z, float dscale,
                 float doffset, int ndf, std::vector<int2> &out) {
  auto coord = [](int i, int n) {
    return (float)((double)i - 0.5 * (double)n + 0.5);
  };
  const float xs[2] = {coord(0, W), coord(W - 1, W)};
  const float zs[2] = {coord(0, Z), coord(Z - 1, Z)};
  out.assign(T, make_int2(0, 0));
  for (size_t t = 0; t < T; ++t) {
    const float c = coef[4 * t + 0], s = coef[4 * t + 1];
    int lo = ndf - 1, hi = 0;
    for (int q = 0; q < 4; ++q) {
      const float x = xs[q & 1], z = zs[q >> 1];
      const float h = -x * s - z * c;
      const float d = h * dscale + doffset;
      float n = std::nearbyint(d);
      if (d - n == 0.5f) n += 1.0f;
      const int k = std::min(std::max((int)n, 0), ndf - 1);
      lo = std::min(lo, k);
      hi = std::max(hi, k);
    }
    out[t] = make_int2(lo, hi);
  }
}

// Slabs the LDS-staged backprojector must hold per tilt: a 32 x 8 voxel tile
// spans at most 31.79 units of h (31 |s| + 7 |c|), i.e. 31.79 |dscale| slab
// widths, so at most ceil of that plus one consecutive slabs (31.9 leaves a
// margin for rounding). Rounded up to a compiled form (1, 2 or 4); 0 when
// more are needed and the direct kernel must run.
int bp_slots(size_t ndf, float dscale) {
  if (ndf <= 1) return 1;
  const double span = 31.9 * std::fabs((double)dscale);
  size_t need = ndf;
  if (std::isfinite(span) && span < (double)ndf)
    need = std::min(ndf, (size_t)std::ceil(span) + 1);
  return need <= 1 ? 1 : need <= 2 ? 2 : need <= 4 ? 4 : 0;
}

// Ramp filter spectrum for rows of N samples zero-padded to L, including the
// 1/L of the unnormalised inverse: r[j] for j = 0..L/2.
// L = 2N: the reference's table float(j)/float(N) (filter.cpp:30-45) times
// 1/(2N). L > 2N (N not a power of two): the reference's 2N-point filter as a
// space kernel h (|d| <= N-1, all the zero-padded product can reach), then
// its L-point DFT. Both give the reference's linear convolution.
std::vector<float> ramp_table(size_t N, size_t L) {
  std::vector<float> r(L / 2 + 1, 0.0f);
  if (L == 2 * N) {
    const float inv_l = 1.0f / (float)L;
    for (size_t j = 0; j <= N; ++j) r[j] = ((float)j / (float)N) * inv_l;
    return r;
  }
  const size_t M = 2 * N;
  std::vector<double> cosM(M), cosL(L), h(N);
  for (size_t i = 0; i < M; ++i) cosM[i] = std::cos(2.0 * M_PI * i / M);
  for (size_t i = 0; i < L; ++i) cosL[i] = std::cos(2.0 * M_PI * i / L);
  for (size_t d = 0; d < N; ++d) {
    double acc = 0.0;
    for (size_t k = 0; k < M; ++k) {
      const size_t kk = k <= N ? k : M - k;
      const double R = (double)((float)kk / (float)N);
      acc += R * cosM[(k * d) % M];
    }
    h[d] = acc / (double)M;
  }
  for (size_t j = 0; j <= L / 2; ++j) {
    double acc = h[0];
    for (size_t d = 1; d < N; ++d) acc += 2.0 * h[d] * cosL[(j * d) % L];
    r[j] = (float)(acc / (double)L);
  }
  return r;
}

size_t filter_length(size_t N) {
  size_t L = next_pow2(2 * N - 1);
  return L < 8 ? 8 : L;
}

}  // namespace gx

namespace {

using gx::defocus_coefficients;
using gx::filter_length;
using gx::ramp_table;
using gx::tilt_geometry;

// The power-of-two correction kernels take 8 <= X <= 8192 and 4 <= Y <=
// 8192; other sizes use Bluestein transforms of length filter_length(n).
bool rows_pow2(size_t X) { return is_pow2(X) && X >= 8 && X <= 8192; }
bool cols_pow2(size_t Y) { return is_pow2(Y) && Y >= 4 && Y <= 8192; }

// In-place radix-2 FFT (sign -1) of n = 2^k complex doubles.
void fft_f64(std::vector<std::complex<double>> &a) {
  const size_t n = a.size();
  for (size_t i = 1, j = 0; i < n; ++i) {
    size_t bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) std::swap(a[i], a[j]);
  }
  for (size_t len = 2; len <= n; len <<= 1) {
    for (size_t i = 0; i < n; i += len) {
      for (size_t k = 0; k < len / 2; ++k) {
        const double ang = -2.0 * M_PI * (double)k / (double)len;
        const std::complex<double> w(std::cos(ang), std::sin(ang));
        const std::complex<double> u = a[i + k], v = a[i + k + len / 2] * w;
        a[i + k] = u + v;
        a[i + k + len / 2] = u - v;
      }
    }
  }
}

// Bluestein table entries of an X x Y correction (0 for power-of-two sizes).
size_t bs_entries(size_t X, size_t Y) {
  size_t n = 0;
  if (!rows_pow2(X)) n += X + 2 * filter_length(X);
  if (!cols_pow2(Y)) n += Y + 2 * filter_length(Y);
  return n;
}

// Tables of an n-point Bluestein transform on M-point FFTs (M >= 2n - 1),
// appended to out: chirp w_m = exp(-i pi m^2 / n) (n values, m^2 mod 2n in
// integers), then FFT_M(b) / M and FFT_M(conj(b)) / M (M values each) with
// b_m = conj(w_|m|) for |m| < n placed circularly.
void bluestein_tables(size_t n, size_t M, std::vector<float2> &out) {
  std::vector<std::complex<double>> w(n), b(M), bc(M);
  for (size_t m = 0; m < n; ++m) {
    const unsigned long long r =
        ((unsigned long long)m * m) % (2ull * (unsigned long long)n);
    const double a = -M_PI * (double)r / (double)n;
    w[m] = std::complex<double>(std::cos(a), std::sin(a));
  }
  for (size_t m = 0; m < n; ++m) {
    b[m] = std::conj(w[m]);
    if (m) b[M - m] = std::conj(w[m]);
  }
  for (size_t m = 0; m < M; ++m) bc[m] = std::conj(b[m]);
  fft_f64(b);
  fft_f64(bc);
  for (size_t m = 0; m < n; ++m)
    out.push_back(make_float2((float)w[m].real(), (float)w[m].imag()));
  for (const auto *v : {&b, &bc})
    for (size_t m = 0; m < M; ++m)
      out.push_back(make_float2((float)((*v)[m].real() / (double)M),
                                (float)((*v)[m].imag() / (double)M)));
}

// BluesteinTable views of tables laid out by bluestein_tables at d.
gx::BluesteinTable bluestein_view(const float2 *d, size_t n, size_t M) {
  return gx::BluesteinTable{d, d + n, d + n + M, (int)n};
}

}  // namespace

namespace gx {

std::vector<float2> twiddle_host_table() {
  // W^j, then the per-thread blocks of every transform shape (gx_fft.hpp)
  std::vector<float2> h(tw_table_len());
  for (int j = 0; j < kTwN; ++j) {
    const double a = 2.0 * M_PI * (double)j / (double)kTwN;
    h[j] = make_float2((float)std::cos(a), (float)-std::sin(a));
  }
  for (int ln = 1; ln <= 14; ++ln)
    for (int e = 2; e <= 16; e *= 2)
      if (tw_count(1 << ln, e) > 0)
        tw_fill_block(1 << ln, e, h.data(),
                      h.data() + kTwN + tw_block_offset(1 << ln, e));
  return h;
}

// One table per GPU in ordinary (cached) device memory, allocated on the
// first use on that device and held for the life of the process. (A device
// global array measured 12-34 % slower filter passes: the FFT kernels read
// their twiddles on every stage.)
const float2 *twiddles(hipError_t *err) {
  static std::mutex mu;
  static float2 *tables[64] = {};
  int dev = 0;
  *err = hipGetDevice(&dev);
  if (*err != hipSuccess) return nullptr;
  if (dev < 0 || dev >= 64) {
    *err = hipErrorInvalidDevice;
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(mu);
  if (tables[dev] == nullptr) {
    const std::vector<float2> h = twiddle_host_table();
    float2 *d = nullptr;
    *err = hipMalloc(&d, sizeof(float2) * h.size());
    if (*err != hipSuccess) return nullptr;
    *err = hipMemcpy(d, h.data(), sizeof(float2) * h.size(),
                     hipMemcpyHostToDevice);
    if (*err != hipSuccess) {
      (void)hipFree(d);
      return nullptr;
    }
    tables[dev] = d;
  }
  *err = hipSuccess;
  return tables[dev];
}

}  // namespace gx

using namespace gx;

// ------------------------------------------------------------ status ------

extern "C" const char *gx_last_error(void) { return g_err.c_str(); }

extern "C" int gx_version(void) { return 1; }

extern "C" int gx_device_count(int *count) {
  GX_CHECK(count != nullptr);
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) n = 0;
  *count = n;
  return GX_OK;
}

extern "C" int gx_set_device(int device) {
  GX_HIP(hipSetDevice(device));
  return GX_OK;
}

extern "C" int gx_get_device(int *device) {
  GX_CHECK(device != nullptr);
  GX_HIP(hipGetDevice(device));
  return GX_OK;
}

extern "C" int gx_device_info(int device, char *name, size_t name_len,
                              size_t *total_mem, int *num_cus) {
  hipDeviceProp_t p;
  GX_HIP(hipGetDeviceProperties(&p, device));
  if (name && name_len) {
    std::snprintf(name, name_len, "%s (%s)", p.name, p.gcnArchName);
  }
  if (total_mem) *total_mem = p.totalGlobalMem;
  if (num_cus) *num_cus = p.multiProcessorCount;
  return GX_OK;
}

extern "C" int gx_mem_get_info(size_t *free_bytes, size_t *total_bytes) {
  GX_CHECK(free_bytes && total_bytes);
  GX_HIP(hipMemGetInfo(free_bytes, total_bytes));
  return GX_OK;
}

// ------------------------------------------------- memory and streams -----

extern "C" int gx_malloc(void **ptr, size_t bytes) {
  GX_CHECK(ptr != nullptr);
  GX_HIP(hipMalloc(ptr, bytes ? bytes : 16));
  return GX_OK;
}

extern "C" int gx_free(void *ptr) {
  if (ptr) GX_HIP(hipFree(ptr));
  return GX_OK;
}

extern "C" int gx_memcpy_htod(void *dst, const void *src, size_t bytes,
                              void *stream) {
  GX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice,
                        (hipStream_t)stream));
  if (stream == nullptr) GX_HIP(hipStreamSynchronize(nullptr));
  return GX_OK;
}

extern "C" int gx_memcpy_dtoh(void *dst, const void *src, size_t bytes,
                              void *stream) {
  GX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost,
                        (hipStream_t)stream));
  if (stream == nullptr) GX_HIP(hipStreamSynchronize(nullptr));
  return GX_OK;
}

extern "C" int gx_memcpy_dtod(void *dst, const void *src, size_t bytes,
                              void *stream) {
  GX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice,
                        (hipStream_t)stream));
  return GX_OK;
}

extern "C" int gx_fill_u32(void *dst, uint32_t value, size_t count,
                           void *stream) {
  GX_HIP(launch_fill_u32((uint32_t *)dst, value, count, (hipStream_t)stream));
  return GX_OK;
}

extern "C" int gx_copy_u32(void *dst, const void *src, size_t count,
                           void *stream) {
  GX_HIP(launch_copy_u32((uint32_t *)dst, (const uint32_t *)src, count,
                         (hipStream_t)stream));
  return GX_OK;
}

extern "C" int gx_stream_create(void **stream) {
  GX_CHECK(stream != nullptr);
  hipStream_t s;
  GX_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = s;
  return GX_OK;
}

extern "C" int gx_stream_destroy(void *stream) {
  if (stream) GX_HIP(hipStreamDestroy((hipStream_t)stream));
  return GX_OK;
}

extern "C" int gx_stream_synchronize(void *stream) {
  GX_HIP(hipStreamSynchronize((hipStream_t)stream));
  return GX_OK;
}

extern "C" int gx_device_synchronize(void) {
  GX_HIP(hipDeviceSynchronize());
  return GX_OK;
}

extern "C" int gx_event_create(void **event) {
  GX_CHECK(event != nullptr);
  hipEvent_t e;
  GX_HIP(hipEventCreate(&e));
  *event = e;
  return GX_OK;
}

extern "C" int gx_event_destroy(void *event) {
  if (event) GX_HIP(hipEventDestroy((hipEvent_t)event));
  return GX_OK;
}

extern "C" int gx_event_record(void *event, void *stream) {
  GX_HIP(hipEventRecord((hipEvent_t)event, (hipStream_t)stream));
  return GX_OK;
}

extern "C" int gx_event_synchronize(void *event) {
  GX_HIP(hipEventSynchronize((hipEvent_t)event));
  return GX_OK;
}

extern "C" int gx_event_elapsed_ms(void *start, void *stop, float *ms) {
  GX_CHECK(ms != nullptr);
  GX_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
  return GX_OK;
}

extern "C" int gx_graph_capture_begin(void *stream) {
  GX_CHECK(stream != nullptr);
  GX_HIP(hipStreamBeginCapture((hipStream_t)stream,
                               hipStreamCaptureModeThreadLocal));
  return GX_OK;
}

extern "C" int gx_graph_capture_end(void *stream, void **graph_exec) {
  GX_CHECK(stream != nullptr && graph_exec != nullptr);
  hipGraph_t g = nullptr;
  GX_HIP(hipStreamEndCapture((hipStream_t)stream, &g));
  hipGraphExec_t ge = nullptr;
  hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  GX_HIP(e);
  *graph_exec = ge;
  return GX_OK;
}

extern "C" int gx_graph_launch(void *graph_exec, void *stream) {
  GX_HIP(hipGraphLaunch((hipGraphExec_t)graph_exec, (hipStream_t)stream));
  return GX_OK;
}

extern "C" int gx_graph_destroy(void *graph_exec) {
  if (graph_exec) GX_HIP(hipGraphExecDestroy((hipGraphExec_t)graph_exec));
  return GX_OK;
}

// ------------------------------------ reference-equivalent entry points ---

namespace gx {
// config.h:66-75, and the filter's size limit
int check_recon_cfg(const gx_recon_cfg *c);
}  // namespace gx

int gx::check_recon_cfg(const gx_recon_cfg *c) {
  GX_CHECK(c != nullptr);
  // config.h:66-75
  GX_CHECK(c->num_pixels > 0);
  GX_CHECK(c->num_angles > 0);
  GX_CHECK(c->num_defocus > 0);
  GX_CHECK(c->grid_height > 0);
  GX_CHECK(c->grid_width > 0);
  GX_CHECK(c->pixel_size > 0);
  GX_CHECK(c->min_defocus <= c->max_defocus);
  GX_CHECK(c->angles != nullptr);
  if (filter_length(c->num_pixels) > 16384)
    return fail(GX_EUNSUPPORTED,
                "num_pixels %zu exceeds the gfx950 filter limit (8192)",
                c->num_pixels);
  return GX_OK;
}

extern "C" int gx_tilt_geometry(const gx_recon_cfg *cfg, float *coef) {
  GX_CHECK(cfg != nullptr && cfg->angles != nullptr && coef != nullptr);
  GX_CHECK(cfg->num_angles > 0);
  tilt_geometry(cfg->num_angles, cfg->angles, cfg->centre, coef);
  return GX_OK;
}

// Shared by gx_recon and gx_project: slab (ndf x T x N, filtered, device)
// -> rec (H x W, device).
static int bp_slice(const gx_recon_cfg *c, const float *d_slab, float *d_rec,
                    hipStream_t s) {
  const size_t T = c->num_angles;
  std::vector<float> coef(4 * T);
  tilt_geometry(T, c->angles, c->centre, coef.data());
  DevBuf d_coef;
  GX_HIP(d_coef.alloc(sizeof(float) * 4 * T));
  GX_HIP(hipMemcpyAsync(d_coef.p, coef.data(), sizeof(float) * 4 * T,
                        hipMemcpyHostToDevice, s));
  BpArgs a{};
  a.slab = d_slab;
  a.coef = d_coef.as<float4>();
  a.win = nullptr;
  defocus_coefficients(c->num_defocus, c->pixel_size, c->min_defocus,
                       c->max_defocus, &a.dscale, &a.doffset);
  a.ndf = (int)c->num_defocus;
  a.N = (int)c->num_pixels;
  a.T = (int)T;
  a.NYB = 1;
  a.W = (int)c->grid_width;
  a.Z = (int)c->grid_height;
  a.Y = 1;
  a.yb0 = 0;
  a.yb1 = 1;
  a.y_base = 0;
  a.vol = d_rec;
  GX_HIP(launch_bp(a, 1, kF32, s));
  GX_HIP(hipStreamSynchronize(s));
  return GX_OK;
}

extern "C" int gx_project(const gx_recon_cfg *cfg, const float *filtered,
                          float *rec) {
  int rc = check_recon_cfg(cfg);
  if (rc) return rc;
  GX_CHECK(filtered != nullptr && rec != nullptr);
  rc = select_device(cfg->gpu_index);
  if (rc) return rc;
  const size_t n_sino = cfg->num_defocus * cfg->num_angles * cfg->num_pixels;
  const size_t n_rec = cfg->grid_height * cfg->grid_width;
  DevBuf d_sino, d_rec;
  GX_HIP(d_sino.alloc(sizeof(float) * n_sino));
  GX_HIP(d_rec.alloc(sizeof(float) * n_rec));
  GX_HIP(hipMemcpy(d_sino.p, filtered, sizeof(float) * n_sino,
                   hipMemcpyHostToDevice));
  rc = bp_slice(cfg, d_sino.as<float>(), d_rec.as<float>(), nullptr);
  if (rc) return rc;
  GX_HIP(hipMemcpy(rec, d_rec.p, sizeof(float) * n_rec, hipMemcpyDeviceToHost));
  return GX_OK;
}

// Rows (nrows x N, device) -> filtered rows (nrows x N, device), F32.
static int filter_rows_dev(const float *d_in, float *d_out, size_t nrows,
                           size_t N, const float *d_wscale_ones,
                           hipStream_t s) {
  const size_t L = filter_length(N);
  std::vector<float> r = ramp_table(N, L);
  DevBuf d_ramp;
  GX_HIP(d_ramp.alloc(sizeof(float) * r.size()));
  GX_HIP(hipMemcpyAsync(d_ramp.p, r.data(), sizeof(float) * r.size(),
                        hipMemcpyHostToDevice, s));
  // rows as (ndf = 1, T = nrows, Y = 1): slab layout [t][x] = input layout
  GX_HIP(launch_filter_rows(d_in, N, 0, 0, (int)N, (int)L, 1, (int)nrows, 1,
                            d_ramp.as<float>(), d_wscale_ones, nullptr, 1.0f,
                            kF32, d_out, 1, 1, s));
  GX_HIP(hipStreamSynchronize(s));
  return GX_OK;
}

extern "C" int gx_filter(const float *rows, float *out, size_t nrows,
                         size_t num_pixels, int gpu_index) {
  GX_CHECK(rows != nullptr && out != nullptr);
  GX_CHECK(num_pixels > 0 && nrows > 0);
  if (filter_length(num_pixels) > 16384)
    return fail(GX_EUNSUPPORTED,
                "num_pixels %zu exceeds the gfx950 filter limit (8192)",
                num_pixels);
  int rc = select_device(gpu_index);
  if (rc) return rc;
  const size_t n = nrows * num_pixels;
  DevBuf d_in, d_out, d_w;
  GX_HIP(d_in.alloc(sizeof(float) * n));
  GX_HIP(d_out.alloc(sizeof(float) * n));
  GX_HIP(d_w.alloc(sizeof(float) * nrows));
  std::vector<float> ones(nrows, 1.0f);
  GX_HIP(hipMemcpy(d_in.p, rows, sizeof(float) * n, hipMemcpyHostToDevice));
  GX_HIP(hipMemcpy(d_w.p, ones.data(), sizeof(float) * nrows,
                   hipMemcpyHostToDevice));
  rc = filter_rows_dev(d_in.as<float>(), d_out.as<float>(), nrows, num_pixels,
                       d_w.as<float>(), nullptr);
  if (rc) return rc;
  GX_HIP(hipMemcpy(out, d_out.p, sizeof(float) * n, hipMemcpyDeviceToHost));
  return GX_OK;
}

extern "C" int gx_recon(const gx_recon_cfg *cfg, const float *sino,
                        float *rec) {
  int rc = check_recon_cfg(cfg);
  if (rc) return rc;
  GX_CHECK(sino != nullptr && rec != nullptr);
  rc = select_device(cfg->gpu_index);
  if (rc) return rc;
  const size_t T = cfg->num_angles, N = cfg->num_pixels;
  const size_t nrows = cfg->num_defocus * T;
  const size_t n_sino = nrows * N;
  const size_t n_rec = cfg->grid_height * cfg->grid_width;
  DevBuf d_sino, d_filt, d_rec, d_w;
  GX_HIP(d_sino.alloc(sizeof(float) * n_sino));
  GX_HIP(d_filt.alloc(sizeof(float) * n_sino));
  GX_HIP(d_rec.alloc(sizeof(float) * n_rec));
  GX_HIP(d_w.alloc(sizeof(float) * nrows));
  std::vector<float> ones(nrows, 1.0f);
  // reconstructor.cpp:105: the input is copied, never modified
  GX_HIP(hipMemcpy(d_sino.p, sino, sizeof(float) * n_sino,
                   hipMemcpyHostToDevice));
  GX_HIP(hipMemcpy(d_w.p, ones.data(), sizeof(float) * nrows,
                   hipMemcpyHostToDevice));
  rc = filter_rows_dev(d_sino.as<float>(), d_filt.as<float>(), nrows, N,
                       d_w.as<float>(), nullptr);
  if (rc) return rc;
  rc = bp_slice(cfg, d_filt.as<float>(), d_rec.as<float>(), nullptr);
  if (rc) return rc;
  GX_HIP(hipMemcpy(rec, d_rec.p, sizeof(float) * n_rec, hipMemcpyDeviceToHost));
  return GX_OK;
}

// Any size from 1 to 8192 in each dimension (Bluestein transforms of up to
// 16384 points for the sizes the power-of-two kernels do not take).
static int check_corr_sizes(size_t X, size_t Y) {
  if (X < 1 || X > 8192 || Y < 1 || Y > 8192)
    return fail(GX_EUNSUPPORTED,
                "CTF correction on gfx950 takes 1 <= xsize, ysize <= 8192 "
                "(got %zu x %zu)", X, Y);
  return GX_OK;
}

// The synthetic data generator uses the power-of-two correction kernels only.
static int check_pow2_corr_sizes(size_t X, size_t Y) {
  if (!rows_pow2(X) || !cols_pow2(Y))
    return fail(GX_EUNSUPPORTED,
                "synthetic CTF data needs power-of-two sizes, 8 <= xsize <= "
                "8192 and 4 <= ysize <= 8192 (got %zu x %zu)", X, Y);
  return GX_OK;
}

extern "C" int gx_corr(const float *image, const float *ctf, float *rec,
                       size_t xsize, size_t ysize, size_t num_ctf,
                       int gpu_index) {
  // ext/correct.h:40-57 and correct.cpp:37-39
  GX_CHECK(image != nullptr && ctf != nullptr && rec != nullptr);
  GX_CHECK(xsize > 0);
  GX_CHECK(ysize > 0);
  GX_CHECK(num_ctf > 0);
  int rc = check_corr_sizes(xsize, ysize);
  if (rc) return rc;
  rc = select_device(gpu_index);
  if (rc) return rc;
  const int X = (int)xsize, Y = (int)ysize, ndf = (int)num_ctf;
  const int XB = X / 2 / CB + 1;
  const size_t npix = xsize * ysize;
  const size_t nspec = (size_t)XB * CB * Y;
  DevBuf d_img, d_ctf, d_mask, d_F, d_G, d_out;
  GX_HIP(d_img.alloc(sizeof(float) * npix));
  GX_HIP(d_ctf.alloc(sizeof(float2) * npix * num_ctf));
  GX_HIP(d_mask.alloc(nspec * num_ctf));
  GX_HIP(d_F.alloc(sizeof(float2) * nspec));
  GX_HIP(d_G.alloc(sizeof(float2) * nspec * num_ctf));
  GX_HIP(d_out.alloc(sizeof(float) * npix * num_ctf));
  GX_HIP(hipMemcpy(d_img.p, image, sizeof(float) * npix, hipMemcpyHostToDevice));
  GX_HIP(hipMemcpy(d_ctf.p, ctf, sizeof(float2) * npix * num_ctf,
                   hipMemcpyHostToDevice));
  hipStream_t s = nullptr;
  const bool rp = rows_pow2(xsize), cp = cols_pow2(ysize);
  const size_t MX = filter_length(xsize), MY = filter_length(ysize);
  std::vector<float2> tab;
  if (!rp) bluestein_tables(xsize, MX, tab);
  const size_t ytab = tab.size();
  if (!cp) bluestein_tables(ysize, MY, tab);
  DevBuf d_tab;
  if (!tab.empty()) {
    GX_HIP(d_tab.alloc(sizeof(float2) * tab.size()));
    GX_HIP(hipMemcpy(d_tab.p, tab.data(), sizeof(float2) * tab.size(),
                     hipMemcpyHostToDevice));
  }
  const gx::BluesteinTable bx = bluestein_view(d_tab.as<float2>(), xsize, MX);
  const gx::BluesteinTable by =
      bluestein_view(d_tab.as<float2>() + ytab, ysize, MY);
  GX_HIP(launch_mask_from_ctf(d_mask.as<int8_t>(), d_ctf.as<float2>(), X, Y, XB,
                              ndf, cp, s));
  if (rp) {
    GX_HIP(launch_rows_r2c(d_img.as<float>(), npix, X, Y, XB, 1,
                           d_F.as<float2>(), nullptr, 0, s));
  } else {
    GX_HIP(launch_rows_r2c_bs(d_img.as<float>(), npix, X, Y, XB, 1,
                              d_F.as<float2>(), nullptr, 0, (int)MX, bx, s));
  }
  if (cp) {
    GX_HIP(launch_cols(d_F.as<float2>(), d_mask.as<int8_t>(),
                       d_G.as<float2>(), Y, XB, ndf, 1, s));
  } else {
    GX_HIP(launch_cols_bs(d_F.as<float2>(), d_mask.as<int8_t>(),
                          d_G.as<float2>(), Y, XB, ndf, 1, (int)MY, by, s));
  }
  if (rp) {
    GX_HIP(launch_rows_c2r(d_G.as<float2>(), X, Y, XB, ndf, 1,
                           d_out.as<float>(), s));
  } else {
    GX_HIP(launch_rows_c2r_bs(d_G.as<float2>(), X, Y, XB, ndf, 1,
                              d_out.as<float>(), (int)MX, bx, s));
  }
  GX_HIP(hipStreamSynchronize(s));
  GX_HIP(hipMemcpy(rec, d_out.p, sizeof(float) * npix * num_ctf,
                   hipMemcpyDeviceToHost));
  return GX_OK;
}

// ------------------------------------------------- whole-volume plans ----

struct gx_volume_plan {
  int device = 0;
  int X = 0, Y = 0, T = 0, ndf = 0, Z = 0, W = 0;
  int storage = 0, layout = 0;
  // NYB: row blocks of the slab buffer, which holds `band` detector rows
  // (Y unless the plan is banded): rows [slab_y0, slab_y1) after the last
  // filter pass
  int YB = 0, NYB = 0, XB = 0, L = 0, chunk = 0;
  int band = 0, slab_y0 = 0, slab_y1 = 0;
  int bp_kind = GX_BP_STAGED;
  int bp_staged = 1;  // slabs per tilt of the staged kernel; 0 = direct
  bool ctf = false;
  float dscale = 0, doffset = 0, nrm_mul = 1;
  size_t slab_bytes = 0, total_bytes = 0;
  // coef: the reference's per-tilt constants; coef_r: the reduced forms'
  // table (bp_coefficients), rebuilt by every backprojection; win: the staged
  // kernel's per-(tile, tilt) windows. Tilt tables hold bp_tilt_pad(T) rows.
  DevBuf coef, coef_r, win, wscale, tmax, ramp, mask, defocus, F, G, slab;
  // krange[t]: the first and last defocus slab any voxel of the volume reads
  // at tilt t (slab plans); the filter passes compute only those. pairs:
  // every (tilt, defocus) they compute, tilt-major, pair_off[t] the first of
  // tilt t (pair_off[T] = their number).
  DevBuf krange, pairs;
  std::vector<int2> h_krange, h_pairs;
  std::vector<int> pair_off;
  // per-row rotation centres: row_off[t * Y + y], the tilt's offset for row
  // y (reduced storage: offset - 1/2); empty when one centre serves all rows
  DevBuf row_off;
  bool per_row = false;
  std::vector<float> h_row_off;
  // Bluestein tables for sizes the power-of-two correction kernels do not
  // take: rows (X points on MX) then columns (Y points on MY)
  bool rows_bs = false, cols_bs = false;
  int MX = 0, MY = 0;
  size_t bs_y = 0;  // offset of the column tables
  DevBuf bs;
  std::vector<float2> h_bs;
  // host copies of the uploaded tables, alive until the plan is destroyed so
  // that stream-ordered uploads may read them
  std::vector<float> h_coef, h_wscale, h_ramp;
  std::vector<double> h_defocus;
  double wavelength = 0, Cs = 0, Ca = 0, Pa = 0, phi = 0, pixel_size = 0;
};

// ------------------------------------------------------------ autotuning --
//
// Tile choices per (device model, plan shape), measured by
// gx_volume_plan_autotune. The candidates never exceed the row blocking a
// plan's slab buffer is sized for (16 rows for float32 slabs, 32 for 16-bit
// ones), so a choice applies to a plan without reallocation.

namespace {

struct TuneChoice {
  int kind, rows;
};

std::mutex g_tune_mu;
std::map<std::string, TuneChoice> g_tune;

// "gfx950:256": architecture and compute units of the current device
std::string device_model() {
  static std::mutex mu;
  static std::string names[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return "?";
  std::lock_guard<std::mutex> lock(mu);
  if (names[dev].empty()) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return "?";
    std::string arch = prop.gcnArchName;
    arch = arch.substr(0, arch.find(':'));
    names[dev] = arch + ":" + std::to_string(prop.multiProcessorCount);
  }
  return names[dev];
}

std::string tune_key(const gx_volume_plan *p) {
  char b[256];
  std::snprintf(b, sizeof b, "%s %d %d %d %d %d %d %d %d %d %d",
                device_model().c_str(), p->X, p->Y, p->T, p->ndf, p->Z, p->W,
                p->storage, p->layout, p->ctf ? 1 : 0, p->bp_staged);
  return b;
}

// the largest candidate comes first: it is the default and sizes the slabs
void tile_candidates(int storage, int rows[2]) {
  rows[0] = storage == GX_STORE_F32 ? 16 : 32;
  rows[1] = rows[0] / 2;
}

bool tile_valid(const gx_volume_plan *p, int rows, int kind) {
  int c[2];
  tile_candidates(p->storage, c);
  if (rows != c[0] && rows != c[1]) return false;
  if (kind == GX_BP_STAGED) return p->bp_staged != 0;
  return kind == GX_BP_DIRECT;
}

void set_tile(gx_volume_plan *p, int rows, int kind) {
  p->YB = rows;
  p->NYB = (p->band + rows - 1) / rows;
  p->bp_kind = kind;
}

void apply_tuned(gx_volume_plan *p) {
  const std::string key = tune_key(p);
  std::lock_guard<std::mutex> lock(g_tune_mu);
  auto it = g_tune.find(key);
  if (it != g_tune.end() && tile_valid(p, it->second.rows, it->second.kind))
    set_tile(p, it->second.rows, it->second.kind);
}

struct EventPair {
  hipEvent_t a = nullptr, b = nullptr;
  ~EventPair() {
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
  }
};

}  // namespace

namespace {

constexpr size_t kPlanAlign = 256;
constexpr int kPlanBuffers = 15;

struct PlanSizes {
  int YB, NYB, XB, L, chunk, band;
  bool ctf;
  size_t slab_bytes, mask_bytes, f_bytes, g_bytes, small_bytes, total;
};

int plan_sizes(const gx_volume_desc *d, PlanSizes *ps) {
  GX_CHECK(d != nullptr);
  GX_CHECK(d->num_pixels > 0 && d->num_rows > 0 && d->num_angles > 0);
  GX_CHECK(d->num_defocus > 0 && d->depth > 0 && d->width > 0);
  GX_CHECK(d->pixel_size > 0);
  GX_CHECK(d->min_defocus <= d->max_defocus);
  GX_CHECK(d->angles != nullptr);
  for (size_t t = 0; t < d->num_angles; ++t)
    if (!std::isfinite(d->angles[t]))
      return fail(GX_EINVAL, "angle %zu is not finite", t);
  if (d->centres)
    for (size_t y = 0; y < d->num_rows; ++y)
      if (!std::isfinite(d->centres[y]))
        return fail(GX_EINVAL, "centre of row %zu is not finite", y);
  GX_CHECK(d->storage >= GX_STORE_F32 && d->storage <= GX_STORE_BF16);
  GX_CHECK(d->input_layout >= GX_INPUT_PROJECTIONS &&
           d->input_layout <= GX_INPUT_SINOGRAMS);
  const size_t X = d->num_pixels, Y = d->num_rows, T = d->num_angles;
  const size_t ndf = d->num_defocus;
  ps->ctf = d->input_layout == GX_INPUT_PROJECTIONS && d->defocus != nullptr;
  if (d->input_layout == GX_INPUT_PROJECTIONS && !ps->ctf && ndf != 1)
    return fail(GX_EINVAL,
                "num_defocus must be 1 for uncorrected projections without "
                "CTF parameters");
  if (ps->ctf) {
    int rc = check_corr_sizes(X, Y);
    if (rc) return rc;
  } else if (filter_length(X) > 16384) {
    return fail(GX_EUNSUPPORTED,
                "num_pixels %zu exceeds the gfx950 filter limit (8192)", X);
  }
  ps->YB = d->storage == GX_STORE_F32 ? 16 : 32;
  // A banded plan holds the filtered slabs of band_rows detector rows at a
  // time (a multiple of 32, so of every tile's row count); 0: every row.
  size_t band = Y;
  if (d->band_rows > 0 && d->band_rows < Y) {
    band = (d->band_rows + 31) / 32 * 32;
    if (band > Y) band = Y;
  }
  ps->band = (int)band;
  ps->NYB = (int)((band + ps->YB - 1) / ps->YB);
  ps->XB = (int)(X / 2 / CB + 1);
  ps->L = (int)(ps->ctf && rows_pow2(X) ? 2 * X : filter_length(X));
  const size_t es = d->storage == GX_STORE_F32 ? 4 : 2;
  ps->slab_bytes = ndf * T * (size_t)ps->NYB * ps->YB * X * es;
  const size_t nspec = (size_t)ps->XB * CB * Y;
  ps->mask_bytes = ps->ctf ? nspec * ndf : 0;
  size_t chunk = 0;
  if (ps->ctf) {
    const size_t per_tilt = nspec * sizeof(float2) * (1 + ndf);
    chunk = d->tilt_chunk;
    if (chunk == 0) {
      // the tilts whose spectra fit in 4 GB; for long columns (two
      // workgroups per CU, 512 at once) raised, within 16 GB, until the
      // column pass's grid is 16 such rounds, so that its last, partly
      // idle round costs little
      const size_t budget = (size_t)4 << 30, cap = (size_t)16 << 30;
      chunk = std::max<size_t>(1, budget / per_tilt);
      if (Y >= 2048 && cols_pow2(Y)) {
        const size_t wg = (Y >= 4096 ? 2 : 1) * (size_t)ps->XB;
        const size_t want = (16 * 512 + wg - 1) / wg;
        chunk = std::max(chunk, std::min(want, cap / per_tilt));
      }
    }
    if (chunk > T) chunk = T;
    ps->f_bytes = nspec * sizeof(float2) * chunk;
    ps->g_bytes = nspec * sizeof(float2) * ndf * chunk;
  } else {
    ps->f_bytes = ps->g_bytes = 0;
  }
  ps->chunk = (int)chunk;
  ps->small_bytes = (size_t)bp_tilt_pad((int)T) * 2 * sizeof(float4) +
                    T * 2 * sizeof(float) +
                    bp_window_entries((int)d->width, (int)d->depth, (int)T) *
                        sizeof(int2) +
                    sizeof(float) * (ps->L / 2 + 1) + sizeof(double) * ndf +
                    (d->centres ? sizeof(float) * T * Y : 0) +
                    (ps->ctf ? sizeof(float2) * bs_entries(X, Y) : 0) +
                    (ndf > 1 ? sizeof(int2) * T * (1 + ndf) : 0);
  // every buffer starts on a kPlanAlign boundary inside a workspace
  ps->total = ps->slab_bytes + ps->mask_bytes + ps->f_bytes + ps->g_bytes +
              ps->small_bytes + kPlanBuffers * kPlanAlign;
  return GX_OK;
}

}  // namespace

extern "C" int gx_volume_plan_bytes(const gx_volume_desc *desc,
                                    size_t *bytes) {
  GX_CHECK(bytes != nullptr);
  PlanSizes ps;
  int rc = plan_sizes(desc, &ps);
  if (rc) return rc;
  *bytes = ps.total;
  return GX_OK;
}

namespace {

// Host fields and host tables of a plan.
void plan_setup(gx_volume_plan *p, const gx_volume_desc *d,
                const PlanSizes &ps) {
  p->X = (int)d->num_pixels;
  p->Y = (int)d->num_rows;
  p->T = (int)d->num_angles;
  p->ndf = (int)d->num_defocus;
  p->Z = (int)d->depth;
  p->W = (int)d->width;
  p->storage = d->storage;
  p->layout = d->input_layout;
  p->YB = ps.YB;
  p->NYB = ps.NYB;
  p->band = ps.band;
  p->slab_y0 = 0;
  p->slab_y1 = ps.band == p->Y ? p->Y : 0;
  p->XB = ps.XB;
  p->L = ps.L;
  p->chunk = ps.chunk;
  p->ctf = ps.ctf;
  p->slab_bytes = ps.slab_bytes;
  p->total_bytes = ps.total;
  p->nrm_mul = ps.ctf ? (float)d->num_pixels * (float)d->num_rows : 1.0f;
  defocus_coefficients(d->num_defocus, d->pixel_size, d->min_defocus,
                       d->max_defocus, &p->dscale, &p->doffset);
  p->per_row = d->centres != nullptr;
  p->bp_staged = p->per_row ? 0 : bp_slots(d->num_defocus, p->dscale);
  p->bp_kind = p->bp_staged ? GX_BP_STAGED : GX_BP_DIRECT;
  apply_tuned(p);
  const size_t T = d->num_angles;
  if (p->per_row) {
    // each row's offsets with the reference's arithmetic (tilt_geometry)
    const size_t Y = d->num_rows;
    std::vector<float> cf(4 * T);
    p->h_row_off.assign(T * Y, 0.0f);
    for (size_t y = 0; y < Y; ++y) {
      tilt_geometry(T, d->angles, d->centres[y], cf.data());
      for (size_t t = 0; t < T; ++t)
        p->h_row_off[t * Y + y] = d->storage == GX_STORE_F32
                                      ? cf[4 * t + 2]
                                      : cf[4 * t + 2] - 0.5f;
    }
  }
  p->h_coef.assign(4 * (size_t)bp_tilt_pad((int)T), 0.0f);
  tilt_geometry(T, d->angles, d->centre, p->h_coef.data());
  if (p->ndf > 1) {
    slab_ranges(T, p->h_coef.data(), p->W, p->Z, p->dscale, p->doffset,
                p->ndf, p->h_krange);
    p->h_pairs.clear();
    p->pair_off.assign(T + 1, 0);
    for (size_t t = 0; t < T; ++t) {
      p->pair_off[t] = (int)p->h_pairs.size();
      for (int k = p->h_krange[t].x; k <= p->h_krange[t].y; ++k)
        p->h_pairs.push_back(make_int2((int)t, k));
    }
    p->pair_off[T] = (int)p->h_pairs.size();
  }
  p->h_wscale.assign(T, 1.0f);
  if (d->weights)
    for (size_t t = 0; t < T; ++t) p->h_wscale[t] = d->weights[t];
  p->h_ramp = ramp_table(d->num_pixels, (size_t)ps.L);
  if (ps.ctf) p->h_defocus.assign(d->defocus, d->defocus + d->num_defocus);
  if (ps.ctf) {
    p->rows_bs = !rows_pow2(d->num_pixels);
    p->cols_bs = !cols_pow2(d->num_rows);
    p->MX = (int)filter_length(d->num_pixels);
    p->MY = (int)filter_length(d->num_rows);
    if (p->rows_bs) bluestein_tables(d->num_pixels, p->MX, p->h_bs);
    p->bs_y = p->h_bs.size();
    if (p->cols_bs) bluestein_tables(d->num_rows, p->MY, p->h_bs);
  }
  p->wavelength = d->wavelength;
  p->Cs = d->Cs;
  p->Ca = d->Ca;
  p->Pa = d->Pa;
  p->phi = d->phi;
  p->pixel_size = d->pixel_size;
}

// The plan's device buffers and their sizes, in workspace order.
void plan_buffers(gx_volume_plan *p, const PlanSizes &ps,
                  std::vector<std::pair<DevBuf *, size_t>> &out) {
  const size_t T = p->T, Tpad = (size_t)bp_tilt_pad(p->T);
  out = {{&p->slab, ps.slab_bytes},
         {&p->coef, sizeof(float4) * Tpad},
         {&p->coef_r, sizeof(float4) * Tpad},
         {&p->win, sizeof(int2) * bp_window_entries(p->W, p->Z, p->T)},
         {&p->wscale, sizeof(float) * T},
         {&p->tmax, sizeof(uint32_t) * T},
         {&p->ramp, sizeof(float) * p->h_ramp.size()},
         {&p->defocus, ps.ctf ? sizeof(double) * p->ndf : 0},
         {&p->mask, ps.mask_bytes},
         {&p->F, ps.f_bytes},
         {&p->G, ps.g_bytes},
         {&p->row_off, sizeof(float) * p->h_row_off.size()},
         {&p->bs, sizeof(float2) * p->h_bs.size()},
         {&p->krange, p->ndf > 1 ? sizeof(int2) * T : 0},
         {&p->pairs, p->ndf > 1 ? sizeof(int2) * T * p->ndf : 0}};
}

// Uploads the tables and initialises the device state on stream s.
int plan_upload(gx_volume_plan *p, hipStream_t s) {
  const size_t T = p->T, Tpad = (size_t)bp_tilt_pad(p->T);
  GX_HIP(hipMemcpyAsync(p->coef.p, p->h_coef.data(), sizeof(float4) * Tpad,
                        hipMemcpyHostToDevice, s));
  // the staged kernel's windows, with the arithmetic of the plan's kernels
  if (p->storage == GX_STORE_F32) {
    GX_HIP(launch_bp_windows(p->coef.as<float4>(), 1, p->T, p->W, p->Z,
                             p->dscale, p->doffset, p->ndf, p->win.as<int2>(),
                             s));
  } else {
    GX_HIP(launch_bp_coefficients(p->coef.as<float4>(), nullptr, 1.0f, p->T,
                                  p->coef_r.as<float4>(), s));
    GX_HIP(launch_bp_windows(p->coef_r.as<float4>(), 0, p->T, p->W, p->Z,
                             p->dscale, p->doffset, p->ndf, p->win.as<int2>(),
                             s));
  }
  // the slabs each tilt's voxels read, and the (tilt, defocus) list
  if (p->ndf > 1) {
    GX_HIP(hipMemcpyAsync(p->krange.p, p->h_krange.data(),
                          sizeof(int2) * p->h_krange.size(),
                          hipMemcpyHostToDevice, s));
    GX_HIP(hipMemcpyAsync(p->pairs.p, p->h_pairs.data(),
                          sizeof(int2) * p->h_pairs.size(),
                          hipMemcpyHostToDevice, s));
  }
  GX_HIP(hipMemcpyAsync(p->wscale.p, p->h_wscale.data(), sizeof(float) * T,
                        hipMemcpyHostToDevice, s));
  if (!p->h_bs.empty())
    GX_HIP(hipMemcpyAsync(p->bs.p, p->h_bs.data(),
                          sizeof(float2) * p->h_bs.size(),
                          hipMemcpyHostToDevice, s));
  if (p->per_row)
    GX_HIP(hipMemcpyAsync(p->row_off.p, p->h_row_off.data(),
                          sizeof(float) * p->h_row_off.size(),
                          hipMemcpyHostToDevice, s));
  GX_HIP(hipMemcpyAsync(p->ramp.p, p->h_ramp.data(),
                        sizeof(float) * p->h_ramp.size(),
                        hipMemcpyHostToDevice, s));
  // The slab buffer is not cleared: the filter writes every row below Y of
  // every slab the volume reads, and what else the backprojector reads (the
  // rows of the last row block past Y, the runs of slabs a tile does not use)
  // reaches only accumulators of rows it discards or LDS it never reads.
  if (p->ctf) {
    GX_HIP(hipMemcpyAsync(p->defocus.p, p->h_defocus.data(),
                          sizeof(double) * p->ndf, hipMemcpyHostToDevice, s));
    GX_HIP(launch_ctf_mask(p->mask.as<int8_t>(), p->X, p->Y, p->XB, p->ndf,
                           p->defocus.as<double>(), p->wavelength, p->Cs,
                           p->Ca, p->Pa, p->phi, p->pixel_size, !p->cols_bs,
                           s));
  }
  return GX_OK;
}

}  // namespace

extern "C" int gx_volume_plan_create(const gx_volume_desc *d,
                                     gx_volume_plan **out) {
  GX_CHECK(out != nullptr);
  *out = nullptr;
  PlanSizes ps;
  int rc = plan_sizes(d, &ps);
  if (rc) return rc;
  rc = select_device(d->gpu_index);
  if (rc) return rc;
  hipError_t terr;
  twiddles(&terr);  // build the table outside any capture
  GX_HIP(terr);

  auto *p = new gx_volume_plan();
  std::unique_ptr<gx_volume_plan> guard(p);
  GX_HIP(hipGetDevice(&p->device));
  plan_setup(p, d, ps);
  std::vector<std::pair<DevBuf *, size_t>> bufs;
  plan_buffers(p, ps, bufs);
  for (auto &b : bufs)
    if (b.second) GX_HIP(b.first->alloc(b.second));
  rc = plan_upload(p, nullptr);
  if (rc) return rc;
  GX_HIP(hipDeviceSynchronize());
  *out = guard.release();
  return GX_OK;
}

namespace gx {

// A plan whose buffers are slices of `ws` (gx_volume_plan_bytes bytes,
// kPlanAlign-aligned), initialised on `s` with no allocation and no host
// synchronisation (for the serving loop, which owns the memory).
int plan_create_in(const gx_volume_desc *d, void *ws, size_t bytes,
                   hipStream_t s, gx_volume_plan **out) {
  GX_CHECK(out != nullptr && ws != nullptr);
  GX_CHECK(((uintptr_t)ws % kPlanAlign) == 0);
  *out = nullptr;
  PlanSizes ps;
  int rc = plan_sizes(d, &ps);
  if (rc) return rc;
  GX_CHECK(bytes >= ps.total);
  hipError_t terr;
  twiddles(&terr);
  GX_HIP(terr);
  auto *p = new gx_volume_plan();
  std::unique_ptr<gx_volume_plan> guard(p);
  GX_HIP(hipGetDevice(&p->device));
  plan_setup(p, d, ps);
  std::vector<std::pair<DevBuf *, size_t>> bufs;
  plan_buffers(p, ps, bufs);
  GX_CHECK((int)bufs.size() == kPlanBuffers);
  char *q = (char *)ws;
  for (auto &b : bufs) {
    if (b.second) b.first->borrow(q);
    q += (b.second + kPlanAlign - 1) / kPlanAlign * kPlanAlign;
  }
  GX_CHECK((size_t)(q - (char *)ws) <= bytes);
  rc = plan_upload(p, s);
  if (rc) return rc;
  *out = guard.release();
  return GX_OK;
}

}  // namespace gx

extern "C" int gx_volume_plan_destroy(gx_volume_plan *plan) {
  delete plan;
  return GX_OK;
}

extern "C" int gx_volume_filter_rows(gx_volume_plan *p, const float *d_in,
                                     size_t y0, size_t y1, void *stream) {
  GX_CHECK(p != nullptr && d_in != nullptr);
  GX_CHECK(y0 < y1 && y1 <= (size_t)p->Y);
  GX_CHECK(y0 % (size_t)p->YB == 0);
  GX_CHECK(y1 - y0 <= (size_t)p->band);
  hipStream_t s = (hipStream_t)stream;
  const bool reduced = p->storage != GX_STORE_F32;
  uint32_t *tmax = reduced ? p->tmax.as<uint32_t>() : nullptr;
  if (reduced) GX_HIP(launch_fill_u32(tmax, 0u, (size_t)p->T, s));
  const size_t X = p->X, Y = p->Y, T = p->T, ndf = p->ndf;
  const int by0 = (int)y0, bny = (int)(y1 - y0);
  const gx::BluesteinTable bx =
      bluestein_view(p->bs.as<float2>(), (size_t)p->X, (size_t)p->MX);
  const gx::BluesteinTable by = bluestein_view(
      p->bs.as<float2>() + p->bs_y, (size_t)p->Y, (size_t)p->MY);
  const int2 *kr = p->ndf > 1 ? p->krange.as<int2>() : nullptr;
  if (p->ctf) {
    // The 2D correction needs whole projections: every row is transformed,
    // the column pass writes and the row pass reads only the band's rows.
    // balanced launches of at most `chunk` tilts
    const int nch = (p->T + p->chunk - 1) / p->chunk;
    for (int c = 0; c < nch; ++c) {
      const int t0 = (int)((long)p->T * c / nch);
      const int tc = (int)((long)p->T * (c + 1) / nch) - t0;
      const float *img = d_in + (size_t)t0 * Y * X;
      if (p->rows_bs) {
        GX_HIP(launch_rows_r2c_bs(img, Y * X, p->X, p->Y, p->XB, tc,
                                  p->F.as<float2>(), tmax, t0, p->MX, bx, s));
      } else {
        GX_HIP(launch_rows_r2c(img, Y * X, p->X, p->Y, p->XB, tc,
                               p->F.as<float2>(), tmax, t0, s));
      }
      // only the slabs the volume reads (per tilt of the launch for the
      // column pass, per tilt of the plan for the row pass)
      if (p->cols_bs) {
        GX_HIP(launch_cols_bs(p->F.as<float2>(), p->mask.as<int8_t>(),
                              p->G.as<float2>(), p->Y, p->XB, p->ndf, tc,
                              p->MY, by, s, by0, by0 + bny,
                              kr ? kr + t0 : nullptr));
      } else {
        GX_HIP(launch_cols(p->F.as<float2>(), p->mask.as<int8_t>(),
                           p->G.as<float2>(), p->Y, p->XB, p->ndf, tc, s,
                           by0, by0 + bny, kr ? kr + t0 : nullptr));
      }
      if (p->rows_bs) {
        GX_HIP(launch_rows_c2r_filter_bs(
            p->G.as<float2>(), p->X, p->Y, p->XB, p->ndf, t0, tc, p->T,
            p->ramp.as<float>(), p->wscale.as<float>(), tmax, p->nrm_mul,
            p->storage, p->slab.p, p->NYB, p->YB, p->MX, bx, s, by0, bny,
            kr));
      } else {
        const int2 *pl = kr ? p->pairs.as<int2>() + p->pair_off[t0] : nullptr;
        const int np = kr ? p->pair_off[t0 + tc] - p->pair_off[t0] : 0;
        GX_HIP(launch_rows_c2r_filter(
            p->G.as<float2>(), p->X, p->Y, p->XB, p->ndf, t0, tc, p->T,
            p->wscale.as<float>(), tmax, p->nrm_mul,
            p->storage, p->slab.p, p->NYB, p->YB, s, by0, bny, kr, pl, np));
      }
    }
  } else {
    size_t st, sk, sy;
    if (p->layout == GX_INPUT_SINOGRAMS) {
      st = X;
      sk = T * X;
      sy = ndf * T * X;
    } else {  // (T, [ndf,] Y, X)
      st = ndf * Y * X;
      sk = Y * X;
      sy = X;
    }
    // rows are independent: the band's rows are a view of the input
    const float *src = d_in + y0 * sy;
    if (reduced)
      GX_HIP(launch_absmax_rows(src, st, sk, sy, p->X, bny, p->T, p->ndf,
                                tmax, s));
    GX_HIP(launch_filter_rows(src, st, sk, sy, p->X, p->L, bny, p->T, p->ndf,
                              p->ramp.as<float>(), p->wscale.as<float>(), tmax,
                              p->nrm_mul, p->storage, p->slab.p, p->NYB, p->YB,
                              s, nullptr, kr));
  }
  p->slab_y0 = (int)y0;
  p->slab_y1 = (int)y1;
  return GX_OK;
}

extern "C" int gx_volume_filter(gx_volume_plan *p, const float *d_in,
                                void *stream) {
  GX_CHECK(p != nullptr);
  if (p->band < p->Y)
    return fail(GX_EINVAL,
                "the plan holds %d of %d rows at a time: filter bands with "
                "gx_volume_filter_rows or run the whole path with "
                "gx_volume_run", p->band, p->Y);
  return gx_volume_filter_rows(p, d_in, 0, (size_t)p->Y, stream);
}

extern "C" int gx_volume_backproject_rows(gx_volume_plan *p, size_t y0,
                                          size_t y1, float *d_vol,
                                          void *stream) {
  GX_CHECK(p != nullptr && d_vol != nullptr);
  GX_CHECK(y0 <= y1 && y1 <= (size_t)p->Y);
  GX_CHECK(y0 % (size_t)p->YB == 0);
  GX_CHECK(y1 % (size_t)p->YB == 0 || y1 == (size_t)p->Y);
  if (y0 < y1 && (y0 < (size_t)p->slab_y0 || y1 > (size_t)p->slab_y1))
    return fail(GX_EINVAL,
                "rows [%zu, %zu) are not in the filtered band [%d, %d)", y0,
                y1, p->slab_y0, p->slab_y1);
  hipStream_t s = (hipStream_t)stream;
  BpArgs a{};
  a.slab = p->slab.p;
  if (p->storage == GX_STORE_F32) {
    a.coef = p->coef.as<float4>();
  } else {
    GX_HIP(launch_bp_coefficients(p->coef.as<float4>(), p->tmax.as<uint32_t>(),
                                  p->nrm_mul, p->T, p->coef_r.as<float4>(), s));
    a.coef = p->coef_r.as<float4>();
  }
  a.win = p->win.as<int2>();
  a.dscale = p->dscale;
  a.doffset = p->doffset;
  a.ndf = p->ndf;
  a.N = p->X;
  a.T = p->T;
  a.NYB = p->NYB;
  a.W = p->W;
  a.Z = p->Z;
  a.Y = (int)y1;
  a.yb0 = (int)(y0 / p->YB);
  a.yb1 = (int)((y1 + p->YB - 1) / p->YB);
  a.y_base = (int)y0;
  a.slab_yb0 = p->slab_y0 / p->YB;
  a.vol = d_vol;
  a.staged = p->bp_kind == GX_BP_STAGED ? p->bp_staged : 0;
  a.row_off = p->per_row ? p->row_off.as<float>() : nullptr;
  a.Y_tab = p->Y;
  GX_HIP(launch_bp(a, p->YB, p->storage, s));
  return GX_OK;
}

extern "C" int gx_volume_backproject(gx_volume_plan *p, float *d_vol,
                                     void *stream) {
  GX_CHECK(p != nullptr);
  if (p->band < p->Y)
    return fail(GX_EINVAL,
                "the plan holds %d of %d rows at a time: backproject bands "
                "with gx_volume_backproject_rows or run the whole path with "
                "gx_volume_run", p->band, p->Y);
  return gx_volume_backproject_rows(p, 0, (size_t)p->Y, d_vol, stream);
}

extern "C" int gx_volume_plan_set_backprojector(gx_volume_plan *p, int kind) {
  GX_CHECK(p != nullptr);
  GX_CHECK(kind == GX_BP_DIRECT || kind == GX_BP_STAGED);
  if (kind == GX_BP_STAGED && p->per_row)
    return fail(GX_EUNSUPPORTED,
                "per-row rotation centres need the direct backprojector");
  if (kind == GX_BP_STAGED && p->bp_staged == 0)
    return fail(GX_EUNSUPPORTED,
                "a voxel tile spans more than 4 defocus slabs; the staged "
                "backprojector holds at most 4");
  p->bp_kind = kind;
  return GX_OK;
}

extern "C" int gx_volume_plan_get_backprojector(gx_volume_plan *p, int *kind,
                                                int *slabs_per_tilt) {
  GX_CHECK(p != nullptr);
  if (kind) *kind = p->bp_kind;
  if (slabs_per_tilt) *slabs_per_tilt = p->bp_staged;
  return GX_OK;
}

extern "C" int gx_volume_rows_per_block(gx_volume_plan *p, size_t *rows) {
  GX_CHECK(p != nullptr && rows != nullptr);
  *rows = (size_t)p->YB;
  return GX_OK;
}

extern "C" int gx_volume_run(gx_volume_plan *p, const float *d_in,
                             float *d_vol, void *stream) {
  GX_CHECK(p != nullptr && d_vol != nullptr);
  // one band of rows at a time: filter it into the slabs, backproject it
  // into its rows of the volume
  for (size_t y0 = 0; y0 < (size_t)p->Y; y0 += (size_t)p->band) {
    const size_t y1 = std::min((size_t)p->Y, y0 + (size_t)p->band);
    int rc = gx_volume_filter_rows(p, d_in, y0, y1, stream);
    if (rc) return rc;
    rc = gx_volume_backproject_rows(p, y0, y1,
                                    d_vol + y0 * (size_t)p->Z * p->W, stream);
    if (rc) return rc;
  }
  return GX_OK;
}

extern "C" int gx_volume_plan_band(gx_volume_plan *p, size_t *band_rows,
                                   size_t *y0, size_t *y1) {
  GX_CHECK(p != nullptr);
  if (band_rows) *band_rows = (size_t)p->band;
  if (y0) *y0 = (size_t)p->slab_y0;
  if (y1) *y1 = (size_t)p->slab_y1;
  return GX_OK;
}

extern "C" int gx_volume_slabs(gx_volume_plan *p, void **d_slab, size_t *bytes,
                               int *rows_per_block) {
  GX_CHECK(p != nullptr);
  if (d_slab) *d_slab = p->slab.p;
  if (bytes) *bytes = p->slab_bytes;
  if (rows_per_block) *rows_per_block = p->YB;
  return GX_OK;
}

extern "C" int gx_volume_plan_slab_range(gx_volume_plan *p, int *first,
                                         int *last) {
  GX_CHECK(p != nullptr && first != nullptr && last != nullptr);
  for (int t = 0; t < p->T; ++t) {
    first[t] = p->ndf > 1 ? p->h_krange[t].x : 0;
    last[t] = p->ndf > 1 ? p->h_krange[t].y : 0;
  }
  return GX_OK;
}

extern "C" int gx_volume_plan_get_tile(gx_volume_plan *p, int *rows,
                                       int *kind) {
  GX_CHECK(p != nullptr);
  if (rows) *rows = p->YB;
  if (kind) *kind = p->bp_kind;
  return GX_OK;
}

extern "C" int gx_volume_plan_set_tile(gx_volume_plan *p, int rows,
                                       int kind) {
  GX_CHECK(p != nullptr);
  if (!tile_valid(p, rows, kind))
    return fail(GX_EUNSUPPORTED,
                "tile of %d rows with backprojector %d is not a candidate for "
                "this plan", rows, kind);
  set_tile(p, rows, kind);
  return GX_OK;
}

extern "C" int gx_volume_plan_autotune(gx_volume_plan *p, const float *d_in,
                                       float *d_vol, void *stream,
                                       gx_tune_result *out) {
  GX_CHECK(p != nullptr && d_in != nullptr && d_vol != nullptr);
  GX_HIP(hipSetDevice(p->device));
  hipStream_t s = (hipStream_t)stream;
  EventPair ev;
  GX_HIP(hipEventCreate(&ev.a));
  GX_HIP(hipEventCreate(&ev.b));
  gx_tune_result r{};
  int rows[2];
  tile_candidates(p->storage, rows);
  const int kinds[2] = {GX_BP_STAGED, GX_BP_DIRECT};
  float best = 0.0f;
  int rc;
  // a banded plan is timed on its first band (rows [0, band))
  const size_t yt = (size_t)p->band;
  for (int ri = 0; ri < 2; ++ri) {
    set_tile(p, rows[ri], p->bp_kind);
    // the first pass warms up; the second is timed
    if ((rc = gx_volume_filter_rows(p, d_in, 0, yt, s))) return rc;
    GX_HIP(hipEventRecord(ev.a, s));
    if ((rc = gx_volume_filter_rows(p, d_in, 0, yt, s))) return rc;
    GX_HIP(hipEventRecord(ev.b, s));
    GX_HIP(hipEventSynchronize(ev.b));
    float fms = 0.0f;
    GX_HIP(hipEventElapsedTime(&fms, ev.a, ev.b));
    for (int kind : kinds) {
      if (!tile_valid(p, rows[ri], kind)) continue;
      p->bp_kind = kind;
      if ((rc = gx_volume_backproject_rows(p, 0, yt, d_vol, s))) return rc;
      GX_HIP(hipEventRecord(ev.a, s));
      for (int i = 0; i < 2; ++i)
        if ((rc = gx_volume_backproject_rows(p, 0, yt, d_vol, s))) return rc;
      GX_HIP(hipEventRecord(ev.b, s));
      GX_HIP(hipEventSynchronize(ev.b));
      float bms = 0.0f;
      GX_HIP(hipEventElapsedTime(&bms, ev.a, ev.b));
      bms *= 0.5f;
      GX_CHECK(r.n < GX_TUNE_MAX);
      r.cand[r.n++] = gx_tune_entry{kind, rows[ri], fms, bms};
      if (r.n == 1 || fms + bms < best) {
        best = fms + bms;
        r.kind = kind;
        r.rows = rows[ri];
      }
    }
  }
  set_tile(p, r.rows, r.kind);
  if ((rc = gx_volume_run(p, d_in, d_vol, s))) return rc;
  GX_HIP(hipStreamSynchronize(s));
  {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    g_tune[tune_key(p)] = TuneChoice{r.kind, r.rows};
  }
  if (out) *out = r;
  return GX_OK;
}

extern "C" int gx_tune_save(const char *path) {
  GX_CHECK(path != nullptr);
  std::string text =
      "# guanaco_amd tile choices: model X Y T ndf Z W storage layout ctf "
      "slabs -> kind rows\n";
  {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    for (const auto &e : g_tune)
      text += e.first + " " + std::to_string(e.second.kind) + " " +
              std::to_string(e.second.rows) + "\n";
  }
  FILE *f = std::fopen(path, "w");
  if (!f) return fail(GX_EINVAL, "cannot write %s", path);
  const bool ok = std::fwrite(text.data(), 1, text.size(), f) == text.size();
  if (std::fclose(f) != 0 || !ok) return fail(GX_EINVAL, "cannot write %s", path);
  return GX_OK;
}

extern "C" int gx_tune_load(const char *path) {
  GX_CHECK(path != nullptr);
  FILE *f = std::fopen(path, "r");
  if (!f) return fail(GX_EINVAL, "cannot read %s", path);
  std::map<std::string, TuneChoice> got;
  char line[512];
  int lineno = 0;
  while (std::fgets(line, sizeof line, f)) {
    ++lineno;
    std::string l(line);
    while (!l.empty() && (l.back() == '\n' || l.back() == '\r')) l.pop_back();
    if (l.empty() || l[0] == '#') continue;
    // key (11 fields) then kind and rows
    const size_t p2 = l.find_last_of(' ');
    const size_t p1 = p2 == std::string::npos ? p2 : l.find_last_of(' ', p2 - 1);
    int kind = -1, rows = -1, nfield = 1;
    for (char c : l) nfield += c == ' ';
    if (p1 == std::string::npos || nfield != 13 ||
        std::sscanf(l.c_str() + p1, "%d %d", &kind, &rows) != 2 ||
        (kind != GX_BP_DIRECT && kind != GX_BP_STAGED) ||
        (rows != 8 && rows != 16 && rows != 32)) {
      std::fclose(f);
      return fail(GX_EINVAL, "%s:%d: not a tile choice line", path, lineno);
    }
    got[l.substr(0, p1)] = TuneChoice{kind, rows};
  }
  std::fclose(f);
  std::lock_guard<std::mutex> lock(g_tune_mu);
  for (const auto &e : got) g_tune[e.first] = e.second;
  return GX_OK;
}

extern "C" int gx_tune_clear(void) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  g_tune.clear();
  return GX_OK;
}

// ------------------------------------------------------- CTF images -------

extern "C" int gx_ctf_grid(const gx_ctf_desc *d, float *out) {
  GX_CHECK(d != nullptr && out != nullptr && d->defocus != nullptr);
  GX_CHECK(d->width > 0 && d->height > 0 && d->num_defocus > 0);
  GX_CHECK(d->pixel_size > 0);
  GX_CHECK(d->width <= (1u << 16) && d->height <= (1u << 16));
  for (size_t k = 0; k < d->num_defocus; ++k)
    if (!std::isfinite(d->defocus[k]))
      return fail(GX_EINVAL, "defocus %zu is not finite", k);
  int rc = select_device(d->gpu_index);
  if (rc) return rc;
  const size_t n = d->width * d->height;
  // defoci per launch: about 256 MiB of output, at least one
  size_t batch = ((size_t)256 << 20) / (n * sizeof(float2));
  if (batch < 1) batch = 1;
  if (batch > d->num_defocus) batch = d->num_defocus;
  GX_CHECK(batch <= 65535);
  DevBuf d_out, d_df;
  GX_HIP(d_out.alloc(batch * n * sizeof(float2)));
  GX_HIP(d_df.alloc(d->num_defocus * sizeof(double)));
  GX_HIP(hipMemcpy(d_df.p, d->defocus, d->num_defocus * sizeof(double),
                   hipMemcpyHostToDevice));
  const CtfParams c{d->wavelength, d->Cs, d->Ca, d->Pa,
                    d->dd, d->theta_c, d->phi};
  for (size_t k0 = 0; k0 < d->num_defocus; k0 += batch) {
    const size_t nk = std::min(batch, d->num_defocus - k0);
    GX_HIP(launch_ctf_grid(d_out.as<float2>(), (int)d->width, (int)d->height,
                           d->pixel_size, d_df.as<double>() + k0, (int)nk, c,
                           d->envelope != 0, d->recentre != 0, nullptr));
    GX_HIP(hipMemcpy(out + 2 * k0 * n, d_out.p, nk * n * sizeof(float2),
                     hipMemcpyDeviceToHost));
  }
  return GX_OK;
}

// ------------------------------------------------------- synthetic data ---

extern "C" int gx_synth_projections(const gx_synth_desc *d, float *d_out,
                                    void *stream) {
  GX_CHECK(d != nullptr && d_out != nullptr && d->angles != nullptr);
  GX_CHECK(d->num_pixels > 0 && d->num_rows > 0 && d->num_angles > 0);
  GX_CHECK(d->num_ellipsoids <= 512);
  GX_CHECK(d->noise >= 0 && d->noise <= 2);
  if (d->ctf) {
    int rc = check_pow2_corr_sizes(d->num_pixels, d->num_rows);
    if (rc) return rc;
    GX_CHECK(d->pixel_size > 0);
  }
  SynthArgs a{};
  a.num_pixels = d->num_pixels;
  a.num_rows = d->num_rows;
  a.num_angles = d->num_angles;
  a.depth = d->depth;
  a.width = d->width;
  a.centre = d->centre;
  a.angles = d->angles;
  a.num_ellipsoids = d->num_ellipsoids;
  a.seed = d->seed;
  a.semi_min = d->semi_min;
  a.semi_max = d->semi_max;
  a.density_min = d->density_min;
  a.density_max = d->density_max;
  a.noise = d->noise;
  a.noise_level = d->noise_level;
  a.contrast = d->contrast;
  a.ctf = d->ctf;
  a.wavelength = d->wavelength;
  a.Cs = d->Cs;
  a.defocus = d->defocus;
  a.phi = d->phi;
  a.pixel_size = d->pixel_size;
  hipError_t terr;
  twiddles(&terr);
  GX_HIP(terr);
  GX_HIP(synth_projections(&a, d_out, (hipStream_t)stream));
  return GX_OK;
}

namespace gx {

// The error channel for the library's other translation units.
int fail_status(int code, const char *msg) { return fail(code, "%s", msg); }

}  // namespace gx
