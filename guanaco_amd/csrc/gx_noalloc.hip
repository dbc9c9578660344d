// gx_recon_noalloc: gx_recon on device buffers with no device allocation.
//
// For callers that cannot allocate device memory (code under graph capture,
// tight serving loops, harnesses that forbid hipMalloc): the filtered copy
// of the sinogram and the per-call tables live in `scratch`, a device global
// of kScratchBytes, and the filter's twiddle factors in `noalloc_twiddles`,
// another device global filled on the first call on a device. Every GPU the
// library runs on holds its own copy of the arrays until the process ends,
// and every call on that GPU shares them, so calls that overlap on one GPU
// race. The kernels are gx_recon's.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <mutex>
#include <vector>

#include "guanaco_hip.h"
#include "gx_common.hpp"
#include "gx_fft.hpp"
#include "gx_kernels.hpp"

using namespace gx;

constexpr size_t kScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
static __device__ __align__(256) float2 noalloc_twiddles[tw_table_len()];

namespace {

std::mutex g_mu;
void *g_scratch[64] = {};   // per device: the address of its copy
void *g_twiddles[64] = {};  // per device: its filled twiddle table

int fail_hip(hipError_t e, const char *what) {
  char buf[256];
  std::snprintf(buf, sizeof buf, "gx_recon_noalloc: %s: %s", what,
                hipGetErrorString(e));
  return fail_status(e == hipErrorOutOfMemory ? GX_ENOMEM : GX_EDEVICE, buf);
}

// This GPU's copies of `scratch` and `noalloc_twiddles`: the arrays' names
// are host-side stand-ins, their device addresses come from
// hipGetSymbolAddress, once per device with that device current; the
// twiddle table is written then (a synchronous copy, no allocation).
int device_arrays(void **scr, const float2 **tw) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail_hip(e, "hipGetDevice");
  if (dev < 0 || dev >= 64)
    return fail_status(GX_EDEVICE, "gx_recon_noalloc: device index");
  std::lock_guard<std::mutex> lock(g_mu);
  if (g_scratch[dev] == nullptr) {
    void *p = nullptr, *q = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return fail_hip(e, "hipGetSymbolAddress(scratch)");
    e = hipGetSymbolAddress(&q, HIP_SYMBOL(noalloc_twiddles));
    if (e != hipSuccess)
      return fail_hip(e, "hipGetSymbolAddress(noalloc_twiddles)");
    if (p == nullptr || q == nullptr)
      return fail_status(GX_EDEVICE, "gx_recon_noalloc: no device address");
    const std::vector<float2> h = twiddle_host_table();
    e = hipMemcpy(q, h.data(), sizeof(float2) * h.size(),
                  hipMemcpyHostToDevice);
    if (e != hipSuccess) return fail_hip(e, "twiddle table");
    g_scratch[dev] = p;
    g_twiddles[dev] = q;
  }
  *scr = g_scratch[dev];
  *tw = (const float2 *)g_twiddles[dev];
  return GX_OK;
}

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

#define NA_HIP(call)                                \
  do {                                              \
    hipError_t e_ = (call);                         \
    if (e_ != hipSuccess) return fail_hip(e_, #call); \
  } while (0)

}  // namespace

extern "C" int gx_recon_noalloc(const gx_recon_cfg *cfg, const float *d_sino,
                                float *d_rec, void *stream) {
  int rc = check_recon_cfg(cfg);
  if (rc) return rc;
  if (d_sino == nullptr || d_rec == nullptr)
    return fail_status(GX_EINVAL, "gx_recon_noalloc: NULL buffer");
  const size_t T = cfg->num_angles, N = cfg->num_pixels;
  const size_t nrows = cfg->num_defocus * T;
  const size_t L = filter_length(N);
  // scratch: filtered rows | per-tilt constants | ramp table | row weights
  const size_t b_filt = align256(sizeof(float) * nrows * N);
  const size_t b_coef = align256(sizeof(float4) * T);
  const size_t b_ramp = align256(sizeof(float) * (L / 2 + 1));
  const size_t b_ones = align256(sizeof(float) * nrows);
  const size_t need = b_filt + b_coef + b_ramp + b_ones;
  if (need > kScratchBytes) {
    char buf[256];
    std::snprintf(buf, sizeof buf,
                  "gx_recon_noalloc needs %zu bytes of scratch (num_defocus "
                  "x num_angles x num_pixels floats and tables); it has %zu",
                  need, kScratchBytes);
    return fail_status(GX_EINVAL, buf);
  }
  if (cfg->gpu_index >= 0) NA_HIP(hipSetDevice(cfg->gpu_index));
  void *base = nullptr;
  const float2 *tw = nullptr;
  rc = device_arrays(&base, &tw);
  if (rc) return rc;
  char *q = (char *)base;
  float *d_filt = (float *)q;
  float4 *d_coef = (float4 *)(q + b_filt);
  float *d_ramp = (float *)(q + b_filt + b_coef);
  float *d_ones = (float *)(q + b_filt + b_coef + b_ramp);
  hipStream_t s = (hipStream_t)stream;
  // earlier work on this stream may still read the scratch
  NA_HIP(hipStreamSynchronize(s));
  std::vector<float> coef(4 * T), ones(nrows, 1.0f);
  tilt_geometry(T, cfg->angles, cfg->centre, coef.data());
  const std::vector<float> ramp = ramp_table(N, L);
  // synchronous copies: the host tables end with this call
  NA_HIP(hipMemcpy(d_coef, coef.data(), sizeof(float) * coef.size(),
                   hipMemcpyHostToDevice));
  NA_HIP(hipMemcpy(d_ramp, ramp.data(), sizeof(float) * ramp.size(),
                   hipMemcpyHostToDevice));
  NA_HIP(hipMemcpy(d_ones, ones.data(), sizeof(float) * nrows,
                   hipMemcpyHostToDevice));
  // reconstructor.cpp:92-115: filter a copy (the input is not modified),
  // then the backprojection, which writes every voxel of rec
  NA_HIP(launch_filter_rows(d_sino, N, 0, 0, (int)N, (int)L, 1, (int)nrows, 1,
                            d_ramp, d_ones, nullptr, 1.0f, kF32, d_filt, 1, 1,
                            s, tw));
  BpArgs a{};
  a.slab = d_filt;
  a.coef = d_coef;
  a.win = nullptr;
  defocus_coefficients(cfg->num_defocus, cfg->pixel_size, cfg->min_defocus,
                       cfg->max_defocus, &a.dscale, &a.doffset);
  a.ndf = (int)cfg->num_defocus;
  a.N = (int)N;
  a.T = (int)T;
  a.NYB = 1;
  a.W = (int)cfg->grid_width;
  a.Z = (int)cfg->grid_height;
  a.Y = 1;
  a.yb0 = 0;
  a.yb1 = 1;
  a.y_base = 0;
  a.vol = d_rec;
  NA_HIP(launch_bp(a, 1, kF32, s));
  NA_HIP(hipStreamSynchronize(s));
  return GX_OK;
}
