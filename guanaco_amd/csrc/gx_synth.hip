// Seeded synthetic tilt series on the GPU (benchmark and test workloads).
//
// A phantom of random ellipsoids is projected analytically in the
// reconstruction geometry of reconstructor.cpp:14-90 (detector coordinate
// u = x cos(theta) - z sin(theta), ray direction (sin(theta), 0, cos(theta))
// in (x, y, z)), optionally multiplied in Fourier space by the CTF
// sin(chi - phi) (chi as ctf.h:233-238), and given Gaussian or Poisson noise
// from a counter-based generator, so every pixel is reproducible from the
// seed alone.
#include <cmath>
#include <vector>

#include "gx_common.hpp"
#include "gx_fft.hpp"
#include "gx_kernels.hpp"

namespace gx {

struct Ellipsoid {
  float cx, cy, cz;              // centre (pixels, volume-centred)
  float a00, a01, a02, a11, a12, a22;  // R diag(1/a^2) R^T
  float density, pad;
};

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

// uniform in (0, 1]
__device__ __forceinline__ float unif(uint64_t seed, uint64_t idx, int j) {
  const uint64_t h = mix64(seed ^ mix64(idx * 4 + (uint64_t)j));
  return ((float)(h >> 40) + 1.0f) * (1.0f / 16777216.0f);
}

__global__ void k_phantom_project(const Ellipsoid *__restrict__ el, int ne,
                                  const float2 *__restrict__ cs, int X, int Y,
                                  float centre, float *__restrict__ out) {
  __shared__ Ellipsoid sh[512];
  for (int i = threadIdx.x; i < ne && i < 512; i += blockDim.x) sh[i] = el[i];
  lds_barrier();
  const int t = blockIdx.z;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y;
  if (i >= X) return;
  const float c = cs[t].x, s = cs[t].y;
  const float u = (float)i + 0.5f - centre;
  const float y = (float)r + 0.5f - 0.5f * (float)Y;
  const float r0x = u * c, r0z = -u * s;
  const float dx = s, dz = c;  // unit ray direction, dy = 0
  float acc = 0.0f;
  for (int e = 0; e < ne; ++e) {
    const Ellipsoid q = sh[e];
    const float qx = r0x - q.cx, qy = y - q.cy, qz = r0z - q.cz;
    // alpha = d^T A d, beta = d^T A q, gamma = q^T A q
    const float Adx = q.a00 * dx + q.a02 * dz;
    const float Ady = q.a01 * dx + q.a12 * dz;
    const float Adz = q.a02 * dx + q.a22 * dz;
    const float alpha = dx * Adx + dz * Adz;
    const float beta = qx * Adx + qy * Ady + qz * Adz;
    const float gamma = qx * (q.a00 * qx + q.a01 * qy + q.a02 * qz) +
                        qy * (q.a01 * qx + q.a11 * qy + q.a12 * qz) +
                        qz * (q.a02 * qx + q.a12 * qy + q.a22 * qz);
    const float disc = beta * beta - alpha * (gamma - 1.0f);
    if (disc > 0.0f) acc += q.density * 2.0f * sqrtf(disc) / alpha;
  }
  out[((size_t)t * Y + r) * X + i] = acc;
}

// CTF transfer sin(chi - phi) on the half-spectrum layout [cb][y][ci].
__global__ void k_ctf_transfer(float *__restrict__ tr, int X, int Y, int XB,
                               double l, double Cs, double df, double phi,
                               double inv_ps) {
  const size_t total = (size_t)XB * Y * CB;
  const size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (e >= total) return;
  const int ci = (int)(e % CB);
  const int ky = (int)((e / CB) % Y);
  const int cb = (int)(e / ((size_t)CB * Y));
  const int kx = cb * CB + ci;
  float v = 0.0f;
  if (kx <= X / 2) {
    const size_t w2 = (size_t)X / 2, h2 = (size_t)Y / 2;
    const double fx = ((double)(((size_t)kx + w2) % (size_t)X) - (double)w2) / X;
    const double fy = ((double)(((size_t)ky + h2) % (size_t)Y) - (double)h2) / Y;
    const double q2 = (fx * fx + fy * fy) * inv_ps * inv_ps;
    const double chi = M_PI * (Cs * l * l * l * q2 * q2 / 2 - l * df * q2);
    v = (float)sin(chi - phi);
  }
  tr[e] = v;
}

// mode 1: out = img + N(0, (level * max|img|)^2);
// mode 2: out = Poisson(level * max(1 + contrast * img / max|img|, 0)).
__global__ void k_noise(float *__restrict__ img, size_t n, int mode,
                        float level, float contrast,
                        const uint32_t *__restrict__ gmax, uint64_t seed) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float v = img[i];
  if (mode == 1) {
    const float sigma = level * __uint_as_float(*gmax);
    const float u1 = unif(seed, i, 0), u2 = unif(seed, i, 1);
    const float z = sqrtf(-2.0f * logf(u1)) * cospif(2.0f * u2);
    img[i] = v + sigma * z;
  } else if (mode == 2) {
    const float m = __uint_as_float(*gmax);
    const float lam =
        level * fmaxf(1.0f + (m > 0.0f ? contrast * v / m : 0.0f), 0.0f);
    // inversion by sequential search; ends by k_max for any lambda used
    const float u = unif(seed, i, 2);
    float p = expf(-lam), F = p;
    int k = 0;
    const int k_max = (int)(lam + 12.0f * sqrtf(lam) + 40.0f);
    while (u > F && k < k_max) {
      ++k;
      p *= lam / (float)k;
      F += p;
    }
    img[i] = (float)k;
  }
}

__global__ void k_max_of(const uint32_t *__restrict__ v, int n,
                         uint32_t *__restrict__ out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    uint32_t m = 0;
    for (int i = 0; i < n; ++i) m = v[i] > m ? v[i] : m;
    *out = m;
  }
}

namespace {

uint64_t host_mix(uint64_t z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

struct HostRng {
  uint64_t s;
  double next() { return (double)(host_mix(s++) >> 11) * (1.0 / 9007199254740992.0); }
  double uniform(double a, double b) { return a + (b - a) * next(); }
};

}  // namespace

hipError_t synth_projections(const SynthArgs *d, float *d_out, hipStream_t s) {
  const int X = (int)d->num_pixels, Y = (int)d->num_rows;
  const int T = (int)d->num_angles;
  const int W = (int)(d->width ? d->width : d->num_pixels);
  const int Z = (int)(d->depth ? d->depth : d->num_pixels);

  // phantom: centres uniform in the volume, semi-axes and densities uniform,
  // random orientation (three Euler angles)
  HostRng rng{d->seed * 0x100000001b3ull + 1};
  std::vector<Ellipsoid> el(d->num_ellipsoids);
  for (auto &e : el) {
    e.cx = (float)rng.uniform(-0.5 * W, 0.5 * W);
    e.cy = (float)rng.uniform(-0.5 * Y, 0.5 * Y);
    e.cz = (float)rng.uniform(-0.5 * Z, 0.5 * Z);
    double ax[3];
    for (double &a : ax) a = rng.uniform(d->semi_min, d->semi_max);
    const double p = rng.uniform(0, 2 * M_PI), q = rng.uniform(0, M_PI),
                 r = rng.uniform(0, 2 * M_PI);
    const double cp = std::cos(p), sp = std::sin(p), cq = std::cos(q),
                 sq = std::sin(q), cr = std::cos(r), sr = std::sin(r);
    const double R[3][3] = {{cp * cr - sp * cq * sr, -cp * sr - sp * cq * cr, sp * sq},
                            {sp * cr + cp * cq * sr, -sp * sr + cp * cq * cr, -cp * sq},
                            {sq * sr, sq * cr, cq}};
    double A[3][3];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        A[i][j] = 0;
        for (int k = 0; k < 3; ++k) A[i][j] += R[i][k] * R[j][k] / (ax[k] * ax[k]);
      }
    e.a00 = (float)A[0][0];
    e.a01 = (float)A[0][1];
    e.a02 = (float)A[0][2];
    e.a11 = (float)A[1][1];
    e.a12 = (float)A[1][2];
    e.a22 = (float)A[2][2];
    e.density = (float)rng.uniform(d->density_min, d->density_max);
    e.pad = 0;
  }
  std::vector<float2> cs(T);
  for (int t = 0; t < T; ++t)
    cs[t] = make_float2(std::cos((double)d->angles[t]),
                        std::sin((double)d->angles[t]));

  void *d_el = nullptr, *d_cs = nullptr, *d_max = nullptr;
  void *d_F = nullptr, *d_G = nullptr, *d_tr = nullptr;
  hipError_t rc = hipSuccess;
  auto chk = [&](hipError_t e) {
    if (e != hipSuccess && rc == hipSuccess) rc = e;
    return rc == hipSuccess;
  };
  const size_t n = (size_t)T * Y * X;
  if (chk(hipMalloc(&d_el, sizeof(Ellipsoid) * (el.empty() ? 1 : el.size()))) &&
      chk(hipMalloc(&d_cs, sizeof(float2) * T)) &&
      chk(hipMalloc(&d_max, sizeof(uint32_t) * (T + 1))) &&
      chk(hipMemcpyAsync(d_el, el.data(), sizeof(Ellipsoid) * el.size(),
                         hipMemcpyHostToDevice, s)) &&
      chk(hipMemcpyAsync(d_cs, cs.data(), sizeof(float2) * T,
                         hipMemcpyHostToDevice, s))) {
    hipLaunchKernelGGL(k_phantom_project, dim3((X + 255) / 256, Y, T),
                       dim3(256), 0, s, (const Ellipsoid *)d_el,
                       (int)el.size(), (const float2 *)d_cs, X, Y,
                       d->centre, d_out);
    chk(hipGetLastError());
  }
  if (rc == hipSuccess && d->ctf) {
    const int XB = X / 2 / CB + 1;
    const size_t nspec = (size_t)XB * CB * Y;
    if (rc == hipSuccess && chk(hipMalloc(&d_F, sizeof(float2) * nspec)) &&
        chk(hipMalloc(&d_G, sizeof(float2) * nspec)) &&
        chk(hipMalloc(&d_tr, sizeof(float) * nspec))) {
      hipLaunchKernelGGL(k_ctf_transfer, dim3((unsigned)((nspec + 255) / 256)),
                         dim3(256), 0, s, (float *)d_tr, X, Y, XB,
                         d->wavelength, d->Cs, d->defocus, d->phi,
                         1.0 / d->pixel_size);
      chk(hipGetLastError());
      for (int t = 0; t < T && rc == hipSuccess; ++t) {
        float *img = d_out + (size_t)t * Y * X;
        chk(launch_rows_r2c(img, (size_t)Y * X, X, Y, XB, 1, (float2 *)d_F,
                            nullptr, 0, s));
        chk(launch_cols_transfer((const float2 *)d_F, (const float *)d_tr,
                                 (float2 *)d_G, Y, XB, 1, s));
        chk(launch_rows_c2r((const float2 *)d_G, X, Y, XB, 1, 1, img, s));
      }
    }
  }
  if (rc == hipSuccess && d->noise != 0) {
    uint32_t *tm = (uint32_t *)d_max;
    chk(launch_fill_u32(tm, 0u, (size_t)T + 1, s));
    chk(launch_absmax_rows(d_out, (size_t)Y * X, 0, (size_t)X, X, Y, T, 1, tm,
                           s));
    hipLaunchKernelGGL(k_max_of, dim3(1), dim3(64), 0, s, tm, T, tm + T);
    chk(hipGetLastError());
    hipLaunchKernelGGL(k_noise, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       s, d_out, n, d->noise, d->noise_level, d->contrast,
                       (const uint32_t *)(tm + T), d->seed);
    chk(hipGetLastError());
  }
  chk(hipStreamSynchronize(s));
  (void)hipFree(d_el);
  (void)hipFree(d_cs);
  (void)hipFree(d_max);
  (void)hipFree(d_F);
  (void)hipFree(d_G);
  (void)hipFree(d_tr);
  return rc;
}

}  // namespace gx
