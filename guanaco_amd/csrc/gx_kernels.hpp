// Kernel launchers shared by the C-ABI layer (host side, compiled by hipcc).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <vector>

#include "guanaco_hip.h"

namespace gx {

// Tables of per-tilt values are padded with zero entries to a multiple of
// kBpTiltPad, so that kernels can read whole batches of tilts.
constexpr int kBpTiltPad = 8;
__host__ __device__ constexpr int bp_tilt_pad(int T) {
  return (T + kBpTiltPad - 1) / kBpTiltPad * kBpTiltPad;
}

struct BpArgs {
  const void *slab;          // [k][t][y/YB][x][y%YB]
  // per tilt (bp_tilt_pad(T) entries):
  //   kF32:    c, s, offset, scale (the reference's arithmetic)
  //   reduced: c, s, offset - 1/2, scale * normalisation (bp_coefficients)
  const float4 *coef;
  const int2 *win;           // staged kernel: per (tile, tilt) window base and
                             // first slab, bit 30 set when the tile lies in
                             // that one slab (bp_windows)
  float dscale, doffset;     // slab selection (reconstructor.cpp:29-32)
  int ndf, N, T, NYB, W, Z, Y;
  int yb0, yb1;              // row blocks [yb0, yb1) to reconstruct
  int y_base;                // vol holds rows [y_base, ...)
  int slab_yb0;              // slab holds row blocks [slab_yb0, ...)
  float *vol;                // [y - y_base][z][x]
  // per-row rotation centres: row_off[t * Y_tab + y] is the tilt's offset for
  // row y (kF32: the reference's offset; reduced: offset - 1/2), or null
  const float *row_off;
  int Y_tab;
  int staged;                // 0: direct kernel; 2 or 4 (or 1 with one slab):
                             // LDS-staged kernel holding that many slabs
};


// Each returns hipSuccess or the launch error.
hipError_t launch_fill_u32(uint32_t *p, uint32_t v, size_t n, hipStream_t s);
hipError_t launch_copy_u32(uint32_t *dst, const uint32_t *src, size_t n,
                           hipStream_t s);

// Tables of an n-point Bluestein transform on M-point FFTs (device
// pointers): chirp w_m = exp(-i pi m^2 / n) (m < n), and the spectra of the
// convolution kernels for the forward and inverse transforms, divided by M.
struct BluesteinTable {
  const float2 *chirp;
  const float2 *bf;
  const float2 *bi;
  int n;
};
// Correction passes for any X, Y (<= 4096): M = the Bluestein length,
// a power of two >= 2n - 1 (and >= 8).
hipError_t launch_rows_r2c_bs(const float *img, size_t tstride, int X, int Y,
                              int XB, int ntilts, float2 *F, uint32_t *tmax,
                              int t0, int M, const BluesteinTable &bx,
                              hipStream_t s);
// ylo, yhi: the rows of G written (the band the row pass reads; yhi < 0:
// every row).
// krange (optional): per tilt of the launch, the first and last defocus to
// compute; the others are skipped.
hipError_t launch_cols_bs(const float2 *F, const int8_t *mask, float2 *G,
                          int Y, int XB, int ndf, int ntilts, int M,
                          const BluesteinTable &by, hipStream_t s,
                          int ylo = 0, int yhi = -1,
                          const int2 *krange = nullptr);
hipError_t launch_rows_c2r_bs(const float2 *G, int X, int Y, int XB, int ndf,
                              int ntilts, float *out, int M,
                              const BluesteinTable &bx, hipStream_t s);
hipError_t launch_rows_c2r_filter_bs(const float2 *G, int X, int Y, int XB,
                                     int ndf, int t0, int ntilts, int T,
                                     const float *ramp, const float *wscale,
                                     const uint32_t *tmax, float nrm_mul,
                                     int store, void *slab, int NYB, int YB,
                                     int M, const BluesteinTable &bx,
                                     hipStream_t s, int y0 = 0, int ny = -1,
                                     const int2 *krange = nullptr);

// CTF parameters in the reference's units (ctf.h:167-178): A and rad
struct CtfParams {
  double l, Cs, Ca, Pa, dd, theta_c, phi;
};
hipError_t launch_ctf_grid(float2 *out, int w, int h, double ps,
                           const double *d_defocus, int ndf,
                           const CtfParams &c, bool envelope, bool recentre,
                           hipStream_t s);
// thread_order: the layout of the power-of-two column pass (k_cols), else
// the Bluestein column pass's [k][cb][ky][ci]
hipError_t launch_ctf_mask(int8_t *mask, int X, int Y, int XB, int ndf,
                           const double *d_defocus, double l, double Cs,
                           double Ca, double Pa, double phi, double ps,
                           bool thread_order, hipStream_t s);
hipError_t launch_mask_from_ctf(int8_t *mask, const float2 *ctf, int X, int Y,
                                int XB, int ndf, bool thread_order,
                                hipStream_t s);

hipError_t launch_rows_r2c(const float *img, size_t tstride, int X, int Y,
                           int XB, int ntilts, float2 *F, uint32_t *tmax,
                           int t0, hipStream_t s);
hipError_t launch_cols(const float2 *F, const int8_t *mask, float2 *G, int Y,
                       int XB, int ndf, int ntilts, hipStream_t s,
                       int ylo = 0, int yhi = -1,
                       const int2 *krange = nullptr);
// Same column pass with a real transfer function (one "defocus").
hipError_t launch_cols_transfer(const float2 *F, const float *transfer,
                                float2 *G, int Y, int XB, int ntilts,
                                hipStream_t s);
hipError_t launch_rows_c2r_filter(const float2 *G, int X, int Y, int XB,
                                  int ndf, int t0, int ntilts, int T,
                                  const float *wscale,
                                  const uint32_t *tmax, float nrm_mul,
                                  int store, void *slab, int NYB, int YB,
                                  hipStream_t s, int y0 = 0, int ny = -1,
                                  const int2 *krange = nullptr,
                                  const int2 *pairs = nullptr,
                                  int npairs = 0);
hipError_t launch_rows_c2r(const float2 *G, int X, int Y, int XB, int ndf,
                           int ntilts, float *out, hipStream_t s);
// tw: a twiddle table (twiddle_host_table() values) to use instead of the
// device's own (twiddles()).
hipError_t launch_filter_rows(const float *src, size_t st, size_t sk,
                              size_t sy, int N, int L, int Y, int T, int ndf,
                              const float *ramp, const float *wscale,
                              const uint32_t *tmax, float nrm_mul, int store,
                              void *slab, int NYB, int YB, hipStream_t s,
                              const float2 *tw = nullptr,
                              const int2 *krange = nullptr);
hipError_t launch_absmax_rows(const float *src, size_t st, size_t sk,
                              size_t sy, int N, int Y, int T, int ndf,
                              uint32_t *tmax, hipStream_t s);

// store: kF32 (exact reference arithmetic) or kF16/kBF16 (FP32 accumulate).
hipError_t launch_bp(const BpArgs &a, int YB, int store, hipStream_t s);
// Reduced-storage coefficient table from the reference's (c, s, offset,
// scale) and the per-tilt normalisation of the filter pass.
hipError_t launch_bp_coefficients(const float4 *coef, const uint32_t *tmax,
                                  float nrm_mul, int T, float4 *out,
                                  hipStream_t s);
// Window table of the staged kernel: bp_window_entries(W, Z, T) int2, from
// the coefficient table the kernels read (exact != 0: kF32 arithmetic).
size_t bp_window_entries(int W, int Z, int T);
hipError_t launch_bp_windows(const float4 *coef, int exact, int T, int W,
                             int Z, float dscale, float doffset, int ndf,
                             int2 *win, hipStream_t s);

// Seeded synthetic tilt series (ellipsoid phantom, optional CTF, noise).
struct SynthArgs {
  size_t num_pixels, num_rows, num_angles, depth, width;
  float centre;
  const float *angles;  // host
  size_t num_ellipsoids;
  uint64_t seed;
  float semi_min, semi_max, density_min, density_max;
  int noise;  // 0 none, 1 Gaussian, 2 Poisson
  float noise_level, contrast;
  int ctf;
  double wavelength, Cs, defocus, phi, pixel_size;
};
hipError_t synth_projections(const SynthArgs *a, float *d_out, hipStream_t s);

// Twiddle table W[j] = exp(-2 pi i j / kTwN) for the current device
// (allocated on its first use there), and its host values.
const float2 *twiddles(hipError_t *err);
std::vector<float2> twiddle_host_table();

// ---- host tables shared by the C-ABI translation units (gx_capi.hip) ----

// Per-tilt constants {c, s, offset, scale} of reconstructor.cpp:29-47.
void tilt_geometry(size_t T, const float *angles, float centre, float *coef);
void slab_ranges(size_t T, const float *coef, int W, int Z, float dscale,
                 float doffset, int ndf, std::vector<int2> &out);
// Slab selection factors of reconstructor.cpp:29-32.
void defocus_coefficients(size_t ndf, float ps, float mn, float mx,
                          float *dscale, float *doffset);
// The reference's ramp filter (filter.cpp:30-45) for N samples on the
// L-point grid, with the 1/L of the unnormalised inverse: L/2 + 1 values.
std::vector<float> ramp_table(size_t N, size_t L);
// Filter transform length of N samples: a power of two >= 2N - 1, >= 8.
size_t filter_length(size_t N);
// Config checks of config.h:66-75 and the filter's size limit (GX_OK or a
// status with the message set).
int check_recon_cfg(const gx_recon_cfg *c);
// Sets the thread's error message; returns code.
int fail_status(int code, const char *msg);

}  // namespace gx
