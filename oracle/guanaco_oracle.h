/*
 * CPU oracle for the guanaco reconstruction hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This library restates the reference's CPU
 * algorithm (FFTW + Reconstructor_t<e_host>) so that the HIP product can be
 * checked against it. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it. The product never links it.
 *
 * Every function cites the reference file:line it follows. FFTW is absent
 * from this image, so the FFT here is a project-written mixed-radix FFT with
 * FFTW's conventions (forward e^{-2 pi i jk/n}, backward e^{+...}, both
 * unnormalised; reference src/libguanaco/fft.cpp:38-71).
 */
#ifndef GUANACO_ORACLE_H
#define GUANACO_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- CTF physics, FP64 (reference include/guanaco/ctf.h) ---------------- */

/* ctf.h:39-48 with constants.h:28-36 (h = 6.62607004e-34, CODATA 2014). */
double or_electron_wavelength(double V);

/* ctf.h:88-101: frequency radius and polar angle of pixel (x, y) of a
 * w x h grid in FFT order. */
void or_r_and_theta(size_t x, size_t y, size_t w, size_t h, double inv_ps,
                    double *r, double *theta);

/* ctf.h:233-238 (De Graef eq. 10.9). */
double or_chi(double l, double df, double Cs, double Ca, double Pa, double q,
              double theta);

/* ctf.h:269-273 + 489-503: exp(-i(chi - phi)) over a w x h grid, written as
 * interleaved (re, im) doubles, row-major (h rows of w). */
void or_ctf_simple_grid(double l, double df, double Cs, double Ca, double Pa,
                        double phi, size_t w, size_t h, double ps,
                        double *out);

/* ctf.h:188-223: log spatial / temporal incoherence envelopes and the
 * envelope scale A (De Graef eq. 10.53). */
double or_log_es(double l, double df, double Cs, double Ca, double Pa,
                 double dd, double theta_c, double q, double theta);
double or_log_et(double l, double dd, double theta_c, double q);
double or_envelope_a(double dd, double theta_c, double q);

/* ctf.h:250-258 + 466-478: A exp(log Es + log Et) exp(-i(chi - phi)) over a
 * w x h grid, interleaved (re, im) doubles. */
void or_ctf_grid(double l, double df, double Cs, double Ca, double Pa,
                 double dd, double theta_c, double phi, size_t w, size_t h,
                 double ps, double *out);

/* ---- FFT (stand-in for FFTW, fft.cpp:38-71) ------------------------------ */

/* In-place complex FFT of n interleaved (re, im) values. sign = -1 forward,
 * +1 backward; unnormalised. */
void or_fft_f32(size_t n, float *data, int sign);
void or_fft_f64(size_t n, double *data, int sign);

/* In-place 2D complex FFT of a row-major ny x nx array (fftw_plan_dft_2d
 * ordering, fft.cpp:49-59). */
void or_fft2_f32(size_t nx, size_t ny, float *data, int sign);
void or_fft2_f64(size_t nx, size_t ny, double *data, int sign);

/* ---- CTF phase-flip correction (correct.cpp:30-90, correct.h:30-39) ------ */

/* image: ysize x xsize float; ctf: num_ctf x ysize x xsize complex64
 * (interleaved float pairs); rec: num_ctf x ysize x xsize float, fully
 * overwritten with Re(IFFT2_unnorm(FFT2(image) * -sign(Im ctf))). */
void or_correct_f32(const float *image, const float *ctf, float *rec,
                    size_t xsize, size_t ysize, size_t num_ctf);
/* FP64 twin: same algorithm in double; the flip sign is still read from the
 * complex64 CTF, as the reference does. */
void or_correct_f64(const float *image, const float *ctf, double *rec,
                    size_t xsize, size_t ysize, size_t num_ctf);

/* ---- Ramp filter (filter.cpp:30-97) -------------------------------------- */

/* In place over num_defocus * num_angles rows of num_pixels samples. */
void or_filter_f32(float *data, size_t num_pixels, size_t num_angles,
                   size_t num_defocus);
void or_filter_f64(double *data, size_t num_pixels, size_t num_angles,
                   size_t num_defocus);

/* ---- Backprojection (reconstructor.cpp:14-115) --------------------------- */

typedef struct {
  size_t num_pixels;   /* N */
  size_t num_angles;   /* T */
  size_t num_defocus;  /* ndf */
  size_t grid_height;  /* H (depth) */
  size_t grid_width;   /* W */
  float pixel_size;
  float centre;
  float min_defocus;
  float max_defocus;
  const float *angles; /* T, radians */
} or_config;

/* config.h:66-75 */
int or_config_is_valid(const or_config *cfg);

/* reconstructor.cpp:36-47: per-tilt {cos/d, sin/d, offset, scale}. */
void or_tilt_coefficients(const or_config *cfg, float *coef);

/* reconstructor.cpp:14-90: rec (H x W) += BP of the filtered sinogram
 * (ndf x T x N). rec is NOT cleared here, matching project(). */
void or_project_f32(const or_config *cfg, const float *sino, float *rec);

/* FP64 twin of project(). When f32_decisions is non-zero, the discrete
 * choices (slab index, left tap index, edge-tap validity) are taken from the
 * FP32 coordinate arithmetic and only the interpolation weight and the
 * accumulation are FP64; otherwise everything is FP64. */
void or_project_f64(const or_config *cfg, const double *sino, double *rec,
                    int f32_decisions);

/* reconstructor.cpp:92-115: copy sino, filter, zero rec, project.
 * Returns 0, or -1 for an invalid configuration (the reference throws). */
int or_recon_f32(const or_config *cfg, const float *sino, float *rec);
int or_recon_f64(const or_config *cfg, const double *sino, double *rec,
                 int f32_decisions);

#ifdef __cplusplus
}
#endif

#endif
