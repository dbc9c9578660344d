/*
 * CPU oracle for the guanaco reconstruction hot path: a restatement of the
 * reference's CPU algorithm in plain C.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker. The product never links or
 * calls it.
 *
 * Build flags (oracle/Makefile) keep the reference's x86-64 arithmetic:
 * no FMA contraction, no fast-math, SSE single precision for float.
 *
 * Pins (tests/test_oracle.py): CTF closed forms from the reference's own
 * tests/test_ctf.py; FFT against numpy.fft and brute-force DFT; the ramp
 * filter against its closed-form spatial kernel; correction special cases;
 * backprojection known answers (single tilt, linear sinogram, hand-computed
 * slab indices, crop invariance).
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "guanaco_oracle.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

#define REAL float
#define SFX _f32
#include "oracle_body.inc"
#undef REAL
#undef SFX

#define REAL double
#define SFX _f64
#include "oracle_body.inc"
#undef REAL
#undef SFX

/* ---- CTF physics ---------------------------------------------------------- */

/* constants.h:28-36 */
static const double k_pi = M_PI;
static const double k_c = 2.99792458e8;
static const double k_e = 1.602176634e-19;
static const double k_m_e = 9.1093837015e-31;
static const double k_h = 6.62607004e-34;
static const double k_m_to_A = 1e10;

/* ctf.h:39-48 */
double or_electron_wavelength(double V) {
  double eV = k_e * V;
  return (k_h * k_c / sqrt(eV * eV + 2 * k_m_e * eV * k_c * k_c)) * k_m_to_A;
}

/* ctf.h:88-101 (size_t modular index, then double) */
void or_r_and_theta(size_t x, size_t y, size_t w, size_t h, double inv_ps,
                    double *r, double *theta) {
  size_t w2 = w / 2, h2 = h / 2;
  double dx = ((double)((x + w2) % w) - (double)w2) / (double)w;
  double dy = ((double)((y + h2) % h) - (double)h2) / (double)h;
  *r = sqrt((dx * dx + dy * dy) * inv_ps * inv_ps);
  *theta = atan2(dy, dx);
}

/* ctf.h:233-238 */
double or_chi(double l, double df, double Cs, double Ca, double Pa, double q,
              double theta) {
  double d = df + Ca * cos(2 * (theta - Pa));
  return k_pi * (Cs * l * l * l * q * q * q * q / 2 - l * d * q * q);
}

/* ctf.h:269-273 (std::exp of the complex (0, -(chi - phi)), which is
 * glibc's cexp: it differs from separate sin()/cos() calls by an ulp at some
 * arguments) and ctf.h:489-503 (grid loop, inv_ps = 1 / ps). */
void or_ctf_simple_grid(double l, double df, double Cs, double Ca, double Pa,
                        double phi, size_t w, size_t h, double ps,
                        double *out) {
  double inv_ps = 1.0 / ps;
  size_t i, j;
  for (j = 0; j < h; ++j) {
    for (i = 0; i < w; ++i) {
      double r, th, chi;
      double complex e;
      or_r_and_theta(i, j, w, h, inv_ps, &r, &th);
      chi = or_chi(l, df, Cs, Ca, Pa, r, th);
      e = cexp(CMPLX(0.0, -(chi - phi)));
      *out++ = creal(e);
      *out++ = cimag(e);
    }
  }
}

/* ctf.h:188-200 (spatial incoherence, De Graef eq. 10.53), 204-214 (temporal)
 * and 218-223 (scale), with u = 1 + 2 pi^2 theta_c^2 dd^2 q^2. */
double or_log_es(double l, double df, double Cs, double Ca, double Pa,
                 double dd, double theta_c, double q, double theta) {
  double d = df + Ca * cos(2 * (theta - Pa));
  double u = 1 + 2 * k_pi * k_pi * theta_c * theta_c * dd * dd * q * q;
  double v = Cs * l * l * l * q * q * q - d * l * q;
  return -(k_pi * k_pi * theta_c * theta_c * v * v / (l * l * u));
}

double or_log_et(double l, double dd, double theta_c, double q) {
  double u = 1 + 2 * k_pi * k_pi * theta_c * theta_c * dd * dd * q * q;
  return -(k_pi * k_pi * l * l * dd * dd * q * q * q * q / (2 * u));
}

double or_envelope_a(double dd, double theta_c, double q) {
  double u = 1 + 2 * k_pi * k_pi * theta_c * theta_c * dd * dd * q * q;
  return 1 / sqrt(u);
}

/* ctf.h:250-258 (A * exp(complex(log Es + log Et, -(chi - phi)))) over the
 * grid of ctf.h:466-478. */
void or_ctf_grid(double l, double df, double Cs, double Ca, double Pa,
                 double dd, double theta_c, double phi, size_t w, size_t h,
                 double ps, double *out) {
  double inv_ps = 1.0 / ps;
  size_t i, j;
  for (j = 0; j < h; ++j) {
    for (i = 0; i < w; ++i) {
      double r, th, chi, les, let, a;
      double complex e;
      or_r_and_theta(i, j, w, h, inv_ps, &r, &th);
      chi = or_chi(l, df, Cs, Ca, Pa, r, th);
      let = or_log_et(l, dd, theta_c, r);
      les = or_log_es(l, df, Cs, Ca, Pa, dd, theta_c, r, th);
      a = or_envelope_a(dd, theta_c, r);
      e = a * cexp(CMPLX(les + let, -(chi - phi)));
      *out++ = creal(e);
      *out++ = cimag(e);
    }
  }
}

/* ---- correction / filter exports ----------------------------------------- */

void or_correct_f32(const float *image, const float *ctf, float *rec,
                    size_t xsize, size_t ysize, size_t num_ctf) {
  correct_impl_f32(image, ctf, rec, xsize, ysize, num_ctf);
}

void or_correct_f64(const float *image, const float *ctf, double *rec,
                    size_t xsize, size_t ysize, size_t num_ctf) {
  correct_impl_f64(image, ctf, rec, xsize, ysize, num_ctf);
}

/* ---- backprojection ------------------------------------------------------- */

/* config.h:66-75 */
int or_config_is_valid(const or_config *c) {
  return c->num_pixels > 0 && c->num_angles > 0 && c->num_defocus > 0 &&
         c->grid_height > 0 && c->grid_width > 0 && c->pixel_size > 0 &&
         c->min_defocus <= c->max_defocus && c->angles != NULL;
}

/* Per-tilt constants, reconstructor.cpp:36-47, in the reference's float /
 * double mix. */
typedef struct {
  float c, s, offset, scale;
} tilt_coef;

static tilt_coef tilt_coefficients(const or_config *cfg, size_t angle) {
  tilt_coef tc;
  double output_scale = k_pi / (double)(2 * cfg->num_angles);
  float centre = cfg->centre;
  float theta = cfg->angles[angle];
  float cos_angle = cosf(theta);
  float sin_angle = sinf(theta);
  float det_x0 = -centre * cos_angle;
  float det_y0 = -centre * sin_angle;
  float ray_length = sqrtf(cos_angle * cos_angle + sin_angle * sin_angle);
  float d = cos_angle * cos_angle + sin_angle * sin_angle;
  tc.offset = (det_y0 * (-sin_angle) - det_x0 * cos_angle) / d;
  tc.scale = (float)(output_scale * (double)ray_length / (double)fabsf(d));
  tc.c = cos_angle / d;
  tc.s = sin_angle / d;
  return tc;
}

void or_tilt_coefficients(const or_config *cfg, float *coef) {
  size_t t;
  for (t = 0; t < cfg->num_angles; ++t) {
    tilt_coef tc = tilt_coefficients(cfg, t);
    coef[4 * t + 0] = tc.c;
    coef[4 * t + 1] = tc.s;
    coef[4 * t + 2] = tc.offset;
    coef[4 * t + 3] = tc.scale;
  }
}

/* reconstructor.cpp:29-32: dscale = ndf * ps / (max - min) (float),
 * doffset = -dscale * (min / ps). */
static void defocus_coefficients(const or_config *cfg, float *dscale,
                                 float *doffset) {
  float ds = cfg->num_defocus > 1
                 ? (float)cfg->num_defocus * cfg->pixel_size /
                       (cfg->max_defocus - cfg->min_defocus)
                 : 0.0f;
  *dscale = ds;
  *doffset = -ds * (cfg->min_defocus / cfg->pixel_size);
}

/* The FP32 coordinate arithmetic of reconstructor.cpp:56-79 for one voxel
 * and tilt: slab index, left tap index, weight t and validity. */
static int f32_decision(const or_config *cfg, const tilt_coef *tc,
                        float dscale, float doffset, size_t i, size_t j,
                        int *k_out, int *ind_out, float *t_out) {
  size_t W = cfg->grid_width, H = cfg->grid_height;
  size_t N = cfg->num_pixels, ndf = cfg->num_defocus;
  float x = (float)((double)i - 0.5 * (double)W + 0.5);
  float y = (float)((double)j - 0.5 * (double)H + 0.5);
  float pixel = x * tc->c - y * tc->s + tc->offset;
  int k = 0, ind;
  float t;
  if (ndf > 1) {
    float height = -x * tc->s - y * tc->c;
    float defocus = height * dscale + doffset;
    k = (int)floor((double)defocus + 0.5);
    if (k < 0) k = 0;
    if (k > (int)ndf - 1) k = (int)ndf - 1;
  }
  pixel = (float)((double)pixel - 0.5);
  ind = (int)floorf(pixel);
  t = pixel - (float)ind;
  *k_out = k;
  *ind_out = ind;
  *t_out = t;
  return ind >= 0 && (size_t)ind < N - 1;
}

/* reconstructor.cpp:14-90 */
void or_project_f32(const or_config *cfg, const float *sino, float *rec) {
  size_t W = cfg->grid_width, H = cfg->grid_height;
  size_t N = cfg->num_pixels, T = cfg->num_angles;
  size_t angle, i, j;
  float dscale, doffset;
  defocus_coefficients(cfg, &dscale, &doffset);
  for (angle = 0; angle < T; ++angle) {
    tilt_coef tc = tilt_coefficients(cfg, angle);
    for (j = 0; j < H; ++j) {
      for (i = 0; i < W; ++i) {
        int k, ind;
        float t;
        if (f32_decision(cfg, &tc, dscale, doffset, i, j, &k, &ind, &t)) {
          const float *row = sino + angle * N + (size_t)k * N * T;
          float v0 = row[ind];
          float v1 = row[ind + 1];
          float value = (1 - t) * v0 + t * v1;
          rec[i + j * W] += tc.scale * value;
        }
      }
    }
  }
}

/* FP64 twin of reconstructor.cpp:14-90. */
void or_project_f64(const or_config *cfg, const double *sino, double *rec,
                    int f32_decisions) {
  size_t W = cfg->grid_width, H = cfg->grid_height;
  size_t N = cfg->num_pixels, T = cfg->num_angles, ndf = cfg->num_defocus;
  size_t angle, i, j;
  float dscale_f, doffset_f;
  double dscale = ndf > 1 ? (double)ndf * (double)cfg->pixel_size /
                                ((double)cfg->max_defocus -
                                 (double)cfg->min_defocus)
                          : 0.0;
  double doffset = -dscale * ((double)cfg->min_defocus /
                              (double)cfg->pixel_size);
  double output_scale = k_pi / (double)(2 * T);
  defocus_coefficients(cfg, &dscale_f, &doffset_f);
  for (angle = 0; angle < T; ++angle) {
    tilt_coef tc = tilt_coefficients(cfg, angle);
    double theta = (double)cfg->angles[angle];
    double c = cos(theta), s = sin(theta);
    double d = c * c + s * s;
    double offset = ((-cfg->centre * s) * (-s) - (-cfg->centre * c) * c) / d;
    double scale = output_scale * sqrt(d) / fabs(d);
    c /= d;
    s /= d;
    for (j = 0; j < H; ++j) {
      for (i = 0; i < W; ++i) {
        double x = (double)i - 0.5 * (double)W + 0.5;
        double y = (double)j - 0.5 * (double)H + 0.5;
        double pixel = x * c - y * s + offset - 0.5;
        int k = 0, ind, valid;
        double t;
        if (f32_decisions) {
          float tf;
          valid = f32_decision(cfg, &tc, dscale_f, doffset_f, i, j, &k, &ind,
                               &tf);
        } else {
          if (ndf > 1) {
            double height = -x * s - y * c;
            k = (int)floor(height * dscale + doffset + 0.5);
            if (k < 0) k = 0;
            if (k > (int)ndf - 1) k = (int)ndf - 1;
          }
          ind = (int)floor(pixel);
          valid = ind >= 0 && (size_t)ind < N - 1;
        }
        if (valid) {
          const double *row = sino + angle * N + (size_t)k * N * T;
          t = pixel - (double)ind;
          rec[i + j * W] += scale * ((1 - t) * row[ind] + t * row[ind + 1]);
        }
      }
    }
  }
}

/* reconstructor.cpp:92-115 */
int or_recon_f32(const or_config *cfg, const float *sino, float *rec) {
  size_t n = cfg->num_pixels * cfg->num_angles * cfg->num_defocus;
  float *f;
  if (!or_config_is_valid(cfg) || sino == NULL || rec == NULL) return -1;
  f = (float *)malloc(sizeof(float) * n);
  memcpy(f, sino, sizeof(float) * n);
  or_filter_f32(f, cfg->num_pixels, cfg->num_angles, cfg->num_defocus);
  memset(rec, 0, sizeof(float) * cfg->grid_height * cfg->grid_width);
  or_project_f32(cfg, f, rec);
  free(f);
  return 0;
}

int or_recon_f64(const or_config *cfg, const double *sino, double *rec,
                 int f32_decisions) {
  size_t n = cfg->num_pixels * cfg->num_angles * cfg->num_defocus;
  double *f;
  if (!or_config_is_valid(cfg) || sino == NULL || rec == NULL) return -1;
  f = (double *)malloc(sizeof(double) * n);
  memcpy(f, sino, sizeof(double) * n);
  or_filter_f64(f, cfg->num_pixels, cfg->num_angles, cfg->num_defocus);
  memset(rec, 0, sizeof(double) * cfg->grid_height * cfg->grid_width);
  or_project_f64(cfg, f, rec, f32_decisions);
  free(f);
  return 0;
}
