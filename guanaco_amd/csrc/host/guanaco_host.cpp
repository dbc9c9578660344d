// C++ host layer of guanaco_amd: the reference's libguanaco interface over
// the C ABI of libguanaco_hip.so.
#include "guanaco_host.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

#include "guanaco_hip.h"

namespace guanaco {

void check_status(int status) {
  if (status != GX_OK) throw Error(std::string(gx_last_error()));
}

namespace {

[[noreturn]] void host_device_unsupported() {
  throw Error(
      "guanaco_amd implements the GPU device (device='gpu' / e_device); the "
      "CPU device is the reference's own FFTW implementation");
}

// Owning device buffer over the C ABI.
struct Dev {
  void *p = nullptr;
  explicit Dev(std::size_t bytes) { check_status(gx_malloc(&p, bytes)); }
  Dev(const Dev &) = delete;
  Dev &operator=(const Dev &) = delete;
  ~Dev() {
    if (p) gx_free(p);
  }
};

struct Plan {
  gx_volume_plan *p = nullptr;
  explicit Plan(const gx_volume_desc &d) {
    check_status(gx_volume_plan_create(&d, &p));
  }
  Plan(const Plan &) = delete;
  Plan &operator=(const Plan &) = delete;
  ~Plan() {
    if (p) gx_volume_plan_destroy(p);
  }
};

}  // namespace

// ---- Reconstructor (reconstructor.cpp:117-132) ------------------------------

Reconstructor::Reconstructor(const Config &config) : config_(config) {}

void Reconstructor::operator()(const float *sinogram,
                               float *reconstruction) const {
  GUANACO_ASSERT(sinogram != nullptr);
  GUANACO_ASSERT(reconstruction != nullptr);
  GUANACO_ASSERT(config_.is_valid());
  if (config_.device != e_device) host_device_unsupported();
  gx_recon_cfg c{};
  c.num_pixels = config_.num_pixels;
  c.num_angles = config_.num_angles;
  c.num_defocus = config_.num_defocus;
  c.grid_height = config_.grid_height;
  c.grid_width = config_.grid_width;
  c.pixel_size = config_.pixel_size;
  c.centre = config_.centre;
  c.min_defocus = config_.min_defocus;
  c.max_defocus = config_.max_defocus;
  c.angles = config_.angles.data();
  c.gpu_index = config_.gpu_index;
  check_status(gx_recon(&c, sinogram, reconstruction));
}

// ---- correct (correct.h:77-95) -----------------------------------------------

void correct(const float *image, const std::complex<float> *ctf, float *rec,
             std::size_t xsize, std::size_t ysize, std::size_t num_ctf,
             eDevice device, int gpu_index) {
  if (device != e_device) host_device_unsupported();
  check_status(gx_corr(image, reinterpret_cast<const float *>(ctf), rec, xsize,
                       ysize, num_ctf, gpu_index));
}

// ---- CTF physics (ctf.h), FP64 ----------------------------------------------

namespace {
// constants.h:28-36 (h is the CODATA 2014 value the reference uses)
const double kPi = M_PI;
const double kC = 2.99792458e8;
const double kE = 1.602176634e-19;
const double kMe = 9.1093837015e-31;
const double kH = 6.62607004e-34;
const double kMToA = 1e10;
}  // namespace

double get_electron_wavelength(double V) {
  const double eV = kE * V;
  return (kH * kC / std::sqrt(eV * eV + 2 * kMe * eV * kC * kC)) * kMToA;
}

void get_r_and_theta(std::size_t x, std::size_t y, std::size_t w,
                     std::size_t h, double inv_ps, double *r, double *theta) {
  const std::size_t w2 = w / 2, h2 = h / 2;
  const double dx = ((double)((x + w2) % w) - (double)w2) / (double)w;
  const double dy = ((double)((y + h2) % h) - (double)h2) / (double)h;
  *r = std::sqrt((dx * dx + dy * dy) * inv_ps * inv_ps);
  *theta = std::atan2(dy, dx);
}

double get_chi(const CTF &c, double q, double theta) {
  const double df = c.df + c.Ca * std::cos(2 * (theta - c.Pa));
  return kPi * (c.Cs * c.l * c.l * c.l * q * q * q * q / 2 - c.l * df * q * q);
}

std::complex<double> get_ctf_simple(const CTF &c, double q, double theta) {
  // the complex exponential (glibc cexp), not separate sin / cos calls:
  // they differ by an ulp at some arguments
  return std::exp(std::complex<double>(0, -(get_chi(c, q, theta) - c.phi)));
}

void get_ctf_n_simple(const CTF &c, std::complex<double> *ctf, std::size_t w,
                      std::size_t h, double ps) {
  GUANACO_ASSERT(ps > 0);
  const double inv_ps = 1.0 / ps;
  for (std::size_t j = 0; j < h; ++j) {
    for (std::size_t i = 0; i < w; ++i) {
      double r, theta;
      get_r_and_theta(i, j, w, h, inv_ps, &r, &theta);
      *ctf++ = get_ctf_simple(c, r, theta);
    }
  }
}

double get_defocus_spread(double Cc, double dEE, double dII, double dVV) {
  return Cc * std::sqrt(dEE * dEE + (2 * dII) * (2 * dII) + dVV * dVV);
}

double q_to_Q(double q, double l, double Cs) {
  return q * std::pow(Cs * l * l * l, 1.0 / 4.0);
}

double Q_to_q(double Q, double l, double Cs) {
  return Q / std::pow(Cs * l * l * l, 1.0 / 4.0);
}

double df_to_D(double df, double l, double Cs) { return df / std::sqrt(Cs * l); }

double D_to_df(double D, double l, double Cs) { return D * std::sqrt(Cs * l); }

namespace {
// 1 + 2 pi^2 theta_c^2 dd^2 q^2, shared by the envelopes and their scale
double envelope_u(const CTF &c, double q) {
  return 1 + 2 * kPi * kPi * c.theta_c * c.theta_c * c.dd * c.dd * q * q;
}
}  // namespace

double get_log_Es(const CTF &c, double q, double theta) {
  const double df = c.df + c.Ca * std::cos(2 * (theta - c.Pa));
  const double u = envelope_u(c, q);
  const double v = c.Cs * c.l * c.l * c.l * q * q * q - df * c.l * q;
  return -(kPi * kPi * c.theta_c * c.theta_c * v * v / (c.l * c.l * u));
}

double get_log_Et(const CTF &c, double q) {
  const double u = envelope_u(c, q);
  return -(kPi * kPi * c.l * c.l * c.dd * c.dd * q * q * q * q / (2 * u));
}

double get_A(const CTF &c, double q) { return 1 / std::sqrt(envelope_u(c, q)); }

std::complex<double> get_ctf(const CTF &c, double q, double theta) {
  const double chi = get_chi(c, q, theta);
  const double e = get_log_Es(c, q, theta) + get_log_Et(c, q);
  return get_A(c, q) * std::exp(std::complex<double>(e, -(chi - c.phi)));
}

double get_ctf_simple_real(const CTF &c, double q, double theta) {
  return std::cos(-(get_chi(c, q, theta) - c.phi));
}

double get_ctf_simple_imag(const CTF &c, double q, double theta) {
  return std::sin(-(get_chi(c, q, theta) - c.phi));
}

namespace {
template <typename F>
void for_grid(std::size_t w, std::size_t h, double ps, F f) {
  GUANACO_ASSERT(ps > 0);
  const double inv_ps = 1.0 / ps;
  for (std::size_t j = 0; j < h; ++j) {
    for (std::size_t i = 0; i < w; ++i) {
      double r, theta;
      get_r_and_theta(i, j, w, h, inv_ps, &r, &theta);
      f(r, theta);
    }
  }
}
}  // namespace

void get_ctf_n(const CTF &c, std::complex<double> *ctf, std::size_t w,
               std::size_t h, double ps) {
  for_grid(w, h, ps, [&](double r, double t) { *ctf++ = get_ctf(c, r, t); });
}

void get_ctf_n_simple_real(const CTF &c, double *ctf, std::size_t w,
                           std::size_t h, double ps) {
  for_grid(w, h, ps,
           [&](double r, double t) { *ctf++ = get_ctf_simple_real(c, r, t); });
}

void get_ctf_n_simple_imag(const CTF &c, double *ctf, std::size_t w,
                           std::size_t h, double ps) {
  for_grid(w, h, ps,
           [&](double r, double t) { *ctf++ = get_ctf_simple_imag(c, r, t); });
}

// ---- whole-volume reconstruction --------------------------------------------

namespace {

// A descriptor of rows [y0, y0 + rows) of the request.
gx_volume_desc make_desc(const VolumeRequest &r, std::size_t rows,
                         std::size_t y0 = 0) {
  gx_volume_desc d{};
  d.num_pixels = r.num_pixels;
  d.num_rows = rows;
  d.num_angles = r.num_angles;
  d.num_defocus = r.num_defocus;
  d.depth = r.depth;
  d.width = r.width;
  d.pixel_size = r.pixel_size;
  d.centre = r.centre;
  d.min_defocus = r.min_defocus;
  d.max_defocus = r.max_defocus;
  d.angles = r.angles.data();
  d.weights = r.weights.empty() ? nullptr : r.weights.data();
  d.storage = (int)r.storage;
  d.input_layout = (int)r.layout;
  d.defocus = r.defocus.empty() ? nullptr : r.defocus.data();
  d.wavelength = r.wavelength;
  d.Cs = r.Cs;
  d.Ca = r.Ca;
  d.Pa = r.Pa;
  d.phi = r.phi;
  d.tilt_chunk = 0;
  d.gpu_index = r.gpu_index;
  d.centres = r.centres.empty() ? nullptr : r.centres.data() + y0;
  return d;
}

void validate(const VolumeRequest &r) {
  GUANACO_ASSERT(r.num_pixels > 0);
  GUANACO_ASSERT(r.num_rows > 0);
  GUANACO_ASSERT(r.num_angles > 0);
  GUANACO_ASSERT(r.num_defocus > 0);
  GUANACO_ASSERT(r.depth > 0);
  GUANACO_ASSERT(r.width > 0);
  GUANACO_ASSERT(r.angles.size() == r.num_angles);
  GUANACO_ASSERT(r.weights.empty() || r.weights.size() == r.num_angles);
  GUANACO_ASSERT(r.defocus.empty() || r.defocus.size() == r.num_defocus);
  GUANACO_ASSERT(r.centres.empty() || r.centres.size() == r.num_rows);
}

// Copy detector rows [y0, y1) of the host input into a contiguous band in
// the same layout.
void gather_rows(const VolumeRequest &r, const float *in, std::size_t y0,
                 std::size_t y1, std::vector<float> &band) {
  const std::size_t X = r.num_pixels, Y = r.num_rows, T = r.num_angles;
  const std::size_t K = r.num_defocus, n = y1 - y0;
  band.resize(T * K * n * X);
  if (r.layout == Layout::sinograms) {  // (Y, K, T, X): contiguous
    std::memcpy(band.data(), in + y0 * K * T * X, sizeof(float) * n * K * T * X);
    return;
  }
  // (T, [K,] Y, X)
  for (std::size_t tk = 0; tk < T * K; ++tk)
    std::memcpy(band.data() + tk * n * X, in + (tk * Y + y0) * X,
                sizeof(float) * n * X);
}

}  // namespace

std::size_t volume_input_bytes(const VolumeRequest &r) {
  validate(r);
  const bool ctf = r.layout == Layout::projections && !r.defocus.empty();
  // the device corrects (T, Y, X) projections into num_defocus slabs;
  // corrected and sinogram input holds every slab
  const std::size_t planes = ctf ? r.num_angles : r.num_angles * r.num_defocus;
  return planes * r.num_rows * r.num_pixels * sizeof(float);
}

std::size_t volume_plan_bytes(const VolumeRequest &r) {
  validate(r);
  gx_volume_desc d = make_desc(r, r.num_rows);
  std::size_t n = 0;
  check_status(gx_volume_plan_bytes(&d, &n));
  return n;
}

void reconstruct_volume(const VolumeRequest &r, const float *input,
                        float *volume) {
  validate(r);
  GUANACO_ASSERT(input != nullptr && volume != nullptr);
  if (r.gpu_index >= 0) check_status(gx_set_device(r.gpu_index));
  const std::size_t X = r.num_pixels, Y = r.num_rows, T = r.num_angles;
  const std::size_t K = r.num_defocus, Z = r.depth, W = r.width;
  const bool ctf = r.layout == Layout::projections && !r.defocus.empty();
  std::size_t free_b = 0, total_b = 0;
  check_status(gx_mem_get_info(&free_b, &total_b));
  const std::size_t budget = free_b / 10 * 9;
  // input per detector row
  const std::size_t in_row = volume_input_bytes(r) / Y;
  const std::size_t vol_row = Z * W * sizeof(float);  // volume per row

  if (ctf) {
    // The 2D correction needs whole projections: they stay on the device,
    // and one plan holds the filtered slabs of as many rows as fit (every
    // row, else bands of about Y / n rows, multiples of 32); each band is
    // backprojected into volume bands that fit beside it.
    const std::size_t in_b = Y * in_row;
    GUANACO_ASSERT(in_b < budget);
    gx_volume_desc d = make_desc(r, Y);
    std::size_t band = Y, plan_b = 0;
    for (std::size_t n = 1;; ++n) {
      band = std::min<std::size_t>(Y, ((Y + n - 1) / n + 31) / 32 * 32);
      d.band_rows = band == Y ? 0 : band;
      check_status(gx_volume_plan_bytes(&d, &plan_b));
      if (plan_b + in_b + std::min<std::size_t>(band, 32) * vol_row <= budget)
        break;
      GUANACO_ASSERT(band > 32);
    }
    Plan plan(d);
    Dev d_in(in_b);
    check_status(gx_memcpy_htod(d_in.p, input, in_b, nullptr));
    std::size_t yb = 0;
    check_status(gx_volume_rows_per_block(plan.p, &yb));
    // volume rows per copy-out: the whole band, else multiples of 32 rows
    std::size_t vband = (budget - plan_b - in_b) / vol_row;
    vband = vband >= band ? band : vband / 32 * 32;
    GUANACO_ASSERT(vband > 0 && (vband % yb == 0 || vband == Y));
    Dev d_vol(vband * vol_row);
    for (std::size_t s0 = 0; s0 < Y; s0 += band) {
      const std::size_t s1 = std::min(Y, s0 + band);
      check_status(gx_volume_filter_rows(plan.p, (const float *)d_in.p, s0,
                                         s1, nullptr));
      for (std::size_t y0 = s0; y0 < s1; y0 += vband) {
        const std::size_t y1 = std::min(s1, y0 + vband);
        check_status(gx_volume_backproject_rows(plan.p, y0, y1,
                                                (float *)d_vol.p, nullptr));
        check_status(gx_memcpy_dtoh(volume + y0 * Z * W, d_vol.p,
                                    (y1 - y0) * vol_row, nullptr));
      }
    }
    return;
  }

  // Row-local path: bands of detector rows, each a complete run.
  VolumeRequest one = r;
  one.num_rows = 1;
  if (!one.centres.empty()) one.centres.resize(1);
  std::size_t per_row_plan = volume_plan_bytes(one);
  std::size_t band = budget / (per_row_plan + in_row + vol_row + 1);
  band = std::max<std::size_t>(1, std::min(Y, band));
  if (band < Y && band >= 16) band = band / 16 * 16;
  std::vector<float> staging;
  for (std::size_t y0 = 0; y0 < Y; y0 += band) {
    const std::size_t y1 = std::min(Y, y0 + band), n = y1 - y0;
    gx_volume_desc d = make_desc(r, n, y0);
    Plan plan(d);
    const float *src = input;
    if (n != Y) {
      gather_rows(r, input, y0, y1, staging);
      src = staging.data();
    }
    Dev d_in(n * in_row), d_vol(n * vol_row);
    check_status(gx_memcpy_htod(d_in.p, src, n * in_row, nullptr));
    check_status(gx_volume_run(plan.p, (const float *)d_in.p, (float *)d_vol.p,
                               nullptr));
    check_status(gx_memcpy_dtoh(volume + y0 * Z * W, d_vol.p, n * vol_row,
                                nullptr));
  }
}

}  // namespace guanaco
