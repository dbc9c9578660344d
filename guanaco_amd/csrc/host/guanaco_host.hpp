// C++ host interface of guanaco_amd, mirroring the reference's libguanaco
// (include/guanaco/{config,constants,ctf,correct,reconstructor,error}.h) so
// that the pybind11 layer reads like the reference's. Device work goes
// through the C ABI of libguanaco_hip.so (include/guanaco_hip.h); this file
// has no HIP dependency.
#pragma once

#include <complex>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace guanaco {

// constants.h:41-44
enum eDevice { e_host = 0, e_device = 1 };

// error.h:31-39: "file:line message"
class Error : public std::runtime_error {
 public:
  explicit Error(const std::string &what) : std::runtime_error(what) {}
  Error(const std::string &file, std::size_t line, const std::string &msg)
      : std::runtime_error(file + ":" + std::to_string(line) + " " + msg) {}
};

#define GUANACO_ASSERT(assertion)                                            \
  if (!(assertion)) {                                                        \
    throw guanaco::Error(__FILE__, __LINE__, "ASSERT (" #assertion ") failed"); \
  }

// Throws guanaco::Error with the C ABI's message for a non-zero status.
void check_status(int status);

// config.h:29-76
struct Config {
  using size_type = std::size_t;
  eDevice device = e_host;
  int gpu_index = 0;
  size_type num_pixels = 0;
  size_type num_angles = 0;
  size_type num_defocus = 0;
  size_type grid_height = 0;
  size_type grid_width = 0;
  float pixel_size = 1;
  float centre = 0;
  std::vector<float> angles;
  float min_defocus = 0;
  float max_defocus = 0;

  size_type sino_size() const {
    return num_pixels * num_angles * num_defocus;
  }
  size_type grid_size() const { return grid_height * grid_width; }
  bool is_valid() const {
    return (device == e_host || device == e_device) && num_pixels > 0 &&
           num_angles > 0 && num_defocus > 0 && grid_height > 0 &&
           grid_width > 0 && pixel_size > 0 && min_defocus <= max_defocus &&
           angles.size() == num_angles;
  }
};

// reconstructor.h:41-55. e_device runs the gfx950 path; e_host is the
// reference's own CPU implementation and is not part of this library.
class Reconstructor {
 public:
  explicit Reconstructor(const Config &config);
  void operator()(const float *sinogram, float *reconstruction) const;

 private:
  Config config_;
};

inline Reconstructor make_reconstructor(const Config &config) {
  return Reconstructor(config);
}

// correct.h:77-95
void correct(const float *image, const std::complex<float> *ctf, float *rec,
             std::size_t xsize, std::size_t ysize, std::size_t num_ctf,
             eDevice device, int gpu_index = -1);

// ---- CTF physics, FP64 (ctf.h) ---------------------------------------------

// ctf.h:167-178
struct CTF {
  double l = 0, df = 0, Cs = 0, Ca = 0, Pa = 0, dd = 0, theta_c = 0, phi = 0;
};

double get_electron_wavelength(double V);              // ctf.h:39-48
void get_r_and_theta(std::size_t x, std::size_t y, std::size_t w,
                     std::size_t h, double inv_ps, double *r,
                     double *theta);                    // ctf.h:88-101
double get_chi(const CTF &c, double q, double theta);  // ctf.h:233-238
std::complex<double> get_ctf_simple(const CTF &c, double q,
                                    double theta);      // ctf.h:269-273
void get_ctf_n_simple(const CTF &c, std::complex<double> *ctf, std::size_t w,
                      std::size_t h, double ps);        // ctf.h:489-503
// ctf.h:61-64: Cc * sqrt(dEE^2 + (2 dII)^2 + dVV^2)
double get_defocus_spread(double Cc, double dEE, double dII, double dVV);
// ctf.h:114-162: dimensionless frequency and defocus (De Graef 10.10-10.11)
double q_to_Q(double q, double l, double Cs);
double Q_to_q(double Q, double l, double Cs);
double df_to_D(double df, double l, double Cs);
double D_to_df(double D, double l, double Cs);
// ctf.h:188-223, 250-258: log envelopes (De Graef 10.53), scale and the full
// CTF A exp(log Es + log Et) exp(-i (chi - phi))
double get_log_Es(const CTF &c, double q, double theta);
double get_log_Et(const CTF &c, double q);
double get_A(const CTF &c, double q);
std::complex<double> get_ctf(const CTF &c, double q, double theta);
double get_ctf_simple_real(const CTF &c, double q, double theta);  // :285-288
double get_ctf_simple_imag(const CTF &c, double q, double theta);  // :300-303
// ctf.h:466-478, 516-541, 543-561: the FFT-ordered w x h grids
void get_ctf_n(const CTF &c, std::complex<double> *ctf, std::size_t w,
               std::size_t h, double ps);
void get_ctf_n_simple_real(const CTF &c, double *ctf, std::size_t w,
                           std::size_t h, double ps);
void get_ctf_n_simple_imag(const CTF &c, double *ctf, std::size_t w,
                           std::size_t h, double ps);

// ---- whole-volume reconstruction (no reference counterpart) ----------------

enum class Layout { projections = 0, corrected = 1, sinograms = 2 };
enum class Storage { f32 = 0, f16 = 1, bf16 = 2 };

struct VolumeRequest {
  std::size_t num_pixels = 0, num_rows = 0, num_angles = 0, num_defocus = 1;
  std::size_t depth = 0, width = 0;
  float pixel_size = 1, centre = 0, min_defocus = 0, max_defocus = 0;
  std::vector<float> angles;
  std::vector<float> weights;   // empty: ones
  Layout layout = Layout::projections;
  Storage storage = Storage::f32;
  std::vector<double> defocus;  // empty: no CTF correction on device
  std::vector<float> centres;   // per-row centres; empty: `centre` for all
  double wavelength = 0, Cs = 0, Ca = 0, Pa = 0, phi = 0;
  int gpu_index = -1;
};

// Bytes of host input the request reads: T x Y x X floats of projections
// corrected on the device, T x num_defocus x Y x X of corrected or sinogram
// input.
std::size_t volume_input_bytes(const VolumeRequest &r);

// Device bytes a plan for this request holds (admission footprint).
std::size_t volume_plan_bytes(const VolumeRequest &r);

// Host-to-host whole-volume reconstruction: input in `layout` (host), volume
// (num_rows x depth x width, host) overwritten. Splits the volume into bands
// of detector rows when it does not fit the device; each band is one
// device-resident run.
void reconstruct_volume(const VolumeRequest &r, const float *input,
                        float *volume);

}  // namespace guanaco
