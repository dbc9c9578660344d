// Probe that compiles the reference's header-only CTF math where it lies
// (include/guanaco/ctf.h, include/guanaco/correct.h) and exports it through
// a C ABI, so tests can pin the oracle's restatement against the reference's
// own code. TEST INFRASTRUCTURE ONLY; built by `make ref` into oracle/_ref/.
//
// The rest of the reference path (reconstructor.cpp, filter.cpp, correct.cpp)
// needs FFTW, which this image lacks, so it is not built.
#include <complex>
#include <cstddef>
#include <guanaco/correct.h>
#include <guanaco/ctf.h>

extern "C" {

double ref_electron_wavelength(double V) {
  return guanaco::get_electron_wavelength<double>(V);
}

void ref_r_and_theta(size_t x, size_t y, size_t w, size_t h, double inv_ps,
                     double *r, double *theta) {
  auto p = guanaco::get_r_and_theta<double>(x, y, w, h, inv_ps);
  *r = p.r;
  *theta = p.theta;
}

double ref_chi(double l, double df, double Cs, double Ca, double Pa, double q,
               double theta) {
  guanaco::CTF c;
  c.l = l;
  c.df = df;
  c.Cs = Cs;
  c.Ca = Ca;
  c.Pa = Pa;
  return guanaco::get_chi<double>(c, q, theta);
}

// get_ctf_n_simple(c, ctf, w, h, ps) as called by the pybind wrapper
// get_ctf_simple_wh<double> (ext/ctf.h:141-149).
void ref_ctf_simple_grid(double l, double df, double Cs, double Ca, double Pa,
                         double phi, size_t w, size_t h, double ps,
                         double *out) {
  guanaco::CTF c;
  c.l = l;
  c.df = df;
  c.Cs = Cs;
  c.Ca = Ca;
  c.Pa = Pa;
  c.phi = phi;
  guanaco::get_ctf_n_simple<double>(
    c, reinterpret_cast<std::complex<double> *>(out), w, h, ps);
}

// get_ctf_n(c, ctf, w, h, ps) (the full CTF with envelopes) as called by the
// pybind wrapper get_ctf_wh<double> (ext/ctf.h:131-139).
void ref_ctf_grid(double l, double df, double Cs, double Ca, double Pa,
                  double dd, double theta_c, double phi, size_t w, size_t h,
                  double ps, double *out) {
  guanaco::CTF c;
  c.l = l;
  c.df = df;
  c.Cs = Cs;
  c.Ca = Ca;
  c.Pa = Pa;
  c.dd = dd;
  c.theta_c = theta_c;
  c.phi = phi;
  guanaco::get_ctf_n<double>(
    c, reinterpret_cast<std::complex<double> *>(out), w, h, ps);
}

double ref_defocus_spread(double Cc, double dEE, double dII, double dVV) {
  return guanaco::get_defocus_spread<double>(Cc, dEE, dII, dVV);
}

// phase_flip(x, h) from correct.h:30-39 on complex64 values.
void ref_phase_flip(const float *x, const float *h, float *out, size_t n) {
  auto xc = reinterpret_cast<const std::complex<float> *>(x);
  auto hc = reinterpret_cast<const std::complex<float> *>(h);
  auto oc = reinterpret_cast<std::complex<float> *>(out);
  for (size_t i = 0; i < n; ++i) oc[i] = guanaco::phase_flip(xc[i], hc[i]);
}
}
