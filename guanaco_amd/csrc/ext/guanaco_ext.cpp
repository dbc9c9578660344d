// pybind11 module guanaco_ext: the reference's Python binding surface
// (src/guanaco/detail/ext/{reconstruct,correct,enums,ctf}.h and
// guanaco_ext.cpp:32-37) over the guanaco_amd C++ host layer. Device calls
// release the GIL.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <complex>
#include <string>
#include <vector>

#include "../host/guanaco_host.hpp"

namespace py = pybind11;
using guanaco::eDevice;

namespace {

// ext/reconstruct.h:35-84. The reference also requires a square grid equal
// to the detector width; any grid is accepted here (the slice of a thinner
// volume is the centre rows of the square grid).
void recon(const py::array_t<float> &sinogram,
           py::array_t<float> &reconstruction, const py::array_t<float> &angles,
           float centre, float pixel_size, float min_defocus,
           float max_defocus, eDevice device, int gpu_index) {
  GUANACO_ASSERT(sinogram.ndim() == 2 || sinogram.ndim() == 3);
  GUANACO_ASSERT(reconstruction.ndim() == 2);
  GUANACO_ASSERT(angles.ndim() == 1);
  GUANACO_ASSERT(sinogram.shape()[sinogram.ndim() - 2] == angles.size());
  guanaco::Config c;
  c.device = device;
  c.gpu_index = gpu_index;
  c.num_pixels = sinogram.shape()[sinogram.ndim() - 1];
  c.num_angles = angles.size();
  c.num_defocus = sinogram.ndim() == 2 ? 1 : sinogram.shape()[0];
  c.grid_width = reconstruction.shape()[1];
  c.grid_height = reconstruction.shape()[0];
  c.pixel_size = pixel_size;
  c.min_defocus = min_defocus;
  c.max_defocus = max_defocus;
  c.centre = centre;
  c.angles.assign(angles.data(), angles.data() + angles.size());
  const float *s = sinogram.data();
  float *r = reconstruction.mutable_data();
  py::gil_scoped_release release;
  guanaco::make_reconstructor(c)(s, r);
}

// ext/correct.h:32-64
void corr(const py::array_t<float> &image,
          const py::array_t<std::complex<float>> &ctf, py::array_t<float> &rec,
          eDevice device) {
  GUANACO_ASSERT(image.ndim() == 2);
  GUANACO_ASSERT(ctf.ndim() == 2 || ctf.ndim() == 3);
  GUANACO_ASSERT(rec.ndim() == image.ndim() + (ctf.ndim() - 2));
  const std::size_t num_ctf = ctf.ndim() == 2 ? 1 : ctf.shape()[0];
  const std::size_t ysize = image.shape()[0];
  const std::size_t xsize = image.shape()[1];
  GUANACO_ASSERT(num_ctf > 0);
  GUANACO_ASSERT(ysize > 0);
  GUANACO_ASSERT(xsize > 0);
  GUANACO_ASSERT(ctf.shape()[ctf.ndim() - 2] == (py::ssize_t)ysize);
  GUANACO_ASSERT(ctf.shape()[ctf.ndim() - 1] == (py::ssize_t)xsize);
  GUANACO_ASSERT(rec.shape()[rec.ndim() - 2] == (py::ssize_t)ysize);
  GUANACO_ASSERT(rec.shape()[rec.ndim() - 1] == (py::ssize_t)xsize);
  const float *im = image.data();
  const std::complex<float> *c = ctf.data();
  float *r = rec.mutable_data();
  py::gil_scoped_release release;
  guanaco::correct(im, c, r, xsize, ysize, num_ctf, device);
}

py::array_t<std::complex<double>> ctf_simple_wh(const guanaco::CTF &c,
                                                std::size_t w, std::size_t h,
                                                double ps) {
  py::array_t<std::complex<double>> out({h, w});
  guanaco::get_ctf_n_simple(c, out.mutable_data(), w, h, ps);
  return out;
}

py::array_t<std::complex<double>> ctf_simple_q(const guanaco::CTF &c,
                                               const py::array_t<double> &q,
                                               const py::array_t<double> &theta) {
  GUANACO_ASSERT(q.size() == theta.size());
  std::vector<py::ssize_t> shape(q.shape(), q.shape() + q.ndim());
  py::array_t<std::complex<double>> out(shape);
  auto *o = out.mutable_data();
  for (py::ssize_t i = 0; i < q.size(); ++i)
    o[i] = guanaco::get_ctf_simple(c, q.data()[i], theta.data()[i]);
  return out;
}

using F64Array = py::array_t<double, py::array::c_style | py::array::forcecast>;

// ext/ctf.h:78-129: the CTF functions mapped over arrays of (q, theta)
template <typename T, typename F>
py::array_t<T> map_q_theta(const F64Array &q, const F64Array &theta, F f) {
  GUANACO_ASSERT(q.size() == theta.size());
  std::vector<py::ssize_t> shape(q.shape(), q.shape() + q.ndim());
  py::array_t<T> out(shape);
  T *o = out.mutable_data();
  const double *qp = q.data(), *tp = theta.data();
  for (py::ssize_t i = 0; i < q.size(); ++i) o[i] = f(qp[i], tp[i]);
  return out;
}

py::array_t<double> Es_q(const guanaco::CTF &c, const F64Array &q,
                         const F64Array &theta) {
  return map_q_theta<double>(q, theta, [&](double a, double t) {
    return std::exp(guanaco::get_log_Es(c, a, t));
  });
}

py::array_t<double> Et_q(const guanaco::CTF &c, const F64Array &q) {
  return map_q_theta<double>(
      q, q, [&](double a, double) { return std::exp(guanaco::get_log_Et(c, a)); });
}

py::array_t<std::complex<double>> ctf_q(const guanaco::CTF &c,
                                        const F64Array &q,
                                        const F64Array &theta) {
  return map_q_theta<std::complex<double>>(
      q, theta, [&](double a, double t) { return guanaco::get_ctf(c, a, t); });
}

py::array_t<double> ctf_simple_real_q(const guanaco::CTF &c,
                                      const F64Array &q,
                                      const F64Array &theta) {
  return map_q_theta<double>(q, theta, [&](double a, double t) {
    return guanaco::get_ctf_simple_real(c, a, t);
  });
}

py::array_t<double> ctf_simple_imag_q(const guanaco::CTF &c,
                                      const F64Array &q,
                                      const F64Array &theta) {
  return map_q_theta<double>(q, theta, [&](double a, double t) {
    return guanaco::get_ctf_simple_imag(c, a, t);
  });
}

// ext/ctf.h:131-169: the FFT-ordered (h, w) grids
py::array_t<std::complex<double>> ctf_wh(const guanaco::CTF &c, std::size_t w,
                                         std::size_t h, double ps) {
  py::array_t<std::complex<double>> out({h, w});
  guanaco::get_ctf_n(c, out.mutable_data(), w, h, ps);
  return out;
}

py::array_t<double> ctf_simple_real_wh(const guanaco::CTF &c, std::size_t w,
                                       std::size_t h, double ps) {
  py::array_t<double> out({h, w});
  guanaco::get_ctf_n_simple_real(c, out.mutable_data(), w, h, ps);
  return out;
}

py::array_t<double> ctf_simple_imag_wh(const guanaco::CTF &c, std::size_t w,
                                       std::size_t h, double ps) {
  py::array_t<double> out({h, w});
  guanaco::get_ctf_n_simple_imag(c, out.mutable_data(), w, h, ps);
  return out;
}

py::tuple q_and_theta(std::size_t w, std::size_t h, double ps) {
  py::array_t<double> q({h, w}), a({h, w});
  GUANACO_ASSERT(ps > 0);
  const double inv_ps = 1.0 / ps;
  auto *qp = q.mutable_data();
  auto *ap = a.mutable_data();
  for (std::size_t j = 0; j < h; ++j)
    for (std::size_t i = 0; i < w; ++i)
      guanaco::get_r_and_theta(i, j, w, h, inv_ps, qp++, ap++);
  return py::make_tuple(q, a);
}

guanaco::VolumeRequest make_request(
    py::ssize_t num_pixels, py::ssize_t num_rows, py::ssize_t num_defocus,
    py::ssize_t depth, py::ssize_t width, const py::array_t<float> &angles,
    int layout, const py::object &centre, float pixel_size,
    float min_defocus, float max_defocus, const py::object &weights,
    const std::string &storage, const py::object &defocus, double wavelength,
    double Cs, double Ca, double Pa, double phi, int gpu_index) {
  guanaco::VolumeRequest r;
  r.num_pixels = num_pixels;
  r.num_rows = num_rows;
  r.num_angles = angles.size();
  r.num_defocus = num_defocus;
  r.depth = depth;
  r.width = width;
  r.pixel_size = pixel_size;
  // one centre, or one per row (reconstruct.py:32-52)
  auto c = py::array_t<float, py::array::c_style | py::array::forcecast>(
      centre);
  if (c.size() == 1) {
    r.centre = c.data()[0];
  } else {
    GUANACO_ASSERT(c.size() == num_rows);
    r.centre = c.data()[0];
    r.centres.assign(c.data(), c.data() + c.size());
  }
  r.min_defocus = min_defocus;
  r.max_defocus = max_defocus;
  r.angles.assign(angles.data(), angles.data() + angles.size());
  if (!weights.is_none()) {
    auto w = py::array_t<float, py::array::c_style | py::array::forcecast>(
        weights);
    r.weights.assign(w.data(), w.data() + w.size());
  }
  r.layout = (guanaco::Layout)layout;
  if (storage == "f32") {
    r.storage = guanaco::Storage::f32;
  } else if (storage == "f16") {
    r.storage = guanaco::Storage::f16;
  } else if (storage == "bf16") {
    r.storage = guanaco::Storage::bf16;
  } else {
    throw py::value_error("storage must be 'f32', 'f16' or 'bf16'");
  }
  if (!defocus.is_none()) {
    auto d = py::array_t<double, py::array::c_style | py::array::forcecast>(
        defocus);
    r.defocus.assign(d.data(), d.data() + d.size());
  }
  r.wavelength = wavelength;
  r.Cs = Cs;
  r.Ca = Ca;
  r.Pa = Pa;
  r.phi = phi;
  r.gpu_index = gpu_index;
  return r;
}

// Layout codes: 0 (T, Y, X) projections, 1 (T, ndf, Y, X) corrected,
// 2 (Y, ndf, T, X) sinograms.
void reconstruct_volume(
    const py::array_t<float, py::array::c_style> &input,
    py::array_t<float, py::array::c_style> &volume,
    const py::array_t<float> &angles, int layout, const py::object &centre,
    float pixel_size, float min_defocus, float max_defocus,
    const py::object &weights, const std::string &storage,
    const py::object &defocus, double wavelength, double Cs, double Ca,
    double Pa, double phi, int gpu_index) {
  GUANACO_ASSERT(volume.ndim() == 3);
  GUANACO_ASSERT(layout >= 0 && layout <= 2);
  GUANACO_ASSERT(input.ndim() == 3 || input.ndim() == 4);
  const py::ssize_t T = angles.size();
  py::ssize_t X = input.shape()[input.ndim() - 1], Y, K = 1;
  if (layout == 2) {
    Y = input.shape()[0];
    if (input.ndim() == 4) K = input.shape()[1];
    GUANACO_ASSERT(input.shape()[input.ndim() - 2] == T);
  } else {
    GUANACO_ASSERT(input.shape()[0] == T);
    Y = input.shape()[input.ndim() - 2];
    if (input.ndim() == 4) K = input.shape()[1];
  }
  if (!defocus.is_none()) K = py::len(defocus);
  GUANACO_ASSERT(volume.shape()[0] == Y);
  auto r = make_request(X, Y, K, volume.shape()[1], volume.shape()[2], angles,
                        layout, centre, pixel_size, min_defocus, max_defocus,
                        weights, storage, defocus, wavelength, Cs, Ca, Pa, phi,
                        gpu_index);
  // the caller's input holds exactly what the request reads
  GUANACO_ASSERT((std::size_t)input.nbytes() ==
                 guanaco::volume_input_bytes(r));
  const float *in = input.data();
  float *out = volume.mutable_data();
  py::gil_scoped_release release;
  guanaco::reconstruct_volume(r, in, out);
}

// Host bytes reconstruct_volume reads for these sizes (no device work).
std::size_t volume_input_bytes(py::ssize_t num_pixels, py::ssize_t num_rows,
                               py::ssize_t num_angles, py::ssize_t num_defocus,
                               int layout, bool device_ctf) {
  guanaco::VolumeRequest r;
  r.num_pixels = num_pixels;
  r.num_rows = num_rows;
  r.num_angles = num_angles;
  r.num_defocus = num_defocus;
  r.depth = r.width = 1;
  r.angles.assign(num_angles, 0.0f);
  r.layout = (guanaco::Layout)layout;
  if (device_ctf) r.defocus.assign(num_defocus, 0.0);
  return guanaco::volume_input_bytes(r);
}

}  // namespace

PYBIND11_MODULE(guanaco_ext, m) {
  m.doc() = "guanaco reconstruction hot path on AMD MI355X (gfx950)";

  // ext/enums.h:35-67: eDevice, constructible from "cpu"/"gpu" and ints
  py::enum_<eDevice> dev(m, "eDevice");
  dev.value("cpu", guanaco::e_host).value("gpu", guanaco::e_device)
      .export_values();
  dev.def(py::init([](const std::string &s) {
    if (s == "cpu") return guanaco::e_host;
    if (s == "gpu") return guanaco::e_device;
    throw py::value_error("\"" + s +
                          "\" is not a valid value for enum type eDevice");
  }));
  dev.def(py::init([](int i) {
    if (i == 0) return guanaco::e_host;
    if (i == 1) return guanaco::e_device;
    throw py::value_error("invalid eDevice value " + std::to_string(i));
  }));
  py::implicitly_convertible<py::str, eDevice>();
  py::implicitly_convertible<py::int_, eDevice>();

  // ext/ctf.h:172-279
  m.def("get_electron_wavelength", &guanaco::get_electron_wavelength,
        py::arg("V"));
  m.def("get_defocus_spread", &guanaco::get_defocus_spread, py::arg("Cc"),
        py::arg("dEE"), py::arg("dII"), py::arg("dVV"));
  m.def("q_to_Q", &guanaco::q_to_Q, py::arg("q"), py::arg("l"), py::arg("Cs"));
  m.def("Q_to_q", &guanaco::Q_to_q, py::arg("Q"), py::arg("l"), py::arg("Cs"));
  m.def("df_to_D", &guanaco::df_to_D, py::arg("df"), py::arg("l"),
        py::arg("Cs"));
  m.def("D_to_df", &guanaco::D_to_df, py::arg("D"), py::arg("l"),
        py::arg("Cs"));
  m.def("get_q_and_theta", &q_and_theta, py::arg("w"), py::arg("h"),
        py::arg("ps"));
  py::class_<guanaco::CTF>(m, "CTF")
      .def(py::init([](double l, double df, double Cs, double Ca, double Pa,
                       double dd, double theta_c, double phi) {
             guanaco::CTF c;
             c.l = l;
             c.df = df;
             c.Cs = Cs;
             c.Ca = Ca;
             c.Pa = Pa;
             c.dd = dd;
             c.theta_c = theta_c;
             c.phi = phi;
             return c;
           }),
           py::arg("l"), py::arg("df") = 0, py::arg("Cs") = 0,
           py::arg("Ca") = 0, py::arg("Pa") = 0, py::arg("dd") = 0,
           py::arg("theta_c") = 0, py::arg("phi") = 0)
      .def_readwrite("l", &guanaco::CTF::l)
      .def_readwrite("df", &guanaco::CTF::df)
      .def_readwrite("Cs", &guanaco::CTF::Cs)
      .def_readwrite("Ca", &guanaco::CTF::Ca)
      .def_readwrite("Pa", &guanaco::CTF::Pa)
      .def_readwrite("dd", &guanaco::CTF::dd)
      .def_readwrite("theta_c", &guanaco::CTF::theta_c)
      .def_readwrite("phi", &guanaco::CTF::phi)
      .def("get_chi", &guanaco::get_chi, py::arg("q"), py::arg("theta"))
      .def("get_Es", &Es_q, py::arg("q"), py::arg("theta"))
      .def("get_Et", &Et_q, py::arg("q"))
      .def("get_ctf", &ctf_q, py::arg("q"), py::arg("theta"))
      .def("get_ctf_simple", &ctf_simple_q, py::arg("q"), py::arg("theta"))
      .def("get_ctf_simple_real", &ctf_simple_real_q, py::arg("q"),
           py::arg("theta"))
      .def("get_ctf_simple_imag", &ctf_simple_imag_q, py::arg("q"),
           py::arg("theta"))
      .def("get_ctf", &ctf_wh, py::arg("w"), py::arg("h"), py::arg("ps"))
      .def("get_ctf_simple", &ctf_simple_wh, py::arg("w"), py::arg("h"),
           py::arg("ps"))
      .def("get_ctf_simple_real", &ctf_simple_real_wh, py::arg("w"),
           py::arg("h"), py::arg("ps"))
      .def("get_ctf_simple_imag", &ctf_simple_imag_wh, py::arg("w"),
           py::arg("h"), py::arg("ps"));

  // ext/correct.h:71-76
  m.def("corr", &corr, py::arg("image"), py::arg("ctf"), py::arg("rec"),
        py::arg("device") = guanaco::e_host);

  // ext/reconstruct.h:88-101
  m.def("recon", &recon, py::arg("sinogram"), py::arg("reconstruction"),
        py::arg("angles"), py::arg("centre"), py::arg("pixel_size") = 1.0,
        py::arg("min_defocus") = 0.0, py::arg("max_defocus") = 0.0,
        py::arg("device") = guanaco::e_host, py::arg("gpu_index") = -1);

  // Additive: the whole volume in one device-resident run (bands of rows
  // when it does not fit).
  m.def("reconstruct_volume", &reconstruct_volume, py::arg("input"),
        py::arg("volume"), py::arg("angles"), py::arg("layout"),
        py::arg("centre"), py::arg("pixel_size") = 1.0f,
        py::arg("min_defocus") = 0.0f, py::arg("max_defocus") = 0.0f,
        py::arg("weights") = py::none(), py::arg("storage") = "f32",
        py::arg("defocus") = py::none(), py::arg("wavelength") = 0.0,
        py::arg("Cs") = 0.0, py::arg("Ca") = 0.0, py::arg("Pa") = 0.0,
        py::arg("phi") = 0.0, py::arg("gpu_index") = -1);
  m.def("volume_input_bytes", &volume_input_bytes, py::arg("num_pixels"),
        py::arg("num_rows"), py::arg("num_angles"), py::arg("num_defocus"),
        py::arg("layout"), py::arg("device_ctf"));

  py::register_exception<guanaco::Error>(m, "Error", PyExc_RuntimeError);
}
