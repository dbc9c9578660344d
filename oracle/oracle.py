"""CPU oracle for the guanaco reconstruction hot path (Python side).

TEST INFRASTRUCTURE ONLY. tests/, ``__graft_entry__.smoke()`` and the CPU
legs of ``bench.py`` (the ``cpu_baseline`` and the check of the GPU volumes
against this oracle) import this module as the checker; the product
(``guanaco_amd``) and the bench's GPU legs never import it.

It wraps ``liboracle.so`` (the C restatement of the reference's CPU path, see
guanaco_oracle.h) with ctypes, and restates the reference's pure-Python
helpers with numpy:

- ``get_defocus_steps``      src/guanaco/detail/correct.py:44-93
- ``get_ctf_array``          src/guanaco/detail/correct.py:116-214
- ``correct_projections``    src/guanaco/detail/correct.py:217-318
- ``get_centre``             src/guanaco/detail/reconstruct.py:32-52
- ``compute_angular_weights`` src/guanaco/detail/reconstruct.py:55-105
- ``get_weights``            src/guanaco/detail/reconstruct.py:108-124
- ``reconstruct``            src/guanaco/detail/reconstruct.py:127-213 with
                             mp.reconstruction_worker (mp.py:106-154)
"""

import ctypes
import os
import subprocess

import numpy

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)
_sz = ctypes.c_size_t


class OrConfig(ctypes.Structure):
    _fields_ = [
        ("num_pixels", _sz),
        ("num_angles", _sz),
        ("num_defocus", _sz),
        ("grid_height", _sz),
        ("grid_width", _sz),
        ("pixel_size", ctypes.c_float),
        ("centre", ctypes.c_float),
        ("min_defocus", ctypes.c_float),
        ("max_defocus", ctypes.c_float),
        ("angles", _f32p),
    ]


def build():
    """Compile liboracle.so (and oracle/_ref when the reference is present)."""
    subprocess.check_call(["make", "-s", "-C", _HERE, "all", "ref"])


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.or_electron_wavelength.restype = ctypes.c_double
        L.or_electron_wavelength.argtypes = [ctypes.c_double]
        L.or_chi.restype = ctypes.c_double
        L.or_chi.argtypes = [ctypes.c_double] * 7
        L.or_r_and_theta.argtypes = [_sz, _sz, _sz, _sz, ctypes.c_double,
                                     _f64p, _f64p]
        L.or_ctf_simple_grid.argtypes = [ctypes.c_double] * 6 + [
            _sz, _sz, ctypes.c_double, _f64p]
        L.or_ctf_grid.argtypes = [ctypes.c_double] * 8 + [
            _sz, _sz, ctypes.c_double, _f64p]
        L.or_log_es.restype = ctypes.c_double
        L.or_log_es.argtypes = [ctypes.c_double] * 9
        L.or_log_et.restype = ctypes.c_double
        L.or_log_et.argtypes = [ctypes.c_double] * 4
        L.or_envelope_a.restype = ctypes.c_double
        L.or_envelope_a.argtypes = [ctypes.c_double] * 3
        L.or_fft_f32.argtypes = [_sz, _f32p, ctypes.c_int]
        L.or_fft_f64.argtypes = [_sz, _f64p, ctypes.c_int]
        L.or_fft2_f32.argtypes = [_sz, _sz, _f32p, ctypes.c_int]
        L.or_fft2_f64.argtypes = [_sz, _sz, _f64p, ctypes.c_int]
        L.or_correct_f32.argtypes = [_f32p, _f32p, _f32p, _sz, _sz, _sz]
        L.or_correct_f64.argtypes = [_f32p, _f32p, _f64p, _sz, _sz, _sz]
        L.or_filter_f32.argtypes = [_f32p, _sz, _sz, _sz]
        L.or_filter_f64.argtypes = [_f64p, _sz, _sz, _sz]
        L.or_project_f32.argtypes = [ctypes.POINTER(OrConfig), _f32p, _f32p]
        L.or_project_f64.argtypes = [ctypes.POINTER(OrConfig), _f64p, _f64p,
                                     ctypes.c_int]
        L.or_recon_f32.argtypes = [ctypes.POINTER(OrConfig), _f32p, _f32p]
        L.or_recon_f64.argtypes = [ctypes.POINTER(OrConfig), _f64p, _f64p,
                                   ctypes.c_int]
        L.or_config_is_valid.argtypes = [ctypes.POINTER(OrConfig)]
        L.or_tilt_coefficients.argtypes = [ctypes.POINTER(OrConfig), _f32p]
        _LIB = L
    return _LIB


def ref_lib():
    """The reference's own header-only CTF math (oracle/_ref), or None."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libref_ctf.so")
        if not os.path.exists(path):
            return None
        R = ctypes.CDLL(path)
        R.ref_electron_wavelength.restype = ctypes.c_double
        R.ref_electron_wavelength.argtypes = [ctypes.c_double]
        R.ref_chi.restype = ctypes.c_double
        R.ref_chi.argtypes = [ctypes.c_double] * 7
        R.ref_r_and_theta.argtypes = [_sz, _sz, _sz, _sz, ctypes.c_double,
                                      _f64p, _f64p]
        R.ref_ctf_simple_grid.argtypes = [ctypes.c_double] * 6 + [
            _sz, _sz, ctypes.c_double, _f64p]
        R.ref_phase_flip.argtypes = [_f32p, _f32p, _f32p, _sz]
        R.ref_ctf_grid.argtypes = [ctypes.c_double] * 8 + [
            _sz, _sz, ctypes.c_double, _f64p]
        R.ref_defocus_spread.restype = ctypes.c_double
        R.ref_defocus_spread.argtypes = [ctypes.c_double] * 4
        _REF = R
    return _REF


def _p32(a):
    assert a.dtype == numpy.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(_f32p)


def _p64(a):
    assert a.dtype == numpy.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_f64p)


# ---- CTF -----------------------------------------------------------------


def electron_wavelength(V):
    return lib().or_electron_wavelength(float(V))


def chi(l, df, Cs, Ca, Pa, q, theta):
    return lib().or_chi(l, df, Cs, Ca, Pa, q, theta)


def r_and_theta(x, y, w, h, inv_ps):
    r = ctypes.c_double()
    t = ctypes.c_double()
    lib().or_r_and_theta(x, y, w, h, inv_ps, ctypes.byref(r), ctypes.byref(t))
    return r.value, t.value


def ctf_simple_grid(l, df=0.0, Cs=0.0, Ca=0.0, Pa=0.0, phi=0.0, w=1, h=1,
                    ps=1.0):
    """complex128 (h, w) grid of exp(-i(chi - phi)) (ctf.h:489-503)."""
    out = numpy.empty((h, w, 2), dtype=numpy.float64)
    lib().or_ctf_simple_grid(l, df, Cs, Ca, Pa, phi, w, h, float(ps), _p64(out))
    return out[..., 0] + 1j * out[..., 1]


def ctf_grid(l, df=0.0, Cs=0.0, Ca=0.0, Pa=0.0, dd=0.0, theta_c=0.0,
             phi=0.0, w=1, h=1, ps=1.0):
    """complex128 (h, w) grid of the full CTF with envelopes (ctf.h:250-258,
    466-478)."""
    out = numpy.empty((h, w, 2), dtype=numpy.float64)
    lib().or_ctf_grid(l, df, Cs, Ca, Pa, dd, theta_c, phi, w, h, float(ps),
                      _p64(out))
    return out[..., 0] + 1j * out[..., 1]


def log_es(l, df, Cs, Ca, Pa, dd, theta_c, q, theta):
    return lib().or_log_es(l, df, Cs, Ca, Pa, dd, theta_c, q, theta)


def log_et(l, dd, theta_c, q):
    return lib().or_log_et(l, dd, theta_c, q)


# ---- FFT -----------------------------------------------------------------


def fft(x, sign=-1):
    x = numpy.ascontiguousarray(x)
    if x.dtype == numpy.complex64:
        y = x.copy()
        lib().or_fft_f32(y.size, y.view(numpy.float32).ctypes.data_as(_f32p),
                         sign)
    else:
        y = x.astype(numpy.complex128)
        lib().or_fft_f64(y.size, y.view(numpy.float64).ctypes.data_as(_f64p),
                         sign)
    return y


def fft2(x, sign=-1):
    x = numpy.ascontiguousarray(x)
    ny, nx = x.shape
    if x.dtype == numpy.complex64:
        y = x.copy()
        lib().or_fft2_f32(nx, ny, y.view(numpy.float32).ctypes.data_as(_f32p),
                          sign)
    else:
        y = x.astype(numpy.complex128)
        lib().or_fft2_f64(nx, ny, y.view(numpy.float64).ctypes.data_as(_f64p),
                          sign)
    return y


# ---- correction ------------------------------------------------------------


def corr(image, ctf, rec=None, precision="f32"):
    """correct.cpp:30-90 for one image. ctf: complex64 (Y, X) or (ndf, Y, X)."""
    image = numpy.ascontiguousarray(image, dtype=numpy.float32)
    ctf = numpy.ascontiguousarray(ctf, dtype=numpy.complex64)
    squeeze = ctf.ndim == 2
    if squeeze:
        ctf = ctf[None]
    ndf, Y, X = ctf.shape
    assert image.shape == (Y, X)
    dt = numpy.float32 if precision == "f32" else numpy.float64
    out = numpy.empty((ndf, Y, X), dtype=dt)
    cptr = ctf.view(numpy.float32).ctypes.data_as(_f32p)
    if precision == "f32":
        lib().or_correct_f32(_p32(image), cptr, _p32(out), X, Y, ndf)
    else:
        lib().or_correct_f64(_p32(image), cptr, _p64(out), X, Y, ndf)
    if squeeze:
        out = out[0]
    if rec is not None:
        rec[...] = out
        return rec
    return out


def flip_mask(ctf):
    """The phase flip of correct.h:30-39 as a real factor per frequency:
    phase_flip(x, H) = -x sign(Im H), the sign read from the complex64 CTF as
    correct.cpp:57-89 reads it (sign(0) = 0). int8 in {-1, 0, 1}, the shape
    of `ctf`."""
    c = numpy.asarray(ctf, dtype=numpy.complex64)
    return (-numpy.sign(c.imag)).astype(numpy.int8)


def corr_rows(image, masks, rows):
    """correct.cpp:30-90 in FP64 at a few detector rows only.

    The correction of one image with flip factors m_k (flip_mask) is
    Re(IFFT2_unnorm(FFT2(image) m_k)) (correct.cpp:57-89: forward FFT,
    phase flip, unnormalised inverse, real part). Its rows y_r are
    Re(IFFT_X_unnorm(E F m_k)) with F = FFT2(image) and
    E[r, ky] = exp(2 pi i ky y_r / Y) the inverse column DFT evaluated at
    those rows only, so the sampled rows cost one 2D FFT plus a
    (rows x Y) (Y x X) product per defocus instead of a 2D inverse each.
    Same arithmetic as corr(precision="f64") up to FP64 rounding (numpy's
    pocketfft in place of the oracle's FFT).

    masks: (ndf, Y, X) int8; returns (ndf, len(rows), X) float64."""
    img = numpy.asarray(image, dtype=numpy.float64)
    masks = numpy.asarray(masks)
    ndf, Y, X = masks.shape
    assert img.shape == (Y, X)
    rows = numpy.asarray(rows, dtype=numpy.int64)
    F = numpy.fft.fft2(img)
    ky = numpy.arange(Y, dtype=numpy.int64)
    # exact phases: (ky * y) mod Y in integers before the exponential
    E = numpy.exp(2j * numpy.pi * ((rows[:, None] * ky[None, :]) % Y) / Y)
    out = numpy.empty((ndf, rows.size, X), numpy.float64)
    for k in range(ndf):
        A = E @ (F * masks[k])
        out[k] = (numpy.fft.ifft(A, axis=1) * X).real
    return out


# ---- ramp filter -----------------------------------------------------------


def filter_rows(data, precision="f32"):
    """filter.cpp:61-97 applied to every row of `data` (..., N)."""
    dt = numpy.float32 if precision == "f32" else numpy.float64
    a = numpy.array(data, dtype=dt, order="C", copy=True)
    N = a.shape[-1]
    nrows = a.size // N
    if precision == "f32":
        lib().or_filter_f32(_p32(a), N, nrows, 1)
    else:
        lib().or_filter_f64(_p64(a), N, nrows, 1)
    return a


# ---- backprojection --------------------------------------------------------


def make_config(N, T, ndf, H, W, pixel_size, centre, min_defocus, max_defocus,
                angles):
    angles = numpy.ascontiguousarray(angles, dtype=numpy.float32)
    cfg = OrConfig(N, T, ndf, H, W, float(pixel_size), float(centre),
                   float(min_defocus), float(max_defocus), _p32(angles))
    cfg._angles_keepalive = angles
    return cfg


def _sino_cfg(sino, grid_shape, angles, centre, pixel_size, min_defocus,
              max_defocus):
    s3 = sino if sino.ndim == 3 else sino[None]
    ndf, T, N = s3.shape
    H, W = grid_shape
    return make_config(N, T, ndf, H, W, pixel_size, centre, min_defocus,
                       max_defocus, angles)


def tilt_coefficients(angles, centre):
    """reconstructor.cpp:36-47: (T, 4) float32 {cos/d, sin/d, offset, scale}."""
    angles = numpy.ascontiguousarray(angles, dtype=numpy.float32)
    cfg = make_config(1, angles.size, 1, 1, 1, 1.0, centre, 0.0, 0.0, angles)
    out = numpy.empty((angles.size, 4), numpy.float32)
    lib().or_tilt_coefficients(ctypes.byref(cfg), _p32(out))
    return out


def project(sino, grid_shape, angles, centre, pixel_size=1.0, min_defocus=0.0,
            max_defocus=0.0, precision="f32", f32_decisions=True, rec=None):
    """reconstructor.cpp:14-90 on an already filtered sinogram."""
    cfg = _sino_cfg(sino, grid_shape, angles, centre, pixel_size, min_defocus,
                    max_defocus)
    if precision == "f32":
        s = numpy.ascontiguousarray(sino, dtype=numpy.float32)
        out = numpy.zeros(grid_shape, numpy.float32) if rec is None else rec
        lib().or_project_f32(ctypes.byref(cfg), _p32(s), _p32(out))
    else:
        s = numpy.ascontiguousarray(sino, dtype=numpy.float64)
        out = numpy.zeros(grid_shape, numpy.float64) if rec is None else rec
        lib().or_project_f64(ctypes.byref(cfg), _p64(s), _p64(out),
                             1 if f32_decisions else 0)
    return out


def recon(sino, grid_shape, angles, centre, pixel_size=1.0, min_defocus=0.0,
          max_defocus=0.0, precision="f32", f32_decisions=True):
    """reconstructor.cpp:92-115: filter a copy, zero the grid, project."""
    cfg = _sino_cfg(sino, grid_shape, angles, centre, pixel_size, min_defocus,
                    max_defocus)
    if precision == "f32":
        s = numpy.ascontiguousarray(sino, dtype=numpy.float32)
        out = numpy.empty(grid_shape, numpy.float32)
        rc = lib().or_recon_f32(ctypes.byref(cfg), _p32(s), _p32(out))
    else:
        s = numpy.ascontiguousarray(sino, dtype=numpy.float64)
        out = numpy.empty(grid_shape, numpy.float64)
        rc = lib().or_recon_f64(ctypes.byref(cfg), _p64(s), _p64(out),
                                1 if f32_decisions else 0)
    if rc != 0:
        raise RuntimeError("invalid reconstruction configuration")
    return out


# ---- Python helpers (restated from the reference's Python layer) -----------


def get_centre(shape, centre, sinogram_order=True):
    """reconstruct.py:32-52."""
    if not sinogram_order:
        shape = list(shape)
        shape[0], shape[-2] = shape[-2], shape[0]
    if centre is None:
        centre = numpy.ones(shape[0], dtype="float32") * (shape[-1] * 0.5)
    elif numpy.array(centre).size == 1:
        centre = numpy.ones(shape[0], dtype="float32") * centre
    return numpy.asarray(centre).astype(dtype="float32", copy=False)


def compute_angular_weights(angles):
    """reconstruct.py:55-105."""
    angles = numpy.asarray(angles)
    index = numpy.argsort(angles)
    weights = numpy.zeros(angles.size)
    if angles.size > 1:
        mean_interval = abs(angles[index[-1]] - angles[index[0]]) / (
            angles.size - 1)
        assert mean_interval != 0
    else:
        mean_interval = 1.0
    for i in range(angles.size):
        si = 0.0
        sw = 0.0
        if i - 2 >= 0:
            sw += 1.0 / 1.5
            si += (1.0 / 1.5) * (angles[index[i - 1]] - angles[index[i - 2]])
        if i - 1 >= 0:
            sw += 2.0
            si += 2.0 * (angles[index[i]] - angles[index[i - 1]])
        if i + 1 < angles.size:
            sw += 2.0
            si += 2.0 * (angles[index[i + 1]] - angles[index[i]])
        if i + 2 < angles.size:
            sw += 1.0 / 1.5
            si += (1.0 / 1.5) * (angles[index[i + 2]] - angles[index[i + 1]])
        weights[index[i]] = si / sw if sw != 0 else 1.0
    weights /= mean_interval
    return weights


def get_weights(angles, angular_weights=False):
    """reconstruct.py:108-124."""
    if angular_weights:
        return compute_angular_weights(angles)
    return numpy.ones(numpy.asarray(angles).shape)


def get_defocus_steps(defocus, num_defocus=None, step_defocus=None,
                      width=None, centre=None, pixel_size=None):
    """correct.py:44-93."""
    if num_defocus is None or num_defocus <= 0:
        num_defocus = 1
    if num_defocus == 1:
        min_defocus = defocus
        max_defocus = defocus
        step_defocus = 0
    else:
        if step_defocus is not None and step_defocus > 0:
            min_defocus = defocus - 0.5 * (num_defocus - 1) * step_defocus
            max_defocus = defocus + 0.5 * (num_defocus - 1) * step_defocus
        else:
            centre = numpy.asarray(centre)
            z0 = defocus - centre * pixel_size
            z1 = defocus + centre * pixel_size
            z2 = defocus - (width - centre) * pixel_size
            z3 = defocus + (width - centre) * pixel_size
            min_defocus = min([z0.min(), z1.min(), z2.min(), z3.min()])
            max_defocus = max([z0.max(), z1.max(), z2.max(), z3.max()])
            step_defocus = (max_defocus - min_defocus) / (num_defocus - 1)
    return min_defocus, max_defocus, step_defocus, num_defocus


def get_ctf_array(width, height, pixel_size, energy, min_defocus, num_defocus,
                  step_defocus, spherical_aberration=0, astigmatism=0,
                  astigmatism_angle=0, phase_shift=0):
    """correct.py:116-214: complex64 (ndf, height, width) CTF stack."""
    out = numpy.zeros((num_defocus, height, width), dtype="complex64")
    l = electron_wavelength(energy * 1000)
    for d in range(num_defocus):
        df = min_defocus + d * step_defocus
        out[d] = ctf_simple_grid(
            l, df, spherical_aberration * 1e7, astigmatism,
            astigmatism_angle * numpy.pi / 180, phase_shift * numpy.pi / 180,
            width, height, pixel_size)
    return out


def correct_projections(projections, centre=None, pixel_size=1, energy=None,
                        defocus=None, num_defocus=None, step_defocus=None,
                        spherical_aberration=0, astigmatism=0,
                        astigmatism_angle=0, phase_shift=0, precision="f32",
                        tilts=None):
    """correct.py:217-318 (in memory). `tilts` optionally restricts the loop
    to a subset of projections (a bounded sample)."""
    if defocus is None:
        return projections, 0, 0
    mn, mx, step, ndf = get_defocus_steps(defocus, num_defocus, step_defocus,
                                          projections.shape[2], centre,
                                          pixel_size)
    assert abs(mn + (ndf - 1) * step - mx) < 1e-7
    ctf = get_ctf_array(projections.shape[2], projections.shape[1], pixel_size,
                        energy, mn, ndf, step, spherical_aberration,
                        astigmatism, astigmatism_angle, phase_shift)
    dt = numpy.float32 if precision == "f32" else numpy.float64
    idx = range(projections.shape[0]) if tilts is None else tilts
    out = numpy.zeros((len(idx), ndf) + projections.shape[1:], dtype=dt)
    for n, z in enumerate(idx):
        out[n] = corr(projections[z], ctf, precision=precision)
    if ndf == 1:
        out = out.reshape((len(idx),) + projections.shape[1:])
    return out, mn - defocus, mx - defocus


def reconstruct(tomogram, angles, centre=None, pixel_size=1, min_defocus=0,
                max_defocus=0, sinogram_order=False, transform=None,
                angular_weights=False, depth=None, slices=None,
                precision="f32", f32_decisions=True):
    """reconstruct.py:127-213 + mp.py:106-154, one slice at a time.

    `depth` keeps the central `depth` rows of the square grid (identical to
    the square result there); `slices` restricts to a subset of y-slices.
    """
    assert tomogram.ndim in (3, 4)
    # reconstruct.py:160 casts to float32; the FP64 twin keeps the input's
    # precision (float64 input stays float64 through the weights and filter)
    dt = numpy.float32 if precision == "f32" else numpy.float64
    sino = tomogram.astype(dt, copy=False)
    if not sinogram_order:
        sino = numpy.swapaxes(sino, 0, -2)
    N = sino.shape[-1]
    H = N if depth is None else depth
    centre = get_centre(sino.shape, centre)
    weights = get_weights(angles, angular_weights)
    ys = range(sino.shape[0]) if slices is None else slices
    out = numpy.zeros((len(ys), H, N), dtype=dt)
    for n, i in enumerate(ys):
        s = sino[i]
        if transform == "minus":
            s = -s
        # The reference's ascontiguousarray can alias a contiguous caller
        # array and then scale it in place; the oracle always copies.
        s = numpy.array(s, dtype=dt, order="C", copy=True)
        if weights is not None:
            if s.ndim == 3:
                for j in range(s.shape[1]):
                    s[:, j, :] *= weights[j]
            else:
                for j in range(s.shape[0]):
                    s[j, :] *= weights[j]
        out[n] = recon(s, (H, N), angles, centre[i], pixel_size, min_defocus,
                       max_defocus, precision=precision,
                       f32_decisions=f32_decisions)
    return out
