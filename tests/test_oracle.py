"""Pins for the CPU oracle (runs without a GPU).

Each oracle function is checked against something other than its own
formula: the reference's own tests (tests/test_ctf.py), the reference's own
header-only CTF code compiled in place (oracle/_ref), library routines
(numpy.fft), closed forms and hand-computed known answers.
"""
import math

import numpy
import pytest

from oracle import oracle as O


# ---------------------------------------------------------------- CTF --------

def _wavelength_scipy(V):
    # reference tests/test_ctf.py:8-24 (scipy CODATA constants)
    from scipy.constants import c, e, h, m_e
    return h * c / math.sqrt((e * V) ** 2 + 2 * m_e * e * V * c ** 2) * 1e10


def test_wavelength_matches_reference_test():
    # reference tests/test_ctf.py:337-340, pytest.approx default rel 1e-6
    assert O.electron_wavelength(300000) == pytest.approx(
        _wavelength_scipy(300000))
    # known value at 300 kV
    assert O.electron_wavelength(300000) == pytest.approx(0.0196875, rel=1e-5)


def test_chi_matches_reference_test():
    # reference tests/test_ctf.py:103-121 and 387-393
    l = O.electron_wavelength(300000)
    q = 10.0
    chi1 = O.chi(l, 20000, 2.7e7, 0, 0, q, 0.0)
    chi2 = math.pi * (2.7e7 * l ** 3 * q ** 4 / 2 - l * 20000 * q ** 2)
    assert chi1 == pytest.approx(chi2)


def test_q_theta_grid_matches_reference_test():
    # reference tests/test_ctf.py:396-406
    w = h = 100
    q1 = numpy.zeros((h, w))
    a1 = numpy.zeros((h, w))
    for j in range(h):
        for i in range(w):
            q1[j, i], a1[j, i] = O.r_and_theta(i, j, w, h, 1.0 / 2)
    y, x = numpy.mgrid[0:100, 0:100]
    y = ((y - 50) / 100) / 2
    x = ((x - 50) / 100) / 2
    q2 = numpy.fft.fftshift(numpy.sqrt(x * x + y * y))
    a2 = numpy.fft.fftshift(numpy.arctan2(y, x))
    assert numpy.abs(q1 - q2).max() < 1e-5
    assert numpy.abs(a1 - a2).max() < 1e-5


def _ctf2d_simple(shape, ps, df, Cs_mm, l, phase_shift):
    # reference tests/test_ctf.py:222-334 with zero envelopes: chi on an
    # fftshifted grid, exp(-i chi) * exp(i phi)
    Y, X = numpy.mgrid[0:shape[0], 0:shape[1]]
    Y = (1 / ps) * (Y - shape[0] // 2) / shape[0]
    X = (1 / ps) * (X - shape[1] // 2) / shape[1]
    q = numpy.sqrt(X ** 2 + Y ** 2)
    Cs = Cs_mm * 1e7
    chi = math.pi * (Cs * l ** 3 * q ** 4 / 2 - l * df * q ** 2)
    return numpy.fft.fftshift(numpy.exp(-1j * chi) * numpy.exp(1j * phase_shift))


def test_ctf_simple_matches_reference_test():
    # reference tests/test_ctf.py:424-460 (complex, real and imaginary parts)
    l = O.electron_wavelength(300000)
    c1 = O.ctf_simple_grid(l, 20000, 2.7e7, 0, 0, 0.1, 100, 100, 2)
    c2 = _ctf2d_simple((100, 100), 2, 20000, 2.7, l, 0.1)
    assert numpy.abs(c1 - c2).max() < 1e-5
    assert numpy.abs(c1.real - c2.real).max() < 1e-5
    assert numpy.abs(c1.imag - c2.imag).max() < 1e-5


@pytest.mark.parametrize("params", [
    (20000.0, 2.7e7, 0.0, 0.0, 0.0, 64, 48, 2.0),
    (30000.0, 2.7e7, 500.0, 0.3, 0.1, 33, 20, 1.5),
    (0.0, 0.0, 0.0, 0.0, math.pi / 2, 16, 16, 1.0),
])
def test_ctf_grid_bitwise_vs_reference_headers(params):
    ref = O.ref_lib()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    import ctypes
    df, Cs, Ca, Pa, phi, w, h, ps = params
    l = O.electron_wavelength(300000)
    assert ref.ref_electron_wavelength(300000.0) == l
    out = numpy.empty((h, w, 2))
    ref.ref_ctf_simple_grid(l, df, Cs, Ca, Pa, phi, w, h, ps,
                            out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    mine = O.ctf_simple_grid(l, df, Cs, Ca, Pa, phi, w, h, ps)
    numpy.testing.assert_array_equal(mine.real, out[..., 0])
    numpy.testing.assert_array_equal(mine.imag, out[..., 1])


def test_phase_flip_sign_vs_reference_headers():
    ref = O.ref_lib()
    if ref is None:
        pytest.skip("oracle/_ref not built")
    import ctypes
    rng = numpy.random.default_rng(5)
    x = (rng.standard_normal(64) + 1j * rng.standard_normal(64)).astype(
        numpy.complex64)
    h = numpy.array([1j, -1j, 0, -0.0j, 1 + 1e-30j, 2 - 3j] * 10 + [0j] * 4,
                    dtype=numpy.complex64)
    out = numpy.empty_like(x)
    p = lambda a: a.view(numpy.float32).ctypes.data_as(
        ctypes.POINTER(ctypes.c_float))
    ref.ref_phase_flip(p(x), p(h), p(out), x.size)
    expect = -x * numpy.sign(h.imag)
    numpy.testing.assert_array_equal(out, expect.astype(numpy.complex64))


# ---------------------------------------------------------------- FFT --------

@pytest.mark.parametrize("n", [1, 2, 3, 5, 8, 12, 16, 97, 128, 1000, 1024,
                               1500, 4096])
def test_fft_vs_numpy(n):
    rng = numpy.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    ref = numpy.fft.fft(x)
    y64 = O.fft(x.astype(numpy.complex128), -1)
    assert numpy.abs(y64 - ref).max() <= 1e-12 * max(1, numpy.abs(ref).max())
    y32 = O.fft(x.astype(numpy.complex64), -1)
    err = numpy.linalg.norm(y32 - ref) / numpy.linalg.norm(ref)
    assert err < 5e-7
    inv = O.fft(x.astype(numpy.complex128), +1)
    assert numpy.allclose(inv, numpy.fft.ifft(x) * n, rtol=0, atol=1e-10 * n)


@pytest.mark.parametrize("n", [2, 3, 4, 6, 7, 9, 10])
def test_fft_vs_bruteforce_dft(n):
    rng = numpy.random.default_rng(100 + n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    j = numpy.arange(n)
    M = numpy.exp(-2j * numpy.pi * numpy.outer(j, j) / n)
    assert numpy.allclose(O.fft(x, -1), M @ x, atol=1e-12)


@pytest.mark.parametrize("shape", [(12, 16), (16, 16), (20, 15), (64, 128)])
def test_fft2_vs_numpy(shape):
    rng = numpy.random.default_rng(7)
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    assert numpy.allclose(O.fft2(x, -1), numpy.fft.fft2(x), atol=1e-9)
    assert numpy.allclose(O.fft2(x, +1), numpy.fft.ifft2(x) * x.size,
                          atol=1e-9)


# ------------------------------------------------------------- filter --------

def ramp_kernel(N):
    """Closed-form space kernel of the reference filter: h[0] = 1/2,
    h[even != 0] = 0, h[odd n] = -1 / (2 N^2 sin^2(pi n / 2N))."""
    n = numpy.arange(-(N - 1), N)
    h = numpy.zeros(n.size)
    odd = (n % 2) != 0
    h[odd] = -1.0 / (2.0 * N * N * numpy.sin(numpy.pi * n[odd] / (2 * N)) ** 2)
    h[n == 0] = 0.5
    return h


@pytest.mark.parametrize("N", [8, 9, 64, 256, 1500])
def test_filter_vs_closed_form_kernel(N):
    rng = numpy.random.default_rng(N)
    rows = rng.standard_normal((3, N))
    h = ramp_kernel(N)
    expect = numpy.stack([numpy.convolve(r, h)[N - 1:2 * N - 1] for r in rows])
    got64 = O.filter_rows(rows, "f64")
    assert numpy.abs(got64 - expect).max() < 1e-12 * numpy.abs(expect).max()
    got32 = O.filter_rows(rows.astype(numpy.float32), "f32")
    rel = numpy.linalg.norm(got32 - expect) / numpy.linalg.norm(expect)
    assert rel < 1e-6


# --------------------------------------------------------- correction -------

def _flip_from(ctf64):
    return -numpy.sign(ctf64.astype(numpy.complex64).imag)


@pytest.mark.parametrize("Y,X", [(16, 12), (32, 32), (20, 30)])
def test_correction_composition_vs_numpy(Y, X):
    rng = numpy.random.default_rng(Y * X)
    img = rng.standard_normal((Y, X)).astype(numpy.float32)
    l = O.electron_wavelength(300000)
    ctf = numpy.stack([O.ctf_simple_grid(l, df, 2.7e7, 0, 0, 0.1, X, Y, 2.0)
                       for df in (15000.0, 20000.0, 25000.0)])
    c64 = ctf.astype(numpy.complex64)
    got = O.corr(img, c64, precision="f64")
    for j in range(3):
        F = numpy.fft.fft2(img.astype(numpy.float64))
        expect = X * Y * numpy.real(numpy.fft.ifft2(F * _flip_from(ctf[j])))
        assert numpy.allclose(got[j], expect, rtol=0,
                              atol=1e-9 * numpy.abs(expect).max())
    got32 = O.corr(img, c64, precision="f32")
    assert numpy.linalg.norm(got32 - got) / numpy.linalg.norm(got) < 1e-6


def test_correction_special_cases():
    rng = numpy.random.default_rng(3)
    Y, X = 16, 24
    img = rng.standard_normal((Y, X)).astype(numpy.float32)
    l = O.electron_wavelength(300000)
    # df = Cs = 0, phi = 90 deg  ->  out = -(X Y) img
    c = O.ctf_simple_grid(l, 0, 0, 0, 0, math.pi / 2, X, Y, 1.0)
    out = O.corr(img, c.astype(numpy.complex64), precision="f64")
    assert numpy.allclose(out, -X * Y * img.astype(numpy.float64), atol=1e-9)
    # df = Cs = phi = 0  ->  out == 0
    c = O.ctf_simple_grid(l, 0, 0, 0, 0, 0, X, Y, 1.0)
    out = O.corr(img, c.astype(numpy.complex64), precision="f32")
    assert numpy.all(out == 0)
    # phi = 0: DC bin has sign(-0) = 0, so the correction removes the mean
    c = O.ctf_simple_grid(l, 20000, 2.7e7, 0, 0, 0, X, Y, 2.0)
    out = O.corr(img, c.astype(numpy.complex64), precision="f64")
    assert abs(out.mean()) < 1e-9 * numpy.abs(out).max()


def test_correction_is_an_involution_without_zero_bins():
    # phi != 0 leaves no zero in the flip mask, and the mask is Hermitian
    # for Ca = 0, so corr(corr(x) / XY) = XY x.
    rng = numpy.random.default_rng(11)
    Y, X = 32, 32
    img = rng.standard_normal((Y, X))
    l = O.electron_wavelength(300000)
    c = O.ctf_simple_grid(l, 20000, 2.7e7, 0, 0, 0.1, X, Y, 2.0)
    c = c.astype(numpy.complex64)
    once = O.corr(img.astype(numpy.float32), c, precision="f64")
    twice = O.corr((once / (X * Y)).astype(numpy.float32), c, precision="f64")
    assert numpy.allclose(twice, X * Y * img, atol=1e-3)


# ------------------------------------------------------ backprojection ------

def test_bp_single_tilt_zero_known_answer():
    # one tilt at theta = 0, centre N/2, W = N: rec[j, i] = (pi/2) row[i]
    # for i <= N-2 and 0 at i = N-1 (both taps must be inside).
    N, H = 16, 5
    row = numpy.arange(1, N + 1, dtype=numpy.float32)
    rec = O.project(row[None, :], (H, N), [0.0], N / 2)
    expect = numpy.float32(math.pi / 2) * row
    expect[-1] = 0
    for j in range(H):
        numpy.testing.assert_allclose(rec[j], expect, rtol=1e-7)


def test_bp_single_tilt_ninety_known_answer():
    # theta = 90 deg maps the detector onto depth: rec[j, i] ~ (pi/2) row[N-1-j]
    # for interior rows (at j = 0 and j = N-1 one tap falls outside, and
    # float cos(pi/2) ~ -4e-8 decides which side).
    N = 16
    row = numpy.arange(1, N + 1, dtype=numpy.float32) ** 2
    rec = O.project(row[None, :], (N, N), [math.pi / 2], N / 2)
    assert numpy.all(rec[0] == 0)
    for j in range(1, N - 1):
        numpy.testing.assert_allclose(rec[j, 2:-2],
                                      (math.pi / 2) * row[N - 1 - j], rtol=1e-5)


def test_bp_linear_sinogram_closed_form():
    # row[i] = a + b i is interpolated exactly, so for voxels whose taps are
    # all inside, rec = sum_t pi/(2T) (a + b (x cos - z sin + c - 1/2)).
    N, T = 64, 7
    a, b = 0.25, 0.5
    angles = numpy.linspace(-1.0, 1.0, T).astype(numpy.float32)
    centre = N / 2
    row = (a + b * numpy.arange(N)).astype(numpy.float32)
    rec = O.project(numpy.tile(row, (T, 1)), (N, N), angles, centre)
    j, i = numpy.mgrid[0:N, 0:N]
    x = i - N / 2 + 0.5
    z = j - N / 2 + 0.5
    expect = numpy.zeros((N, N))
    for th in angles.astype(numpy.float64):
        p = x * math.cos(th) - z * math.sin(th) + centre
        expect += math.pi / (2 * T) * (a + b * (p - 0.5))
    inner = (x ** 2 + z ** 2) < (N / 2 - 4) ** 2
    assert numpy.allclose(rec[inner], expect[inner], rtol=2e-5)


def test_bp_slab_index_hand_computed():
    # H = 4, ndf = 3, defocus range [-2, 2], ps = 1, one tilt at 0:
    # dscale = 3/4, doffset = 1.5; z = {-1.5, -0.5, 0.5, 1.5} -> height -z ->
    # defocus {2.625, 1.875, 1.125, 0.375} -> slab {2, 2, 1, 0}.
    N, H = 8, 4
    sino = numpy.stack([numpy.full((1, N), k + 1.0, numpy.float32)
                        for k in range(3)])
    rec = O.project(sino, (H, N), [0.0], N / 2, 1.0, -2.0, 2.0)
    s = numpy.float32(math.pi / 2)
    for j, k in enumerate([2, 2, 1, 0]):
        numpy.testing.assert_array_equal(rec[j, :-1], s * numpy.float32(k + 1))
        assert rec[j, -1] == 0


def test_bp_thin_grid_is_crop_of_square_grid():
    rng = numpy.random.default_rng(1)
    N, T, ndf = 32, 9, 3
    angles = numpy.linspace(-1.0, 1.0, T).astype(numpy.float32)
    sino = rng.standard_normal((ndf, T, N)).astype(numpy.float32)
    full = O.project(sino, (N, N), angles, 15.3, 1.0, -40.0, 30.0)
    thin = O.project(sino, (8, N), angles, 15.3, 1.0, -40.0, 30.0)
    numpy.testing.assert_array_equal(thin, full[12:20])


def test_bp_linearity_and_f64_twin():
    rng = numpy.random.default_rng(2)
    N, T, ndf = 24, 5, 2
    angles = numpy.linspace(-1.0, 1.0, T).astype(numpy.float32)
    s1 = rng.standard_normal((ndf, T, N))
    s2 = rng.standard_normal((ndf, T, N))
    args = ((N, N), angles, 12.0, 2.0, -20.0, 20.0)
    r1 = O.project(s1, *args, precision="f64")
    r2 = O.project(s2, *args, precision="f64")
    r12 = O.project(2 * s1 - 3 * s2, *args, precision="f64")
    assert numpy.allclose(r12, 2 * r1 - 3 * r2, atol=1e-12)
    r32 = O.project(s1.astype(numpy.float32), *args)
    assert numpy.linalg.norm(r32 - r1) / numpy.linalg.norm(r1) < 1e-6


def test_recon_disk_phantom_density():
    # Analytic Radon transform of a centred disk of radius R and density 1:
    # p(s) = 2 sqrt(R^2 - s^2). Full 180-degree coverage; FBP ~ 1 inside.
    N, T, R = 128, 180, 30.0
    angles = (numpy.arange(T) * math.pi / T).astype(numpy.float32)
    s = numpy.arange(N) + 0.5 - N / 2
    p = 2 * numpy.sqrt(numpy.clip(R * R - s * s, 0, None))
    sino = numpy.tile(p.astype(numpy.float32), (T, 1))
    rec = O.recon(sino, (N, N), angles, N / 2)
    j, i = numpy.mgrid[0:N, 0:N]
    r = numpy.hypot(i - N / 2 + 0.5, j - N / 2 + 0.5)
    assert abs(rec[r < R - 4].mean() - 1.0) < 0.03
    assert abs(rec[(r > R + 4) & (r < N / 2 - 4)].mean()) < 0.03


# ------------------------------------------------------ Python helpers ------

def test_defocus_steps_known_answers():
    # ndf = 1
    assert O.get_defocus_steps(20000.0, 1) == (20000.0, 20000.0, 0, 1)
    # explicit step: df -/+ (ndf-1)/2 step
    mn, mx, st, n = O.get_defocus_steps(30000.0, 5, 100.0)
    assert (mn, mx, st, n) == (29800.0, 30200.0, 100.0, 5)
    # automatic: C3 geometry, X = 2048, centre 1024, ps 1.5 -> +-1536, 204.8
    c = numpy.full(4, 1024.0, numpy.float32)
    mn, mx, st, n = O.get_defocus_steps(30000.0, 16, None, 2048, c, 1.5)
    assert mn == pytest.approx(30000 - 1536) and mx == pytest.approx(31536)
    assert st == pytest.approx(204.8)
    # num_defocus <= 0 or None -> 1
    assert O.get_defocus_steps(5.0, 0)[3] == 1


def test_angular_weights_known_answers():
    # uniform spacing -> all ones; two angles -> ones; one angle -> one
    assert numpy.allclose(O.compute_angular_weights(numpy.linspace(-1, 1, 9)),
                          1.0)
    assert numpy.allclose(O.compute_angular_weights(numpy.array([0.0, 0.3])),
                          1.0)
    assert numpy.allclose(O.compute_angular_weights(numpy.array([0.2])), 1.0)
    # hand computation for angles 0, 1, 3 (sorted), mean interval 1.5:
    # i=0: (2*1 + (2/3)*2) / (2 + 2/3) = 1.25
    # i=1: (2*1 + 2*2) / 4 = 1.5
    # i=2: ((2/3)*1 + 2*2) / (2/3 + 2) = 1.75
    w = O.compute_angular_weights(numpy.array([3.0, 0.0, 1.0]))
    assert numpy.allclose(w, numpy.array([1.75, 1.25, 1.5]) / 1.5)


def test_get_centre():
    c = O.get_centre((10, 5, 64), None)
    assert c.shape == (10,) and numpy.all(c == 32.0) and c.dtype == numpy.float32
    c = O.get_centre((5, 10, 64), 20.5, sinogram_order=False)
    assert c.shape == (10,) and numpy.all(c == 20.5)


@pytest.mark.parametrize("Y,X,rows", [(16, 12, [0, 5, 15]),
                                      (64, 48, [0, 31, 63]),
                                      (30, 40, [1, 2, 29])])
def test_corr_rows_equals_full_fp64_correction(Y, X, rows):
    # the sampled-row FP64 correction is the full one at those rows
    rng = numpy.random.default_rng(Y * X)
    img = rng.standard_normal((Y, X)).astype(numpy.float32) + 1.0
    l = O.electron_wavelength(300000)
    ctf = numpy.stack([O.ctf_simple_grid(l, df, 2.7e7, 300.0, 0.3, 0.1, X, Y,
                                         1.5) for df in (15000.0, 20000.0)]
                      ).astype(numpy.complex64)
    full = O.corr(img, ctf, precision="f64")
    got = O.corr_rows(img, O.flip_mask(ctf), rows)
    assert numpy.allclose(got, full[:, rows], rtol=0, atol=1e-9 * abs(
        full).max())


def test_flip_mask_special_cases():
    # df = Cs = 0, phi = 90 deg: H = exp(i pi / 2) = i, flip -1 everywhere;
    # df = Cs = phi = 0: H = 1, Im H = 0, flip 0 (the reference's sign(0))
    l = O.electron_wavelength(300000)
    h = O.ctf_simple_grid(l, 0.0, 0.0, 0, 0, math.pi / 2, 8, 4, 1.0)
    assert (O.flip_mask(h) == -1).all()
    h0 = O.ctf_simple_grid(l, 0.0, 0.0, 0, 0, 0.0, 8, 4, 1.0)
    assert (O.flip_mask(h0) == 0).all()


def test_bp_all_fp64_decisions_closed_form_and_slabs():
    # With FP64 decisions (slab choice and taps from FP64 coordinates) the
    # linear sinogram is interpolated exactly: the closed form to FP64
    # rounding on every voxel whose taps are inside; and the hand-computed
    # slabs of test_bp_slab_index_hand_computed hold (no tie is involved).
    N, T = 64, 7
    a, b = 0.25, 0.5
    angles = numpy.linspace(-1.0, 1.0, T).astype(numpy.float32)
    row = a + b * numpy.arange(N, dtype=numpy.float64)
    rec = O.project(numpy.tile(row, (T, 1)), (N, N), angles, N / 2,
                    precision="f64", f32_decisions=False)
    j, i = numpy.mgrid[0:N, 0:N]
    x = i - N / 2 + 0.5
    z = j - N / 2 + 0.5
    expect = numpy.zeros((N, N))
    for th in angles.astype(numpy.float64):
        p = x * math.cos(th) - z * math.sin(th) + N / 2
        expect += math.pi / (2 * T) * (a + b * (p - 0.5))
    inner = (x ** 2 + z ** 2) < (N / 2 - 4) ** 2
    assert numpy.allclose(rec[inner], expect[inner], rtol=1e-12)
    sino = numpy.stack([numpy.full((1, 8), k + 1.0) for k in range(3)])
    rec = O.project(sino, (4, 8), [0.0], 4.0, 1.0, -2.0, 2.0,
                    precision="f64", f32_decisions=False)
    for jj, k in enumerate([2, 2, 1, 0]):
        assert numpy.allclose(rec[jj, :-1], math.pi / 2 * (k + 1))
        assert rec[jj, -1] == 0


def test_fp64_reconstruct_keeps_fp64_input():
    # a float64 tomogram is not rounded to float32 on the FP64 path
    N, T = 32, 5
    angles = numpy.linspace(-1.0, 1.0, T).astype(numpy.float32)
    rng = numpy.random.default_rng(8)
    tomo = rng.standard_normal((T, 2, N)) * (1 + 1e-9)
    a = O.reconstruct(tomo, angles, precision="f64")
    b = O.reconstruct(tomo.astype(numpy.float32).astype(numpy.float64),
                      angles, precision="f64")
    direct = O.recon(numpy.ascontiguousarray(tomo[:, 1, :]), (N, N), angles,
                     N / 2, precision="f64")
    assert numpy.array_equal(a[1], direct)
    assert not numpy.array_equal(a, b)

