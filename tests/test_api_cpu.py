"""The drop-in Python/pybind11 surface without a GPU: helper semantics,
host CTF math, argument checks and error behaviour."""
import ctypes
import math

import numpy
import pytest

import guanaco_amd as G
from oracle import oracle as O


def test_public_names_match_reference_surface():
    # src/guanaco/__init__.py:22-28, detail/__init__.py:21,
    # ext/guanaco_ext.cpp:32-37
    for name in ["reconstruct", "correct_projections", "recon", "corr",
                 "eDevice", "CTF", "get_electron_wavelength",
                 "reconstruct_file", "correct_file", "reconstruct_tomopy",
                 "get_centre", "get_defocus_steps", "get_corrected_array",
                 "get_ctf", "get_ctf_array", "compute_angular_weights",
                 "get_weights", "get_defocus_spread", "get_q_and_theta",
                 "q_to_Q", "Q_to_q", "df_to_D", "D_to_df"]:
        assert hasattr(G, name), name
    import guanaco_amd.plot as P
    assert callable(P.plot_ctf) and callable(P.generate_ctf)
    with pytest.raises(RuntimeError, match="requires tomopy"):
        G.reconstruct_tomopy(numpy.zeros((3, 2, 8), numpy.float32),
                             numpy.zeros(3), numpy.zeros((2, 8, 8)))
    assert hasattr(G.detail, "recon") and hasattr(G.detail, "corr")


def test_edevice_accepts_strings_and_ints():
    # ext/enums.h:35-67
    assert G.eDevice("gpu") == G.eDevice.gpu == G.eDevice(1)
    assert G.eDevice("cpu") == G.eDevice.cpu == G.eDevice(0)
    with pytest.raises(ValueError):
        G.eDevice("tpu")


def test_wavelength_and_ctf_bitwise_vs_reference_headers():
    ref = O.ref_lib()
    if ref is None:
        pytest.skip("oracle/_ref not built")
    for V in (80e3, 200e3, 300e3):
        assert G.get_electron_wavelength(V) == ref.ref_electron_wavelength(V)
    l = G.get_electron_wavelength(300e3)
    for df, Cs, Ca, Pa, phi, w, h, ps in [
            (20000.0, 2.7e7, 0.0, 0.0, 0.1, 40, 30, 2.0),
            (30000.0, 2.7e7, 300.0, 0.5, 0.0, 33, 17, 1.5)]:
        mine = G.CTF(l, df, Cs, Ca, Pa, phi=phi).get_ctf_simple(w, h, ps)
        out = numpy.empty((h, w, 2))
        ref.ref_ctf_simple_grid(l, df, Cs, Ca, Pa, phi, w, h, ps,
                                out.ctypes.data_as(
                                    ctypes.POINTER(ctypes.c_double)))
        numpy.testing.assert_array_equal(mine.real, out[..., 0])
        numpy.testing.assert_array_equal(mine.imag, out[..., 1])


def test_q_and_theta_matches_reference_test():
    # reference tests/test_ctf.py:396-406
    q1, a1 = G.get_q_and_theta(100, 100, 2)
    y, x = numpy.mgrid[0:100, 0:100]
    y = ((y - 50) / 100) / 2
    x = ((x - 50) / 100) / 2
    assert numpy.abs(q1 - numpy.fft.fftshift(numpy.hypot(x, y))).max() < 1e-5
    assert numpy.abs(a1 - numpy.fft.fftshift(numpy.arctan2(y, x))).max() \
        < 1e-5


def test_get_ctf_array_matches_oracle():
    a = G.get_ctf_array(32, 16, 2.0, 300, 19000.0, 3, 500.0, 2.7, 100.0,
                        30.0, 5.0)
    b = O.get_ctf_array(32, 16, 2.0, 300, 19000.0, 3, 500.0, 2.7, 100.0,
                        30.0, 5.0)
    assert a.dtype == numpy.complex64 and a.shape == (3, 16, 32)
    numpy.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("args", [
    (20000.0, 1, None, 1024, 512.0, 2.0),
    (20000.0, 0, None, 1024, 512.0, 2.0),
    (30000.0, 5, 100.0, 2048, 1024.0, 1.5),
    (30000.0, 16, None, 2048, numpy.full(8, 1024.0, numpy.float32), 1.5),
    (25000.0, 32, None, 4096, numpy.full(4, 2048.0, numpy.float32), 1.0),
    (25000.0, 7, None, 1500, numpy.linspace(700, 800, 5).astype(
        numpy.float32), 1.0),
])
def test_defocus_steps_match_oracle(args):
    got = G.get_defocus_steps(*args)
    exp = O.get_defocus_steps(*args)
    assert len(got) == 4
    for g, e in zip(got, exp):
        assert g == e and type(g) == type(e)


def test_defocus_steps_known_values():
    # C4 geometry: 4096 px, centre 2048, 1 A/px -> +-2048 A, step 4096/31
    mn, mx, st, n = G.get_defocus_steps(
        25000.0, 32, None, 4096, numpy.full(1, 2048.0, numpy.float32), 1.0)
    assert (mn, mx, n) == (25000 - 2048, 25000 + 2048, 32)
    assert st == pytest.approx(4096 / 31)


def test_angular_weights_match_oracle_and_hand_values():
    rng = numpy.random.default_rng(3)
    for n in (1, 2, 3, 7, 61):
        a = numpy.sort(rng.uniform(-1, 1, n))
        rng.shuffle(a)
        numpy.testing.assert_array_equal(G.compute_angular_weights(a),
                                         O.compute_angular_weights(a))
    w = G.compute_angular_weights(numpy.array([3.0, 0.0, 1.0]))
    assert numpy.allclose(w, numpy.array([1.75, 1.25, 1.5]) / 1.5)
    assert numpy.allclose(G.get_weights(numpy.zeros(4)), 1.0)


def test_centre_helpers():
    c = G.get_centre((5, 7, 64), None, sinogram_order=False)
    assert c.shape == (7,) and numpy.all(c == 32.0)
    c = G.get_centre((7, 5, 64), 30.5)
    assert c.shape == (7,) and numpy.all(c == 30.5)


def test_cpu_device_is_refused_with_a_clear_error():
    sino = numpy.zeros((3, 16), numpy.float32)
    rec = numpy.zeros((16, 16), numpy.float32)
    with pytest.raises(RuntimeError, match="GPU device"):
        G.recon(sino, rec, numpy.zeros(3, numpy.float32), 8.0)
    with pytest.raises(RuntimeError, match="GPU device"):
        G.corr(numpy.zeros((8, 8), numpy.float32),
               numpy.zeros((8, 8), numpy.complex64),
               numpy.zeros((8, 8), numpy.float32), "cpu")


def test_recon_shape_assertions():
    # ext/reconstruct.h:46-59 raise before any device work
    rec = numpy.zeros((16, 16), numpy.float32)
    with pytest.raises(RuntimeError, match="ASSERT"):
        G.recon(numpy.zeros((4, 16), numpy.float32), rec,
                numpy.zeros(3, numpy.float32), 8.0, device="gpu")
    with pytest.raises(RuntimeError, match="ASSERT"):
        G.recon(numpy.zeros(16, numpy.float32), rec,
                numpy.zeros(3, numpy.float32), 8.0, device="gpu")


def test_corr_shape_assertions():
    with pytest.raises(RuntimeError, match="ASSERT"):
        G.corr(numpy.zeros((8, 8), numpy.float32),
               numpy.zeros((8, 4), numpy.complex64),
               numpy.zeros((8, 8), numpy.float32), "gpu")


def test_volume_input_bytes_of_device_corrected_projections():
    # The fused path reads the (T, Y, X) projections once; num_defocus only
    # multiplies corrected (T, ndf, Y, X) and sinogram (Y, ndf, T, X) input.
    ext = G.detail
    T, Y, X, K = 7, 16, 64, 8
    assert ext.volume_input_bytes(X, Y, T, K, 0, True) == 4 * T * Y * X
    assert ext.volume_input_bytes(X, Y, T, K, 1, False) == 4 * T * K * Y * X
    assert ext.volume_input_bytes(X, Y, T, K, 2, False) == 4 * T * K * Y * X
    assert ext.volume_input_bytes(X, Y, T, 1, 0, False) == 4 * T * Y * X


def test_reconstruct_volume_checks_array_sizes():
    # a wrong-sized input or volume is refused before any device work
    T, Y, X, K = 5, 8, 32, 4
    angles = numpy.zeros(T, numpy.float32)
    vol = numpy.zeros((Y, 8, X), numpy.float32)
    # (T, K, Y, X) handed to the path that corrects (T, Y, X) projections
    wide = numpy.zeros((T, K, Y, X), numpy.float32)
    with pytest.raises(RuntimeError, match="volume_input_bytes"):
        G.detail.reconstruct_volume(wide, vol, angles, 0, X / 2.0,
                                    defocus=numpy.full(K, 1e4),
                                    wavelength=0.02, Cs=2.7e7)


def test_detectors_above_8192_are_refused_before_any_work(tmp_path):
    # the gfx950 kernels take up to 8192 rows and columns; larger detectors
    # raise a RuntimeError naming the limit (no GPU needed to refuse them)
    from guanaco_amd.detail.correct import check_detector_size
    check_detector_size(8192, 8192)
    check_detector_size(5760, 4092)
    check_detector_size(8192)
    with pytest.raises(RuntimeError, match="width 8193 exceeds the 8192"):
        check_detector_size(8193, 4092)
    with pytest.raises(RuntimeError, match="height 9000"):
        check_detector_size(1000, 9000)
    with pytest.raises(RuntimeError, match="width 8193"):
        check_detector_size(8193)
    proj = numpy.zeros((3, 8, 8200), numpy.float32)
    with pytest.raises(RuntimeError, match="8192"):
        G.reconstruct_projections(proj, numpy.zeros(3, numpy.float32),
                                  defocus=20000.0, num_defocus=2)
