"""The CTF with envelopes (ctf.h:188-258, 466-561) on the host: the oracle's
restatement against the reference's own header compiled in place
(oracle/_ref) and against the reference's test formulas (tests/test_ctf.py,
re-expressed with the same inputs and tolerances), and the product's host
CTF (guanaco_ext.CTF) against the oracle. The CTF tool commands' options."""
import math

import numpy
import pytest

import guanaco_amd as G
from oracle import oracle as O

L300 = O.electron_wavelength(300000)


def defocus_spread_py(Cc, dEE, dII, dVV):
    # tests/test_ctf.py:27-43
    return Cc * math.sqrt(dEE ** 2 + (2 * dII) ** 2 + dVV ** 2)


def ctf1d_py(q, defocus, Cs, l, d, theta_c):
    # tests/test_ctf.py:204-251 (chi, envelopes and scale)
    chi = math.pi * (Cs * l ** 3 * q ** 4 / 2 - l * defocus * q ** 2)
    u = 1 + 2 * (math.pi * theta_c * d) ** 2 * q ** 2
    Et = numpy.exp(-((math.pi * l * d) ** 2) * q ** 4 / (2 * u))
    Es = numpy.exp(-((math.pi * theta_c) ** 2) / (l ** 2 * u)
                   * (Cs * l ** 3 * q ** 3 - defocus * l * q) ** 2)
    A = 1 / numpy.sqrt(u)
    return A * Es * Et * numpy.exp(-1j * chi), Es, Et


def ctf2d_py(shape, ps, df, Cs, l, d, theta_c, phase_shift):
    # tests/test_ctf.py:254-334 with wavelength, spread and units given
    Y, X = numpy.mgrid[0:shape[0], 0:shape[1]]
    Y = (1 / ps) * (Y - shape[0] // 2) / shape[0]
    X = (1 / ps) * (X - shape[1] // 2) / shape[1]
    q = numpy.sqrt(X ** 2 + Y ** 2)
    ctf = ctf1d_py(q, df, Cs, l, d, theta_c)[0]
    ctf = ctf * numpy.exp(1j * phase_shift)
    return numpy.fft.fftshift(ctf)


def test_oracle_full_ctf_matches_the_reference_test_formula():
    # tests/test_ctf.py:409-421: 100 x 100, ps 2, df 20000, Cs 2.7 mm, dd 1,
    # theta_c 0.1 mrad, phi 0.1; |difference| <= 1e-5
    got = O.ctf_grid(L300, 20000.0, 2.7e7, 0, 0, 1.0, 0.1 / 1e3, 0.1, 100,
                     100, 2.0)
    expect = ctf2d_py((100, 100), 2.0, 20000.0, 2.7e7, L300, 1.0, 1e-4, 0.1)
    assert numpy.abs(got - expect).max() < 1e-5
    # with strong envelopes the comparison is not trivially |ctf| = 1
    got = O.ctf_grid(L300, 20000.0, 2.7e7, 0, 0, 60.0, 2e-5, 0.1, 64, 48, 1.0)
    expect = ctf2d_py((48, 64), 1.0, 20000.0, 2.7e7, L300, 60.0, 2e-5, 0.1)
    assert numpy.abs(expect).min() < 0.5
    assert numpy.abs(got - expect).max() < 1e-12


def test_oracle_full_ctf_bitwise_vs_the_reference_header():
    R = O.ref_lib()
    if R is None:
        pytest.skip("oracle/_ref not built (the reference is absent)")
    for args in [(20000.0, 2.7e7, 0.0, 0.0, 1.0, 1e-4, 0.1, 100, 100, 2.0),
                 (15000.0, 2.7e7, 300.0, 0.7, 60.0, 2e-5, 0.0, 33, 20, 1.3)]:
        df, Cs, Ca, Pa, dd, tc, phi, w, h, ps = args
        got = O.ctf_grid(L300, df, Cs, Ca, Pa, dd, tc, phi, w, h, ps)
        ref = numpy.empty((h, w, 2))
        R.ref_ctf_grid(L300, df, Cs, Ca, Pa, dd, tc, phi, w, h, ps,
                       O._p64(ref))
        numpy.testing.assert_array_equal(got.real, ref[..., 0])
        numpy.testing.assert_array_equal(got.imag, ref[..., 1])
    assert R.ref_defocus_spread(2.7e7, 1e-6, 1e-6, 1e-6) == \
        G.get_defocus_spread(2.7e7, 1e-6, 1e-6, 1e-6)


def test_host_ctf_functions_match_reference_tests():
    # tests/test_ctf.py:343-385
    assert G.get_defocus_spread(2.7e7, 1e-6, 1e-6, 1e-6) == pytest.approx(
        defocus_spread_py(2.7e7, 1e-6, 1e-6, 1e-6))
    Q = G.q_to_Q(10, L300, 2.7e7)
    assert Q == pytest.approx(10 * (2.7e7 * L300 ** 3) ** 0.25)
    assert G.Q_to_q(Q, L300, 2.7e7) == pytest.approx(10)
    D = G.df_to_D(20000, L300, 2.7e7)
    assert D == pytest.approx(20000 / math.sqrt(2.7e7 * L300))
    assert G.D_to_df(D, L300, 2.7e7) == pytest.approx(20000)
    c = G.CTF(L300, 20000, 2.7e7, 0, 0, 1, 1e-3, 0)
    assert c.get_Es([10], [0])[0] == pytest.approx(0.0)
    assert c.get_Et([10])[0] == pytest.approx(
        ctf1d_py(10.0, 20000, 2.7e7, L300, 1, 1e-3)[2])
    # and where the envelopes are between 0 and 1
    q = numpy.linspace(0.0, 0.4, 41)
    c = G.CTF(L300, 20000, 2.7e7, 0, 0, 60.0, 2e-5, 0)
    _, Es, Et = ctf1d_py(q, 20000, 2.7e7, L300, 60.0, 2e-5)
    assert 0.01 < Es.min() < 0.99 and 0.01 < Et.min() < 0.99
    numpy.testing.assert_allclose(c.get_Es(q, numpy.zeros_like(q)), Es,
                                  rtol=1e-12)
    numpy.testing.assert_allclose(c.get_Et(q), Et, rtol=1e-12)
    full = c.get_ctf(q, numpy.zeros_like(q))
    numpy.testing.assert_allclose(full, ctf1d_py(q, 20000, 2.7e7, L300, 60.0,
                                                 2e-5)[0], atol=1e-12)


def test_host_ctf_grids_match_the_oracle():
    for (df, Cs, Ca, Pa, dd, tc, phi, w, h, ps) in [
            (20000.0, 2.7e7, 0.0, 0.0, 1.0, 1e-4, 0.1, 100, 100, 2.0),
            (15000.0, 2.7e7, 300.0, 0.7, 60.0, 2e-5, 0.3, 33, 20, 1.3)]:
        c = G.CTF(L300, df, Cs, Ca, Pa, dd, tc, phi)
        numpy.testing.assert_array_equal(
            c.get_ctf(w, h, ps), O.ctf_grid(L300, df, Cs, Ca, Pa, dd, tc,
                                            phi, w, h, ps))
        simple = O.ctf_simple_grid(L300, df, Cs, Ca, Pa, phi, w, h, ps)
        # cos / sin of the phase, as ctf.h:285-303 (not the complex exp)
        numpy.testing.assert_allclose(c.get_ctf_simple_real(w, h, ps),
                                      simple.real, atol=1e-15)
        numpy.testing.assert_allclose(c.get_ctf_simple_imag(w, h, ps),
                                      simple.imag, atol=1e-15)
        q, t = G.get_q_and_theta(w, h, ps)
        numpy.testing.assert_allclose(c.get_ctf_simple_imag(q, t),
                                      simple.imag, atol=1e-15)


def test_ctf_tool_commands(monkeypatch, tmp_path):
    from guanaco_amd import command_line as cl
    seen = {}
    monkeypatch.setattr(cl.plot, "generate_ctf", lambda **kw: seen.update(kw))
    cl.generate_ctf(["--size", "64,32", "--ndf", "3", "--sdf", "100", "--px",
                     "1.5", "--df", "20000", "--recentre", "1"])
    assert seen["image_size"] == [64, 32] and seen["pixel_size"] == 1.5
    assert seen["num_defocus"] == 3 and seen["step_defocus"] == 100
    assert seen["recentre"] is True and seen["defocus"] == 20000.0
    assert seen["source_spread"] == 0.02 and seen["chromatic_aberration"] == 2.7
    assert seen["filename"] == "ctf.mrc"
    with pytest.raises(SystemExit):
        cl.generate_ctf([])   # --size is required
    out = tmp_path / "ctf.png"
    cl.plot_ctf(["-o", str(out), "--df", "20000", "--envelope", "all",
                 "--component", "abs"])
    assert out.stat().st_size > 1000
