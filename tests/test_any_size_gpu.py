"""CTF correction at sizes the power-of-two kernels do not take (Bluestein
transforms on power-of-two FFTs): gx_corr, the fused volume path and the
file driver against the oracle, whose FFT is a mixed-radix DFT of any
length."""
import math

import numpy
import pytest

import guanaco_amd as G
from guanaco_amd import mrc
from oracle import oracle as O

pytestmark = pytest.mark.gpu

L300 = O.electron_wavelength(300000)


def rel_l2(a, b):
    a = numpy.asarray(a, numpy.float64)
    b = numpy.asarray(b, numpy.float64)
    return float(numpy.linalg.norm(a - b) / numpy.linalg.norm(b))


@pytest.mark.parametrize("X,Y,ndf", [
    (30, 20, 1), (33, 17, 3), (100, 64, 2), (64, 48, 2), (1, 5, 1),
    (4, 2, 1), (600, 250, 2),
    # beyond 4096: 16384-point Bluestein rows (one slot per workgroup) and
    # columns
    (4100, 6, 1), (10, 5000, 2),
])
def test_corr_any_size_matches_oracle(capi, X, Y, ndf):
    rng = numpy.random.default_rng(X * 1000 + Y)
    img = (rng.standard_normal((Y, X)) + 3.0).astype(numpy.float32)
    ctf = numpy.stack([
        O.ctf_simple_grid(L300, 15000.0 + 700.0 * k, 2.7e7, 0, 0,
                          math.asin(0.1), X, Y, 1.3)
        for k in range(ndf)]).astype(numpy.complex64)
    got = numpy.zeros((ndf, Y, X), numpy.float32)
    G.corr(img, ctf, got, "gpu")
    expect = O.corr(img, ctf)
    assert rel_l2(got, expect) < 1e-5, (X, Y)


def test_corr_any_size_closed_form(capi):
    # df = Cs = 0, phi = 90 degrees: H = exp(i pi / 2) at every frequency,
    # so the flip -sign(Im H) is -1 everywhere (correct.h:30-39) and the
    # unnormalised transforms give out = -X Y img up to rounding
    X, Y = 45, 27
    img = numpy.random.default_rng(3).standard_normal((Y, X)).astype(
        numpy.float32)
    ctf = O.ctf_simple_grid(L300, 0.0, 0.0, 0, 0, math.pi / 2, X, Y,
                            1.0).astype(numpy.complex64)
    got = numpy.zeros((Y, X), numpy.float32)
    G.corr(img, ctf, got, "gpu")
    assert rel_l2(got, -(X * Y) * img.astype(numpy.float64)) < 1e-6


@pytest.mark.parametrize("X,Y,ndf", [(96, 40, 3), (100, 50, 1), (64, 30, 2),
                                     (5000, 6, 2)])
def test_fused_volume_path_any_size(capi, X, Y, ndf):
    T = 9
    rng = numpy.random.default_rng(X + Y + ndf)
    proj = (rng.standard_normal((T, Y, X)) + 4.0).astype(numpy.float32)
    angles = numpy.deg2rad(numpy.linspace(-60, 60, T)).astype(numpy.float32)
    kw = dict(pixel_size=1.5, energy=300, defocus=25000.0, num_defocus=ndf,
              spherical_aberration=2.7, phase_shift=5.7392)
    got = G.reconstruct_projections(proj, angles, depth=24, **kw)
    centre = O.get_centre(proj.shape, None, False)
    c, mn, mx = O.correct_projections(proj, centre, 1.5, 300, 25000.0, ndf,
                                      None, 2.7, 0, 0, 5.7392)
    expect = O.reconstruct(c, angles, None, 1.5, mn, mx, depth=24)
    assert rel_l2(got, expect) < 1e-5
    c64, _, _ = O.correct_projections(proj, centre, 1.5, 300, 25000.0, ndf,
                                      None, 2.7, 0, 0, 5.7392,
                                      precision="f64")
    e64 = O.reconstruct(c64, angles, None, 1.5, mn, mx, depth=24,
                        precision="f64")
    r = G.reconstruct_projections(proj, angles, depth=24, storage="f16", **kw)
    assert rel_l2(r, e64) < 1e-3


def test_reconstruct_file_any_size(capi, tmp_path):
    T, Y, X = 7, 6, 48
    rng = numpy.random.default_rng(35)
    proj = (rng.standard_normal((T, Y, X)) + 5.0).astype(numpy.float32)
    ext = numpy.zeros(T, mrc.FEI1_EXTENDED_HEADER_DTYPE)
    ext["Metadata size"] = 768
    ext["Alpha tilt"] = numpy.linspace(-45.0, 45.0, T)
    src, out = tmp_path / "t.mrc", tmp_path / "r.mrc"
    mrc.write(src, proj, voxel_size=1.5, extended_header=ext, exttyp=b"FEI1")
    G.reconstruct_file(str(src), str(out), energy=300, defocus=25000.0,
                       num_defocus=2, spherical_aberration=2.7,
                       phase_shift=5.7392, device="gpu")
    angles = numpy.deg2rad(numpy.linspace(-45.0, 45.0, T)).astype(
        numpy.float32)
    centre = O.get_centre(proj.shape, None, False)
    c, mn, mx = O.correct_projections(proj, centre, 1.5, 300, 25000.0, 2,
                                      None, 2.7, 0, 0, 5.7392)
    expect = O.reconstruct(c, angles, None, 1.5, mn, mx)
    with mrc.open(out) as f:
        assert rel_l2(f.data, expect) < 1e-5
