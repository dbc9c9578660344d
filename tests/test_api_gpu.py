"""The drop-in API on the GPU (pybind11 -> C++ host -> C ABI -> gfx950)
against the CPU oracle chain: the same calls a user of the reference makes
with device="gpu"."""
import math

import numpy
import pytest

import guanaco_amd as G
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a = numpy.asarray(a, numpy.float64)
    b = numpy.asarray(b, numpy.float64)
    return float(numpy.linalg.norm(a - b) / numpy.linalg.norm(b))


def tilt_series(T, Y, X, seed, mean=5.0):
    rng = numpy.random.default_rng(seed)
    proj = (rng.standard_normal((T, Y, X)) + mean).astype(numpy.float32)
    angles = numpy.deg2rad(numpy.linspace(-60, 60, T)).astype(numpy.float32)
    return proj, angles


def test_recon_device_gpu_matches_oracle(capi):
    T, N = 11, 96
    rng = numpy.random.default_rng(1)
    sino = rng.standard_normal((3, T, N)).astype(numpy.float32)
    angles = numpy.deg2rad(numpy.linspace(-60, 60, T)).astype(numpy.float32)
    rec = numpy.zeros((N, N), numpy.float32)
    G.recon(sino, rec, angles, 47.0, 1.5, -60.0, 60.0, "gpu")
    expect = O.recon(sino, (N, N), angles, 47.0, 1.5, -60.0, 60.0)
    assert rel_l2(rec, expect) < 1e-5


def test_corr_device_gpu_matches_oracle(capi):
    Y, X = 64, 128
    img = numpy.random.default_rng(2).poisson(40, (Y, X)).astype(
        numpy.float32)
    ctf = G.get_ctf_array(X, Y, 2.0, 300, 19000.0, 3, 1000.0, 2.7, 0, 0, 5.7)
    out = numpy.zeros((3, Y, X), numpy.float32)
    G.corr(img, ctf, out, "gpu")
    expect = O.corr(img, ctf)
    for k in range(3):
        assert rel_l2(out[k], expect[k]) < 1e-5


@pytest.mark.parametrize("ndf", [1, 4])
def test_correct_then_reconstruct_matches_oracle_chain(capi, ndf):
    # reconstruct_file's FBP_CTF branch (reconstruct.py:348-388)
    T, Y, X = 9, 16, 64
    proj, angles = tilt_series(T, Y, X, 10 + ndf)
    corrected, mn, mx = G.correct_projections(
        proj, None, 1.5, 300, 25000.0, ndf, None, 2.7, 0, 0, 5.7,
        device="gpu")
    rec = G.reconstruct(corrected, angles, pixel_size=1.5, min_defocus=mn,
                        max_defocus=mx, device="gpu")
    centre = O.get_centre(proj.shape, None, False)
    c2, mn2, mx2 = O.correct_projections(proj, centre, 1.5, 300, 25000.0, ndf,
                                         None, 2.7, 0, 0, 5.7)
    assert (mn, mx) == (mn2, mx2)
    assert rel_l2(corrected, c2) < 1e-5
    expect = O.reconstruct(c2, angles, None, 1.5, mn2, mx2)
    assert rec.shape == (Y, X, X)
    assert rel_l2(rec, expect) < 1e-5


def test_fused_projection_path_matches_two_step_path(capi):
    T, Y, X, ndf = 11, 32, 128, 6
    proj, angles = tilt_series(T, Y, X, 21)
    kw = dict(pixel_size=1.0, energy=300, defocus=30000.0, num_defocus=ndf,
              spherical_aberration=2.7, phase_shift=5.7392)
    fused = G.reconstruct_projections(proj, angles, depth=32, **kw)
    centre = O.get_centre(proj.shape, None, False)
    c, mn, mx = O.correct_projections(proj, centre, 1.0, 300, 30000.0, ndf,
                                      None, 2.7, 0, 0, 5.7392)
    expect = O.reconstruct(c, angles, None, 1.0, mn, mx, depth=32)
    assert rel_l2(fused, expect) < 1e-5
    c64, _, _ = O.correct_projections(proj, centre, 1.0, 300, 30000.0, ndf,
                                      None, 2.7, 0, 0, 5.7392,
                                      precision="f64")
    e64 = O.reconstruct(c64, angles, None, 1.0, mn, mx, depth=32,
                        precision="f64")
    for storage, bound in (("f16", 1e-3), ("bf16", 4e-3)):
        r = G.reconstruct_projections(proj, angles, depth=32, storage=storage,
                                      **kw)
        assert rel_l2(r, e64) < bound, storage


def test_reconstruct_orders_weights_and_minus(capi):
    T, Y, X = 13, 10, 80
    proj, angles = tilt_series(T, Y, X, 5, mean=0.0)
    angles = (angles + numpy.random.default_rng(6).uniform(-0.01, 0.01, T)
              ).astype(numpy.float32)
    a = G.reconstruct(proj, angles, transform="minus", angular_weights=True,
                      device="gpu")
    e = O.reconstruct(proj, angles, transform="minus", angular_weights=True)
    assert rel_l2(a, e) < 1e-5
    sino = numpy.ascontiguousarray(numpy.swapaxes(proj, 0, 1))
    b = G.reconstruct(sino, angles, sinogram_order=True, device="gpu",
                      transform="minus", angular_weights=True)
    numpy.testing.assert_array_equal(a, b)
    # the caller's array is not modified (the reference scales a contiguous
    # slice in place)
    before = sino.copy()
    G.reconstruct(sino, angles, sinogram_order=True, device="gpu",
                  angular_weights=True)
    numpy.testing.assert_array_equal(sino, before)


def test_reconstruct_per_row_centres(capi):
    T, Y, X = 9, 20, 64
    proj, angles = tilt_series(T, Y, X, 8, mean=0.0)
    centre = numpy.linspace(30.0, 34.0, Y).astype(numpy.float32)
    rec = G.reconstruct(proj, angles, centre=centre, device="gpu")
    expect = O.reconstruct(proj, angles, centre=centre)
    assert rel_l2(rec, expect) < 1e-5
    e64 = O.reconstruct(proj, angles, centre=centre, precision="f64")
    for storage, bound in (("f16", 1e-3), ("bf16", 4e-3)):
        r = G.reconstruct(proj, angles, centre=centre, device="gpu",
                          storage=storage)
        assert rel_l2(r, e64) < bound, storage


def test_fused_path_with_per_row_centres(capi):
    T, Y, X, ndf = 11, 32, 128, 4
    proj, angles = tilt_series(T, Y, X, 22)
    centre = (64.0 + numpy.linspace(-3.0, 2.5, Y)).astype(numpy.float32)
    kw = dict(pixel_size=1.0, energy=300, defocus=30000.0, num_defocus=ndf,
              spherical_aberration=2.7, phase_shift=5.7392)
    got = G.reconstruct_projections(proj, angles, centre=centre, depth=32,
                                    **kw)
    c, mn, mx = O.correct_projections(proj, centre, 1.0, 300, 30000.0, ndf,
                                      None, 2.7, 0, 0, 5.7392)
    expect = O.reconstruct(c, angles, centre, 1.0, mn, mx, depth=32)
    assert rel_l2(got, expect) < 1e-5


def test_reconstruct_into_memmap_output(capi, tmp_path):
    T, Y, X = 7, 8, 32
    proj, angles = tilt_series(T, Y, X, 9, mean=0.0)
    out = numpy.memmap(tmp_path / "rec.dat", mode="w+", dtype="float32",
                       shape=(Y, X, X))
    G.reconstruct(proj, angles, reconstruction=out, device="gpu")
    expect = O.reconstruct(proj, angles)
    assert rel_l2(numpy.asarray(out), expect) < 1e-5


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_fused_path_4096_point_rows(capi, storage):
    # the widest rows: 4096-point corrections and 8192-point filter
    # transforms (one row pair per slot per workgroup)
    T, Y, X, ndf = 3, 8, 4096, 2
    proj, angles = tilt_series(T, Y, X, 41)
    kw = dict(pixel_size=1.0, energy=300, defocus=25000.0, num_defocus=ndf,
              spherical_aberration=2.7, phase_shift=5.7392)
    got = G.reconstruct_projections(proj, angles, depth=8, storage=storage,
                                    **kw)
    centre = O.get_centre(proj.shape, None, False)
    prec = "f32" if storage == "f32" else "f64"
    c, mn, mx = O.correct_projections(proj, centre, 1.0, 300, 25000.0, ndf,
                                      None, 2.7, 0, 0, 5.7392,
                                      precision=prec)
    expect = O.reconstruct(c, angles, None, 1.0, mn, mx, depth=8,
                           precision=prec)
    assert rel_l2(got, expect) < (1e-5 if storage == "f32" else 1e-3)
