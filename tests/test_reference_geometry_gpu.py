"""The reference's own test geometries through both GPU routes.

The four geometries of the reference's tests/data/*.yaml (detector 1500 or
1750 x 200, 90 tilts, rotation axes at 500, 750 and 875 pixels, a 200-pixel
object centred or off centre, df 20000 A, Cs 2.7 mm, 1 A/px): the data files
the reference's tests open (*.mrc) are not in its repository, so the images
here are analytic projections of a uniform sphere at the yaml's position,
with the CTF at its df applied in Fourier space (numpy). Tilts: 90 evenly
spaced over +-60 degrees. Detectors that are not powers of two take the
Bluestein correction and the zero-padded filter.

Each geometry runs through the fused route (reconstruct_projections, 8
defoci) and the reference's two steps (correct_projections, then
reconstruct with the returned range), in FP32 against the FP32 oracle chain
(1e-5) and with FP16 slabs against the FP64 chain (1e-3), at sampled slices
of a 300-row-deep volume (the depth the object's rows need).
"""
import math
import os

import numpy
import pytest

import guanaco_amd as G
from oracle import oracle as O
from oracle import sampled

pytestmark = pytest.mark.gpu

# tests/data/*.yaml of the reference: xsize, ysize, nangles, centre,
# position (x, y, z), object size, df, Cs, pixel size
GEOMETRIES = {
    "centred_axis_centred_object": (1500, 200, 90, 750, (750, 100, 750)),
    "centred_axis_off_centre_object": (1750, 200, 90, 875, (375, 100, 875)),
    "off_centre_axis_centred_object": (1500, 200, 90, 500, (500, 100, 750)),
    "off_centre_axis_off_centre_object": (1750, 200, 90, 750,
                                          (250, 100, 875)),
}
SIZE, DF, CS, PS = 200, 20000.0, 2.7, 1.0
NDF, DEPTH = 8, 300
ROWS = [0, 60, 100, 140, 199]


@pytest.fixture(scope="module")
def pool():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    p = sampled.make_pool(max(2, min(16, n)))
    yield p
    p.close()
    p.join()


def sphere_projections(X, Y, T, centre, position):
    """Line integrals of a unit-density sphere (radius SIZE / 2) centred at
    the volume position (x, y, z) in pixels, in the reconstruction geometry
    (detector position x cos t - z sin t + centre for the voxel at x, z
    relative to the grid centre), then the CTF sin(chi - phi) of df, Cs at
    300 kV applied to each image in Fourier space."""
    angles = numpy.deg2rad(numpy.linspace(-60.0, 60.0, T)).astype(
        numpy.float32)
    R = SIZE / 2.0
    px, py, pz = position
    x0, z0 = px - X / 2.0, pz - X / 2.0  # relative to the square grid centre
    u = numpy.arange(X) + 0.5
    v = numpy.arange(Y) + 0.5
    proj = numpy.zeros((T, Y, X))
    for t, th in enumerate(angles.astype(numpy.float64)):
        c = x0 * math.cos(th) - z0 * math.sin(th) + centre
        r2 = (u[None, :] - c) ** 2 + (v[:, None] - py) ** 2
        proj[t] = 2.0 * numpy.sqrt(numpy.clip(R * R - r2, 0.0, None))
    l = O.electron_wavelength(300000)
    h = O.ctf_simple_grid(l, DF, CS * 1e7, 0, 0, 0.0, X, Y, PS)
    ctf = -h.imag  # sin(chi - phi)
    proj = numpy.fft.ifft2(numpy.fft.fft2(proj) * ctf[None]).real
    return (proj + 1.0).astype(numpy.float32), angles


@pytest.mark.parametrize("name", sorted(GEOMETRIES))
def test_reference_geometry_fused_and_two_step(capi, pool, name):
    X, Y, T, centre, position = GEOMETRIES[name]
    print("%s: %d x %d, %d tilts, axis %d" % (name, X, Y, T, centre),
          flush=True)
    proj, angles = sphere_projections(X, Y, T, centre, position)
    ref = sampled.reference_rows(pool, proj, angles, ROWS, DEPTH, PS, DF, NDF,
                                 0.0, CS * 1e7, fp32=True, centre=centre)
    kw = dict(centre=centre, pixel_size=PS, energy=300, defocus=DF,
              num_defocus=NDF, spherical_aberration=CS, depth=DEPTH)
    errs = {}
    for storage in ("f32", "f16"):
        fused = G.reconstruct_projections(proj, angles, storage=storage, **kw)
        errs["fused", storage] = fused[ROWS]
    centres = numpy.full(Y, centre, numpy.float32)  # one per row
    corrected, mn, mx = G.correct_projections(
        proj, centres, PS, 300, DF, NDF, None, CS, 0, 0, 0, device="gpu")
    o_range = ref["range"]
    assert abs(mn - o_range[0]) < 1e-6 and abs(mx - o_range[1]) < 1e-6
    for storage in ("f32", "f16"):
        vol = numpy.zeros((Y, DEPTH, X), numpy.float32)
        G.reconstruct(corrected, angles, vol, centre=centre, pixel_size=PS,
                      min_defocus=mn, max_defocus=mx, device="gpu",
                      storage=storage, depth=DEPTH)
        errs["two-step", storage] = vol[ROWS]
    for (route, storage), got in errs.items():
        e = sampled.rel_l2(got, ref["f32" if storage == "f32" else "f64"])
        print("%s %s %s: %.3g" % (name, route, storage, e))
        assert e < (1e-5 if storage == "f32" else 1e-3), (route, storage, e)
