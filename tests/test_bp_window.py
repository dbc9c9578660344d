"""The bounds the LDS-staged backprojector relies on, checked on the CPU with
the kernel's float32 arithmetic (numpy float32 operations are correctly
rounded, as the kernel's are; it is compiled without FMA contraction).

For every 32 x 8 voxel tile and tilt:
- the smallest tap index i0 over the tile is the smallest over its four
  corners (each operation producing p is monotone in x and in z);
- every i0 lies within kWPOS - 2 = 38 positions of that corner minimum, so
  both taps are inside the kWPOS = 40 staged positions;
- the slab indices k of the tile's voxels lie within the slot count the host
  picks, min(ndf, ceil(31.9 |dscale|) + 1).
Both the exact (FP32 storage) and the fused multiply-add (FP16/BF16
storage) forms of p are checked; a float32 fused multiply-add is emulated in
float64 (the product is exact there, the sum is rounded twice).
Coefficients come from the oracle's restatement of reconstructor.cpp:36-47.
"""
import math

import numpy
import pytest

from oracle import oracle as O

KTX, KTZ, KWPOS = 32, 8, 40
f32 = numpy.float32


def voxel_coord(i, n):
    # (float)((double)i - 0.5 * (double)n + 0.5), reconstructor.cpp:56-58
    return (numpy.asarray(i, numpy.float64) - 0.5 * n + 0.5).astype(f32)


def fma32(a, b, c):
    return (numpy.asarray(a, numpy.float64) * numpy.asarray(b, numpy.float64)
            + numpy.asarray(c, numpy.float64)).astype(f32)


def tap_index(x, z, cf, exact=True):
    c, s, off = f32(cf[0]), f32(cf[1]), f32(cf[2])
    if exact:
        p = (x * c - z * s) + off
        p = p - f32(0.5)
    else:  # the reduced forms' table holds offset - 1/2
        p = fma32(x, c, fma32(-z, s, off - f32(0.5)))
    return numpy.floor(p).astype(numpy.int64)


def slab_of(x, z, cf, dscale, doffset, ndf):
    c, s = f32(cf[0]), f32(cf[1])
    h = (-x) * s - z * c
    d = h * f32(dscale) + f32(doffset)
    n = numpy.rint(d)
    n = numpy.where(d - n == f32(0.5), n + f32(1.0), n)
    return numpy.clip(n.astype(numpy.int64), 0, ndf - 1)


def host_slots(ndf, dscale):
    # gx_capi.hip bp_slots before rounding up to a compiled form
    if ndf <= 1:
        return 1
    span = 31.9 * abs(float(dscale))
    if not math.isfinite(span) or span >= ndf:
        return ndf
    return min(ndf, math.ceil(span) + 1)


def tiles(a, W, Z, fill):
    """(Z, W) -> (nbz, nbx, KTZ * KTX) with inactive voxels set to fill."""
    Zp, Wp = -(-Z // KTZ) * KTZ, -(-W // KTX) * KTX
    out = numpy.full((Zp, Wp), fill, a.dtype)
    out[:Z, :W] = a
    return out.reshape(Zp // KTZ, KTZ, Wp // KTX, KTX).transpose(
        0, 2, 1, 3).reshape(Zp // KTZ, Wp // KTX, KTZ * KTX)


def corner_min(fn, W, Z):
    """Per tile: min of fn over the four corners of its active voxels."""
    nbz, nbx = -(-Z // KTZ), -(-W // KTX)
    bx = numpy.arange(nbx)
    bz = numpy.arange(nbz)
    x0 = voxel_coord(bx * KTX, W)
    x1 = voxel_coord(numpy.minimum(bx * KTX + KTX, W) - 1, W)
    z0 = voxel_coord(bz * KTZ, Z)
    z1 = voxel_coord(numpy.minimum(bz * KTZ + KTZ, Z) - 1, Z)
    m = None
    for xc in (x0, x1):
        for zc in (z0, z1):
            v = fn(xc[None, :], zc[:, None])
            m = v if m is None else numpy.minimum(m, v)
    return m


GEOMETRIES = [
    # N, W, Z, centre, tilt range (deg), T
    (1024, 1024, 256, 512.0, (-60, 60), 121),
    (4096, 4096, 64, 2048.0, (-60, 60), 9),
    (300, 333, 77, 151.3, (-180, 180), 73),
    (64, 200, 64, 31.7, (-90, 90), 37),
]


@pytest.mark.parametrize("exact", [True, False])
@pytest.mark.parametrize("geom", GEOMETRIES)
def test_staged_window_holds_every_tap(geom, exact):
    N, W, Z, centre, (a0, a1), T = geom
    rng = numpy.random.default_rng(N + T)
    angles = numpy.deg2rad(numpy.linspace(a0, a1, T)) + rng.uniform(
        -0.01, 0.01, T)
    coef = O.tilt_coefficients(angles.astype(f32), centre)
    x = voxel_coord(numpy.arange(W), W)[None, :]
    z = voxel_coord(numpy.arange(Z), Z)[:, None]
    big = numpy.iinfo(numpy.int64).max
    worst = 0
    for t in range(T):
        i0 = tap_index(x, z, coef[t], exact)
        wb = corner_min(lambda xc, zc: tap_index(xc, zc, coef[t], exact),
                        W, Z)
        lo = tiles(i0, W, Z, big).min(axis=2)
        hi = tiles(i0, W, Z, -big).max(axis=2)
        numpy.testing.assert_array_equal(lo, wb)
        worst = max(worst, int((hi - wb).max()))
    assert worst <= 32, worst
    assert worst <= KWPOS - 2


@pytest.mark.parametrize("ndf,ps,rng_df", [
    (16, 1.5, 3072.0), (32, 1.0, 4096.0), (5, 2.0, 600.0), (8, 1.0, 1024.0),
    (16, 1.0, 512.0), (3, 1.0, 20.0), (16, 1.0, 64.0)])
def test_slab_span_within_host_slot_count(ndf, ps, rng_df):
    N = W = 512
    Z, T = 128, 41
    centre = N / 2.0
    mn, mx = -rng_df / 2, rng_df / 2
    dscale = f32(f32(ndf) * f32(ps) / (f32(mx) - f32(mn)))
    doffset = f32(-dscale * (f32(mn) / f32(ps)))
    angles = numpy.deg2rad(numpy.linspace(-60, 60, T)).astype(f32)
    coef = O.tilt_coefficients(angles, centre)
    x = voxel_coord(numpy.arange(W), W)[None, :]
    z = voxel_coord(numpy.arange(Z), Z)[:, None]
    slots = host_slots(ndf, dscale)
    big = numpy.iinfo(numpy.int64).max
    for t in range(T):
        k = slab_of(x, z, coef[t], dscale, doffset, ndf)
        kb = corner_min(
            lambda xc, zc: slab_of(xc, zc, coef[t], dscale, doffset, ndf),
            W, Z)
        lo = tiles(k, W, Z, big).min(axis=2)
        hi = tiles(k, W, Z, -big).max(axis=2)
        numpy.testing.assert_array_equal(lo, kb)
        assert int((hi - kb).max()) < slots, (t, slots)
