"""C-ABI checks that need no GPU: the library loads, exports every function
include/guanaco_hip.h declares, computes the per-tilt constants bit for bit
like the reference, and rejects invalid input before touching a device."""
import ctypes
import os
import re

import numpy
import pytest

from guanaco_amd import _capi
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(gx_\w+)\s*\(", text))
    return names


def test_library_exports_every_declared_function():
    lib = _capi.lib()
    declared = header_functions()
    assert len(declared) >= 30
    for name in declared:
        assert hasattr(lib, name), name
    assert declared == set(_capi.SIGNATURES), (
        declared ^ set(_capi.SIGNATURES))


def test_device_count_without_gpu_is_not_an_error():
    assert _capi.device_count() >= 0


@pytest.mark.parametrize("centre", [0.0, 32.0, 511.5, 750.0, 1023.25])
def test_tilt_geometry_bitwise_vs_oracle(centre):
    rng = numpy.random.default_rng(int(centre * 4))
    angles = numpy.concatenate([
        numpy.deg2rad(numpy.arange(-60, 61, 1.0)),
        rng.uniform(-numpy.pi, numpy.pi, 200),
        [0.0, numpy.pi / 2, -numpy.pi / 2, numpy.pi, 1e-8],
    ]).astype(numpy.float32)
    cfg = _capi.recon_cfg(1, angles.size, 1, 1, 1, 1.0, centre, 0, 0, angles)
    got = numpy.empty((angles.size, 4), numpy.float32)
    _capi.call("gx_tilt_geometry", ctypes.byref(cfg), _capi.f32p(got))
    expect = O.tilt_coefficients(angles, centre)
    numpy.testing.assert_array_equal(got.view(numpy.uint32),
                                     expect.view(numpy.uint32))


def _bad(**kw):
    base = dict(N=16, T=3, ndf=1, H=16, W=16, pixel_size=1.0, centre=8.0,
                min_defocus=0.0, max_defocus=0.0,
                angles=numpy.zeros(3, numpy.float32))
    base.update(kw)
    return _capi.recon_cfg(**base)


@pytest.mark.parametrize("kw,text", [
    (dict(N=0), "num_pixels > 0"),
    (dict(T=0), "num_angles > 0"),
    (dict(ndf=0), "num_defocus > 0"),
    (dict(H=0), "grid_height > 0"),
    (dict(pixel_size=0.0), "pixel_size > 0"),
    (dict(min_defocus=1.0, max_defocus=0.0), "min_defocus <= c->max_defocus"),
])
def test_recon_rejects_invalid_config(kw, text):
    # config.h:66-75, checked at reconstructor.cpp:10-11 (guanaco::Error)
    cfg = _bad(**kw)
    sino = numpy.zeros(16 * 3, numpy.float32)
    rec = numpy.zeros(16 * 16, numpy.float32)
    with pytest.raises(_capi.GxError) as e:
        _capi.call("gx_recon", ctypes.byref(cfg), _capi.f32p(sino),
                   _capi.f32p(rec))
    assert e.value.code == _capi.GX_EINVAL
    assert "ASSERT" in str(e.value) and text in str(e.value)


def test_corr_rejects_unsupported_sizes():
    # any size up to 8192 x 8192 is taken (power-of-two kernels or Bluestein
    # transforms); wider rows are refused before any device work
    img = numpy.zeros((2, 8200), numpy.float32)
    ctf = numpy.zeros((2, 8200), numpy.complex64)
    out = numpy.zeros((2, 8200), numpy.float32)
    with pytest.raises(_capi.GxError) as e:
        _capi.call("gx_corr", _capi.f32p(img),
                   ctf.view(numpy.float32).ctypes.data_as(_capi._f32p),
                   _capi.f32p(out), 8200, 2, 1, -1)
    assert e.value.code == _capi.GX_EUNSUPPORTED


def test_plan_footprint_is_host_only_arithmetic():
    angles = numpy.deg2rad(numpy.arange(-60, 61, 1.0)).astype(numpy.float32)
    f32 = _capi.plan_bytes(num_pixels=1024, num_rows=1024, depth=256,
                           angles=angles, storage="f32")
    f16 = _capi.plan_bytes(num_pixels=1024, num_rows=1024, depth=256,
                           angles=angles, storage="f16")
    slab32 = 121 * 1024 * 1024 * 4
    # the staged backprojector's window table: one int2 per (32 x 8 tile,
    # tilt padded to 8)
    win = (1024 // 32) * (256 // 8) * 128 * 8
    assert slab32 + win <= f32 < slab32 + win + (1 << 16)
    assert slab32 // 2 + win <= f16 < slab32 // 2 + win + (1 << 16)
    ctf = _capi.plan_bytes(num_pixels=1024, num_rows=1024, depth=256,
                           angles=angles, storage="f16", num_defocus=4,
                           defocus=numpy.linspace(1e4, 2e4, 4))
    assert ctf > 4 * f16 - (1 << 20)
    # non-power-of-two corrections carry their Bluestein tables
    odd = _capi.plan_bytes(num_pixels=1000, num_rows=1000, depth=256,
                           angles=angles, defocus=[1e4], storage="f16")
    assert odd > 121 * 1000 * 1000 * 2
    wide = _capi.plan_bytes(num_pixels=5000, num_rows=1024, depth=256,
                            angles=angles, defocus=[1e4], storage="f16")
    assert wide > 121 * 1024 * 5000 * 2
    with pytest.raises(_capi.GxError):
        _capi.plan_bytes(num_pixels=8200, num_rows=1024, depth=256,
                         angles=angles, defocus=[1e4])


def test_tile_choice_file_round_trip_and_errors(tmp_path):
    from guanaco_amd import _capi as C
    C.tune_clear()
    line = "gfx950:256 1024 1024 121 1 256 1024 1 0 1 1 1 32"
    good = tmp_path / "tiles.txt"
    good.write_text("# a comment\n\n" + line + "\n")
    C.tune_load(str(good))
    out = tmp_path / "saved.txt"
    C.tune_save(str(out))
    kept = [l for l in out.read_text().splitlines() if not l.startswith("#")]
    assert kept == [line]
    bad = tmp_path / "bad.txt"
    for text in ("gfx950:256 1 2 3 1 7\n", line.replace(" 32", " 12") + "\n",
                 line[:-4] + " 5 32\n"):
        bad.write_text(text)
        with pytest.raises(C.GxError, match="not a tile choice"):
            C.tune_load(str(bad))
    with pytest.raises(C.GxError, match="cannot read"):
        C.tune_load(str(tmp_path / "missing.txt"))
    C.tune_clear()
    C.tune_save(str(out))
    assert all(l.startswith("#") for l in out.read_text().splitlines())
