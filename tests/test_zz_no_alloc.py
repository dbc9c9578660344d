"""gx_recon_noalloc through the C ABI: the allocation-free form of gx_recon
(device buffers in and out, scratch in a device global) gives gx_recon's
result bit for bit, on two calls in a row (the second smaller), and refuses
sizes beyond its scratch."""
import ctypes

import numpy
import pytest

from guanaco_amd import _capi as C

pytestmark = pytest.mark.gpu


def _lib():
    return C.lib().gx_recon_noalloc  # argtypes from _capi.SIGNATURES


def _case(N, T, ndf, H, W, seed):
    rng = numpy.random.default_rng(seed)
    angles = numpy.deg2rad(numpy.linspace(-60, 60, T)).astype(numpy.float32)
    sino = rng.standard_normal((ndf, T, N)).astype(numpy.float32)
    mn, mx = (-150.0, 150.0) if ndf > 1 else (0.0, 0.0)
    cfg = C.recon_cfg(N, T, ndf, H, W, 1.5, N / 2.0 - 0.3, mn, mx, angles)
    return cfg, sino


def _with_recon(cfg, sino, H, W):
    out = numpy.empty((H, W), numpy.float32)
    C.call("gx_recon", ctypes.byref(cfg), C.f32p(sino), C.f32p(out))
    return out


def test_noalloc_equals_recon_twice_second_smaller(capi):
    f = _lib()
    stream = C.Stream()
    for N, T, ndf, H, W, seed in [(256, 61, 3, 128, 256, 1),
                                  (100, 21, 1, 40, 100, 2)]:
        print("gx_recon_noalloc N=%d T=%d ndf=%d" % (N, T, ndf), flush=True)
        cfg, sino = _case(N, T, ndf, H, W, seed)
        want = _with_recon(cfg, sino, H, W)
        d_sino = C.DeviceArray.from_host(sino)
        d_rec = C.DeviceArray((H, W))
        C.call("gx_fill_u32", d_rec.ptr, 0x7fc00000, H * W, None)
        C.call("gx_device_synchronize")
        C.check(f(ctypes.byref(cfg), d_sino.ptr, d_rec.ptr, stream.handle))
        got = d_rec.to_host()
        numpy.testing.assert_array_equal(got, want)
        # the input is not modified (the reference filters a copy)
        numpy.testing.assert_array_equal(d_sino.to_host(), sino)
        d_sino.free()
        d_rec.free()
    stream.destroy()


def test_noalloc_refuses_sizes_beyond_its_scratch(capi):
    f = _lib()
    N, T, ndf = 4096, 121, 40  # 79 MB of filtered rows
    angles = numpy.zeros(T, numpy.float32)
    cfg = C.recon_cfg(N, T, ndf, 8, 8, 1.0, N / 2.0, -10.0, 10.0, angles)
    d = C.DeviceArray((16,))
    rc = f(ctypes.byref(cfg), d.ptr, d.ptr, None)
    assert rc == C.GX_EINVAL
    assert b"scratch" in C.lib().gx_last_error()
