"""Device-side CTF images (gx_ctf_grid) against the oracle's FP64 CTF, and
generate_ctf / the guanaco.generate_ctf command end to end."""
import math

import numpy
import pytest

from guanaco_amd import _capi as C
from guanaco_amd import command_line, mrc
from oracle import oracle as O

pytestmark = pytest.mark.gpu

L300 = O.electron_wavelength(300000)


@pytest.mark.parametrize("w,h,ps,Ca,Pa,dd,tc,phi", [
    (100, 100, 2.0, 0.0, 0.0, 1.0, 1e-4, 0.1),
    (64, 48, 1.0, 300.0, 0.7, 60.0, 2e-5, 0.3),
    (33, 20, 1.3, 0.0, 0.0, 0.0, 0.0, 0.0),
])
def test_device_ctf_grid_matches_oracle(capi, w, h, ps, Ca, Pa, dd, tc, phi):
    dfs = numpy.array([12000.0, 20000.0, 31000.5])
    full = C.ctf_grid(w, h, ps, dfs, L300, Cs=2.7e7, Ca=Ca, Pa=Pa, dd=dd,
                      theta_c=tc, phi=phi)
    simple = C.ctf_grid(w, h, ps, dfs, L300, Cs=2.7e7, Ca=Ca, Pa=Pa, dd=dd,
                        theta_c=tc, phi=phi, envelope=False)
    for k, df in enumerate(dfs):
        e = O.ctf_grid(L300, df, 2.7e7, Ca, Pa, dd, tc, phi, w, h, ps)
        # FP64 on both sides; the device result is rounded to complex64
        assert numpy.abs(full[k] - e).max() < 2e-7
        e = O.ctf_simple_grid(L300, df, 2.7e7, Ca, Pa, phi, w, h, ps)
        assert numpy.abs(simple[k] - e).max() < 2e-7


@pytest.mark.parametrize("w,h", [(64, 48), (33, 21)])
def test_device_ctf_recentre_is_fftshift(capi, w, h):
    a = C.ctf_grid(w, h, 1.5, [20000.0], L300, Cs=2.7e7, dd=30.0,
                   theta_c=2e-5)
    b = C.ctf_grid(w, h, 1.5, [20000.0], L300, Cs=2.7e7, dd=30.0,
                   theta_c=2e-5, recentre=True)
    numpy.testing.assert_array_equal(b[0], numpy.fft.fftshift(a[0]))


def test_device_ctf_batches_of_defoci(capi):
    # 4096^2 complex64 = 128 MiB per image: three defoci take two launches
    dfs = [10000.0, 20000.0, 30000.0]
    many = C.ctf_grid(4096, 4096, 1.0, dfs, L300, Cs=2.7e7, dd=20.0,
                      theta_c=2e-5)
    for k, df in enumerate(dfs):
        one = C.ctf_grid(4096, 4096, 1.0, [df], L300, Cs=2.7e7, dd=20.0,
                         theta_c=2e-5)
        numpy.testing.assert_array_equal(many[k], one[0])
    row = O.ctf_grid(L300, dfs[2], 2.7e7, 0, 0, 20.0, 2e-5, 0.0, 4096, 2,
                     1.0)[0]
    # the first row of the 4096 x 2 grid has dy = 0, as row 0 of 4096 x 4096
    assert numpy.abs(many[2, 0] - row).max() < 2e-7


def test_generate_ctf_command(capi, tmp_path):
    out = tmp_path / "ctf.mrc"
    command_line.generate_ctf(["-o", str(out), "--size", "48,64", "--px",
                               "1.5", "--df", "20000", "--ndf", "3", "--sdf",
                               "500", "--Cs", "2.7", "--phi", "0.2"])
    with mrc.open(out) as f:
        assert f.data.shape == (3, 48, 64) and f.data.dtype == numpy.complex64
        assert int(f.header["mode"]) == 4
        assert float(f.voxel_size["x"]) == 1.5
        spread = 2.7e7 * math.sqrt(0.33e-6 ** 2 + (2 * 0.33e-6) ** 2 +
                                   0.8e-6 ** 2)
        for i, df in enumerate([19500.0, 20000.0, 20500.0]):
            e = O.ctf_grid(L300, df, 2.7e7, 0, 0, spread, 0.02 / 1e3, 0.2,
                           64, 48, 1.5)
            assert numpy.abs(f.data[i] - e).max() < 2e-7
