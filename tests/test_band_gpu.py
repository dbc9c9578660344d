"""Banded plans: the filtered slabs of a band of detector rows at a time.

A banded plan must give the volume of the plan that holds every row: bit
for bit where the filter of a row does not depend on the band (float32
slabs, and reduced storage of CTF-corrected input, whose per-tilt
normalisation is taken over whole projections); within FP16 rounding for
reduced storage of row-local input, whose normalisation is per band. The
host path (reconstruct_projections) falls back to bands when the slabs of
every row do not fit; here the band is forced through the C ABI.
"""
import numpy
import pytest

from guanaco_amd import _capi as C
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def tilts(T):
    return numpy.deg2rad(numpy.linspace(-60, 60, T)).astype(numpy.float32)


def run_plan(proj, angles, X, Y, Z, ndf, storage, band_rows, ctf, layout,
             ps=1.5, df=30000.0):
    kw = dict(num_defocus=ndf, pixel_size=ps, centre=X / 2.0,
              storage=storage, input_layout=layout, band_rows=band_rows)
    if ctf:
        centre = numpy.full(Y, X / 2.0, numpy.float32)
        mn, mx, st, n = O.get_defocus_steps(df, ndf, None, X, centre, ps)
        kw.update(defocus=numpy.array([mn + d * st for d in range(n)]),
                  min_defocus=mn - df, max_defocus=mx - df,
                  wavelength=O.electron_wavelength(300000), Cs=2.7e7,
                  phi=0.1)
    elif ndf > 1:
        kw.update(min_defocus=-40.0, max_defocus=40.0)
    plan = C.VolumePlan(X, Y, angles, Z, **kw)
    d_in = C.DeviceArray.from_host(proj)
    d_vol = C.DeviceArray((Y, Z, X))
    C.call("gx_fill_u32", d_vol.ptr, 0x7fc00000, Y * Z * X, None)
    plan.run(d_in, d_vol)
    C.call("gx_device_synchronize")
    band = plan.band()
    out = d_vol.to_host()
    plan.destroy()
    return out, band


BAND_CASES = [
    # X, Y, T, ndf, Z, band, ctf, storage
    (64, 96, 7, 3, 24, 32, True, "f32"),     # three whole bands
    (64, 100, 7, 3, 24, 40, True, "f16"),    # band 64: one whole, one partial
    (60, 70, 5, 2, 16, 32, True, "f32"),     # Bluestein rows and columns
    (256, 80, 5, 1, 16, 32, True, "bf16"),   # four-transform row pass
    (96, 72, 9, 1, 16, 32, False, "f32"),    # row-local filter
]


@pytest.mark.parametrize("case", BAND_CASES)
def test_banded_plan_equals_whole_plan(capi, case):
    X, Y, T, ndf, Z, band, ctf, storage = case
    print("case", case, flush=True)
    rng = numpy.random.default_rng(X * Y + T)
    layout = C.INPUT_PROJECTIONS
    if ctf or ndf == 1:
        inp = (rng.standard_normal((T, Y, X)) + 3.0).astype(numpy.float32)
    else:
        inp = rng.standard_normal((T, ndf, Y, X)).astype(numpy.float32)
        layout = C.INPUT_CORRECTED
    angles = tilts(T)
    whole, b0 = run_plan(inp, angles, X, Y, Z, ndf, storage, 0, ctf, layout)
    banded, b1 = run_plan(inp, angles, X, Y, Z, ndf, storage, band, ctf,
                          layout)
    assert b0[0] == Y and b1[0] == (band + 31) // 32 * 32 < Y
    assert numpy.isfinite(whole).all()
    numpy.testing.assert_array_equal(banded, whole)


def test_banded_reduced_row_local_within_rounding(capi):
    X, Y, T, Z = 64, 64, 9, 16
    rng = numpy.random.default_rng(5)
    inp = rng.standard_normal((T, Y, X)).astype(numpy.float32)
    inp[:, 32:] *= 50.0  # the second band's rows are larger
    angles = tilts(T)
    whole, _ = run_plan(inp, angles, X, Y, Z, 1, "f16", 0, False,
                        C.INPUT_PROJECTIONS)
    banded, _ = run_plan(inp, angles, X, Y, Z, 1, "f16", 32, False,
                         C.INPUT_PROJECTIONS)
    expect = O.reconstruct(inp.astype(numpy.float64), angles, X / 2.0, 1.0,
                           depth=Z, precision="f64")
    for v in (whole, banded):
        err = numpy.linalg.norm(v - expect) / numpy.linalg.norm(expect)
        assert err < 1e-3, err


def test_banded_plan_refuses_whole_stage_calls(capi):
    X, Y, T, Z = 64, 64, 5, 8
    angles = tilts(T)
    plan = C.VolumePlan(X, Y, angles, Z, band_rows=32)
    d_in = C.DeviceArray((T, Y, X))
    d_vol = C.DeviceArray((Y, Z, X))
    with pytest.raises(C.GxError, match="rows at a time"):
        plan.filter(d_in)
    with pytest.raises(C.GxError, match="not in the filtered band"):
        plan.backproject_rows(0, 32, d_vol)
    plan.filter_rows(d_in, 32, 64)
    with pytest.raises(C.GxError, match="not in the filtered band"):
        plan.backproject_rows(0, 32, d_vol)
    with pytest.raises(C.GxError):
        plan.filter_rows(d_in, 0, 64)  # wider than the band
    plan.destroy()


def test_host_path_ctf_reads_only_the_projections(capi):
    # The fused host path copies the (T, Y, X) projections, not num_defocus
    # times as many bytes, and checks the array sizes it is given.
    import guanaco_amd as G
    X, Y, T, Z = 64, 16, 7, 16
    rng = numpy.random.default_rng(11)
    proj = (rng.standard_normal((T, Y, X)) + 4.0).astype(numpy.float32)
    angles = tilts(T)
    vol = G.reconstruct_projections(proj, angles, pixel_size=1.5, depth=Z,
                                    defocus=30000.0, num_defocus=8,
                                    spherical_aberration=2.7)
    centre = numpy.full(Y, X / 2.0, numpy.float32)
    corrected, mn, mx = O.correct_projections(proj, centre, 1.5, 300, 30000.0,
                                              8, None, 2.7, 0, 0, 0)
    expect = O.reconstruct(corrected, angles, X / 2.0, 1.5, mn, mx, depth=Z)
    err = numpy.linalg.norm(vol - expect) / numpy.linalg.norm(expect)
    assert err < 1e-5, err
