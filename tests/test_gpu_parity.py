"""GPU parity: every kernel path through the C ABI against the CPU oracle.

Tolerances: the FP32 path within 1e-5 relative L2 of the FP32 oracle
(backprojection of an identical filtered sinogram bit-exact); the FP16/BF16
storage forms within the stated bound of the FP64 oracle.
"""
import ctypes
import math

import numpy
import pytest

from guanaco_amd import _capi as C
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a = numpy.asarray(a, numpy.float64)
    b = numpy.asarray(b, numpy.float64)
    return float(numpy.linalg.norm(a - b) / max(numpy.linalg.norm(b), 1e-300))


def tilts(T, span=60.0):
    if T == 1:
        return numpy.zeros(1, numpy.float32)
    return numpy.deg2rad(numpy.linspace(-span, span, T)).astype(numpy.float32)


# ------------------------------------------------------------- filter ------

@pytest.mark.parametrize("N", [8, 13, 64, 100, 256, 1024, 1500, 4096, 5760,
                               8192])
def test_filter_matches_oracle(capi, N):
    rng = numpy.random.default_rng(N)
    rows = rng.standard_normal((5, N)).astype(numpy.float32)
    out = numpy.empty_like(rows)
    C.call("gx_filter", C.f32p(rows), C.f32p(out), 5, N, -1)
    expect = O.filter_rows(rows, "f32")
    assert rel_l2(out, expect) < 1e-6, N


# ----------------------------------------------------- backprojection -----

BP_CASES = [
    # N, T, ndf, H, W, centre, (min_df, max_df), pixel
    (16, 1, 1, 16, 16, 8.0, (0.0, 0.0), 1.0),
    (64, 7, 1, 64, 64, 32.0, (0.0, 0.0), 1.0),
    (64, 61, 1, 16, 64, 31.3, (0.0, 0.0), 1.0),
    (100, 9, 3, 100, 100, 47.5, (-150.0, 150.0), 1.5),
    (64, 33, 5, 32, 80, 30.0, (-64.0, 64.0), 1.0),
    (128, 17, 16, 128, 128, 64.0, (-96.0, 96.0), 1.5),
]


@pytest.mark.parametrize("case", BP_CASES)
def test_project_bitexact_vs_oracle(capi, case):
    N, T, ndf, H, W, centre, (mn, mx), ps = case
    rng = numpy.random.default_rng(N * T + ndf)
    angles = tilts(T)
    sino = rng.standard_normal((ndf, T, N)).astype(numpy.float32)
    cfg = C.recon_cfg(N, T, ndf, H, W, ps, centre, mn, mx, angles)
    got = numpy.empty((H, W), numpy.float32)
    C.call("gx_project", ctypes.byref(cfg), C.f32p(sino), C.f32p(got))
    expect = O.project(sino if ndf > 1 else sino[0], (H, W), angles, centre,
                       ps, mn, mx)
    numpy.testing.assert_array_equal(got, expect)


@pytest.mark.parametrize("case", BP_CASES[1:4] + [
    # rows wider than 4096: 16384-point filters, one row per row block
    (5760, 5, 2, 24, 5760, 2880.0, (-40.0, 40.0), 1.0),
    (8192, 3, 1, 16, 8192, 4096.0, (0.0, 0.0), 1.0)])
def test_recon_matches_oracle(capi, case):
    N, T, ndf, H, W, centre, (mn, mx), ps = case
    rng = numpy.random.default_rng(7 + N)
    angles = tilts(T)
    sino = rng.standard_normal((ndf, T, N)).astype(numpy.float32)
    cfg = C.recon_cfg(N, T, ndf, H, W, ps, centre, mn, mx, angles)
    got = numpy.empty((H, W), numpy.float32)
    C.call("gx_recon", ctypes.byref(cfg), C.f32p(sino), C.f32p(got))
    expect = O.recon(sino, (H, W), angles, centre, ps, mn, mx)
    assert rel_l2(got, expect) < 1e-5


# ---------------------------------------------------------- correction ----

def ctf_stack(X, Y, ps, dfs, Cs=2.7e7, Ca=0.0, Pa=0.0, phi=0.0):
    l = O.electron_wavelength(300000)
    return numpy.stack([O.ctf_simple_grid(l, df, Cs, Ca, Pa, phi, X, Y, ps)
                        for df in dfs]).astype(numpy.complex64)


CORR_CASES = [
    # Y, X, ps, dfs, Ca, Pa, phi
    (16, 16, 2.0, [20000.0], 0.0, 0.0, 0.0),
    (32, 64, 2.0, [15000.0, 20000.0, 25000.0], 0.0, 0.0, 0.1),
    (64, 32, 1.5, [30000.0, 31000.0], 400.0, 0.4, 0.1),
    (256, 128, 1.0, [25000.0], 0.0, 0.0, math.asin(0.1)),
    (1024, 1024, 2.0, [20000.0], 0.0, 0.0, math.asin(0.1)),
    # long columns: 4-column (2048) and 2-column (4096) workgroups, several
    # defoci through the thread-order flip mask
    (2048, 32, 1.5, [30000.0, 30500.0], 0.0, 0.0, 0.1),
    (4096, 16, 1.0, [20000.0, 21000.0, 22000.0], 300.0, 0.3, 0.1),
    # 8192 points: two-column workgroups of 1024 threads, and 8192-point rows
    (8192, 16, 1.0, [20000.0, 20500.0], 0.0, 0.0, 0.1),
    (8, 8192, 1.0, [25000.0], 200.0, 0.2, 0.1),
]


@pytest.mark.parametrize("case", CORR_CASES)
def test_corr_matches_oracle(capi, case):
    Y, X, ps, dfs, Ca, Pa, phi = case
    rng = numpy.random.default_rng(X + Y)
    img = rng.standard_normal((Y, X)).astype(numpy.float32) + 2.0
    ctf = ctf_stack(X, Y, ps, dfs, Ca=Ca, Pa=Pa, phi=phi)
    out = numpy.empty((len(dfs), Y, X), numpy.float32)
    C.call("gx_corr", C.f32p(img), ctf.view(numpy.float32).ctypes.data_as(
        C._f32p), C.f32p(out), X, Y, len(dfs), -1)
    expect = O.corr(img, ctf, precision="f32")
    for k in range(len(dfs)):
        assert rel_l2(out[k], expect[k]) < 1e-5, (case, k)


def test_corr_special_case_exact_negation(capi):
    # df = Cs = 0, phi = 90 deg: out = -(X Y) image
    Y, X = 32, 64
    img = numpy.random.default_rng(1).standard_normal((Y, X)).astype(
        numpy.float32)
    ctf = ctf_stack(X, Y, 1.0, [0.0], Cs=0.0, phi=math.pi / 2)
    out = numpy.empty((1, Y, X), numpy.float32)
    C.call("gx_corr", C.f32p(img), ctf.view(numpy.float32).ctypes.data_as(
        C._f32p), C.f32p(out), X, Y, 1, -1)
    assert rel_l2(out[0], -X * Y * img.astype(numpy.float64)) < 1e-6


# ------------------------------------------------------ whole volumes -----

def ctf_problem(X, Y, T, ndf, ps, df, seed, phi=0.0, Ca=0.0, Pa=0.0,
                step=None):
    rng = numpy.random.default_rng(seed)
    proj = (rng.standard_normal((T, Y, X)) + 5.0).astype(numpy.float32)
    angles = tilts(T)
    centre = numpy.full(Y, X / 2.0, numpy.float32)
    mn, mx, st, n = O.get_defocus_steps(df, ndf, step, X, centre, ps)
    dfs = numpy.array([mn + d * st for d in range(n)], dtype=numpy.float64)
    return proj, angles, centre, dfs, (mn - df, mx - df)


def gpu_volume(proj, angles, X, Y, Z, ndf, ps, rel, storage, dfs=None,
               layout=C.INPUT_PROJECTIONS, phi=0.0, Ca=0.0, Pa=0.0,
               weights=None, tilt_chunk=0):
    l = O.electron_wavelength(300000)
    plan = C.VolumePlan(X, Y, angles, Z, num_defocus=ndf, pixel_size=ps,
                        centre=X / 2.0, min_defocus=rel[0],
                        max_defocus=rel[1], storage=storage,
                        input_layout=layout, defocus=dfs, wavelength=l,
                        Cs=2.7e7, Ca=Ca, Pa=Pa, phi=phi, weights=weights,
                        tilt_chunk=tilt_chunk)
    d_in = C.DeviceArray.from_host(proj)
    d_vol = C.DeviceArray((Y, Z, X))
    plan.run(d_in, d_vol)
    C.call("gx_device_synchronize")
    out = d_vol.to_host()
    plan.destroy()
    return out


VOL_CASES = [
    # X, Y, T, ndf, Z, ps, df, chunk
    (64, 32, 9, 1, 16, 2.0, 20000.0, 0),
    (64, 16, 9, 3, 32, 1.5, 30000.0, 4),
    (128, 8, 13, 8, 32, 1.0, 25000.0, 0),
]


@pytest.mark.parametrize("case", VOL_CASES)
def test_volume_ctf_f32_matches_oracle_chain(capi, case):
    X, Y, T, ndf, Z, ps, df, chunk = case
    proj, angles, centre, dfs, rel = ctf_problem(X, Y, T, ndf, ps, df, X + T)
    got = gpu_volume(proj, angles, X, Y, Z, ndf, ps, rel, "f32", dfs,
                     tilt_chunk=chunk)
    corrected, mn, mx = O.correct_projections(
        proj, centre, ps, 300, df, ndf, None, 2.7, 0, 0, 0)
    expect = O.reconstruct(corrected, angles, X / 2.0, ps, mn, mx, depth=Z)
    assert rel_l2(got, expect) < 1e-5


@pytest.mark.parametrize("storage,bound", [("f16", 1e-3), ("bf16", 4e-3)])
def test_volume_ctf_reduced_vs_fp64_oracle(capi, storage, bound):
    X, Y, T, ndf, Z, ps, df = 128, 16, 13, 4, 32, 1.0, 25000.0
    proj, angles, centre, dfs, rel = ctf_problem(X, Y, T, ndf, ps, df, 99)
    got = gpu_volume(proj, angles, X, Y, Z, ndf, ps, rel, storage, dfs)
    corrected, mn, mx = O.correct_projections(
        proj, centre, ps, 300, df, ndf, None, 2.7, 0, 0, 0, precision="f64")
    expect = O.reconstruct(corrected, angles, X / 2.0, ps, mn, mx, depth=Z,
                           precision="f64")
    err = rel_l2(got, expect)
    print("rel-L2 %s vs FP64 oracle: %.3g" % (storage, err))
    assert err < bound


def test_volume_plain_fbp_matches_oracle(capi):
    X, Y, T, Z = 100, 12, 11, 40
    rng = numpy.random.default_rng(3)
    proj = rng.standard_normal((T, Y, X)).astype(numpy.float32)
    angles = tilts(T)
    got = gpu_volume(proj, angles, X, Y, Z, 1, 1.0, (0.0, 0.0), "f32")
    expect = O.reconstruct(proj, angles, X / 2.0, 1.0, depth=Z)
    assert rel_l2(got, expect) < 1e-5


def test_volume_sinogram_layout_and_weights(capi):
    X, Y, T, ndf, Z = 64, 10, 9, 2, 20
    rng = numpy.random.default_rng(4)
    sino = rng.standard_normal((Y, ndf, T, X)).astype(numpy.float32)
    angles = (tilts(T) + rng.uniform(-0.01, 0.01, T)).astype(numpy.float32)
    w = O.compute_angular_weights(angles)
    got = gpu_volume(sino, angles, X, Y, Z, ndf, 1.0, (-40.0, 40.0), "f32",
                     layout=C.INPUT_SINOGRAMS, weights=-w)
    expect = O.reconstruct(sino, angles, X / 2.0, 1.0, -40.0, 40.0,
                           sinogram_order=True, transform="minus",
                           angular_weights=True, depth=Z)
    assert rel_l2(got, expect) < 1e-5


# Rows of 256 points and more take the four-transform row passes
# (k_rows_c2r_ramp with the correction, k_filter_ramp without): one, two and
# four row pairs per workgroup (X = 4096, 2048, <= 1024), row counts that
# leave a partial last row group, and a row shorter than its transform
# (N = 300 filtered on the 1024-point grid, 512-point halves).
RAMP_CTF_CASES = [
    # X, Y, T, ndf, Z, ps, df
    (256, 12, 7, 1, 8, 1.0, 20000.0),
    (512, 9, 5, 2, 8, 1.0, 25000.0),
    (1024, 16, 4, 1, 8, 2.0, 20000.0),
    (2048, 6, 3, 2, 4, 1.5, 30000.0),
    (4096, 3, 3, 1, 4, 1.0, 25000.0),
    # 4096-point columns (2-column workgroups) under short rows
    (64, 4096, 2, 2, 4, 1.0, 25000.0),
    # row counts that end inside a forward-transform workgroup (16 rows per
    # workgroup at 1024 points, 8 at 2048)
    (1024, 21, 3, 2, 4, 2.0, 22000.0),
    (2048, 11, 2, 1, 4, 1.5, 28000.0),
    # 8192-point rows (16384-point ramp, two row pairs per slot with 16-bit
    # storage) and 8192-point columns
    (8192, 6, 2, 2, 4, 1.0, 25000.0),
    (32, 8192, 2, 2, 4, 1.0, 25000.0),
]


@pytest.mark.parametrize("case", RAMP_CTF_CASES)
def test_volume_ctf_long_rows_match_oracle_chain(capi, case):
    X, Y, T, ndf, Z, ps, df = case
    proj, angles, centre, dfs, rel = ctf_problem(X, Y, T, ndf, ps, df, X + Y)
    corrected, mn, mx = O.correct_projections(
        proj, centre, ps, 300, df, ndf, None, 2.7, 0, 0, 0)
    expect = O.reconstruct(corrected, angles, X / 2.0, ps, mn, mx, depth=Z)
    got = gpu_volume(proj, angles, X, Y, Z, ndf, ps, rel, "f32", dfs)
    assert rel_l2(got, expect) < 1e-5
    corrected, mn, mx = O.correct_projections(
        proj, centre, ps, 300, df, ndf, None, 2.7, 0, 0, 0, precision="f64")
    expect = O.reconstruct(corrected, angles, X / 2.0, ps, mn, mx, depth=Z,
                           precision="f64")
    for storage, bound in (("f16", 1e-3), ("bf16", 4e-3)):
        got = gpu_volume(proj, angles, X, Y, Z, ndf, ps, rel, storage, dfs)
        assert rel_l2(got, expect) < bound, storage


RAMP_PLAIN_CASES = [
    # X, Y, T, Z
    (256, 13, 7, 16),
    (300, 10, 6, 16),
    (1024, 8, 5, 8),
    (2048, 5, 3, 4),
    (4096, 3, 3, 4),
    (5000, 3, 3, 4),
    (8192, 3, 2, 4),
]


@pytest.mark.parametrize("case", RAMP_PLAIN_CASES)
def test_volume_plain_long_rows_match_oracle(capi, case):
    X, Y, T, Z = case
    rng = numpy.random.default_rng(X + Y)
    proj = rng.standard_normal((T, Y, X)).astype(numpy.float32)
    angles = tilts(T)
    got = gpu_volume(proj, angles, X, Y, Z, 1, 1.0, (0.0, 0.0), "f32")
    expect = O.reconstruct(proj, angles, X / 2.0, 1.0, depth=Z)
    assert rel_l2(got, expect) < 1e-5
    expect = O.reconstruct(proj, angles, X / 2.0, 1.0, depth=Z,
                           precision="f64")
    for storage, bound in (("f16", 1e-3), ("bf16", 4e-3)):
        got = gpu_volume(proj, angles, X, Y, Z, 1, 1.0, (0.0, 0.0), storage)
        assert rel_l2(got, expect) < bound, storage


def test_slab_range_is_every_slab_the_voxels_read(capi):
    # The filter computes, per tilt, only the defocus slabs some voxel reads:
    # the plan's range must be the min and max of the reference's slab
    # choice (reconstructor.cpp:66-73, float arithmetic) over every voxel,
    # here by brute force; and the volume is the oracle chain's, which
    # computes every slab.
    X, Y, T, ndf, Z, ps, df = 256, 4, 61, 16, 64, 1.0, 25000.0
    proj, angles, centre, dfs, rel = ctf_problem(X, Y, T, ndf, ps, df, 11)
    l = O.electron_wavelength(300000)
    plan = C.VolumePlan(X, Y, angles, Z, num_defocus=ndf, pixel_size=ps,
                        centre=X / 2.0, min_defocus=rel[0],
                        max_defocus=rel[1], storage="f16", defocus=dfs,
                        wavelength=l, Cs=2.7e7)
    first, last = plan.slab_range()
    plan.destroy()
    tc = O.tilt_coefficients(angles, X / 2.0)
    f = numpy.float32
    dscale = f(f(ndf) * f(ps) / (f(rel[1]) - f(rel[0])))
    doffset = f(-dscale * f(f(rel[0]) / f(ps)))
    x = (numpy.arange(X) - 0.5 * X + 0.5).astype(f)[None, :]
    z = (numpy.arange(Z) - 0.5 * Z + 0.5).astype(f)[:, None]
    for t in range(T):
        c, s_ = tc[t, 0], tc[t, 1]
        h = (-x * s_) - (z * c)
        d = h * dscale + doffset
        k = numpy.clip(numpy.floor(d.astype(numpy.float64) + 0.5), 0,
                       ndf - 1).astype(int)
        assert (first[t], last[t]) == (k.min(), k.max()), t
    assert (last - first + 1).sum() < T * ndf  # some slabs are skipped
    corrected, mn, mx = O.correct_projections(
        proj, centre, ps, 300, df, ndf, None, 2.7, 0, 0, 0)
    expect = O.reconstruct(corrected, angles, X / 2.0, ps, mn, mx, depth=Z)
    got = gpu_volume(proj, angles, X, Y, Z, ndf, ps, rel, "f32", dfs)
    assert rel_l2(got, expect) < 1e-5


# ------------------------------------------- graphs and synthetic data ----

def test_graph_replays_match_eager(capi):
    # The whole request captured once and replayed several times, with eager
    # work and synchronisations in between; every replay must equal the
    # eager result bit for bit.
    X, Y, T, ndf, Z, ps, df = 64, 16, 9, 3, 16, 1.5, 30000.0
    proj, angles, centre, dfs, rel = ctf_problem(X, Y, T, ndf, ps, df, 5)
    l = O.electron_wavelength(300000)
    for storage in ("f32", "f16"):
        plan = C.VolumePlan(X, Y, angles, Z, num_defocus=ndf, pixel_size=ps,
                            min_defocus=rel[0], max_defocus=rel[1],
                            storage=storage, defocus=dfs, wavelength=l,
                            Cs=2.7e7)
        s = C.Stream()
        d_in = C.DeviceArray.from_host(proj)
        d_eager = C.DeviceArray((Y, Z, X))
        d_graph = C.DeviceArray((Y, Z, X))
        plan.run(d_in, d_eager, s)
        s.synchronize()
        eager = d_eager.to_host()
        C.call("gx_graph_capture_begin", s.handle)
        plan.run(d_in, d_graph, s)
        g = C.ctypes.c_void_p()
        C.call("gx_graph_capture_end", s.handle, C.ctypes.byref(g))
        for rep in range(4):
            C.call("gx_fill_u32", d_graph.ptr, 0x7fc00000, Y * Z * X, s.handle)
            C.call("gx_graph_launch", g.value, s.handle)
            s.synchronize()
            numpy.testing.assert_array_equal(d_graph.to_host(), eager)
            plan.run(d_in, d_eager, s)  # eager work between replays
            s.synchronize()
        C.call("gx_graph_destroy", g.value)
        s.destroy()
        plan.destroy()


def test_synthetic_projections_are_seeded(capi):
    X, Y, T = 64, 32, 5
    angles = tilts(T)
    a = C.DeviceArray((T, Y, X))
    b = C.DeviceArray((T, Y, X))
    ctf = dict(wavelength=O.electron_wavelength(300000), Cs=2.7e7,
               defocus=20000.0, phi=math.asin(0.1), pixel_size=2.0)
    C.synth_projections(a, X, Y, angles, depth=16, num_ellipsoids=10,
                        seed=7, noise="poisson", noise_level=50.0, ctf=ctf)
    C.synth_projections(b, X, Y, angles, depth=16, num_ellipsoids=10,
                        seed=7, noise="poisson", noise_level=50.0, ctf=ctf)
    ha, hb = a.to_host(), b.to_host()
    numpy.testing.assert_array_equal(ha, hb)
    assert numpy.all(ha >= 0) and numpy.all(ha == numpy.round(ha))
    assert 40 < ha.mean() < 60
    # no ellipsoids, no noise: an empty tilt series
    C.synth_projections(a, X, Y, angles, depth=X, num_ellipsoids=0, seed=1)
    assert numpy.all(a.to_host() == 0)


@pytest.mark.parametrize("chunk", [16, 0])
def test_identical_tilts_give_identical_slabs(capi, chunk):
    # Regression: with many workgroups resident per CU, LDS loads that the
    # compiler had moved below a workgroup barrier read data another wave
    # had already overwritten, corrupting one column block of a few tilts
    # per run. Identical images in every tilt must give bit-identical slabs,
    # run after run.
    X = Y = 1024
    T, Z = 48, 8
    img = numpy.random.default_rng(0).poisson(50, (Y, X)).astype(
        numpy.float32)
    d_in = C.DeviceArray.from_host(numpy.broadcast_to(img, (T, Y, X)).copy())
    l = O.electron_wavelength(300000)
    for storage in ("f32", "f16"):
        plan = C.VolumePlan(X, Y, numpy.zeros(T, numpy.float32), Z,
                            pixel_size=2.0, storage=storage,
                            defocus=numpy.array([20000.0]), wavelength=l,
                            Cs=2.7e7, phi=math.asin(0.1), tilt_chunk=chunk)
        s = C.Stream()
        for run in range(4):
            plan.filter(d_in, s)
            s.synchronize()
            p, n, yb = plan.slabs()
            a = numpy.empty(n // 4, numpy.uint32)
            C.call("gx_memcpy_dtoh", a.ctypes.data, p, n, None)
            a = a.reshape(T, -1)
            bad = [t for t in range(T) if not numpy.array_equal(a[t], a[0])]
            assert not bad, (storage, run, bad)
        s.destroy()
        plan.destroy()


# ------------------------------------------ staged and direct backprojection

BP_KIND_CASES = [
    # X, Y, T, ndf, Z, W, (first, last) tilt deg, (min_df, max_df), pixel
    (96, 20, 37, 1, 24, 70, (-60.0, 60.0), (0.0, 0.0), 1.0),
    # every quadrant (negative cosines), T > 64: several tilt chunks
    (64, 16, 150, 1, 13, 64, (-180.0, 178.0), (0.0, 0.0), 1.0),
    # three slabs per tile at most -> four staged
    (128, 24, 130, 16, 40, 128, (-70.0, 70.0), (-128.0, 128.0), 1.0),
    # two slabs per tile at most
    (64, 16, 21, 8, 64, 200, (-60.0, 60.0), (-512.0, 512.0), 1.0),
    # narrow slabs: more than four per tile, direct kernel only
    (64, 16, 21, 16, 64, 64, (-60.0, 60.0), (-32.0, 32.0), 1.0),
    # a detector narrower than a staged run (40 positions) and the volume,
    # a partial last row block
    (24, 10, 29, 1, 20, 56, (-60.0, 60.0), (0.0, 0.0), 1.0),
]


def sinogram_plan(case, storage):
    X, Y, T, ndf, Z, W, (a0, a1), (mn, mx), ps = case
    angles = numpy.deg2rad(numpy.linspace(a0, a1, T)).astype(numpy.float32)
    rng = numpy.random.default_rng(X + T + ndf)
    sino = rng.standard_normal((Y, ndf, T, X)).astype(numpy.float32)
    plan = C.VolumePlan(X, Y, angles, Z, width=W, num_defocus=ndf,
                        pixel_size=ps, centre=X / 2.0 - 0.7, min_defocus=mn,
                        max_defocus=mx, storage=storage,
                        input_layout=C.INPUT_SINOGRAMS)
    return plan, sino, angles


def slab_rows(plan, case, ys):
    """The plan's filtered FP32 slabs, [k][t][y/YB][x][y%YB], as one
    (ndf, T, X) array per detector row in ys."""
    X, Y, T, ndf = case[:4]
    p, n, yb = plan.slabs()
    # the buffer is sized for the largest row blocking; the layout in use
    # holds ceil(Y / yb) blocks per (slab, tilt)
    nyb = (Y + yb - 1) // yb
    used = ndf * T * nyb * X * yb * 4
    assert used <= n
    a = numpy.empty(used // 4, numpy.float32)
    C.call("gx_memcpy_dtoh", a.ctypes.data, p, used, None)
    a = a.reshape(ndf, T, nyb, X, yb)
    return {y: numpy.ascontiguousarray(a[:, :, y // yb, :, y % yb])
            for y in ys}


@pytest.mark.parametrize("case", BP_KIND_CASES)
def test_staged_and_direct_bp_bitexact_vs_oracle(capi, case):
    X, Y, T, ndf, Z, W, _, (mn, mx), ps = case
    plan, sino, angles = sinogram_plan(case, "f32")
    kind, slots = plan.backprojector()
    assert (kind == C.BP_STAGED) == (slots > 0)
    d_in = C.DeviceArray.from_host(sino)
    plan.filter(d_in)
    vols = {}
    for k in ("direct", "staged"):
        if k == "staged" and slots == 0:
            with pytest.raises(C.GxError, match="more than 4 defocus slabs"):
                plan.set_backprojector(k)
            continue
        plan.set_backprojector(k)
        d_vol = C.DeviceArray((Y, Z, W))
        C.call("gx_fill_u32", d_vol.ptr, 0x7fc00000, Y * Z * W, None)
        plan.backproject(d_vol)
        C.call("gx_device_synchronize")
        vols[k] = d_vol.to_host()
    ys = [0, Y // 2, Y - 1]
    rows = slab_rows(plan, case, ys)
    for y in ys:
        s = rows[y] if ndf > 1 else rows[y][0]
        expect = O.project(s, (Z, W), angles, X / 2.0 - 0.7, ps, mn, mx)
        for k, v in vols.items():
            numpy.testing.assert_array_equal(v[y], expect, err_msg=k)
    plan.destroy()


@pytest.mark.parametrize("storage", ["f16", "bf16"])
@pytest.mark.parametrize("case", BP_KIND_CASES[:4] + BP_KIND_CASES[5:])
def test_staged_bp_equals_direct_reduced_storage(capi, case, storage):
    X, Y, T, ndf, Z, W = case[:6]
    plan, sino, angles = sinogram_plan(case, storage)
    d_in = C.DeviceArray.from_host(sino)
    plan.filter(d_in)
    out = {}
    for k in ("direct", "staged"):
        plan.set_backprojector(k)
        d_vol = C.DeviceArray((Y, Z, W))
        plan.backproject(d_vol)
        C.call("gx_device_synchronize")
        out[k] = d_vol.to_host()
    assert numpy.isfinite(out["staged"]).all()
    numpy.testing.assert_array_equal(out["staged"], out["direct"])
    plan.destroy()


@pytest.mark.parametrize("storage", ["f32", "f16", "bf16"])
@pytest.mark.parametrize("case", [BP_KIND_CASES[0], BP_KIND_CASES[2],
                                  BP_KIND_CASES[5]])
def test_unwritten_slab_memory_never_reaches_the_volume(capi, case, storage):
    # A plan does not clear its slab buffer: the rows of the last row block
    # past the detector's rows and the slabs no voxel reads at a tilt keep
    # whatever the buffer held. Filled with NaN before the filter pass, they
    # must leave every voxel, from either backprojector, bit for bit as it is
    # when the buffer held zeros.
    X, Y, T, ndf, Z, W = case[:6]
    out = {}
    for fill in (0, 0x7fc00000):
        plan, sino, angles = sinogram_plan(case, storage)
        p, n, _ = plan.slabs()
        C.call("gx_fill_u32", p, fill, n // 4, None)
        d_in = C.DeviceArray.from_host(sino)
        slots = plan.backprojector()[1]
        for k in (("staged", "direct") if slots else ("direct",)):
            plan.set_backprojector(k)
            d_vol = C.DeviceArray((Y, Z, W))
            C.call("gx_fill_u32", d_vol.ptr, 0x7fc00000, Y * Z * W, None)
            plan.run(d_in, d_vol)
            C.call("gx_device_synchronize")
            out[fill, k] = d_vol.to_host()
        plan.destroy()
    for (fill, k), v in out.items():
        assert numpy.isfinite(v).all(), (fill, k)
        numpy.testing.assert_array_equal(v, out[0, k], err_msg=k)


def slab_rows_16(plan, case, ys, storage):
    """slab_rows for 16-bit slabs, decoded to float32."""
    X, Y, T, ndf = case[:4]
    p, n, yb = plan.slabs()
    nyb = (Y + yb - 1) // yb
    used = ndf * T * nyb * X * yb * 2
    assert used <= n
    a = numpy.empty(used // 2, numpy.uint16)
    C.call("gx_memcpy_dtoh", a.ctypes.data, p, used, None)
    if storage == "f16":
        f = a.view(numpy.float16).astype(numpy.float32)
    else:
        f = (a.astype(numpy.uint32) << 16).view(numpy.float32)
    f = f.reshape(ndf, T, nyb, X, yb)
    return {y: numpy.ascontiguousarray(f[:, :, y // yb, :, y % yb])
            for y in ys}


# C3's geometry at two rows: 2048 wide, 1.5 A/px, +-1536 A in 16 slabs,
# 61 tilts of 2 degrees, where rounding puts voxels exactly on slab ties
SLAB_TIE_CASE = (2048, 2, 61, 16, 512, 2048, (-60.0, 60.0), (-1536.0, 1536.0),
                 1.5)


@pytest.mark.parametrize("storage", ["f16", "bf16"])
@pytest.mark.parametrize("case", [BP_KIND_CASES[2], BP_KIND_CASES[3],
                                  SLAB_TIE_CASE])
def test_reduced_storage_bp_takes_the_reference_slabs(capi, case, storage):
    # The slab of a (voxel, tilt) is a discontinuous choice, and on the
    # half-integer voxel grid whole rows of voxels sit exactly on slab
    # boundaries at some tilts: reduced storage must choose as the reference
    # does. Backprojecting the plan's own decoded slabs (times the per-tilt
    # normalisation the filter divided out: max |input| of the tilt) through
    # the oracle leaves only FP32 rounding differences: below 1e-5 for narrow
    # detectors, and at 2048 positions those of the reduced path's FMA tap
    # geometry (an ulp of p is 1.2e-4): 1.35e-4 measured with the
    # reference's slab choice, against 5.0e-4 for a float-floor FMA one.
    X, Y, T, ndf, Z, W, _, (mn, mx), ps = case
    tol = 1e-5 if X <= 256 else 2.5e-4
    plan, sino, angles = sinogram_plan(case, storage)
    d_in = C.DeviceArray.from_host(sino)
    plan.filter(d_in)
    d_vol = C.DeviceArray((Y, Z, W))
    for k in ("direct", "staged"):
        plan.set_backprojector(k)
        plan.backproject(d_vol)
        C.call("gx_device_synchronize")
        vol = d_vol.to_host()
        nrm = numpy.abs(sino).max(axis=(0, 1, 3)).astype(numpy.float32)
        ys = [0, Y // 2, Y - 1]
        rows = slab_rows_16(plan, case, ys, storage)
        for y in ys:
            s = (rows[y] * nrm[None, :, None]).astype(numpy.float32)
            expect = O.project(s, (Z, W), angles, X / 2.0 - 0.7, ps, mn, mx)
            assert rel_l2(vol[y], expect) < tol, (k, y)
    plan.destroy()


# ------------------------------------------------------------- autotuning

@pytest.mark.parametrize("storage", ["f32", "f16"])
@pytest.mark.parametrize("case", [BP_KIND_CASES[0], BP_KIND_CASES[2],
                                  BP_KIND_CASES[4]])
def test_autotuner_candidates_agree_and_choice_is_remembered(
        capi, case, storage, tmp_path):
    X, Y, T, ndf, Z, W = case[:6]
    C.tune_clear()
    plan, sino, angles = sinogram_plan(case, storage)
    default = plan.tile()
    d_in = C.DeviceArray.from_host(sino)
    ref = C.DeviceArray((Y, Z, W))
    plan.run(d_in, ref)
    C.call("gx_device_synchronize")
    ref = ref.to_host()
    assert numpy.isfinite(ref).all()
    # every candidate computes the same volume, bit for bit
    big = 16 if storage == "f32" else 32
    slots = plan.backprojector()[1]
    kinds = ["staged", "direct"] if slots else ["direct"]
    for rows in (big, big // 2):
        for k in kinds:
            plan.set_tile(rows, k)
            d_vol = C.DeviceArray((Y, Z, W))
            C.call("gx_fill_u32", d_vol.ptr, 0x7fc00000, Y * Z * W, None)
            plan.run(d_in, d_vol)
            C.call("gx_device_synchronize")
            numpy.testing.assert_array_equal(d_vol.to_host(), ref,
                                             err_msg="%d %s" % (rows, k))
    with pytest.raises(C.GxError, match="not a candidate"):
        plan.set_tile(big * 2, "direct")
    d_vol = C.DeviceArray((Y, Z, W))
    res = plan.autotune(d_in, d_vol)
    numpy.testing.assert_array_equal(d_vol.to_host(), ref)
    assert len(res["candidates"]) == 2 * len(kinds)
    best = min(res["candidates"], key=lambda c: c["filter_ms"] + c["bp_ms"])
    assert (res["rows"], res["backprojector"]) == (best["rows"],
                                                   best["backprojector"])
    chosen = (res["rows"], C.BACKPROJECTOR[res["backprojector"]])
    assert plan.tile() == chosen
    # later plans of this shape start with the choice; others do not
    p2, _, _ = sinogram_plan(case, storage)
    assert p2.tile() == chosen
    p2.destroy()
    C.tune_save(str(tmp_path / "tiles.txt"))
    C.tune_clear()
    p3, _, _ = sinogram_plan(case, storage)
    assert p3.tile() == default
    p3.destroy()
    C.tune_load(str(tmp_path / "tiles.txt"))
    p4, _, _ = sinogram_plan(case, storage)
    assert p4.tile() == chosen
    p4.destroy()
    plan.destroy()
    C.tune_clear()


# ------------------------------------------------- per-row rotation centres

@pytest.mark.parametrize("case", [BP_KIND_CASES[0], BP_KIND_CASES[2],
                                  BP_KIND_CASES[4]])
def test_per_row_centre_bp_bitexact_vs_oracle(capi, case):
    X, Y, T, ndf, Z, W, (a0, a1), (mn, mx), ps = case
    angles = numpy.deg2rad(numpy.linspace(a0, a1, T)).astype(numpy.float32)
    rng = numpy.random.default_rng(X + T + ndf + 7)
    sino = rng.standard_normal((Y, ndf, T, X)).astype(numpy.float32)
    centres = (X / 2.0 + rng.uniform(-4.0, 4.0, Y)).astype(numpy.float32)
    plan = C.VolumePlan(X, Y, angles, Z, width=W, num_defocus=ndf,
                        pixel_size=ps, centre=centres, min_defocus=mn,
                        max_defocus=mx, storage="f32",
                        input_layout=C.INPUT_SINOGRAMS)
    assert plan.backprojector() == (C.BP_DIRECT, 0)
    with pytest.raises(C.GxError, match="per-row rotation centres"):
        plan.set_backprojector("staged")
    d_in = C.DeviceArray.from_host(sino)
    ys = [0, Y // 2, Y - 1]
    for rows in (16, 8):
        plan.set_tile(rows, "direct")
        d_vol = C.DeviceArray((Y, Z, W))
        C.call("gx_fill_u32", d_vol.ptr, 0x7fc00000, Y * Z * W, None)
        plan.run(d_in, d_vol)
        C.call("gx_device_synchronize")
        vol = d_vol.to_host()
        srows = slab_rows(plan, case, ys)
        for y in ys:
            s = srows[y] if ndf > 1 else srows[y][0]
            expect = O.project(s, (Z, W), angles, float(centres[y]), ps, mn,
                               mx)
            numpy.testing.assert_array_equal(vol[y], expect,
                                             err_msg="row %d, %d rows" % (
                                                 y, rows))
    plan.destroy()
