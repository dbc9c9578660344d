"""The full-size workloads through the HIP path, checked against the oracle.

- 61 x 256^2 projections, no correction, 256 x 128 x 256 volume: every slice
  (FP32 within 1e-5 of the FP32 oracle, FP16 slabs within 1e-3 of FP64).
- 121 x 1024^2, one defocus, 1024 x 256 x 1024, and 61 x 2048^2, 16 defoci,
  2048 x 512 x 2048: four sampled slices of the whole volume (FP32 within
  1e-5 of the FP32 oracle chain, FP16 within 1e-3 of the FP64 chain).
- 121 x 4096^2, 32 defoci, 4096 x 1024 x 4096: FP16 and BF16 slabs (separate
  runs) and FP32 slabs in bands of rows. Four sampled slices of each whole
  volume against the FP64 chain; the filtered slabs of 8 tilts at those rows
  against the FP64 correction and filter; the backprojection of the plan's
  own decoded slabs against the oracle's.
- 3 x 8192^2, 2 defoci, 8192 x 8 x 8192: the widest detector, four sampled
  slices (FP32 and FP16 as above).
- A 32-request serving mix drawn as the serving workload draws them
  (detector 512, 1024 or 2048 squared, 41-121 tilts over +-60 degrees, 1-16
  defoci, df 10000-40000 A, a seeded phantom each): every result equals its
  plan-path result bit for bit; three are checked against the oracle.

The oracle at sampled rows (oracle/sampled.py) runs on spawned CPU workers
(they hold no GPU state). Input data are the seeded synthetic tilt series of
the library's generator: ellipsoid phantoms projected analytically, the CTF
forward model, Poisson noise.
"""
import math
import os
import time

import numpy
import pytest

from guanaco_amd import _capi as C
from oracle import oracle as O
from oracle import sampled

pytestmark = pytest.mark.gpu

L300 = O.electron_wavelength(300000)
PHI = math.asin(0.1)


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


def series(X, Y, T, step, Z, ps, df, ne, semi, dose, seed):
    """A seeded synthetic tilt series on the device, and its angles."""
    angles = numpy.deg2rad(numpy.arange(T) * step -
                           0.5 * (T - 1) * step).astype(numpy.float32)
    d = C.DeviceArray((T, Y, X))
    ctf = None if df is None else dict(wavelength=L300, Cs=2.7e7, defocus=df,
                                       phi=PHI, pixel_size=ps)
    C.synth_projections(d, X, Y, angles, depth=Z, width=X, num_ellipsoids=ne,
                        seed=seed, semi=semi,
                        noise="poisson" if dose else "gaussian",
                        noise_level=dose if dose else 0.01, ctf=ctf)
    C.call("gx_device_synchronize")
    return d, angles


def plan_kw(X, Y, ndf, ps, df, storage):
    kw = dict(num_defocus=ndf, pixel_size=ps, centre=X / 2.0,
              storage=storage)
    if df is not None:
        dfs, mn, mx = sampled.defocus_list(df, ndf, X, Y, ps)
        kw.update(defocus=numpy.array(dfs), min_defocus=mn, max_defocus=mx,
                  wavelength=L300, Cs=2.7e7, phi=PHI)
    return kw


def run_rows(d_in, angles, X, Y, Z, ndf, ps, df, storage, rows, d_vol,
             band_rows=0):
    """The whole path into d_vol; the sampled slices to the host."""
    plan = C.VolumePlan(X, Y, angles, Z, band_rows=band_rows,
                        **plan_kw(X, Y, ndf, ps, df, storage))
    C.call("gx_fill_u32", d_vol.ptr, 0x7fc00000, Y * Z * X, None)
    plan.run(d_in, d_vol)
    C.call("gx_device_synchronize")
    out = numpy.stack([C.DeviceView(d_vol.ptr + int(y) * Z * X * 4,
                                    (Z, X)).to_host() for y in rows])
    return plan, out


def test_plain_fbp_256_every_slice(capi):
    X, Y, T, Z = 256, 256, 61, 128
    print("256^2 x 61, no correction", flush=True)
    d_in, angles = series(X, Y, T, 2.0, Z, 1.0, None, 20, (5.0, 40.0), None,
                          1000)
    proj = d_in.to_host()
    d_vol = C.DeviceArray((Y, Z, X))
    rows = numpy.arange(Y)
    got = {}
    for st in ("f32", "f16"):
        plan, got[st] = run_rows(d_in, angles, X, Y, Z, 1, 1.0, None, st,
                                 rows, d_vol)
        plan.destroy()
    ref32 = O.reconstruct(proj, angles, X / 2.0, 1.0, depth=Z)
    ref64 = O.reconstruct(proj.astype(numpy.float64), angles, X / 2.0, 1.0,
                          depth=Z, precision="f64")
    e32 = sampled.rel_l2(got["f32"], ref32)
    e16 = sampled.rel_l2(got["f16"], ref64)
    print("FP32 vs FP32 oracle %.3g, FP16 vs FP64 %.3g" % (e32, e16))
    assert e32 < 1e-5
    assert e16 < 1e-3


SAMPLED = {
    # X, Y, T, step, Z, ndf, ps, df, ellipsoids, semi-axes, dose, seed
    "1024": (1024, 1024, 121, 1.0, 256, 1, 2.0, 20000.0, 200, (20.0, 160.0),
             50.0, 1001),
    "2048": (2048, 2048, 61, 2.0, 512, 16, 1.5, 30000.0, 400, (40.0, 320.0),
             30.0, 1002),
}


@pytest.mark.parametrize("name", sorted(SAMPLED))
def test_ctf_workload_sampled_slices(capi, pool, name):
    X, Y, T, step, Z, ndf, ps, df, ne, semi, dose, seed = SAMPLED[name]
    print("%d^2 x %d, %d defoci" % (X, T, ndf), flush=True)
    d_in, angles = series(X, Y, T, step, Z, ps, df, ne, semi, dose, seed)
    d_vol = C.DeviceArray((Y, Z, X))
    rows = numpy.linspace(0, Y - 1, 4).astype(int)
    got = {}
    for st in ("f32", "f16"):
        plan, got[st] = run_rows(d_in, angles, X, Y, Z, ndf, ps, df, st,
                                 rows, d_vol)
        plan.destroy()
    d_vol.free()
    proj = d_in.to_host()
    d_in.free()
    ref = sampled.reference_rows(pool, proj, angles, rows, Z, ps, df, ndf,
                                 PHI)
    e32 = sampled.rel_l2(got["f32"], ref["f32"])
    e32_64 = sampled.rel_l2(got["f32"], ref["f64"])
    e16 = sampled.rel_l2(got["f16"], ref["f64"])
    print("FP32 vs FP32 oracle %.3g (vs FP64 %.3g), FP16 vs FP64 %.3g" % (
        e32, e32_64, e16))
    assert e32 < 1e-5
    assert e16 < 1e-3


# ------------------------------------- the largest detector: 8192 x 8192 --

def test_8192_detector_sampled_slices(capi, pool):
    # The widest detector the path takes, end to end: 8192-point rows,
    # columns and ramp (16384 points) and 32-bit spectrum offsets near their
    # largest, on three tilts with two defoci.
    X = Y = 8192
    T, step, Z, ndf, ps, df = 3, 60.0, 8, 2, 1.0, 25000.0
    print("8192^2 x %d, %d defoci" % (T, ndf), flush=True)
    d_in, angles = series(X, Y, T, step, Z, ps, df, 200, (160.0, 1280.0),
                          30.0, 1006)
    d_vol = C.DeviceArray((Y, Z, X))
    rows = numpy.linspace(0, Y - 1, 4).astype(int)
    got = {}
    for st in ("f32", "f16"):
        plan, got[st] = run_rows(d_in, angles, X, Y, Z, ndf, ps, df, st,
                                 rows, d_vol)
        plan.destroy()
    d_vol.free()
    proj = d_in.to_host()
    d_in.free()
    ref = sampled.reference_rows(pool, proj, angles, rows, Z, ps, df, ndf,
                                 PHI)
    e32 = sampled.rel_l2(got["f32"], ref["f32"])
    e16 = sampled.rel_l2(got["f16"], ref["f64"])
    print("FP32 vs FP32 oracle %.3g, FP16 vs FP64 %.3g" % (e32, e16))
    assert e32 < 1e-5
    assert e16 < 1e-3


# ----------------------------------------------- 121 x 4096^2, 32 defoci --

C4 = dict(X=4096, Y=4096, T=121, step=1.0, Z=1024, ndf=32, ps=1.0,
          df=25000.0, ne=400, semi=(80.0, 640.0), dose=30.0, seed=1003)
C4_ROWS = numpy.linspace(0, 4095, 4).astype(int)
C4_TILTS = [0, 17, 34, 51, 60, 77, 94, 120]


@pytest.fixture(scope="module")
def c4(pool):
    """The 4096^2 series on the device and the oracle at the sampled rows
    (FP64 chain; FP64-corrected rows of C4_TILTS)."""
    w = C4
    print("4096^2 x 121, 32 defoci: synthesis", flush=True)
    d_in, angles = series(w["X"], w["Y"], w["T"], w["step"], w["Z"],
                          w["ps"], w["df"], w["ne"], w["semi"], w["dose"],
                          w["seed"])
    proj = d_in.to_host()
    print("4096^2 x 121: oracle at rows %s" % C4_ROWS.tolist(), flush=True)
    t0 = time.perf_counter()
    ref = sampled.reference_rows(pool, proj, angles, C4_ROWS, w["Z"],
                                 w["ps"], w["df"], w["ndf"], PHI, fp32=False,
                                 tilts=C4_TILTS)
    print("4096^2 x 121: oracle %.1f s" % (time.perf_counter() - t0),
          flush=True)
    # per-tilt normalisation of reduced storage: max |projection|
    nrm = numpy.abs(proj).max(axis=(1, 2)).astype(numpy.float32)
    del proj
    yield d_in, angles, ref, nrm
    d_in.free()


def decoded_slab_rows(plan, storage, X, Y, T, ndf, rows, tilts):
    """The plan's filtered slabs at detector `rows` for `tilts`, decoded to
    float32: (len(tilts), ndf, len(rows), X); layout [k][t][y/YB][x][y%YB]."""
    p, _, yb = plan.slabs()
    nyb = (Y + yb - 1) // yb
    es = 4 if storage == "f32" else 2
    out = numpy.empty((len(tilts), ndf, len(rows), X), numpy.float32)
    for i, t in enumerate(tilts):
        for k in range(ndf):
            for j, y in enumerate(rows):
                off = (((k * T + t) * nyb + y // yb) * X) * yb * es
                raw = C.DeviceView(p + off, (X, yb),
                                   {4: numpy.float32, 2: numpy.uint16}[es]
                                   ).to_host()[:, y % yb]
                if storage == "f32":
                    out[i, k, j] = raw
                elif storage == "f16":
                    out[i, k, j] = raw.view(numpy.float16)
                else:
                    out[i, k, j] = (raw.astype(numpy.uint32) << 16).view(
                        numpy.float32)
    return out


@pytest.mark.parametrize("storage,bound", [("f16", 1e-3), ("bf16", 8e-3)])
def test_4096_reduced_storage_against_fp64(capi, pool, c4, storage, bound):
    # BF16 (8-bit mantissa) does not meet the 1e-3 gate; its bound records
    # what it reaches.
    d_in, angles, ref, nrm = c4
    w = C4
    X, Y, T, Z, ndf = w["X"], w["Y"], w["T"], w["Z"], w["ndf"]
    print("4096^2 x 121, 32 defoci, %s" % storage, flush=True)
    d_vol = C.DeviceArray((Y, Z, X))
    plan, got = run_rows(d_in, angles, X, Y, Z, ndf, w["ps"], w["df"],
                         storage, C4_ROWS, d_vol)
    d_vol.free()
    err = sampled.rel_l2(got, ref["f64"])
    err_d = sampled.rel_l2(got, ref["f64_all_f64_decisions"])
    print("%s vs FP64 chain %.3g (FP64 decisions %.3g)" % (storage, err,
                                                          err_d))
    # filtered slabs of 8 tilts: decoded times the per-tilt normalisation
    # (X Y max|projection|) against the FP64 correction + filter
    # (the slabs some voxel reads at that tilt: the plan computes only those)
    first, last = plan.slab_range()
    assert (first <= last).all() and first.min() >= 0 and last.max() < ndf
    slabs = decoded_slab_rows(plan, storage, X, Y, T, ndf, C4_ROWS, C4_TILTS)
    want = O.filter_rows(ref["corrected64"], "f64")
    for i, t in enumerate(C4_TILTS):
        k0, k1 = int(first[t]), int(last[t]) + 1
        e = sampled.rel_l2(slabs[i][k0:k1] * (float(X) * Y * nrm[t]),
                           want[i][k0:k1])
        assert e < (2e-3 if storage == "f16" else 1e-2), (t, e)
    # the backprojection of the plan's own decoded slabs (every tilt) at the
    # first sampled row, through the oracle's backprojector
    if storage == "f16":
        y = int(C4_ROWS[1])
        allt = decoded_slab_rows(plan, storage, X, Y, T, ndf, [y],
                                 list(range(T)))[:, :, 0, :]
        sino = numpy.ascontiguousarray(numpy.swapaxes(allt, 0, 1))
        sino *= (float(X) * Y * nrm)[None, :, None]
        for t in range(T):  # slabs the plan does not compute: zero
            sino[:first[t], t] = 0
            sino[last[t] + 1:, t] = 0
        _, mn, mx = sampled.defocus_list(w["df"], ndf, X, Y, w["ps"])
        bp = O.project(sino, (Z, X), angles, X / 2.0, w["ps"], mn, mx)
        e = sampled.rel_l2(got[1], bp)
        print("BP of the decoded slabs vs oracle BP %.3g" % e)
        assert e < 1e-3
    plan.destroy()
    assert err < bound


def test_4096_fp32_in_row_bands_against_fp64(capi, c4):
    # FP32 slabs of every row (260 GB) do not fit beside the volume: two
    # bands of 2048 rows, each with the whole 2D correction again
    d_in, angles, ref, _ = c4
    w = C4
    X, Y, Z, ndf = w["X"], w["Y"], w["Z"], w["ndf"]
    print("4096^2 x 121, 32 defoci, f32 in bands of 2048 rows", flush=True)
    d_vol = C.DeviceArray((Y, Z, X))
    plan, got = run_rows(d_in, angles, X, Y, Z, ndf, w["ps"], w["df"], "f32",
                         C4_ROWS, d_vol, band_rows=2048)
    assert plan.band()[0] == 2048
    plan.destroy()
    d_vol.free()
    err = sampled.rel_l2(got, ref["f64"])
    print("f32 (bands) vs FP64 chain %.3g" % err)
    # the reference's FP32 arithmetic against FP64 (5e-5 at 1024^2,
    # 1.1e-4 at 2048^2 x 16 defoci)
    assert err < 5e-4


# ------------------------------------------------------- serving mix -----

def mix_specs(n, seed):
    rng = numpy.random.default_rng(seed)
    out = []
    for i in range(n):
        S = int((512, 1024, 2048)[rng.integers(3)])
        T = int(rng.integers(41, 122))
        ndf = int(rng.integers(1, 17))
        df = float(rng.uniform(10000.0, 40000.0))
        out.append((S, T, ndf, df, seed * 1000 + i))
    return out


def test_serving_mix_equals_plan_path_and_oracle(capi, pool):
    from guanaco_amd.serve import Server
    ps = 1.5
    specs = mix_specs(32, 1005)
    print("serving mix: %s" % [(S, T, k) for S, T, k, _, _ in specs],
          flush=True)
    inputs = []
    for S, T, ndf, df, sd in specs:
        angles = numpy.deg2rad(numpy.linspace(-60.0, 60.0, T)).astype(
            numpy.float32)
        d = C.DeviceArray((T, S, S))
        C.synth_projections(d, S, S, angles, depth=S // 4, width=S,
                            num_ellipsoids=40, seed=sd,
                            semi=(S / 64.0, S / 8.0), noise="poisson",
                            noise_level=40.0,
                            ctf=dict(wavelength=L300, Cs=2.7e7, defocus=df,
                                     phi=PHI, pixel_size=ps))
        kw = plan_kw(S, S, ndf, ps, df, "f16")
        kw.pop("centre")
        inputs.append((d, angles, S // 4, kw))
    C.call("gx_device_synchronize")

    def sample_rows(S):  # every row up to 1024, 64 spread rows at 2048
        return numpy.arange(S) if S <= 1024 else numpy.linspace(
            0, S - 1, 64).astype(int)

    served = []
    with Server(max_inflight=4, backfill=True) as srv:
        reqs = [srv.submit(d, a, Z, **kw) for d, a, Z, kw in inputs]
        infos = [r.wait() for r in reqs]
        assert all(i["state"] == "done" for i in infos), infos
        spans = sorted((i["start_s"], i["end_s"]) for i in infos)
        assert any(spans[k + 1][0] < spans[k][1]
                   for k in range(len(spans) - 1))
        for r, (S, T, ndf, df, _) in zip(reqs, specs):
            Z = S // 4
            ptr = r.info()["volume_ptr"]
            served.append(numpy.stack([
                C.DeviceView(ptr + int(y) * Z * S * 4, (Z, S)).to_host()
                for y in sample_rows(S)]))
            r.release()
    for (d, a, Z, kw), (S, T, ndf, df, _), got in zip(inputs, specs, served):
        plan = C.VolumePlan(S, S, a, Z, **kw)
        d_vol = C.DeviceArray((S, Z, S))
        plan.run(d, d_vol)
        C.call("gx_device_synchronize")
        alone = numpy.stack([C.DeviceView(d_vol.ptr + int(y) * Z * S * 4,
                                          (Z, S)).to_host()
                             for y in sample_rows(S)])
        plan.destroy()
        d_vol.free()
        numpy.testing.assert_array_equal(got, alone, err_msg=str((S, T, ndf)))
    # three requests (the smallest detectors) against the FP64 oracle
    order = sorted(range(len(specs)), key=lambda i: (specs[i][0], specs[i][1]))
    for i in order[:3]:
        S, T, ndf, df, _ = specs[i]
        d, a, Z, _ = inputs[i]
        sr = sample_rows(S)
        pos = numpy.linspace(0, sr.size - 1, 4).astype(int)
        proj = d.to_host()
        ref = sampled.reference_rows(pool, proj, a, sr[pos], Z, ps, df, ndf,
                                     PHI, fp32=False)
        e = sampled.rel_l2(served[i][pos], ref["f64"])
        print("request %d (%d^2, %d tilts, %d defoci) vs FP64 %.3g" % (
            i, S, T, ndf, e))
        assert e < 1e-3
    for d, _, _, _ in inputs:
        d.free()
