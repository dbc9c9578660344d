"""Benchmark of the reconstruction hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c4]

One step reconstructs one whole tilt series that is already resident in HBM:
CTF phase-flip correction (2D R2C, per-defocus flip, C2R), ramp filter and
backprojection of every defocus slab into the volume. The default workload
is the largest one that fits one GPU: 121 x 4096^2 projections, 32 defocus
slabs, a 4096 x 1024 x 4096 volume. Paths on the same input:

- batched: FP16 storage of the filtered slabs, FP32 accumulation, the whole
           request captured in one HIP graph and replayed (the value);
- batched, BF16 slabs (the same path, a separate run);
- plain:   FP32 storage, exact reference arithmetic, kernels launched
           eagerly; at 4096^2 x 32 defoci the FP32 slabs (260 GB) do not fit
           beside the volume, so the plan holds bands of detector rows and
           repeats the 2D correction per band (every row's 2D transform, the
           band's row pass), as a per-slice FP32 run must;
- served:  the request submitted repeatedly to the native serving loop
           (configurations whose requests fit the server twice over).

Before timing, each plan's backprojection tile is chosen by the autotuner on
the plan's own data. The value is the batched path's whole-job throughput in
voxel*tilt updates/s (updates = X * Y * Z * T; ndf does not multiply it).
rel-L2 of every path's volume is measured against the FP64 CPU oracle on
sampled y-slices (the corrected rows from the sampled-row FP64 correction,
then the oracle's filter and backprojection). The default run adds the
2048^2 x 16 and 1024^2 x 1 workloads as extra entries, the serving workload
(256 Poisson requests) and the CPU baseline (the oracle on the host cores,
bounded samples, extrapolated).

The roofline entry is for the backprojection kernel of the batched path,
timed with HIP events on the stream it is launched on: algorithmic HBM bytes
are the compulsory ones (every slab read once, the volume written once);
measured HBM bytes and the instruction counts that bound the kernel come from
the rocprofv3 counter summary committed for the same workload.
"""
import os

# single-threaded BLAS in this process and its CPU workers (each worker is
# one core of the oracle's sampled check and baseline)
os.environ["OPENBLAS_NUM_THREADS"] = "1"
os.environ["MKL_NUM_THREADS"] = "1"

import argparse  # noqa: E402
import json  # noqa: E402
import math  # noqa: E402
import platform  # noqa: E402
import sys  # noqa: E402
import threading  # noqa: E402
import time  # noqa: E402

import numpy  # noqa: E402

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("voxel·tilt updates/s at ≤1e-3 rel-L2 vs FP64, batched vs "
          "plain; p50/p99 latency")
UNIT = "voxel·tilt updates/s"
HBM_PEAK_GBS = 8000.0
CLOCK_HZ = 2.4e9
NUM_CU = 256
# LDS: ds_read_b128 conflict-free, measured on gfx950 with probe kernels at
# 190-200 bytes per clock per CU: 200 x 256 CU x 2.4 GHz
LDS_PEAK_GBS = 200 * NUM_CU * CLOCK_HZ / 1e9
# VALU issue: chains of FP32 FMA / mul / add issue at most 0.245 wave64
# instructions per SIMD per clock (probe kernels, 8 waves per SIMD)
VALU_ISSUE_PEAK = 0.245
WORKERS_MAX = 16  # CPU share of one GPU on the box

# name: sizes, tilt step (deg), defocus, phantom and noise, seed
WORKLOADS = {
    "c4": dict(X=4096, Y=4096, T=121, step=1.0, Z=1024, ndf=32, ps=1.0,
               df=25000.0, phi=math.asin(0.1), ne=400, semi=(80.0, 640.0),
               dose=30.0, seed=1003, storages=("f16", "bf16"),
               name="reduced-precision defocus-range correction, ndf 32: "
                    "121 x 4096^2 projections (+-60 deg, 1 deg), 1 A/px, "
                    "300 kV, Cs 2.7 mm, df 25000 A, amplitude contrast 0.1 "
                    "-> 4096 x 1024 x 4096 volume; FP16 and BF16 slabs in "
                    "separate runs, FP32 accumulation"),
    "c3": dict(X=2048, Y=2048, T=61, step=2.0, Z=512, ndf=16, ps=1.5,
               df=30000.0, phi=math.asin(0.1), ne=400, semi=(40.0, 320.0),
               dose=30.0, seed=1002,
               name="defocus-range CTF correction, ndf 16: 61 x 2048^2 "
                    "projections (+-60 deg, 2 deg), 1.5 A/px, df 30000 A -> "
                    "2048 x 512 x 2048 volume"),
    "c2": dict(X=1024, Y=1024, T=121, step=1.0, Z=256, ndf=1, ps=2.0,
               df=20000.0, phi=math.asin(0.1), ne=200, semi=(20.0, 160.0),
               dose=50.0, seed=1001, served=True,
               name="single-defocus CTF-corrected FBP: 121 x 1024^2 "
                    "projections (+-60 deg, 1 deg), 2 A/px, 300 kV, "
                    "Cs 2.7 mm, df 20000 A, amplitude contrast 0.1 -> "
                    "1024 x 256 x 1024 volume"),
    "c1": dict(X=256, Y=256, T=61, step=2.0, Z=128, ndf=1, ps=1.0,
               df=None, phi=0.0, ne=20, semi=(5.0, 40.0), dose=None,
               seed=1000, served=True,
               name="plain FBP, no CTF: 61 x 256^2 projections -> "
                    "256 x 128 x 256 volume"),
}


def percentile(a, q):
    a = sorted(a)
    if not a:
        return None
    k = (len(a) - 1) * q / 100.0
    f, c = math.floor(k), math.ceil(k)
    return a[f] if f == c else a[f] + (a[c] - a[f]) * (k - f)


def log(msg):
    print("bench: " + msg, file=sys.stderr, flush=True)


def workload_angles(w):
    T = w["T"]
    return numpy.deg2rad(numpy.arange(T) * w["step"] -
                         0.5 * (T - 1) * w["step"]).astype(numpy.float32)


def host_info():
    """CPU model, logical CPUs and the CPUs this process may run on."""
    model = platform.processor() or ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return {"cpu_model": model, "nproc": os.cpu_count() or 1,
            "usable_cpus": usable}


# ---------------------------------------------------------------- GPU -----

def _wavelength():
    """The 300 kV electron wavelength (A), from the package's host code."""
    import guanaco_amd.detail as ext
    return ext.get_electron_wavelength(300000.0)


def _plan_defoci(df, ndf, width, centre, ps):
    """(min, max, step, n) of the automatic defocus range, from the
    package's get_defocus_steps (the GPU legs use no oracle code)."""
    from guanaco_amd.detail.correct import get_defocus_steps
    return get_defocus_steps(df, ndf, None, width, centre, ps)


def plan_kwargs(w, storage):
    kw = dict(num_defocus=w["ndf"], pixel_size=w["ps"], centre=w["X"] / 2.0,
              storage=storage)
    if w["df"] is not None:
        centre = numpy.full(w["Y"], w["X"] / 2.0, numpy.float32)
        mn, mx, st, n = _plan_defoci(w["df"], w["ndf"], w["X"], centre,
                                     w["ps"])
        dfs = numpy.array([mn + d * st for d in range(n)], numpy.float64)
        kw.update(min_defocus=mn - w["df"], max_defocus=mx - w["df"],
                  defocus=dfs, wavelength=_wavelength(), Cs=2.7e7,
                  phi=w["phi"])
    return kw


def band_rows_that_fit(C, w, angles, storage, reserve):
    """0 when the plan of every row fits in the free HBM less `reserve`
    bytes, else the largest band of about Y / n rows (a multiple of 32)
    whose plan does."""
    avail = C.mem_free() - reserve
    kw = plan_kwargs(w, storage)
    base = dict(num_pixels=w["X"], num_rows=w["Y"], angles=angles,
                depth=w["Z"], **{k: v for k, v in kw.items()
                                 if k not in ("wavelength", "Cs", "phi",
                                              "centre")})
    Y = w["Y"]
    for n in range(1, Y // 32 + 2):
        band = min(Y, ((Y + n - 1) // n + 31) // 32 * 32)
        need = C.plan_bytes(band_rows=0 if band == Y else band, **base)
        if need <= 0.95 * avail:
            return 0 if band == Y else band
    raise RuntimeError("no band of rows fits the free HBM")


def time_path(C, plan, d_in, d_vol, stream, steps, warmup, graph):
    """Whole-request steps; per-step device times from HIP events on the
    plan's stream (with a graph, one graph launch per step)."""
    ev = [C.Event() for _ in range(steps + 1)]
    g = None
    if graph:
        C.call("gx_graph_capture_begin", stream.handle)
        plan.run(d_in, d_vol, stream)
        p = C.ctypes.c_void_p()
        C.call("gx_graph_capture_end", stream.handle, C.ctypes.byref(p))
        g = p.value

    def step():
        if g is not None:
            C.call("gx_graph_launch", g, stream.handle)
        else:
            plan.run(d_in, d_vol, stream)

    for _ in range(warmup):
        step()
    stream.synchronize()
    t0 = time.perf_counter()
    ev[0].record(stream)
    for i in range(steps):
        step()
        ev[i + 1].record(stream)
    stream.synchronize()
    wall = time.perf_counter() - t0
    lat = [ev[i].elapsed_ms(ev[i + 1]) for i in range(steps)]
    total = ev[0].elapsed_ms(ev[steps])
    if g is not None:
        C.call("gx_graph_destroy", g)
    for e in ev:
        e.destroy()
    return total / steps, lat, wall


def time_kernel(C, fn, stream, reps):
    """Mean device time of fn()'s launches on `stream` (HIP events on the
    stream the kernels are launched on), after one untimed call."""
    a, b = C.Event(), C.Event()
    fn()
    stream.synchronize()
    a.record(stream)
    for _ in range(reps):
        fn()
    b.record(stream)
    stream.synchronize()
    ms = a.elapsed_ms(b) / reps
    a.destroy()
    b.destroy()
    return ms


def measure_path(C, w, angles, d_in, d_vol, stream, storage, graph, steps,
                 warmup, rows, band_rows=0, bp_kind=None):
    """Plan, autotune, time the whole request and its two stages; returns
    the result entry and the volume's sampled rows (host)."""
    X, Y, Z, T = w["X"], w["Y"], w["Z"], w["T"]
    updates = float(X) * Y * Z * T
    plan = C.VolumePlan(X, Y, angles, Z, band_rows=band_rows,
                        **plan_kwargs(w, storage))
    tuned = None
    if bp_kind:
        plan.set_backprojector(bp_kind)
    else:
        tuned = plan.autotune(d_in, d_vol, stream)
    kind, slots = plan.backprojector()
    band = plan.band()[0]
    ms, lat, wall = time_path(C, plan, d_in, d_vol, stream, steps, warmup,
                              graph)
    if band >= Y:
        filt_ms = time_kernel(C, lambda: plan.filter(d_in, stream), stream,
                              3)
        bp_ms = time_kernel(C, lambda: plan.backproject(d_vol, stream),
                            stream, max(3, min(steps, 10)))
    else:
        # one band's stages, times the number of bands
        nb = -(-Y // band)
        filt_ms = nb * time_kernel(
            C, lambda: plan.filter_rows(d_in, 0, band, stream), stream, 2)
        bp_ms = nb * time_kernel(
            C, lambda: plan.backproject_rows(0, band, d_vol, stream), stream,
            2)
    plan.run(d_in, d_vol, stream)
    stream.synchronize()
    vol_rows = numpy.stack([C.DeviceView(d_vol.ptr + int(y) * Z * X * 4,
                                         (Z, X)).to_host() for y in rows])
    entry = dict(storage=storage, ms_per_step=ms, steps=steps,
                 value=updates / (ms * 1e-3), p50_ms=percentile(lat, 50),
                 p99_ms=percentile(lat, 99), wall_s=wall, filter_ms=filt_ms,
                 bp_ms=bp_ms, bp_updates_per_s=updates / (bp_ms * 1e-3),
                 backprojector=("staged, %d slab(s) per tilt" % slots
                                if kind == C.BP_STAGED else "direct"),
                 tile_rows=plan.tile()[0], launch="graph" if graph
                 else "eager", band_rows=band if band < Y else None,
                 slab_bytes=plan.slabs()[1])
    if tuned is not None:
        entry["autotune"] = tuned
    plan.destroy()
    log("%s %s: %.2f ms/step, BP %.2f ms, filter %.2f ms%s" % (
        "graph" if graph else "eager", storage, ms, bp_ms, filt_ms,
        ", bands of %d rows" % band if band < Y else ""))
    return entry, vol_rows


def served_path(C, w, angles, d_in, storage, n, max_inflight):
    """The same request n times through the native server (gx_server_*),
    max_inflight on the device at once: each request's correction, filter
    and backprojection run on its own stream and overlap with the others'.
    Throughput is n * updates over the wall time from the first submission
    to the last completion (after 2 warm-up requests)."""
    from guanaco_amd.serve import Server
    kw = plan_kwargs(w, storage)
    kw.pop("centre")
    with Server(max_inflight=max_inflight) as srv:
        for r in [srv.submit(d_in, angles, w["Z"], **kw) for _ in range(2)]:
            r.wait()
            r.release()
        t0 = time.perf_counter()
        reqs = [srv.submit(d_in, angles, w["Z"], **kw) for _ in range(n)]
        infos = []
        for r in reqs:
            infos.append(r.wait())
            r.release()
        wall = time.perf_counter() - t0
    assert all(i["state"] == "done" for i in infos), infos
    updates = float(w["X"]) * w["Y"] * w["Z"] * w["T"]
    lat = [i["latency_ms"] for i in infos]
    out = dict(storage=storage, requests=n, max_inflight=max_inflight,
               value=n * updates / wall, ms_per_request=1e3 * wall / n,
               p50_ms=percentile(lat, 50), p99_ms=percentile(lat, 99),
               device_ms_mean=sum(i["device_ms"] for i in infos) / n)
    log("served (%s, %d in flight) %.2f ms/request" % (
        storage, max_inflight, out["ms_per_request"]))
    return out


def host_path(w, angles, proj, storage, reps=3):
    """The drop-in call with host buffers (guanaco_amd.reconstruct_projections
    on a numpy tilt series, numpy volume returned): PCIe copies, plan set-up
    and the device path, as a reference user's call sees it."""
    import guanaco_amd as G
    vol = numpy.empty((w["Y"], w["Z"], w["X"]), numpy.float32)
    kw = dict(pixel_size=w["ps"], depth=w["Z"], storage=storage,
              reconstruction=vol)
    if w["df"] is not None:
        kw.update(energy=300, defocus=w["df"], num_defocus=w["ndf"],
                  spherical_aberration=2.7,
                  phase_shift=math.degrees(w["phi"]))
    G.reconstruct_projections(proj, angles, **kw)
    t0 = time.perf_counter()
    for _ in range(reps):
        G.reconstruct_projections(proj, angles, **kw)
    ms = 1e3 * (time.perf_counter() - t0) / reps
    updates = float(w["X"]) * w["Y"] * w["Z"] * w["T"]
    return dict(storage=storage, ms_per_request=ms,
                value=updates / (ms * 1e-3),
                bytes_in=int(proj.nbytes), bytes_out=int(vol.nbytes))


# ---------------------------------------------------- CPU oracle legs -----
# Run in worker processes forked before this process touches the GPU; the
# oracle (oracle/) is the checker and the CPU baseline here, never the GPU
# legs.

def oracle_check(w, proj, angles, rows, vols, pool, fp32_oracle):
    """rel-L2 of the GPU volumes' sampled rows against the FP64 oracle chain
    (sampled-row FP64 correction, FP64 filter and backprojection), with the
    slab and tap decisions from the FP32 coordinates (as the reference's
    arithmetic takes them) and, as a second figure, from FP64 coordinates.
    fp32_oracle: also the full FP32 oracle chain, for the FP32 path."""
    from oracle import sampled
    ref = sampled.reference_rows(pool, proj, angles, rows, w["Z"], w["ps"],
                                 w["df"], w["ndf"], w["phi"],
                                 fp32=fp32_oracle)
    out = {}
    for name, v in vols.items():
        out[name + "_vs_fp64"] = sampled.rel_l2(v, ref["f64"])
        out[name + "_vs_fp64_all_f64_decisions"] = sampled.rel_l2(
            v, ref["f64_all_f64_decisions"])
        if name == "f32" and "f32" in ref:
            out["f32_vs_fp32_oracle"] = sampled.rel_l2(v, ref["f32"])
    out["fp64_f32_decisions_vs_all_f64_decisions"] = sampled.rel_l2(
        ref["f64"], ref["f64_all_f64_decisions"])
    out["sample"] = "y-slices %s, every tilt and defocus" % [
        int(r) for r in rows]
    return out


def _oracle_defoci(w):
    from oracle import sampled
    return sampled.defocus_list(w["df"], w["ndf"], w["X"], w["Y"], w["ps"])


def _ctf_c64(w, df):
    from oracle import oracle as O
    return O.ctf_simple_grid(O.electron_wavelength(300000), df, 2.7e7, 0, 0,
                             w["phi"], w["X"], w["Y"], w["ps"]).astype(
        numpy.complex64)


def _timed(fn, *a, **kw):
    t0 = time.perf_counter()
    fn(*a, **kw)
    return time.perf_counter() - t0


def _baseline_task(a):
    """Seconds of one oracle unit of work (one core)."""
    from oracle import oracle as O
    kind = a[0]
    if kind == "corr":
        _, img, ctf = a
        return _timed(O.corr, img, ctf, precision="f32")
    if kind == "filter":
        _, rows = a
        return _timed(O.filter_rows, rows, "f32")
    _, sino, Z, X, angles, ps, mn, mx = a
    return _timed(O.project, sino, (Z, X), angles, X / 2.0, ps, mn, mx)


def cpu_baseline(w, proj, angles, pool, nworkers):
    """The oracle (FP32 restatement of the reference's CPU path, project-
    written FFT in place of FFTW) timed on bounded samples of the workload
    and extrapolated linearly, on one core and on `nworkers` processes:
    - correction of one tilt with 1 and 3 defoci (per-tilt cost = forward
      part + ndf inverse parts);
    - the ramp filter of 2 x 64 sinogram rows (N points, 2N transforms);
    - backprojection of 4 slices over a band of the depth (cost per
      voxel-tilt update)."""
    T, Y, X, Z, ndf = w["T"], w["Y"], w["X"], w["Z"], w["ndf"]
    angles = numpy.asarray(angles, numpy.float32)
    mn = mx = 0.0
    t_corr = 0.0
    tasks_c = []
    if w["df"] is not None:
        dfs, mn, mx = _oracle_defoci(w)
        ctf3 = numpy.stack([_ctf_c64(w, df) for df in dfs[:3]])
        tasks_c = [("corr", proj[0], ctf3[:1]), ("corr", proj[0], ctf3)]
        c1, c3 = [pool.apply(_baseline_task, (t,)) for t in tasks_c]
        d = max(0.0, (c3 - c1) / 2.0)
        t_corr = c1 + (ndf - 1) * d
    rows = numpy.ascontiguousarray(proj[:64, Y // 2, :])
    t_filt = min(pool.apply(_baseline_task, (("filter", rows),))
                 for _ in range(2)) / rows.shape[0]
    zb = max(8, min(Z, int(2e8 // (X * T))))  # about 2e8 updates per slice
    rng = numpy.random.default_rng(0)
    sino = rng.standard_normal((ndf, T, X)).astype(numpy.float32)
    if ndf == 1:
        sino = sino[0]
    bp_task = ("bp", sino, zb, X, angles, w["ps"], mn, mx)
    t_bp = pool.apply(_baseline_task, (bp_task,)) / (zb * X * T)
    per_slice = t_filt * ndf * T + t_bp * Z * X * T
    total1 = t_corr * T + per_slice * Y
    sample = ("correction of 1 tilt with 1 and 3 of %d defoci (%.2f s per "
              "tilt, extrapolated), ramp filter of 64 rows of %d (%.4f s "
              "per row), backprojection of %d of %d depth rows of one slice "
              "(%.2f ns per voxel-tilt update); linear in tilts, rows and "
              "updates; project-written FFT in place of FFTW" % (
                  ndf, t_corr, X, t_filt, zb, Z, 1e9 * t_bp))
    # all cores: the same units on every worker at once (wall clock)
    n = nworkers
    t0 = time.perf_counter()
    units = []
    if tasks_c:
        units += [tasks_c[1]] * n
    units += [("filter", rows)] * n + [bp_task] * n
    pool.map(_baseline_task, units, chunksize=1)
    wall = time.perf_counter() - t0
    # one-core seconds of one worker's share; wall / that = the slowdown of
    # every core working at once
    one_round = ((c3 if tasks_c else 0.0) + t_filt * rows.shape[0] +
                 t_bp * zb * X * T)
    total_n = total1 * (wall / one_round) / n
    info = host_info()
    return {
        "value": X * Y * Z * T / total_n,
        "unit": UNIT,
        "cores": n,
        "kind": "port",
        "sample": "%s; the same units run on %d worker processes at once "
                  "(wall clock %.2f s for %.2f s of one-core work each) and "
                  "the whole job extrapolated from that rate; workers = the "
                  "CPUs this process may use, at most the GPU's 16-CPU share "
                  "of the host" % (sample, n, wall, one_round),
        "seconds_estimated_whole_job": total_n,
        **info,
        "one_core": {
            "value": X * Y * Z * T / total1,
            "unit": UNIT,
            "cores": 1,
            "seconds_estimated_whole_job": total1,
        },
    }


# ------------------------------------------------------------- serving ----

SERVE_SIZES = (512, 1024, 2048)
SERVE_DESC = ("Poisson arrivals of whole-volume requests: detector S^2 with S "
              "uniform in {512, 1024, 2048}, T uniform in 41..121 tilts "
              "spread evenly over +-60 deg, ndf uniform in 1..16, df uniform "
              "in 10000..40000 A, Cs 2.7 mm, 300 kV, 1.5 A/px, volume "
              "S x S/4 x S, FP16 slabs; every request its own seeded "
              "ellipsoid phantom (CTF forward model at its df, Poisson "
              "40 e-/px), device-resident before the arrivals start; "
              "arrival rate 0.7 x the capacity measured closed-loop first; "
              "footprint admission with bounded backfill")


def serving_bench(C, n_requests, stream, seed=1005, max_inflight=4,
                  backfill=True):
    """Continuous batching through the native server (gx_server_*): every
    request's tilt series is synthesised first (its own seed), capacity is
    measured closed-loop on 24 of them, then n_requests arrive as a Poisson
    process at 0.7 x capacity; latency is submission to completion."""
    from guanaco_amd.serve import Server
    l = _wavelength()
    ps = 1.5
    rng = numpy.random.default_rng(seed)
    specs = []
    for i in range(n_requests):
        S = int(SERVE_SIZES[rng.integers(len(SERVE_SIZES))])
        T = int(rng.integers(41, 122))
        ndf = int(rng.integers(1, 17))
        df = float(rng.uniform(10000.0, 40000.0))
        specs.append((S, T, ndf, df, seed * 1000 + i))
    in_bytes = sum(4 * S * S * T for S, T, _, _, _ in specs)
    # inputs first, then the server's budget from what is left
    inputs = []
    t0 = time.perf_counter()
    for S, T, ndf, df, sd in specs:
        angles = numpy.deg2rad(numpy.linspace(-60.0, 60.0, T)).astype(
            numpy.float32)
        d = C.DeviceArray((T, S, S))
        C.synth_projections(d, S, S, angles, depth=S // 4, width=S,
                            num_ellipsoids=40, seed=sd,
                            semi=(S / 64.0, S / 8.0), noise="poisson",
                            noise_level=40.0,
                            ctf=dict(wavelength=l, Cs=2.7e7, defocus=df,
                                     phi=math.asin(0.1), pixel_size=ps),
                            stream=stream)
        centre = numpy.full(S, S / 2.0, numpy.float32)
        mn, mx, st, n = _plan_defoci(df, ndf, S, centre, ps)
        dfs = numpy.array([mn + k * st for k in range(n)], numpy.float64)
        kw = dict(num_defocus=ndf, pixel_size=ps, min_defocus=mn - df,
                  max_defocus=mx - df, storage="f16", defocus=dfs,
                  wavelength=l, Cs=2.7e7, phi=math.asin(0.1))
        inputs.append((d, angles, S // 4, kw))
    stream.synchronize()
    log("serving: %d seeded tilt series synthesised (%.1f GB) in %.1f s" % (
        n_requests, in_bytes / 1e9, time.perf_counter() - t0))
    budget = int(0.85 * C.mem_free())
    srv = Server(budget_bytes=budget, max_inflight=max_inflight,
                 backfill=backfill)
    try:
        # capacity: the first 24 requests submitted at once
        reqs = [srv.submit(d, a, Z, **kw) for d, a, Z, kw in inputs[:24]]
        infos = []
        for r in reqs:
            infos.append(r.wait())
            r.release()
        span = max(i["end_s"] for i in infos) - min(i["submit_s"]
                                                    for i in infos)
        capacity = len(reqs) / span
        log("serving capacity %.1f requests/s (24 closed-loop, budget "
            "%.1f GB)" % (capacity, budget / 1e9))
        rate = 0.7 * capacity
        arrivals = numpy.cumsum(numpy.random.default_rng(seed + 1)
                                .exponential(1.0 / rate, n_requests))
        handles = []
        done = []
        lock = threading.Lock()
        submitted = threading.Semaphore(0)

        def waiter():
            for i in range(n_requests):
                submitted.acquire()
                with lock:
                    r = handles[i]
                inf = r.wait()
                done.append(inf)
                r.release()

        th = threading.Thread(target=waiter)
        th.start()
        t0 = time.perf_counter()
        for i, (d, a, Z, kw) in enumerate(inputs):
            dt = arrivals[i] - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            r = srv.submit(d, a, Z, **kw)
            with lock:
                handles.append(r)
            submitted.release()
        th.join()
        st = srv.stats()
    finally:
        srv.close()
        for d, _, _, _ in inputs:
            d.free()
    ok = [i for i in done if i["state"] == "done"]
    lat = [i["latency_ms"] for i in ok]
    span = (max(i["end_s"] for i in ok) - min(i["submit_s"] for i in ok)
            if ok else 0.0)
    return {
        "description": SERVE_DESC,
        "requests": n_requests,
        "completed": len(ok),
        "failed": sum(1 for i in done if i["state"] == "failed"),
        "capacity_requests_per_s": capacity,
        "arrival_rate_per_s": rate,
        "p50_ms": percentile(lat, 50),
        "p99_ms": percentile(lat, 99),
        "mean_ms": sum(lat) / max(1, len(lat)),
        "max_ms": max(lat) if lat else None,
        "device_ms_mean": sum(i["device_ms"] for i in ok) / max(1, len(ok)),
        "updates_per_s": (sum(i["updates"] for i in ok) / span
                          if span > 0 else 0.0),
        "max_inflight": max_inflight,
        "budget_bytes": int(st["budget_bytes"]),
        "peak_bytes": int(st["peak_bytes"]),
        "input_bytes": int(in_bytes),
        "backfill": backfill,
        "seed": seed,
    }


# ----------------------------------------------------------- roofline -----

def load_counters(name):
    """Per-dispatch counters of a kernel from a committed rocprofv3 summary:
    the 'key: value' lines after 'kernel: <name>'."""
    path = os.path.join(ROOT, "profiles", name)
    out = {}
    try:
        with open(path) as f:
            cur = None
            for line in f:
                line = line.strip()
                if line.startswith("kernel:"):
                    cur = line.split(":", 1)[1].strip()
                    out[cur] = {}
                elif cur and ":" in line and not line.startswith("#"):
                    k, v = line.split(":", 1)
                    try:
                        out[cur][k.strip()] = float(v)
                    except ValueError:
                        pass
    except OSError:
        pass
    return out


def bp_roofline(w, entry, counters):
    """Roofline of the backprojection kernel of `entry`, timed live.

    HBM: algorithmic bytes = the compulsory traffic of one launch, every
    filtered slab read once and the volume written once; measured bytes from
    the FETCH_SIZE / WRITE_SIZE counters (committed summary of the same
    workload). What binds it is instruction issue: VALU instructions per
    SIMD per clock (counted instructions / this run's launch time) against
    the FP32 issue rate probe kernels reach, and LDS bytes against the
    measured ds_read_b128 peak. The gather model (two taps of s
    bytes per update) is kept as a labelled figure: staging makes its taps
    LDS reads, so it exceeds the HBM peak."""
    X, Y, Z, T = w["X"], w["Y"], w["Z"], w["T"]
    updates = float(X) * Y * Z * T
    s = 2 if entry["storage"] in ("f16", "bf16") else 4
    ms = entry["bp_ms"]
    compulsory = entry["slab_bytes"] + 4.0 * Y * Z * X
    achieved = compulsory / (ms * 1e-3) / 1e9
    roof = {
        "bound": "hbm",
        "kernel": "k_bp_lds (backprojection, %s slabs)" % entry["storage"],
        "achieved": achieved,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS,
        "traffic": None,
        "algorithmic_bytes_per_launch": compulsory,
        "algorithmic_bytes": "filtered slabs read once + FP32 volume "
                             "written once",
        "launch_ms": ms,
        "updates_per_launch": updates,
    }
    c = counters
    if c:
        traffic = c.get("fetch_bytes_scaled", 0) + c.get(
            "write_bytes_scaled", 0)
        roof["traffic"] = traffic or None
        if traffic:
            roof["traffic_vs_algorithmic"] = traffic / compulsory
        simd_clk = NUM_CU * 4 * CLOCK_HZ * ms * 1e-3
        if "valu_insts" in c:
            v = c["valu_insts"] / simd_clk
            roof["binding"] = {
                "resource": "VALU issue",
                "achieved": v, "peak": VALU_ISSUE_PEAK,
                "unit": "wave64 VALU instructions per SIMD per clock",
                "frac": v / VALU_ISSUE_PEAK,
                "instructions_per_update": c["valu_insts"] * 64 / updates,
            }
        if "lds_insts" in c:
            lds = c["lds_insts"] * c.get("lds_bytes_per_inst", 1024) / (
                ms * 1e-3) / 1e9
            roof["lds"] = {"achieved": lds, "peak": LDS_PEAK_GBS,
                           "unit": "GB/s", "frac": lds / LDS_PEAK_GBS}
        roof["counters"] = "profiles/%s" % c.get("_file", "")
    b_alg = 2 * s + 4.0 / T
    roof["gather_model"] = {
        "bytes_per_update": b_alg,
        "achieved": updates * b_alg / (ms * 1e-3) / 1e9, "unit": "GB/s",
        "note": "two taps of s bytes per update counted as HBM traffic; the "
                "staged kernel reads them from LDS",
    }
    return roof


# --------------------------------------------------------------- main -----

def run_workload(C, name, w, args, pool, stream, headline):
    """Paths of one workload; returns its result entry."""
    from guanaco_amd import _capi  # noqa: F401
    X, Y, T, Z = w["X"], w["Y"], w["T"], w["Z"]
    angles = workload_angles(w)
    updates = float(X) * Y * Z * T
    d_in = C.DeviceArray((T, Y, X))
    ctf = None
    if w["df"] is not None:
        ctf = dict(wavelength=_wavelength(), Cs=2.7e7, defocus=w["df"],
                   phi=w["phi"], pixel_size=w["ps"])
    C.synth_projections(d_in, X, Y, angles, depth=Z, width=X,
                        num_ellipsoids=w["ne"], seed=w["seed"],
                        semi=w["semi"],
                        noise="poisson" if w["dose"] else "gaussian",
                        noise_level=w["dose"] if w["dose"] else 0.01,
                        ctf=ctf, stream=stream)
    stream.synchronize()
    d_vol = C.DeviceArray((Y, Z, X))
    rows = numpy.linspace(0, Y - 1, 4).astype(int)
    steps = args.steps if headline else max(3, min(args.steps, 10))
    warmup = args.warmup if headline else 1
    paths, vols = {}, {}
    storages = w.get("storages", ("f16",))
    for i, st in enumerate(storages):
        key = "batched" if i == 0 else "batched_" + st
        paths[key], vols[st] = measure_path(
            C, w, angles, d_in, d_vol, stream, st, True, steps, warmup, rows)
    # plain FP32: eager launches, bands of rows when the slabs do not fit
    band = band_rows_that_fit(C, w, angles, "f32", 1 << 30)
    paths["plain"], vols["f32"] = measure_path(
        C, w, angles, d_in, d_vol, stream, "f32", False,
        steps if band == 0 else max(2, min(steps, 5)), 1 if band else warmup,
        rows, band_rows=band)
    d_vol.free()
    if w.get("served") and not args.no_served:
        paths["served"] = served_path(C, w, angles, d_in, "f16",
                                      args.served, args.served_inflight)
    proj = None
    if not args.no_check or (headline and not args.no_cpu):
        proj = d_in.to_host()
    host = None
    if name == "c2" and not args.no_host and proj is not None:
        host = host_path(w, angles, proj, "f16")
        log("host buffers (f16) %.1f ms/request incl. PCIe" % (
            host["ms_per_request"]))
    d_in.free()
    head = paths["batched"]
    entry = {
        "workload": w["name"],
        "volume": [Y, Z, X], "tilts": T, "num_defocus": w["ndf"],
        "updates_per_step": updates,
        "value": head["value"], "ms_per_step": head["ms_per_step"],
        "paths": paths,
        "speedup_batched_vs_plain": head["value"] / paths["plain"]["value"],
        "seed": w["seed"],
    }
    for st in storages[1:]:
        entry["speedup_batched_%s_vs_plain" % st] = (
            paths["batched_" + st]["value"] / paths["plain"]["value"])
    if "served" in paths:
        entry["speedup_served_vs_plain"] = (paths["served"]["value"] /
                                            paths["plain"]["value"])
    if host:
        entry["host_buffers"] = host
    if not args.no_check:
        t0 = time.perf_counter()
        entry["rel_l2"] = oracle_check(w, proj, angles, rows, vols, pool,
                                       fp32_oracle=name != "c4")
        log("%s oracle check %.1f s %s" % (name, time.perf_counter() - t0,
                                            entry["rel_l2"]))
    return entry, proj, angles


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c4",
                    choices=sorted(WORKLOADS) + ["c5"],
                    help="headline workload (c5: the serving benchmark "
                         "only)")
    ap.add_argument("--extras", default=None,
                    help="comma-separated extra workloads (default c3,c2 "
                         "after c4, none otherwise; 'none' for none)")
    ap.add_argument("--serve-requests", type=int, default=256,
                    help="requests of the serving workload (0: skip)")
    ap.add_argument("--serve-inflight", type=int, default=4)
    ap.add_argument("--served", type=int, default=32,
                    help="requests of the served path of c1/c2 (32: the "
                         "first and last pair, which overlap nothing, weigh "
                         "little)")
    ap.add_argument("--served-inflight", type=int, default=2)
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-host", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--only", choices=["batched", "plain"], default=None,
                    help="with --profile-kernels: one path's plans")
    ap.add_argument("--storage", default="f16", choices=["f16", "bf16"],
                    help="with --profile-kernels: the batched storage")
    ap.add_argument("--profile-kernels", action="store_true",
                    help="no timing: run the filter pass once, the "
                         "backprojection twice and a 256 MiB copy kernel "
                         "twice (for rocprofv3 traces and counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.profile_kernels:
        return profile_kernels(args)
    if args.config == "c5":
        return main_serving(args)
    extras = args.extras
    if extras is None:
        extras = "c3,c2" if args.config == "c4" else "none"
    extras = [] if extras == "none" else [e for e in extras.split(",") if e]
    # CPU workers are forked before anything touches the GPU, so they
    # inherit no device state
    nworkers = min(WORKERS_MAX, host_info()["usable_cpus"])
    pool = None
    if not args.no_check or not args.no_cpu:
        from oracle import sampled
        pool = sampled.make_pool(nworkers, "fork")

    from guanaco_amd import _capi as C
    if C.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    stream = C.Stream()
    log("%s on %s" % (args.config, C.device_info(0)["name"]))
    w = WORKLOADS[args.config]
    entry, proj, angles = run_workload(C, args.config, w, args, pool, stream,
                                       True)
    head = entry["paths"]["batched"]
    counters = load_counters("counters_%s_bp.txt" % args.config).get(
        "k_bp_lds %s" % head["storage"], {})
    if counters:
        counters["_file"] = "counters_%s_bp.txt" % args.config
    line = {
        "metric": METRIC,
        "value": head["value"],
        "unit": UNIT,
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded %d-ellipsoid phantom, analytic "
                "projections%s, %s (seed %d)" % (
                    w["ne"], ", CTF forward model" if w["df"] else "",
                    "Poisson %g e-/px" % w["dose"] if w["dose"]
                    else "Gaussian noise 1%% of max", w["seed"]),
        "config": {
            "workload": w["name"],
            "path": "batched",
            "storage": head["storage"],
            "accumulate": "f32",
            "volume": entry["volume"],
            "tilts": entry["tilts"],
            "num_defocus": entry["num_defocus"],
            "updates_per_step": entry["updates_per_step"],
            "fits_in_driver_run": True,
        },
        "latency_ms": {"p50": head["p50_ms"], "p99": head["p99_ms"]},
        "paths": entry["paths"],
        "roofline": bp_roofline(w, head, counters),
    }
    for k in entry:
        if k.startswith("speedup") or k in ("rel_l2", "host_buffers"):
            line[k] = entry[k]
    extra = {}
    for name in extras:
        e, _, _ = run_workload(C, name, WORKLOADS[name], args, pool, stream,
                               False)
        extra[name] = e
    if extra:
        line["extra_workloads"] = extra
    if args.serve_requests > 0:
        line["serving"] = serving_bench(C, args.serve_requests, stream,
                                        max_inflight=args.serve_inflight)
        log("serving %s" % {k: v for k, v in line["serving"].items()
                            if k != "description"})
    if not args.no_cpu:
        t0 = time.perf_counter()
        line["cpu_baseline"] = cpu_baseline(w, proj, angles, pool, nworkers)
        log("CPU oracle %.3g updates/s on 1 core, %.3g on %d (%.1f s)" % (
            line["cpu_baseline"]["one_core"]["value"],
            line["cpu_baseline"]["value"], nworkers,
            time.perf_counter() - t0))
    if pool is not None:
        pool.close()
    print(json.dumps(line), flush=True)


def profile_kernels(args):
    """One filter pass, two backprojections per plan and a 256 MiB copy
    kernel twice, for rocprofv3 traces and counters (no timing)."""
    from guanaco_amd import _capi as C
    w = WORKLOADS[args.config]
    X, Y, T, Z = w["X"], w["Y"], w["T"], w["Z"]
    angles = workload_angles(w)
    stream = C.Stream()
    d_in = C.DeviceArray((T, Y, X))
    ctf = None
    if w["df"] is not None:
        ctf = dict(wavelength=_wavelength(), Cs=2.7e7, defocus=w["df"],
                   phi=w["phi"], pixel_size=w["ps"])
    C.synth_projections(d_in, X, Y, angles, depth=Z, width=X,
                        num_ellipsoids=w["ne"], seed=w["seed"],
                        semi=w["semi"],
                        noise="poisson" if w["dose"] else "gaussian",
                        noise_level=w["dose"] if w["dose"] else 0.01,
                        ctf=ctf, stream=stream)
    d_vol = C.DeviceArray((Y, Z, X))
    storages = {"batched": (args.storage,), "plain": ("f32",),
                None: (args.storage, "f32")}[args.only]
    for storage in storages:
        band = band_rows_that_fit(C, w, angles, storage, 1 << 30)
        plan = C.VolumePlan(X, Y, angles, Z, band_rows=band,
                            **plan_kwargs(w, storage))
        b = plan.band()[0]
        plan.filter_rows(d_in, 0, b, stream)
        plan.backproject_rows(0, b, d_vol, stream)
        plan.backproject_rows(0, b, d_vol, stream)
        stream.synchronize()
        plan.destroy()
    n = 64 << 20
    a, b = C.DeviceArray((n,), numpy.uint32), C.DeviceArray((n,),
                                                            numpy.uint32)
    for _ in range(2):
        C.call("gx_copy_u32", b.ptr, a.ptr, n, stream.handle)
    stream.synchronize()
    log("profile-kernels done (copy kernel: %d bytes read, %d written per "
        "launch)" % (4 * n, 4 * n))


def main_serving(args):
    """--config c5: the serving benchmark as the headline line."""
    from guanaco_amd import _capi as C
    if C.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    stream = C.Stream()
    n = args.serve_requests or 256
    sv = serving_bench(C, n, stream, max_inflight=args.serve_inflight)
    print(json.dumps({
        "metric": METRIC,
        "value": sv["updates_per_s"],
        "unit": UNIT,
        "n_gpus": 1,
        "steps": n,
        "warmup": 0,
        "ms_per_step": sv["mean_ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: a seeded 40-ellipsoid phantom per request, CTF "
                "forward model, Poisson 40 e-/px",
        "config": {"workload": SERVE_DESC, "path": "served",
                   "storage": "f16", "accumulate": "f32",
                   "requests": n, "max_inflight": sv["max_inflight"]},
        "latency_ms": {"p50": sv["p50_ms"], "p99": sv["p99_ms"]},
        "serving": sv,
    }), flush=True)


if __name__ == "__main__":
    main()
