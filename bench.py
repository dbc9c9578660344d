"""Benchmark of the reconstruction hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c2]

One step reconstructs one whole tilt series that is already resident in HBM:
CTF phase-flip correction (2D R2C, per-defocus flip, C2R), ramp filter and
backprojection of every defocus slab into the volume. Three paths run on
the same input:

- plain:   FP32 storage, exact reference arithmetic, kernels launched eagerly;
- batched: FP16 storage of the filtered slabs, FP32 accumulation, the whole
           request captured in one HIP graph and replayed;
- served:  the batched request submitted repeatedly to the native serving
           loop, two requests on the device at once (their kernels overlap).

Before timing, each plan's backprojection tile (rows per tile, staged or
direct kernel) is chosen by the autotuner (gx_volume_plan_autotune) on the
plan's own data; the candidates' timings are reported.

The value is the batched path's whole-job throughput in voxel*tilt updates/s
(updates = X * Y * Z * T; ndf does not multiply it). rel-L2 of both volumes
is measured against the FP64 CPU oracle on sampled y-slices. The roofline
entry is for the backprojection kernel of the batched path, timed with HIP
events on the stream it is launched on; algorithmic bytes per update are
2 * s + 4 / T (two interpolation taps at storage size s, the FP32 voxel
written once per voxel). The CPU baseline is the oracle (a restatement of the
reference's CPU path) on one host core over a bounded sample, extrapolated.
"""
import argparse
import json
import math
import multiprocessing as mp
import os
import sys
import time

import numpy

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("voxel·tilt updates/s at ≤1e-3 rel-L2 vs FP64, batched vs "
          "plain; p50/p99 latency")
UNIT = "voxel·tilt updates/s"
HBM_PEAK_GBS = 8000.0
# 256 CUs x 128 B/clk x 2.4 GHz
LDS_PEAK_GBS = 256 * 128 * 2.4

# name: (X, Y, T, tilt step deg, Z, ndf, pixel A, defocus A, phi rad,
#        ellipsoids, semi-axes px, noise, seed, description)
WORKLOADS = {
    "c2": dict(X=1024, Y=1024, T=121, step=1.0, Z=256, ndf=1, ps=2.0,
               df=20000.0, phi=math.asin(0.1), ne=200, semi=(20.0, 160.0),
               dose=50.0, seed=1001,
               name="single-defocus CTF-corrected FBP: 121 x 1024^2 "
                    "projections (+-60 deg, 1 deg), 2 A/px, 300 kV, "
                    "Cs 2.7 mm, df 20000 A, amplitude contrast 0.1 -> "
                    "1024 x 256 x 1024 volume"),
    "c3": dict(X=2048, Y=2048, T=61, step=2.0, Z=512, ndf=16, ps=1.5,
               df=30000.0, phi=math.asin(0.1), ne=400, semi=(40.0, 320.0),
               dose=30.0, seed=1002,
               name="defocus-range CTF correction, ndf 16: 61 x 2048^2 "
                    "projections -> 2048 x 512 x 2048 volume"),
    "c4": dict(X=4096, Y=4096, T=121, step=1.0, Z=1024, ndf=32, ps=1.0,
               df=25000.0, phi=math.asin(0.1), ne=400, semi=(80.0, 640.0),
               dose=30.0, seed=1003, paths=("batched",), check=False,
               name="reduced-precision defocus-range correction, ndf 32: "
                    "121 x 4096^2 projections (+-60 deg, 1 deg), 1 A/px, "
                    "df 25000 A -> 4096 x 1024 x 4096 volume (FP32 slabs, "
                    "260 GB, do not fit beside the volume: batched path "
                    "only)"),
    "c1": dict(X=256, Y=256, T=61, step=2.0, Z=128, ndf=1, ps=1.0,
               df=None, phi=0.0, ne=20, semi=(5.0, 40.0), dose=None,
               seed=1000,
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
    l = _wavelength()
    kw = dict(num_defocus=w["ndf"], pixel_size=w["ps"], centre=w["X"] / 2.0,
              storage=storage)
    if w["df"] is not None:
        centre = numpy.full(w["Y"], w["X"] / 2.0, numpy.float32)
        mn, mx, st, n = _plan_defoci(w["df"], w["ndf"], w["X"], centre,
                                     w["ps"])
        dfs = numpy.array([mn + d * st for d in range(n)], numpy.float64)
        kw.update(min_defocus=mn - w["df"], max_defocus=mx - w["df"],
                  defocus=dfs, wavelength=l, Cs=2.7e7, phi=w["phi"])
    return kw


def make_plans(C, w, angles, storage):
    return C.VolumePlan(w["X"], w["Y"], angles, w["Z"],
                        **plan_kwargs(w, storage))


def served_path(C, w, angles, d_in, storage, n, max_inflight, log):
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
    print("bench: served (%s, %d in flight) %.2f ms/request" % (
        storage, max_inflight, out["ms_per_request"]), file=log, flush=True)
    return out


def _defocus_steps(df, ndf, width, centre, ps):
    # correct.py:44-93 (automatic step); the oracle legs' own copy
    if ndf == 1:
        return df, df, 0, 1
    z0 = df - centre * ps
    z1 = df + centre * ps
    z2 = df - (width - centre) * ps
    z3 = df + (width - centre) * ps
    mn = min([z0.min(), z1.min(), z2.min(), z3.min()])
    mx = max([z0.max(), z1.max(), z2.max(), z3.max()])
    return mn, mx, (mx - mn) / (ndf - 1), ndf


def time_path(C, plan, d_in, d_vol, stream, steps, warmup, graph):
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

def _oracle_defoci(w):
    """Absolute defoci of the slabs and the relative range
    (correct.py:44-93, 263-318)."""
    centre = numpy.full(w["Y"], w["X"] / 2.0, numpy.float32)
    mn, mx, st, n = _defocus_steps(w["df"], w["ndf"], w["X"], centre, w["ps"])
    return [mn + d * st for d in range(n)], mn - w["df"], mx - w["df"]


def _oracle_correct_rows(args):
    """FP32 and FP64 oracle correction of one tilt for every defocus; only
    the sampled rows kept: (ndf, nrows, X) each."""
    from oracle import oracle as O
    img, w, rows = args
    l = O.electron_wavelength(300000)
    dfs, _, _ = _oracle_defoci(w)
    ctf = numpy.stack([O.ctf_simple_grid(l, df, 2.7e7, 0, 0, w["phi"], w["X"],
                                         w["Y"], w["ps"]) for df in dfs])
    ctf = ctf.astype(numpy.complex64)
    c32 = O.corr(img, ctf, precision="f32")
    c64 = O.corr(img, ctf, precision="f64")
    return c32[:, rows], c64[:, rows]


def oracle_check(w, proj, angles, rows, vols, pool, dump=None):
    """rel-L2 of GPU volumes at the sampled rows against the FP32 and FP64
    oracle chains (correct_projections + reconstruct)."""
    from oracle import oracle as O
    T = proj.shape[0]
    mn = mx = 0.0
    if w["df"] is None:
        s32 = numpy.ascontiguousarray(proj[:, None, rows, :])
        s64 = s32.astype(numpy.float64)
    else:
        _, mn, mx = _oracle_defoci(w)
        res = pool.map(_oracle_correct_rows,
                       [(proj[t], w, rows) for t in range(T)])
        s32 = numpy.stack([r[0] for r in res])  # (T, ndf, nrows, X)
        s64 = numpy.stack([r[1] for r in res])
    out = {}

    def sino(s, i):  # (ndf, T, X) of sampled row i, or (T, X) for one slab
        a = numpy.ascontiguousarray(numpy.swapaxes(s[:, :, i, :], 0, 1))
        return a[0] if a.shape[0] == 1 else a

    ref32 = numpy.stack([O.recon(sino(s32, i), (w["Z"], w["X"]), angles,
                                 w["X"] / 2.0, w["ps"], mn, mx)
                         for i in range(len(rows))])
    ref64 = numpy.stack([O.recon(sino(s64, i), (w["Z"], w["X"]), angles,
                                 w["X"] / 2.0, w["ps"], mn, mx,
                                 precision="f64")
                         for i in range(len(rows))])
    if dump:
        os.makedirs(dump, exist_ok=True)
        numpy.save(os.path.join(dump, "s32.npy"), s32)
        numpy.save(os.path.join(dump, "ref32.npy"), ref32)
        for name, v in vols.items():
            numpy.save(os.path.join(dump, "vol_%s.npy" % name), v)
        numpy.save(os.path.join(dump, "proj_t0.npy"), proj[0])
    for name, v in vols.items():
        v = v.astype(numpy.float64)
        out[name + "_vs_fp64"] = float(numpy.linalg.norm(v - ref64) /
                                       numpy.linalg.norm(ref64))
        if name == "f32":
            out["f32_vs_fp32_oracle"] = float(
                numpy.linalg.norm(v - ref32) / numpy.linalg.norm(ref32))
    return out


def _oracle_slice(a):
    """One y-slice of filter + backprojection through the oracle (a pool
    worker's share of the all-cores baseline); returns seconds."""
    from oracle import oracle as O
    sino, Z, X, angles, ps = a
    t0 = time.perf_counter()
    O.recon(sino, (Z, X), angles, X / 2.0, ps)
    return time.perf_counter() - t0


_WORKER_CTF = {}


def _oracle_correction(a):
    """One tilt's correction through the oracle; returns seconds. The CTF
    grid is built once per worker process from its parameters."""
    from oracle import oracle as O
    img, key = a
    if key not in _WORKER_CTF:
        df, phi, X, Y, ps = key
        _WORKER_CTF[key] = O.ctf_simple_grid(
            O.electron_wavelength(300000), df, 2.7e7, 0, 0, phi, X, Y,
            ps).astype(numpy.complex64)
    t0 = time.perf_counter()
    O.corr(img, _WORKER_CTF[key], precision="f32")
    return time.perf_counter() - t0


def cpu_baseline_all_cores(w, proj, angles, pool, n):
    """The oracle slice-parallel on the pool's workers (the upper bound of a
    cpu_count-process run): two rounds of y-slices and one of tilt
    corrections, timed by wall clock and extrapolated to the whole job."""
    T, Y, X, Z = w["T"], w["Y"], w["X"], w["Z"]
    t_corr = 0.0
    if w["df"] is not None:
        key = (w["df"], w["phi"], X, Y, w["ps"])
        k = min(n, T)
        tasks = [(proj[t], key) for t in range(k)]
        # an untimed round first, so that the workers' CTF grids are built
        pool.map(_oracle_correction, tasks, chunksize=1)
        t0 = time.perf_counter()
        pool.map(_oracle_correction, tasks, chunksize=1)
        t_corr = (time.perf_counter() - t0) / k
    ys = numpy.linspace(0, Y - 1, min(2 * n, Y)).astype(int)
    t0 = time.perf_counter()
    pool.map(_oracle_slice, [(numpy.ascontiguousarray(proj[:, y, :]), Z, X,
                              angles, w["ps"]) for y in ys], chunksize=1)
    t_slice = (time.perf_counter() - t0) / len(ys)
    total = t_corr * T + t_slice * Y
    return {
        "value": X * Y * Z * T / total,
        "unit": UNIT,
        "cores": n,
        "sample": "%d tilt corrections and %d y-slices on %d worker "
                  "processes, wall clock, extrapolated linearly" % (
                      min(n, T) if w["df"] is not None else 0, len(ys), n),
        "seconds_estimated_whole_job": total,
    }


def cpu_baseline(w, proj, angles, budget_s):
    """Oracle (one core) on a bounded sample: some tilt corrections and some
    y-slices of filter + backprojection; extrapolated to the whole job."""
    from oracle import oracle as O
    T, Y, X, Z = w["T"], w["Y"], w["X"], w["Z"]
    l = O.electron_wavelength(300000)
    t_corr = 0.0
    n_corr = 0
    if w["df"] is not None:
        ctf = O.ctf_simple_grid(l, w["df"], 2.7e7, 0, 0, w["phi"], X, Y,
                                w["ps"]).astype(numpy.complex64)
        t0 = time.perf_counter()
        while n_corr < T and (n_corr < 2 or
                              time.perf_counter() - t0 < 0.3 * budget_s):
            O.corr(proj[n_corr], ctf, precision="f32")
            n_corr += 1
        t_corr = (time.perf_counter() - t0) / n_corr
    n_sl = 0
    t0 = time.perf_counter()
    ys = numpy.linspace(0, Y - 1, 64).astype(int)
    while n_sl < len(ys) and (n_sl < 2 or
                              time.perf_counter() - t0 < 0.7 * budget_s):
        sino = numpy.ascontiguousarray(proj[:, ys[n_sl], :])
        O.recon(sino, (Z, X), angles, X / 2.0, w["ps"])
        n_sl += 1
    t_slice = (time.perf_counter() - t0) / n_sl
    total = t_corr * T + t_slice * Y
    return {
        "value": X * Y * Z * T / total,
        "unit": UNIT,
        "cores": 1,
        "kind": "port",
        "sample": "%d of %d tilt corrections (%.3f s each) and %d of %d "
                  "y-slices of filter+backprojection (%.3f s each), "
                  "extrapolated linearly; project-written FFT in place of "
                  "FFTW" % (n_corr, T, t_corr, n_sl, Y, t_slice),
        "seconds_estimated_whole_job": total,
    }


# ------------------------------------------------------------- serving ----

SERVE_SIZES = (512, 1024, 2048)
SERVE_DESC = ("Poisson arrivals of whole-volume requests: detector S^2 with S "
              "uniform in {512, 1024, 2048}, T uniform in 41..121 consecutive "
              "1-degree tilts centred on 0 (slices of one seeded 121-tilt "
              "+-60 degree series per size, device-resident), ndf uniform in "
              "1..16, df uniform in 10000..40000 A, Cs 2.7 mm, 300 kV, "
              "1.5 A/px, volume S x S/4 x S, FP16 slabs; arrival rate 0.7 x "
              "the capacity measured closed-loop first; 4 requests in "
              "flight, footprint admission with bounded backfill")


def serving_bench(C, n_requests, stream, seed=1005, max_inflight=4,
                  backfill=True, log=sys.stderr):
    """Continuous batching through the native server (gx_server_*):
    capacity measured closed-loop on 24 requests, then n_requests Poisson
    arrivals at 0.7 x capacity; latency is submission to completion."""
    import threading
    from guanaco_amd.serve import Server
    l = _wavelength()
    ps = 1.5
    angles121 = numpy.deg2rad(numpy.arange(121) - 60.0).astype(numpy.float32)
    series = {}
    for k, S in enumerate(SERVE_SIZES):
        d = C.DeviceArray((121, S, S))
        C.synth_projections(d, S, S, angles121, depth=S // 4, width=S,
                            num_ellipsoids=100, seed=1100 + k,
                            semi=(S / 64.0, S / 8.0), noise="poisson",
                            noise_level=40.0,
                            ctf=dict(wavelength=l, Cs=2.7e7, defocus=25000.0,
                                     phi=math.asin(0.1), pixel_size=ps),
                            stream=stream)
        series[S] = d
    stream.synchronize()

    def make(rng):
        S = int(SERVE_SIZES[rng.integers(len(SERVE_SIZES))])
        T = int(rng.integers(41, 122))
        ndf = int(rng.integers(1, 17))
        df = float(rng.uniform(10000.0, 40000.0))
        start = 60 - (T - 1) // 2
        view = C.DeviceView(series[S].ptr + start * S * S * 4, (T, S, S))
        centre = numpy.full(S, S / 2.0, numpy.float32)
        mn, mx, st, n = _plan_defoci(df, ndf, S, centre, ps)
        dfs = numpy.array([mn + d * st for d in range(n)], numpy.float64)
        kw = dict(num_defocus=ndf, pixel_size=ps, min_defocus=mn - df,
                  max_defocus=mx - df, storage="f16", defocus=dfs,
                  wavelength=l, Cs=2.7e7, phi=math.asin(0.1))
        return view, angles121[start:start + T], S // 4, kw

    srv = Server(max_inflight=max_inflight, backfill=backfill)
    try:
        # capacity: 24 requests submitted at once
        rng = numpy.random.default_rng(seed + 1)
        reqs = []
        for _ in range(24):
            v, a, Z, kw = make(rng)
            reqs.append(srv.submit(v, a, Z, **kw))
        infos = []
        for r in reqs:
            infos.append(r.wait())
            r.release()
        span = max(i["end_s"] for i in infos) - min(i["submit_s"]
                                                    for i in infos)
        capacity = len(reqs) / span
        print("bench: serving capacity %.1f requests/s (%d closed-loop, "
              "budget %.1f GB)" % (capacity, len(reqs),
                                   srv.stats()["budget_bytes"] / 1e9),
              file=log, flush=True)
        rate = 0.7 * capacity
        rng = numpy.random.default_rng(seed)
        work = [make(rng) for _ in range(n_requests)]
        arrivals = numpy.cumsum(rng.exponential(1.0 / rate, n_requests))
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
        for i, (v, a, Z, kw) in enumerate(work):
            dt = arrivals[i] - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            r = srv.submit(v, a, Z, **kw)
            with lock:
                handles.append(r)
            submitted.release()
        th.join()
        st = srv.stats()
    finally:
        srv.close()
        for d in series.values():
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
        "backfill": backfill,
    }


def load_traffic(name):
    """HBM bytes per BP launch from the committed rocprofv3 PMC summary."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


# ------------------------------------------------------------------ main --

def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c2",
                    choices=sorted(WORKLOADS) + ["c5"],
                    help="c5: the serving benchmark only")
    ap.add_argument("--serve-requests", type=int, default=None,
                    help="requests of the serving phase (default 64 after "
                         "the c1-c3 workloads, 256 for c5; 0: skip)")
    ap.add_argument("--storage", default="f16", choices=["f16", "bf16"])
    ap.add_argument("--served", type=int, default=8,
                    help="requests of the served path (the workload n times "
                         "through the server; 0: skip)")
    ap.add_argument("--served-inflight", type=int, default=2)
    ap.add_argument("--serve-inflight", type=int, default=4,
                    help="requests in flight in the serving benchmark")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-host", action="store_true",
                    help="skip the host-buffer (PCIe-inclusive) measurement")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--dump", default=None,
                    help="directory to save the arrays the oracle check "
                         "compares")
    ap.add_argument("--bp", choices=["staged", "direct"], default=None,
                    help="backprojection kernel (default: the plan's "
                         "choice, staged where a tile spans <= 4 slabs)")
    ap.add_argument("--only", choices=["plain", "batched"], default=None,
                    help="time one path only")
    ap.add_argument("--profile-kernels", action="store_true",
                    help="no timing: run the filter pass once, the "
                         "backprojection twice and a 256 MiB copy kernel "
                         "twice (for rocprofv3 traces and counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.config == "c5":
        return main_serving(args)
    w = WORKLOADS[args.config]
    # the oracle chain at C4 (32 corrections of 121 4096^2 images in FP32
    # and FP64) is too large for a bench run: its parity is the smaller
    # configurations' and the tests'
    check = not args.no_check and args.only is None and w.get("check", True)
    # CPU workers for the oracle check are forked before anything touches
    # the GPU, so they inherit no device state
    # (and the all-cores CPU baseline's)
    pool = None
    workers = min(16, os.cpu_count() or 1)
    if check or not args.no_cpu:
        pool = mp.get_context("fork").Pool(workers)

    from guanaco_amd import _capi as C
    if C.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    X, Y, T, Z, ndf = w["X"], w["Y"], w["T"], w["Z"], w["ndf"]
    angles = numpy.deg2rad(
        numpy.arange(T) * w["step"] - 0.5 * (T - 1) * w["step"]).astype(
        numpy.float32)
    updates = float(X) * Y * Z * T
    stream = C.Stream()
    print("bench: %s on %s" % (args.config, C.device_info(0)["name"]),
          file=sys.stderr, flush=True)

    # synthetic input, generated and kept in HBM
    d_in = C.DeviceArray((T, Y, X))
    ctf = None
    if w["df"] is not None:
        ctf = dict(wavelength=_wavelength(), Cs=2.7e7,
                   defocus=w["df"], phi=w["phi"], pixel_size=w["ps"])
    C.synth_projections(d_in, X, Y, angles, depth=Z, width=X,
                        num_ellipsoids=w["ne"], seed=w["seed"], semi=w["semi"],
                        noise="poisson" if w["dose"] else "gaussian",
                        noise_level=w["dose"] if w["dose"] else 0.01,
                        ctf=ctf, stream=stream)
    paths = w.get("paths", ("plain", "batched"))
    if args.only:
        paths = tuple(p for p in paths if p == args.only)
    d_vol = C.DeviceArray((Y, Z, X)) if "plain" in paths else None
    d_vol2 = C.DeviceArray((Y, Z, X)) if "batched" in paths else None
    if args.profile_kernels:
        for storage in (("f32", args.storage) if args.only is None else
                        (("f32",) if args.only == "plain" else
                         (args.storage,))):
            plan = make_plans(C, w, angles, storage)
            if args.bp:
                plan.set_backprojector(args.bp)
            vol = d_vol if d_vol is not None else d_vol2
            plan.filter(d_in, stream)
            plan.backproject(vol, stream)
            plan.backproject(vol, stream)
            stream.synchronize()
            plan.destroy()
        n = 64 << 20
        a, b = C.DeviceArray((n,), numpy.uint32), C.DeviceArray((n,),
                                                                numpy.uint32)
        for _ in range(2):
            C.call("gx_copy_u32", b.ptr, a.ptr, n, stream.handle)
        stream.synchronize()
        print("bench: profile-kernels done (copy kernel: %d bytes read, %d "
              "written per launch)" % (4 * n, 4 * n), file=sys.stderr)
        return

    results = {}
    plans = {}
    for path, storage, graph, vol in (("plain", "f32", False, d_vol),
                                      ("batched", args.storage, True,
                                       d_vol2)):
        if path not in paths:
            continue
        plan = make_plans(C, w, angles, storage)
        tuned = None
        if args.bp:
            plan.set_backprojector(args.bp)
        else:
            # the tile (rows per tile, staged or direct) by measurement
            tuned = plan.autotune(d_in, vol, stream)
        kind, slots = plan.backprojector()
        plans[path] = plan
        ms, lat, wall = time_path(C, plan, d_in, vol, stream, args.steps,
                                  args.warmup, graph)
        filt_ms = time_kernel(C, lambda: plan.filter(d_in, stream), stream, 3)
        bp_ms = time_kernel(C, lambda: plan.backproject(vol, stream), stream,
                            max(3, args.steps))
        other = {}
        if tuned is not None:
            other["autotune"] = tuned
        plan.run(d_in, vol, stream)
        stream.synchronize()
        results[path] = dict(storage=storage, ms_per_step=ms,
                             value=updates / (ms * 1e-3),
                             p50_ms=percentile(lat, 50),
                             p99_ms=percentile(lat, 99), wall_s=wall,
                             filter_ms=filt_ms, bp_ms=bp_ms,
                             bp_updates_per_s=updates / (bp_ms * 1e-3),
                             backprojector=("staged, %d slab(s) per tilt"
                                            % slots if kind == C.BP_STAGED
                                            else "direct"),
                             tile_rows=plan.tile()[0], **other)
        print("bench: %s (%s) %.2f ms/step, BP %.2f ms, filter %.2f ms"
              % (path, storage, ms, bp_ms, filt_ms), file=sys.stderr,
              flush=True)

    if args.served > 0 and args.only is None:
        results["served"] = served_path(C, w, angles, d_in, args.storage,
                                        args.served, args.served_inflight,
                                        sys.stderr)

    head = results.get("batched") or results["plain"]
    s_bytes = 2 if head["storage"] in ("f16", "bf16") else 4
    b_alg = 2 * s_bytes + 4.0 / T
    achieved = updates * b_alg / (head["bp_ms"] * 1e-3) / 1e9
    traffic = load_traffic("bp_%s_%s_pmc.txt" % (args.config, head["storage"]))
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
            "path": "batched" if "batched" in results else "plain",
            "storage": head["storage"],
            "accumulate": "f32",
            "volume": [Y, Z, X],
            "tilts": T,
            "num_defocus": ndf,
            "updates_per_step": updates,
        },
        "latency_ms": {"p50": head["p50_ms"], "p99": head["p99_ms"]},
        "paths": results,
        "roofline": {
            "bound": "hbm",
            "kernel": "k_bp (backprojection, %s slabs)" % head["storage"],
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "algorithmic_bytes_per_update": b_alg,
            "updates_per_launch": updates,
            "launch_ms": head["bp_ms"],
            # taps read from LDS by the staged kernel: 2 * s bytes per update
            "lds_read": {
                "achieved": updates * 2 * s_bytes / (head["bp_ms"] * 1e-3)
                / 1e9,
                "peak": LDS_PEAK_GBS, "unit": "GB/s",
            } if head["backprojector"].startswith("staged") else None,
        },
    }
    if "batched" in results and "plain" in results:
        line["speedup_batched_vs_plain"] = (results["batched"]["value"] /
                                            results["plain"]["value"])
    if "served" in results and "plain" in results:
        line["speedup_served_vs_plain"] = (results["served"]["value"] /
                                           results["plain"]["value"])
    if check:
        rows = numpy.linspace(0, Y - 1, 4).astype(int)
        proj = d_in.to_host()
        v32 = d_vol.to_host()[rows]
        v16 = d_vol2.to_host()[rows]
        t0 = time.perf_counter()
        line["rel_l2"] = oracle_check(
            w, proj, angles, rows, {"f32": v32, args.storage: v16}, pool,
            dump=args.dump)
        line["rel_l2"]["sample"] = "y-slices %s" % rows.tolist()
        print("bench: oracle check %.1f s %s" % (time.perf_counter() - t0,
                                                line["rel_l2"]),
              file=sys.stderr, flush=True)
    n_serve = 64 if args.serve_requests is None else args.serve_requests
    if n_serve > 0 and args.only is None:
        line["serving"] = serving_bench(C, n_serve, stream,
                                        max_inflight=args.serve_inflight)
        print("bench: serving %s" % {k: v for k, v in line["serving"].items()
                                     if k != "description"},
              file=sys.stderr, flush=True)
    if args.only is None and not args.no_host:
        proj = d_in.to_host()
        line["host_buffers"] = host_path(w, angles, proj, args.storage)
        print("bench: host buffers (%s) %.1f ms/request incl. PCIe" % (
            args.storage, line["host_buffers"]["ms_per_request"]),
            file=sys.stderr, flush=True)
    if not args.no_cpu:
        proj = d_in.to_host()
        line["cpu_baseline"] = cpu_baseline(w, proj, angles, args.cpu_budget)
        line["cpu_baseline"]["all_cores"] = cpu_baseline_all_cores(
            w, proj, angles, pool, workers)
        print("bench: CPU oracle %.3g updates/s on 1 core, %.3g on %d" % (
            line["cpu_baseline"]["value"],
            line["cpu_baseline"]["all_cores"]["value"],
            line["cpu_baseline"]["all_cores"]["cores"]), file=sys.stderr,
            flush=True)
    if pool is not None:
        pool.close()
    for p in plans.values():
        p.destroy()
    print(json.dumps(line), flush=True)


def main_serving(args):
    """--config c5: the serving benchmark as the headline line."""
    from guanaco_amd import _capi as C
    if C.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    stream = C.Stream()
    n = 256 if args.serve_requests is None else args.serve_requests
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
        "data": "synthetic: seeded 100-ellipsoid phantoms, CTF forward model, "
                "Poisson 40 e-/px",
        "config": {"workload": SERVE_DESC, "path": "served",
                   "storage": "f16", "accumulate": "f32",
                   "requests": n, "max_inflight": sv["max_inflight"]},
        "latency_ms": {"p50": sv["p50_ms"], "p99": sv["p99_ms"]},
        "serving": sv,
    }), flush=True)


if __name__ == "__main__":
    main()
