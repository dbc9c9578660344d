"""Oracle volumes at sampled detector rows, spread over CPU worker processes.

TEST INFRASTRUCTURE ONLY: used by tests/ and by the oracle check of
bench.py (never by the product or the bench's GPU legs).

Slice y of a reconstruction needs row y of every corrected projection
(reconstructor.cpp:50,72, one slice per call in mp.py:106-154), so the
oracle chain at a few rows costs the corrections' sampled rows plus a few
slices. The corrections' rows come from corr_rows (FP64: one 2D FFT per
tilt, the inverse column DFT at the sampled rows only) and from the oracle's
full FP32 correction (corr), the slices from recon (filter.cpp:61-97 and
reconstructor.cpp:14-115) in FP32 and FP64.

reference_rows(pool, ...) returns the sampled slices of the FP64 chain with
the reference's FP32 slab/tap decisions, of the FP64 chain with FP64
decisions, and (optionally) of the FP32 chain. The pool's workers must not
hold GPU state: fork them before the process first touches the GPU, or use
the spawn context.
"""
import os
import tempfile

import numpy

from oracle import oracle as O


def worker_init():
    """Pool initializer: one BLAS/OpenMP thread per worker (each worker is
    one core; the environment may ask libraries for many threads)."""
    try:
        from threadpoolctl import threadpool_limits
        threadpool_limits(1)
    except ImportError:
        pass


def make_pool(n, context="spawn"):
    """A pool of n single-threaded oracle workers."""
    import multiprocessing as mp
    return mp.get_context(context).Pool(n, initializer=worker_init)


def defocus_list(df, ndf, width, height, ps, centre=None):
    """Absolute defoci of the slabs and the relative range
    (get_defocus_steps, correct.py:44-93; centre: one rotation centre for
    every row, width / 2 by default)."""
    centre = numpy.full(height, width / 2.0 if centre is None else centre,
                        numpy.float32)
    mn, mx, st, n = O.get_defocus_steps(df, ndf, None, width, centre, ps)
    return [mn + d * st for d in range(n)], mn - df, mx - df


def mask_task(a):
    """Flip factors of one defocus from the oracle's complex64 CTF grid."""
    X, Y, ps, phi, Cs, df = a
    l = O.electron_wavelength(300000)
    h = O.ctf_simple_grid(l, df, Cs, 0.0, 0.0, phi, X, Y, ps)
    return O.flip_mask(h.astype(numpy.complex64))


_MASKS = {}


def _masks(path):
    if path not in _MASKS:
        _MASKS.clear()
        _MASKS[path] = numpy.load(path, mmap_mode="r")
    return _MASKS[path]


def corr_rows_task(a):
    """FP64 correction of one tilt at the sampled rows, every defocus."""
    img, mask_path, rows = a
    return O.corr_rows(img, _masks(mask_path), rows)


def corr_f32_task(a):
    """The oracle's FP32 correction of one tilt, sampled rows kept. The
    correction reads only the sign of Im(ctf) (correct.h:30-39), so the CTF
    is handed over as -i times the flip factor, which has that sign."""
    img, mask_path, rows = a
    ctf = (-1j * numpy.asarray(_masks(mask_path), numpy.float32)).astype(
        numpy.complex64)
    return O.corr(img, ctf, precision="f32")[:, rows]


def recon_task(a):
    """Filter and backprojection of one slice's sinogram."""
    sino, Z, X, angles, centre, ps, mn, mx, precision, f32_dec = a
    return O.recon(sino, (Z, X), angles, centre, ps, mn, mx,
                   precision=precision, f32_decisions=f32_dec)


def reference_rows(pool, proj, angles, rows, depth, ps=1.0, df=None, ndf=1,
                   phi=0.0, Cs=2.7e7, fp32=True, tilts=None, centre=None):
    """Oracle slices at `rows` of the reconstruction of (T, Y, X)
    projections, corrected with ndf defoci around df (None: no correction),
    rotation centre `centre` on every row (X / 2 by default).

    Returns {"f64": (nrows, depth, X), "f64_all_f64_decisions": ...,
    "f32": ... (when fp32)} and, with `tilts`, "corrected64" (the FP64
    corrected rows of those tilts, (len(tilts), ndf, nrows, X))."""
    proj = numpy.asarray(proj)
    T, Y, X = proj.shape
    rows = numpy.asarray(rows, dtype=numpy.int64)
    mn = mx = 0.0
    tmp = None
    out = {}
    try:
        if df is None:
            s64 = proj[:, None, rows, :].astype(numpy.float64)
            s32 = numpy.ascontiguousarray(proj[:, None, rows, :],
                                          dtype=numpy.float32)
        else:
            dfs, mn, mx = defocus_list(df, ndf, X, Y, ps, centre)
            masks = pool.map(mask_task, [(X, Y, ps, phi, Cs, d) for d in dfs])
            fd, tmp = tempfile.mkstemp(suffix=".npy")
            os.close(fd)
            numpy.save(tmp, numpy.stack(masks))
            del masks
            s64 = numpy.stack(pool.map(
                corr_rows_task, [(proj[t], tmp, rows) for t in range(T)],
                chunksize=1))  # (T, ndf, nrows, X)
            s32 = None
            if fp32:
                s32 = numpy.stack(pool.map(
                    corr_f32_task, [(proj[t], tmp, rows) for t in range(T)],
                    chunksize=1))
        if tilts is not None:
            out["corrected64"] = s64[list(tilts)]

        def sino(s, i):  # (ndf, T, X) of sampled row i, (T, X) for one slab
            a = numpy.ascontiguousarray(numpy.swapaxes(s[:, :, i, :], 0, 1))
            return a[0] if a.shape[0] == 1 else a

        centre = X / 2.0 if centre is None else float(centre)
        tasks, keys = [], []
        for i in range(rows.size):
            for key, s, prec, dec in (("f64", s64, "f64", True),
                                      ("f64_all_f64_decisions", s64, "f64",
                                       False),
                                      ("f32", s32, "f32", True)):
                if s is None or (key == "f32" and not fp32):
                    continue
                tasks.append((sino(s, i), depth, X, angles, centre, ps, mn,
                              mx, prec, dec))
                keys.append(key)
        res = pool.map(recon_task, tasks, chunksize=1)
        for key in ("f64", "f64_all_f64_decisions", "f32"):
            got = [r for r, k in zip(res, keys) if k == key]
            if got:
                out[key] = numpy.stack(got)
    finally:
        if tmp:
            os.unlink(tmp)
    out["range"] = (mn, mx)
    return out


def rel_l2(a, b):
    a = numpy.asarray(a, numpy.float64)
    b = numpy.asarray(b, numpy.float64)
    return float(numpy.linalg.norm(a - b) / max(numpy.linalg.norm(b), 1e-300))
