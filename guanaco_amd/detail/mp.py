"""Slice dispatch for reconstruct() (reference src/guanaco/detail/mp.py).

For ``device="gpu"`` the whole volume is one device-resident call
(``guanaco_ext.reconstruct_volume``), split into bands of detector rows only
when it does not fit; every slice sees the same filtered data and geometry as
the reference's per-slice ``recon`` calls, including a rotation centre per
slice. The per-slice workers (as the reference dispatches them) serve the
CPU device, which belongs to the reference and raises here.
"""
import concurrent.futures as cf
import multiprocessing as mp

import numpy

import guanaco_amd.detail as ext

__all__ = ["get_ncore_slices", "reconstruction_dispatcher",
           "reconstruction_worker"]


def get_ncore_slices(axis_size, ncore=None, nchunk=None):
    """Split axis_size slices over ncore workers (mp.py:11-35)."""
    if not ncore:
        ncore = mp.cpu_count()
    if nchunk is None:
        chunk = axis_size // ncore
        sizes = numpy.full(ncore, chunk, dtype=int)
        sizes[: axis_size % ncore] += 1
        offsets = numpy.concatenate([[0], numpy.cumsum(sizes)])
        slices = [numpy.s_[offsets[i]:offsets[i + 1]] for i in range(ncore)]
    elif nchunk == 0:
        slices = [numpy.s_[i] for i in range(axis_size)]
    else:
        slices = [numpy.s_[o:o + nchunk] for o in range(0, axis_size, nchunk)]
    return ncore, slices


def _uniform(centre):
    c = numpy.asarray(centre, dtype=numpy.float32)
    return c.size == 0 or bool(numpy.all(c == c.flat[0]))


def reconstruction_dispatcher(sinogram, reconstruction, centre, angles,
                              weights=None, pixel_size=1, min_defocus=0,
                              max_defocus=0, transform=None, device="cpu",
                              ncore=None, nchunk=None, gpu_list=None,
                              tomogram=None, sinogram_order=True,
                              storage="f32"):
    """Reconstruct every slice of `sinogram` (Y, [ndf,] T, X) into
    `reconstruction` (Y, H, W) (mp.py:38-103).

    With device "gpu" the volume is one batched call on the original array
    `tomogram` (avoiding the (T, ..., Y, X) -> (Y, ..., T, X) transpose
    copy), with one centre or one centre per row.
    """
    dev = ext.eDevice(device) if isinstance(device, (str, int)) else device
    if dev == ext.eDevice.gpu:
        gpu = -1 if not gpu_list else int(gpu_list[0])
        src = sinogram if tomogram is None else tomogram
        order = True if tomogram is None else sinogram_order
        c = numpy.asarray(centre, dtype=numpy.float32)
        _volume_call(src, order, reconstruction,
                     float(c.flat[0]) if _uniform(c) else c, angles, weights,
                     pixel_size, min_defocus, max_defocus, transform, storage,
                     gpu)
        return
    ncore, slices = get_ncore_slices(sinogram.shape[0], ncore, nchunk)
    args = (angles, weights, pixel_size, min_defocus, max_defocus, transform,
            device)
    if ncore == 1:
        for s in slices:
            reconstruction_worker(sinogram[s], reconstruction[s], centre[s],
                                  *args, None)
    else:
        with cf.ThreadPoolExecutor(ncore) as e:
            futures = [e.submit(reconstruction_worker, sinogram[s],
                                reconstruction[s], centre[s], *args, None)
                       for s in slices]
            for f in futures:
                f.result()


def _volume_call(src, sinogram_order, reconstruction, centre, angles, weights,
                 pixel_size, min_defocus, max_defocus, transform, storage,
                 gpu):
    src = numpy.ascontiguousarray(src, dtype=numpy.float32)
    if sinogram_order:
        layout = 2          # (Y, [ndf,] T, X)
    else:
        layout = 0 if src.ndim == 3 else 1  # (T, Y, X) / (T, ndf, Y, X)
    w = numpy.ones(len(angles), numpy.float32) if weights is None else \
        numpy.asarray(weights, dtype=numpy.float32)
    if transform == "minus":
        w = -w
    out = reconstruction
    direct = (isinstance(out, numpy.ndarray) and out.dtype == numpy.float32
              and out.flags.c_contiguous and out.flags.writeable)
    if not direct:
        out = numpy.empty(reconstruction.shape, numpy.float32)
    ext.reconstruct_volume(src, out, numpy.asarray(angles, numpy.float32),
                           layout, centre, pixel_size, min_defocus,
                           max_defocus, weights=w, storage=storage,
                           gpu_index=gpu)
    if not direct:
        reconstruction[...] = out


def reconstruction_worker(sinogram, reconstruction, centre, angles, weights,
                          pixel_size, min_defocus, max_defocus, transform,
                          device, gpu_index):
    """Per-slice reconstruction (mp.py:106-154). The slice is copied before
    the per-tilt weights are applied, so the caller's array is never
    modified."""
    if gpu_index is None:
        gpu_index = 0
    for i in range(sinogram.shape[0]):
        sino = numpy.array(sinogram[i], dtype=numpy.float32, order="C")
        if transform == "minus":
            sino = -sino
        if weights is not None:
            # float64 weights, product rounded to float32 (as mp.py:135-141)
            w = numpy.asarray(weights, dtype=numpy.float64)
            if sino.ndim == 3:
                sino *= w[None, :, None]
            else:
                sino *= w[:, None]
        ext.recon(sino, reconstruction[i], angles, centre[i], pixel_size,
                  min_defocus, max_defocus, device, gpu_index)
