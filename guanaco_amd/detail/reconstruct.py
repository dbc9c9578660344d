"""reconstruct() and its helpers (reference src/guanaco/detail/reconstruct.py).

``reconstruct`` keeps the reference's signature and semantics; with
``device="gpu"`` the slices go through the gfx950 path, batched into one
device-resident call. ``reconstruct_projections`` is the fused path: raw
projections in, CTF correction, filtering and backprojection on the device,
without materialising the corrected stack on the host. ``reconstruct_file``
is the file driver behind the ``guanaco`` command.
"""
import time

import numpy

import guanaco_amd.detail as ext
from guanaco_amd import mrc
from guanaco_amd.detail import mp
from guanaco_amd.detail.correct import (check_detector_size,
                                        correct_projections,
                                        get_defocus_steps)

__all__ = ["reconstruct", "reconstruct_projections", "reconstruct_file",
           "reconstruct_tomopy", "get_centre", "compute_angular_weights",
           "get_weights"]


def get_centre(shape, centre, sinogram_order=True):
    """Per-slice rotation centres (reconstruct.py:32-52): X/2 by default, a
    scalar broadcast to every slice, or the given array."""
    if not sinogram_order:
        shape = list(shape)
        shape[0], shape[-2] = shape[-2], shape[0]
    if centre is None:
        centre = numpy.ones(shape[0], dtype="float32") * (shape[-1] * 0.5)
    elif numpy.array(centre).size == 1:
        centre = numpy.ones(shape[0], dtype="float32") * centre
    return numpy.asarray(centre).astype(dtype="float32", copy=False)


def compute_angular_weights(angles):
    """Angular weights (reconstruct.py:55-105): for each tilt in sorted order,
    the mean of its neighbouring intervals weighted 2 for the adjacent ones
    and 2/3 for the next ones, normalised by the mean interval."""
    angles = numpy.asarray(angles)
    order = numpy.argsort(angles)
    a = angles[order]
    n = a.size
    if n > 1:
        mean_interval = abs(a[-1] - a[0]) / (n - 1)
        assert mean_interval != 0
    else:
        mean_interval = 1.0
    weights = numpy.zeros(n)
    for i in range(n):
        terms = []
        if i - 2 >= 0:
            terms.append((1.0 / 1.5, a[i - 1] - a[i - 2]))
        if i - 1 >= 0:
            terms.append((2.0, a[i] - a[i - 1]))
        if i + 1 < n:
            terms.append((2.0, a[i + 1] - a[i]))
        if i + 2 < n:
            terms.append((1.0 / 1.5, a[i + 2] - a[i + 1]))
        sw = sum(w for w, _ in terms)
        si = sum(w * d for w, d in terms)
        weights[order[i]] = si / sw if sw != 0 else 1.0
    return weights / mean_interval


def get_weights(angles, angular_weights=False):
    """Per-tilt sinogram weights (reconstruct.py:108-124)."""
    if angular_weights:
        return compute_angular_weights(angles)
    return numpy.ones(numpy.asarray(angles).shape)


def reconstruct(tomogram, angles, reconstruction=None, centre=None,
                pixel_size=1, min_defocus=0, max_defocus=0,
                sinogram_order=False, transform=None, angular_weights=False,
                device="cpu", ncore=None, nchunk=None, gpu_list=None,
                storage="f32", depth=None):
    """Filtered backprojection of a tilt series (reconstruct.py:127-213).

    tomogram: (T, Y, X) projections or (T, ndf, Y, X) defocus-corrected
    projections, or the same in sinogram order (Y, T, X) / (Y, ndf, T, X)
    with sinogram_order=True. Returns the (Y, X, X) volume (or writes
    `reconstruction`). Extensions: `storage` ("f32", "f16", "bf16") selects
    the filtered-projection precision on the GPU ("f16" stays within 1e-3
    rel-L2 of the FP64 reconstruction; "bf16", with an 8-bit mantissa, does
    not: about 2e-3); `depth` reconstructs only the central `depth` rows of
    each square slice, shape (Y, depth, X).
    """
    assert len(tomogram.shape) in [3, 4], "Dimensions must be 3 or 4"
    sinogram = tomogram.astype(dtype="float32", copy=False)
    if not sinogram_order:
        sinogram = numpy.swapaxes(sinogram, 0, -2)  # a view, as the reference
    X = sinogram.shape[-1]
    if reconstruction is None:
        shape = (sinogram.shape[0], X if depth is None else depth, X)
        reconstruction = numpy.zeros(shape, dtype="float32")
    else:
        assert reconstruction.shape[0] == sinogram.shape[0], \
            "Inconsistent reconstruction/sinogram shape"
        assert reconstruction.shape[2] == X, \
            "Inconsistent reconstruction/sinogram shape"
        if depth is None:
            assert reconstruction.shape[1] == X, \
                "Inconsistent reconstruction/sinogram shape"
    centre = get_centre(sinogram.shape, centre)
    weights = get_weights(angles, angular_weights)
    mp.reconstruction_dispatcher(
        sinogram, reconstruction, centre, angles, weights=weights,
        pixel_size=pixel_size, min_defocus=min_defocus,
        max_defocus=max_defocus, transform=transform, device=device,
        ncore=ncore, nchunk=nchunk, gpu_list=gpu_list,
        tomogram=tomogram, sinogram_order=sinogram_order, storage=storage)
    return reconstruction


def reconstruct_projections(projections, angles, centre=None, pixel_size=1,
                            energy=300, defocus=None, num_defocus=None,
                            step_defocus=None, spherical_aberration=0,
                            astigmatism=0, astigmatism_angle=0,
                            phase_shift=0, transform=None,
                            angular_weights=False, depth=None,
                            storage="f32", reconstruction=None,
                            gpu_index=-1):
    """CTF-corrected FBP of raw projections (T, Y, X) in one device call.

    Equivalent to correct_projections(...) followed by reconstruct(...) with
    the returned relative defocus range (reconstruct.py:341-388), with the
    same defocus steps (correct.py:44-93) and CTF (correct.py:116-214); the
    corrected stack stays on the device. defocus=None reconstructs without
    correction. Units as the reference: energy keV, defoci A, Cs mm, Ca A,
    Pa and phase_shift degrees.
    """
    projections = numpy.ascontiguousarray(projections, dtype=numpy.float32)
    assert projections.ndim == 3, "projections must be (T, Y, X)"
    T, Y, X = projections.shape
    check_detector_size(X, Y if defocus is not None else None)
    c = get_centre(projections.shape, centre, False)
    weights = get_weights(angles, angular_weights).astype(numpy.float32)
    if transform == "minus":
        weights = -weights
    shape = (Y, X if depth is None else depth, X)
    out = reconstruction
    if out is None:
        out = numpy.zeros(shape, numpy.float32)
    kw = {}
    mn_rel = mx_rel = 0.0
    if defocus is not None:
        mn, mx, step, ndf = get_defocus_steps(defocus, num_defocus,
                                              step_defocus, X, c, pixel_size)
        assert abs(mn + (ndf - 1) * step - mx) < 1e-7, \
            "Inconsistent defocus steps"
        dfs = numpy.array([mn + d * step for d in range(ndf)],
                          dtype=numpy.float64)
        kw = dict(defocus=dfs,
                  wavelength=ext.get_electron_wavelength(energy * 1000),
                  Cs=spherical_aberration * 1e7, Ca=astigmatism,
                  Pa=astigmatism_angle * numpy.pi / 180,
                  phi=phase_shift * numpy.pi / 180)
        mn_rel, mx_rel = mn - defocus, mx - defocus
    ext.reconstruct_volume(projections, out,
                           numpy.asarray(angles, numpy.float32), 0,
                           float(c[0]) if numpy.all(c == c[0]) else c,
                           pixel_size, mn_rel, mx_rel,
                           weights=weights, storage=storage,
                           gpu_index=gpu_index, **kw)
    return out


def _require_gpu(device):
    dev = ext.eDevice(device) if isinstance(device, (str, int)) else device
    if dev != ext.eDevice.gpu:
        raise RuntimeError(
            "guanaco_amd implements the GPU device (device='gpu'); the CPU "
            "device is the reference's own FFTW implementation")


def read_projection_metadata(infile):
    """Tilt angles (rad, float32), voxel size and the relative defocus range
    of a projection file (reconstruct.py:245-292): angles from the FEI
    extended header's "Alpha tilt" (degrees); for a corrected (T, ndf, Y, X)
    stack, the defocus range from "Defocus" of the first tilt's blocks."""
    voxel_size = infile.voxel_size
    print("Voxel size: ", voxel_size)
    header = infile.indexed_extended_header
    if header is None:
        raise ValueError(
            "the input has no FEI extended header with tilt angles; give "
            "start_angle, step_angle and pixel_size")
    shape = infile.data.shape
    if len(shape) == 3:
        assert shape[0] == header.shape[0], "Inconsistent data/header shape"
        tilt = header["Alpha tilt"]
        min_defocus = max_defocus = 0
    else:
        assert header.size == shape[0] * shape[1], \
            "Inconsistent data/header shape"
        header = header.reshape(shape[0:2])
        tilt = header[:, 0]["Alpha tilt"]
        defocus = numpy.asarray(header[0, :]["Defocus"], dtype=numpy.float32)
    angles = numpy.deg2rad(numpy.asarray(tilt, numpy.float64)).astype(
        numpy.float32)
    for i, a in enumerate(angles):
        print("Image %d; angle %.4f deg" % (i + 1, numpy.rad2deg(a)))
    if len(shape) == 4:
        for i, d in enumerate(defocus):
            print("Correction %d; defocus %.4f A" % (i + 1, d))
        min_defocus, max_defocus = numpy.min(defocus), numpy.max(defocus)
    return angles, voxel_size, min_defocus, max_defocus


def reconstruct_file(input_filename, output_filename,
                     corrected_filename="GUANACO_CORRECTED.dat",
                     start_angle=None, step_angle=None, pixel_size=None,
                     centre=None, energy=None, defocus=None, num_defocus=None,
                     step_defocus=None, spherical_aberration=None,
                     astigmatism=None, astigmatism_angle=None,
                     phase_shift=None, device="cpu", ncore=None,
                     transform=None, angular_weights=False, chunk_size=None,
                     method="FBP_CTF", num_iter=None, storage="f32",
                     depth=None, fused=None):
    """Reconstructs an MRC tilt series into an MRC volume
    (reconstruct.py:214-402).

    The input is a (T, Y, X) projection stack or a (T, ndf, Y, X) corrected
    stack written by correct_file. Angles and pixel size come from the FEI
    extended header and the voxel size unless start_angle, step_angle and
    pixel_size are all given (degrees, Angstrom). With method "FBP_CTF" an
    uncorrected stack is CTF-corrected first. The output is a float32
    (Y, X, X) volume (Y, depth, X with `depth`) with the input's voxel size.

    On the GPU an uncorrected stack takes the fused device path by default
    (fused=None): correction, filtering and backprojection in one device
    run, so the corrected intermediate `corrected_filename` is not written.
    fused=False keeps the reference's two steps (the corrected stack in
    `corrected_filename`, then reconstruct). Methods other than FBP_CTF
    belong to tomopy in the reference and are not part of this package.
    """
    start_time = time.time()
    _require_gpu(device)
    if fused is None:
        fused = True
    print("Reading %s" % input_filename)
    with mrc.mmap(input_filename) as infile:
        projections = infile.data
        assert projections.ndim in (3, 4), "Dimensions must be 3 or 4"
        if start_angle is None or step_angle is None or pixel_size is None:
            angles, voxel_size, min_defocus, max_defocus = \
                read_projection_metadata(infile)
            assert voxel_size["x"] == voxel_size["y"], \
                "Pixel must be a square"
            pixel_size = float(voxel_size["x"])
        else:
            angles = numpy.deg2rad(
                start_angle + numpy.arange(projections.shape[0]) * step_angle)
            voxel_size = pixel_size
            min_defocus = max_defocus = 0
        print("Pixel size %d A" % pixel_size)
        X = projections.shape[-1]
        output_shape = (projections.shape[-2], X if depth is None else depth,
                        X)
        centre = get_centre(projections.shape, centre, False)
        corrected = projections.ndim == 4
        if corrected:
            method = "FBP_CTF"
        if method != "FBP_CTF":
            # the third-party algorithms: checked before any output is made
            _tomopy()
        run_fused = fused and not corrected and defocus is not None
        # sizes are checked before any output file is made
        check_detector_size(X, projections.shape[-2]
                            if (not corrected and defocus is not None)
                            else None)
        if not corrected and defocus is not None and not run_fused:
            projections, min_defocus, max_defocus = correct_projections(
                projections, centre, pixel_size, energy, defocus,
                num_defocus, step_defocus, spherical_aberration or 0,
                astigmatism or 0, astigmatism_angle or 0, phase_shift or 0,
                corrected_filename, device)
        print("Writing reconstruction to %s" % output_filename)
        with mrc.new_mmap(output_filename, output_shape, mrc_mode=2,
                          overwrite=True) as outfile:
            outfile.voxel_size = voxel_size
            if method != "FBP_CTF":
                reconstruct_tomopy(projections, angles, outfile.data,
                                   centre=centre, transform=transform,
                                   device=device, ncore=ncore, method=method,
                                   num_iter=num_iter)
            elif run_fused:
                reconstruct_projections(
                    projections, angles, centre=centre,
                    pixel_size=pixel_size, energy=energy, defocus=defocus,
                    num_defocus=num_defocus, step_defocus=step_defocus,
                    spherical_aberration=spherical_aberration or 0,
                    astigmatism=astigmatism or 0,
                    astigmatism_angle=astigmatism_angle or 0,
                    phase_shift=phase_shift or 0, transform=transform,
                    angular_weights=angular_weights, depth=depth,
                    storage=storage, reconstruction=outfile.data)
            else:
                reconstruct(projections, angles, outfile.data, centre=centre,
                            pixel_size=pixel_size, min_defocus=min_defocus,
                            max_defocus=max_defocus, sinogram_order=False,
                            transform=transform,
                            angular_weights=angular_weights, device=device,
                            ncore=ncore, storage=storage, depth=depth)
    print("Time: %.2f seconds" % (time.time() - start_time))


def _tomopy():
    try:
        import tomopy
    except ImportError:
        raise RuntimeError(
            "The reconstruction algorithm selected requires tomopy and astra")
    return tomopy


def reconstruct_tomopy(projections, angles, reconstruction, centre=None,
                       transform=None, device="cpu", ncore=None, method="FBP",
                       num_iter=200, output_rot90=False):
    """The reference's hand-over of methods other than FBP_CTF (SIRT, SART,
    CGLS, EM, FBP) to tomopy with the astra toolbox
    (reconstruct.py:405-451). tomopy is a third-party package, not part of
    this one: without it the call raises RuntimeError, as the reference's
    does. With it, the astra options follow the reference's: the "linear"
    projector on the CPU, the "<method>_CUDA" algorithms otherwise."""
    tomopy = _tomopy()
    print("Using %s" % device)
    if device == "cpu":
        options = {"proj_type": "linear", "method": method}
    else:
        options = {"proj_type": "cuda", "method": "%s_CUDA" % method}
        ncore = 1 if ncore is None else ncore
    if method in ("SIRT", "SART", "CGLS"):
        options["num_iter"] = num_iter
    if transform == "minus":
        projections = -projections
    reconstruction[:] = tomopy.recon(projections, angles,
                                     algorithm=tomopy.astra, options=options,
                                     ncore=ncore)
