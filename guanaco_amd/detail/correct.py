"""CTF correction orchestration (reference src/guanaco/detail/correct.py).

``correct_projections`` keeps the reference's signature, units and return
value ``(projections, min_defocus - defocus, max_defocus - defocus)``; with
``device="gpu"`` every tilt is corrected by the gfx950 ``corr`` kernels.
``correct_file`` is the file driver behind the ``guanaco.correct`` command.
"""
import tempfile
import time
from math import pi

import numpy

import guanaco_amd.detail as ext
from guanaco_amd import mrc

__all__ = ["correct_projections", "correct_file", "get_defocus_steps",
           "get_corrected_array", "get_ctf", "get_ctf_array"]


# Detector sizes of the gfx950 kernels: the ramp filter takes rows of up to
# 8192 pixels (2N-point filters up to 16384 points), the 2D CTF correction
# up to 8192 rows and columns (power-of-two transforms, Bluestein transforms
# of up to 16384 points for other sizes).
MAX_DETECTOR = 8192


def check_detector_size(width, height=None):
    """Raises RuntimeError for detectors the gfx950 path does not take:
    width (the filtered rows, and the corrected rows) and, with the 2D
    correction, height above MAX_DETECTOR."""
    too = [(n, v) for n, v in (("width", width), ("height", height))
           if v is not None and v > MAX_DETECTOR]
    if too:
        raise RuntimeError(
            "detector %s exceeds the %d pixels the gfx950 kernels take" % (
                " and ".join("%s %d" % t for t in too), MAX_DETECTOR))


def get_centre(shape, centre, sinogram_order=True):
    """Per-row rotation centres (correct.py:33-41)."""
    if not sinogram_order:
        shape = list(shape)
        shape[0], shape[-2] = shape[-2], shape[0]
    if centre is None:
        centre = numpy.ones(shape[0], dtype="float32") * (shape[-1] / 2.0)
    elif numpy.array(centre).size == 1:
        centre = numpy.ones(shape[0], dtype="float32") * centre
    return numpy.asarray(centre).astype(dtype="float32", copy=False)


def get_defocus_steps(defocus, num_defocus=None, step_defocus=None,
                      width=None, centre=None, pixel_size=None):
    """Defocus range (min, max, step, num) (correct.py:44-93).

    One defocus: min = max = defocus. With a step: defocus -/+ (n-1)/2 step.
    Otherwise the range covers the beam heights of the detector edges about
    the rotation centre, and step = (max - min) / (n - 1).
    """
    if num_defocus is None or num_defocus <= 0:
        num_defocus = 1
    if num_defocus == 1:
        return defocus, defocus, 0, 1
    if step_defocus is not None and step_defocus > 0:
        half = 0.5 * (num_defocus - 1) * step_defocus
        return defocus - half, defocus + half, step_defocus, num_defocus
    centre = numpy.asarray(centre)
    ends = [defocus - centre * pixel_size, defocus + centre * pixel_size,
            defocus - (width - centre) * pixel_size,
            defocus + (width - centre) * pixel_size]
    mn = min(e.min() for e in ends)
    mx = max(e.max() for e in ends)
    return mn, mx, (mx - mn) / (num_defocus - 1), num_defocus


def get_corrected_array(shape, filename_or_array=None):
    """Storage for the corrected stack (correct.py:96-113): a temporary
    memmap, a named memmap, or the given array reshaped."""
    if filename_or_array is None:
        return numpy.memmap(tempfile.TemporaryFile(), mode="w+",
                            dtype="float32", shape=shape)
    if isinstance(filename_or_array, str):
        return numpy.memmap(filename_or_array, mode="w+", dtype="float32",
                            shape=shape)
    return filename_or_array.reshape(shape)


def get_ctf(width, height, pixel_size, energy, defocus,
            spherical_aberration=0, astigmatism=0, astigmatism_angle=0,
            phase_shift=0):
    """exp(-i(chi - phi)) on the FFT-ordered grid (correct.py:116-154);
    energy keV, defocus A, Cs mm, Ca A, Pa and phase_shift degrees."""
    c = ext.CTF(l=ext.get_electron_wavelength(energy * 1000), df=defocus,
                Cs=spherical_aberration * 1e7, Ca=astigmatism,
                Pa=astigmatism_angle * pi / 180, phi=phase_shift * pi / 180)
    return c.get_ctf_simple(width, height, pixel_size)


def get_ctf_array(width, height, pixel_size, energy, min_defocus,
                  num_defocus, step_defocus, spherical_aberration=0,
                  astigmatism=0, astigmatism_angle=0, phase_shift=0):
    """complex64 (ndf, height, width) stack for defoci min + d * step
    (correct.py:157-214)."""
    out = numpy.zeros((num_defocus, height, width), dtype="complex64")
    for d in range(num_defocus):
        out[d] = get_ctf(width, height, pixel_size, energy,
                         min_defocus + d * step_defocus, spherical_aberration,
                         astigmatism, astigmatism_angle, phase_shift)
    return out


def correct_projections(projections, centre=None, pixel_size=1, energy=None,
                        defocus=None, num_defocus=None, step_defocus=None,
                        spherical_aberration=0, astigmatism=0,
                        astigmatism_angle=0, phase_shift=0,
                        corrected_output=None, device="cpu"):
    """Phase-flip every projection for each defocus (correct.py:217-318).

    Returns (corrected, min_defocus - defocus, max_defocus - defocus); the
    corrected stack is (T, ndf, Y, X), or (T, Y, X) for one defocus. With
    defocus None the projections are returned untouched with (0, 0). A
    missing centre means X/2 for every row (the reference raises TypeError
    there when num_defocus > 1).
    """
    if defocus is None:
        return projections, 0, 0
    if centre is None:
        centre = get_centre(projections.shape, None, sinogram_order=False)
    mn, mx, step, ndf = get_defocus_steps(defocus, num_defocus, step_defocus,
                                          projections.shape[2], centre,
                                          pixel_size)
    assert abs(mn + (ndf - 1) * step - mx) < 1e-7, "Inconsistent defocus steps"
    T, Y, X = projections.shape
    corrected = get_corrected_array((T, ndf, Y, X), corrected_output)
    ctf = get_ctf_array(X, Y, pixel_size, energy, mn, ndf, step,
                        spherical_aberration, astigmatism, astigmatism_angle,
                        phase_shift)
    buf = numpy.empty((ndf, Y, X), numpy.float32)
    for z in range(T):
        image = numpy.ascontiguousarray(projections[z], dtype=numpy.float32)
        ext.corr(image, ctf, buf, device)
        corrected[z] = buf
    if ndf == 1:
        corrected = corrected.reshape(projections.shape)
    return corrected, mn - defocus, mx - defocus


def correct_file(input_filename, output_filename, centre=None, energy=None,
                 defocus=None, num_defocus=None, step_defocus=None,
                 spherical_aberration=None, astigmatism=None,
                 astigmatism_angle=None, phase_shift=None, device="cpu"):
    """CTF-corrects an MRC tilt series into a (T, ndf, Y, X) MRC volume
    stack (correct.py:321-473).

    Angles come from the input's FEI extended header ("Alpha tilt",
    degrees), the pixel size from its voxel size. The output carries an FEI1
    extended header with one block per (tilt, defocus): "Alpha tilt" in
    degrees and "Defocus" = the block's defocus relative to `defocus`
    (min_defocus - defocus + i * step), which reconstruct_file reads back as
    the slab range.
    """
    start_time = time.time()
    dev = ext.eDevice(device) if isinstance(device, (str, int)) else device
    if dev != ext.eDevice.gpu:
        raise RuntimeError(
            "guanaco_amd implements the GPU device (device='gpu'); the CPU "
            "device is the reference's own FFTW implementation")
    print("Reading %s" % input_filename)
    with mrc.mmap(input_filename) as infile:
        voxel_size = infile.voxel_size
        print("Voxel size: ", voxel_size)
        header = infile.indexed_extended_header
        if header is None:
            raise ValueError("the input has no FEI extended header with tilt "
                             "angles")
        assert infile.data.shape[0] == header.shape[0], \
            "Inconsistent data/header array shape"
        angles = numpy.deg2rad(numpy.asarray(header["Alpha tilt"],
                                             numpy.float64)).astype(
                                                 numpy.float32)
        for i, a in enumerate(angles):
            print("Image %d; angle %.4f deg" % (i + 1, numpy.rad2deg(a)))
        assert voxel_size["x"] == voxel_size["y"]
        pixel_size = float(voxel_size["x"])
        projections = infile.data
        check_detector_size(projections.shape[2], projections.shape[1])
        centre = get_centre(projections.shape, centre, False)
        if num_defocus is None or num_defocus == 0:
            num_defocus = 1
            step_defocus = 0
        output_shape = (projections.shape[0], num_defocus,
                        projections.shape[1], projections.shape[2])
        extended_header = numpy.zeros(output_shape[0:2],
                                      mrc.FEI1_EXTENDED_HEADER_DTYPE)
        extended_header["Metadata size"] = extended_header.dtype.itemsize
        print("Writing corrected projections to %s" % output_filename)
        with mrc.new_mmap(output_filename, output_shape,
                          mrc_mode=mrc.mode_from_dtype(numpy.float32),
                          overwrite=True, extended_header=extended_header,
                          exttyp=b"FEI1") as outfile:
            outfile.voxel_size = voxel_size
            _, min_defocus, max_defocus = correct_projections(
                projections, centre, pixel_size, energy, defocus,
                num_defocus, step_defocus, spherical_aberration or 0,
                astigmatism or 0, astigmatism_angle or 0, phase_shift or 0,
                outfile.data, device)
            if step_defocus is None:
                if num_defocus == 1:
                    step_defocus = 0
                else:
                    step_defocus = (max_defocus - min_defocus) / (
                        num_defocus - 1)
            blocks = outfile.indexed_extended_header.reshape(
                output_shape[0:2])
            blocks["Alpha tilt"] = numpy.rad2deg(angles)[:, None]
            blocks["Defocus"] = (min_defocus + numpy.arange(num_defocus)
                                 * step_defocus)[None, :]
    print("Time: %.2f seconds" % (time.time() - start_time))
