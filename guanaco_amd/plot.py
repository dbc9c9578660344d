"""CTF tools (reference src/guanaco/plot.py): plot_ctf and generate_ctf.

Same arguments, units and defaults as the reference. generate_ctf computes
the CTF images on the GPU (gx_ctf_grid: FP64 on the device, complex64 out,
every defocus in one launch per 256 MiB of output) and writes them to an
MRC file; plot_ctf evaluates the 1D CTF and its envelopes with the host
CTF class (a thousand points) and draws them with matplotlib.
"""
from math import pi

import numpy

import guanaco_amd.detail as ext
from guanaco_amd import _capi, mrc

__all__ = ["plot_ctf", "generate_ctf"]


def _units(spherical_aberration, chromatic_aberration, astigmatism_angle,
           energy, source_spread, defocus_spread, energy_spread,
           current_spread, acceleration_voltage_spread):
    """The reference's conversions (plot.py:77-91): mm -> A, deg -> rad,
    keV -> eV, mrad -> rad, and the defocus spread from Cc, dE/E, dI/I,
    dV/V when not given."""
    source_spread = source_spread / 1e3
    spherical_aberration = spherical_aberration * 1e7
    chromatic_aberration = chromatic_aberration * 1e7
    astigmatism_angle = astigmatism_angle * pi / 180
    energy = energy * 1e3
    if defocus_spread is None:
        defocus_spread = ext.get_defocus_spread(
            Cc=chromatic_aberration, dEE=energy_spread, dII=current_spread,
            dVV=acceleration_voltage_spread)
    return (spherical_aberration, astigmatism_angle, energy, source_spread,
            defocus_spread)


def plot_ctf(filename=None, q_max=0.5, defocus=0, spherical_aberration=2.7,
             astigmatism=0, astigmatism_angle=0, energy=300,
             defocus_spread=None, source_spread=0.02,
             chromatic_aberration=2.7, energy_spread=0.33e-6,
             current_spread=0.33e-6, acceleration_voltage_spread=0.8e-6,
             phase_shift=0, component="imag", envelope=None, ax=None,
             label=None, Et_label=None, Es_label=None, show=False):
    """Plots the CTF over q in [0, q_max) (plot.py:28-172): component
    "real", "imag" or "abs" of A Es Et exp(-i(chi - phi)), and optionally
    the "spatial", "temporal" or "all" envelopes. phase_shift is in radians
    here, as in the reference's function (its command converts nothing)."""
    Cs, Pa, energy, theta_c, dd = _units(
        spherical_aberration, chromatic_aberration, astigmatism_angle,
        energy, source_spread, defocus_spread, energy_spread, current_spread,
        acceleration_voltage_spread)
    if defocus is None:
        defocus = 0
    c = ext.CTF(l=ext.get_electron_wavelength(energy), df=defocus, Cs=Cs,
                Ca=astigmatism, Pa=Pa, dd=dd, theta_c=theta_c,
                phi=phase_shift)
    q = numpy.arange(0, q_max, q_max / 1000.0)
    theta = numpy.zeros(q.size)
    ctf = c.get_ctf(q, theta)
    if component not in (None, "real", "imag", "abs"):
        raise ValueError(
            "Expected component in ['real', 'imag', 'abs'], got %s" % component)
    from matplotlib import pyplot
    fig = None
    if ax is None:
        fig, ax = pyplot.subplots(figsize=(12, 8))
    ax.set_xlabel("Spatial frequency (1/A)")
    ax.set_ylabel("CTF")
    if component is not None:
        part = {"real": numpy.real, "imag": numpy.imag, "abs": numpy.abs}
        ax.plot(q, part[component](ctf),
                label="CTF (%s)" % component if label is None else label)
    if envelope is not None:
        if envelope in ("spatial", "all"):
            ax.plot(q, c.get_Es(q, theta),
                    label="Es" if Es_label is None else Es_label)
        if envelope in ("temporal", "all"):
            ax.plot(q, c.get_Et(q),
                    label="Et" if Et_label is None else Et_label)
        ax.legend()
    if fig is not None:
        if filename is not None:
            fig.savefig(filename, dpi=300, bbox_inches="tight")
        elif show:
            pyplot.show()
        pyplot.close(fig)


def generate_ctf(filename=None, image_size=None, pixel_size=1, defocus=0,
                 num_defocus=None, step_defocus=None,
                 spherical_aberration=2.7, astigmatism=0, astigmatism_angle=0,
                 energy=300, defocus_spread=None, source_spread=0.02,
                 chromatic_aberration=2.7, energy_spread=0.33e-6,
                 current_spread=0.33e-6, acceleration_voltage_spread=0.8e-6,
                 phase_shift=0, component="imag", envelope=True,
                 recentre=False, gpu_index=-1):
    """Writes the CTF images of num_defocus defoci, defocus + (i - (n-1)/2)
    * step_defocus, to an MRC file of complex64 (num_defocus, image_size[0],
    image_size[1]) (plot.py:175-284). As in the reference, the images are
    always the full CTF with envelopes (`component` and `envelope` are
    accepted and unused), in FFT order unless recentre."""
    Cs, Pa, energy, theta_c, dd = _units(
        spherical_aberration, chromatic_aberration, astigmatism_angle,
        energy, source_spread, defocus_spread, energy_spread, current_spread,
        acceleration_voltage_spread)
    if defocus is None:
        defocus = 0
    if num_defocus is None or num_defocus == 0:
        num_defocus = 1
    if step_defocus is None:
        step_defocus = 0
    h, w = int(image_size[0]), int(image_size[1])
    dfs = numpy.array([defocus + (i - 0.5 * (num_defocus - 1)) * step_defocus
                       for i in range(num_defocus)], numpy.float64)
    with mrc.new_mmap(filename, (num_defocus, h, w),
                      mrc_mode=mrc.mode_from_dtype(numpy.complex64),
                      overwrite=True) as handle:
        handle.voxel_size = pixel_size
        for df in dfs:
            print("Computing CTF for defocus = %d A" % df)
        _capi.ctf_grid(w, h, pixel_size, dfs,
                       ext.get_electron_wavelength(energy), Cs=Cs,
                       Ca=astigmatism, Pa=Pa, dd=dd, theta_c=theta_c,
                       phi=phase_shift, envelope=True, recentre=recentre,
                       out=handle.data, gpu_index=gpu_index)
