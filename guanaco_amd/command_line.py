"""The ``guanaco``, ``guanaco.correct``, ``guanaco.plot_ctf`` and
``guanaco.generate_ctf`` commands (reference src/guanaco/command_line.py;
console scripts at setup.py:64-69).

Same options, destinations and defaults as the reference, so the same
command lines work: ``python -m guanaco_amd -i tilts.mrc -o rec.mrc -d gpu
--df 20000 --ndf 8`` reconstructs, ``python -m guanaco_amd.correct -i ...``
corrects, ``python -m guanaco_amd.plot_ctf`` and ``python -m
guanaco_amd.generate_ctf --size 512,512`` are the CTF tools. The device option keeps the reference's spellings (``-d gpu``;
the reference declares it as the single option string "-d,--device", which
argparse matches by prefix) and also takes ``--device``. Two options are
additive: ``--storage`` (filtered-projection precision on the GPU) and
``--depth`` (reconstruct only the central rows of each slice).
"""
import argparse

from guanaco_amd import plot
from guanaco_amd.detail.correct import correct_file
from guanaco_amd.detail.reconstruct import reconstruct_file

__all__ = ["add_argument", "main", "correct", "plot_ctf", "generate_ctf"]


def add_argument(parser, name):
    """The options the commands share (command_line.py:26-160)."""
    spec = {
        "energy": (["--E"], dict(dest="energy", default=300, type=float,
                   help="The energy of the incident electrons (keV)")),
        "defocus": (["--df"], dict(dest="defocus", default=None, type=float,
                    help="The defocus at the rotation axis (A)")),
        "num_defocus": (["--ndf"], dict(
            dest="num_defocus", default=None, type=int,
            help="The number of defocus steps to use")),
        "step_defocus": (["--sdf"], dict(
            dest="step_defocus", default=None, type=int,
            help="The defocus step to use")),
        "spherical_aberration": (["--Cs"], dict(
            dest="spherical_aberration", default=0, type=float,
            help="The spherical aberration (mm)")),
        "astigmatism": (["--Ca"], dict(
            dest="astigmatism", default=0, type=float,
            help="The 2-fold astigmatism (A)")),
        "astigmatism_angle": (["--Pa"], dict(
            dest="astigmatism_angle", default=0, type=float,
            help="The angle for the 2-fold astigmatism (deg)")),
        "phase_shift": (["--phi"], dict(
            dest="phase_shift", default=0, type=float,
            help="The phase shift to apply to the CTF (deg)")),
        "defocus_spread": (["--d", "--defocus_spread"], dict(
            dest="defocus_spread", type=float, default=None,
            help="The defocus spread (A)")),
        "source_spread": (["--theta_c", "--source_spread"], dict(
            dest="source_spread", type=float, default=0.02,
            help="The source spread (mrad)")),
        "chromatic_aberration": (["--Cc", "--chromatic_aberration"], dict(
            dest="chromatic_aberration", type=float, default=2.7,
            help="The chromatic aberration (mm)")),
        "energy_spread": (["--dEE", "--energy_spread"], dict(
            dest="energy_spread", type=float, default=0.33e-6,
            help="The energy spread dE / E")),
        "acceleration_voltage_spread": (
            ["--dVV", "--acceleration_voltage_spread"], dict(
                dest="acceleration_voltage_spread", type=float,
                default=0.8e-6,
                help="The acceleration voltage spread dV / V")),
        "current_spread": (["--dII", "--current_spread"], dict(
            dest="current_spread", type=float, default=0.33e-6,
            help="The current spread dI / I")),
    }
    flags, kw = spec[name]
    parser.add_argument(*flags, **kw)


def _device_argument(parser):
    parser.add_argument("-d", "--device", "-d,--device", dest="device",
                        default="cpu", choices=["cpu", "gpu"],
                        help="Use either the CPU or GPU (this package "
                        "implements gpu)")


def reconstruct_parser():
    parser = argparse.ArgumentParser(description="Reconstruct")
    parser.add_argument(
        "-o", dest="output", default="rec.mrc",
        help="The output file containing the reconstructed projections.")
    parser.add_argument(
        "-i", dest="input", default=None, required=True,
        help="The input file containing a stack of projection images.")
    parser.add_argument(
        "--corrected-filename", dest="corrected_filename",
        default="GUANACO_CORRECTED.dat",
        help="The intermediate file containing the CTF corrected projections "
        "(written only by the two-step path, --no-fused)")
    _device_argument(parser)
    parser.add_argument(
        "--transform", dest="transform", default=None,
        choices=["none", "minus"],
        help="Set the transform to use on the corrected projections")
    # type=bool as the reference: any non-empty value switches it on
    parser.add_argument(
        "--angular_weights", dest="angular_weights", default=False,
        type=bool, help="Set whether or not to use angular weights")
    parser.add_argument(
        "-n", "--ncore", "-n,--ncore", dest="ncore", default=None, type=int,
        help="Set the number of cores to use")
    parser.add_argument(
        "--chunk_size", dest="chunk_size", default=None, type=int,
        help="Set the number of rows to reconstruct")
    parser.add_argument(
        "--start_angle", dest="start_angle", default=None, type=float,
        help="The starting angle in degrees.")
    parser.add_argument(
        "--step_angle", dest="step_angle", default=None, type=float,
        help="The step angle in degrees.")
    parser.add_argument(
        "--pixel_size", dest="pixel_size", default=None, type=float,
        help="The pixel size in A")
    parser.add_argument(
        "--centre", dest="centre", default=None, type=float,
        help="The rotation centre in pixels.")
    parser.add_argument(
        "--method", dest="method", default="FBP_CTF",
        choices=["FBP_CTF", "FBP", "SIRT", "SART", "CGLS", "EM"],
        help="The reconstruction algorithm; this package implements FBP_CTF "
        "(the others are the reference's tomopy path)")
    parser.add_argument(
        "--num_iter", dest="num_iter", default=50, type=int,
        help="The number of iterations to use in SIRT/SART/CGLS/EM")
    parser.add_argument(
        "--storage", dest="storage", default="f32",
        choices=["f32", "f16", "bf16"],
        help="Precision of the filtered projections on the GPU "
        "(accumulation is float32; f16 stays within 1e-3 rel-L2 of FP64, "
        "bf16 does not)")
    parser.add_argument(
        "--depth", dest="depth", default=None, type=int,
        help="Reconstruct only the central DEPTH rows of each slice")
    parser.add_argument(
        "--no-fused", dest="fused", action="store_false", default=True,
        help="Correct into --corrected-filename first, then reconstruct")
    for name in ("energy", "defocus", "num_defocus", "step_defocus",
                 "spherical_aberration", "astigmatism", "astigmatism_angle",
                 "phase_shift"):
        add_argument(parser, name)
    return parser


def main(args=None):
    """The ``guanaco`` command (command_line.py:167-326)."""
    args = reconstruct_parser().parse_args(args=args)
    reconstruct_file(
        input_filename=args.input,
        output_filename=args.output,
        corrected_filename=args.corrected_filename,
        start_angle=args.start_angle,
        step_angle=args.step_angle,
        pixel_size=args.pixel_size,
        centre=args.centre,
        energy=args.energy,
        defocus=args.defocus,
        num_defocus=args.num_defocus,
        step_defocus=args.step_defocus,
        spherical_aberration=args.spherical_aberration,
        astigmatism=args.astigmatism,
        astigmatism_angle=args.astigmatism_angle,
        phase_shift=args.phase_shift,
        device=args.device,
        ncore=args.ncore,
        transform=args.transform,
        angular_weights=args.angular_weights,
        chunk_size=args.chunk_size,
        method=args.method,
        num_iter=args.num_iter,
        storage=args.storage,
        depth=args.depth,
        fused=args.fused,
    )


_ENVELOPE_ARGS = ("energy", "defocus", "spherical_aberration", "astigmatism",
                  "astigmatism_angle", "phase_shift", "defocus_spread",
                  "source_spread", "chromatic_aberration", "energy_spread",
                  "acceleration_voltage_spread", "current_spread")


def _envelope_kwargs(args):
    return {k: getattr(args, k) for k in _ENVELOPE_ARGS}


def plot_ctf_parser():
    parser = argparse.ArgumentParser(description="Plot the CTF")
    parser.add_argument("-o", dest="output", default="ctf.png",
                        help="The output image file for the CTF plot")
    parser.add_argument("--q_max", dest="q_max", type=float, default=0.5,
                        help="The maximum spatial frequency (1/A)")
    parser.add_argument("--component", dest="component", type=str,
                        choices=["real", "imag", "abs"], default="imag",
                        help="The component to plot")
    parser.add_argument("--envelope", dest="envelope", type=str,
                        choices=["spatial", "temporal", "all"], default=None,
                        help="Plot the CTF envelope")
    for name in _ENVELOPE_ARGS:
        add_argument(parser, name)
    return parser


def plot_ctf(args=None):
    """The ``guanaco.plot_ctf`` command (command_line.py:329-409)."""
    args = plot_ctf_parser().parse_args(args=args)
    plot.plot_ctf(filename=args.output, q_max=args.q_max,
                  component=args.component, envelope=args.envelope,
                  **_envelope_kwargs(args))


def generate_ctf_parser():
    parser = argparse.ArgumentParser(description="Generate the CTF image")
    parser.add_argument("-o", dest="output", default="ctf.mrc",
                        help="The output image file for the CTF image")
    parser.add_argument("--size", "--image_size", dest="image_size",
                        type=lambda x: list(map(int, x.split(","))),
                        default=None, required=True, help="The image size")
    parser.add_argument("--px", "--pixel_size", dest="pixel_size", type=float,
                        default=1, help="The pixel size")
    # type=bool as the reference: any non-empty value switches it on
    parser.add_argument("--recentre", dest="recentre", type=bool,
                        default=False,
                        help="Recentre the CTF in Fourier space")
    for name in _ENVELOPE_ARGS[:2] + ("num_defocus", "step_defocus") + \
            _ENVELOPE_ARGS[2:]:
        add_argument(parser, name)
    return parser


def generate_ctf(args=None):
    """The ``guanaco.generate_ctf`` command (command_line.py:412-494); the
    images are computed on the GPU."""
    args = generate_ctf_parser().parse_args(args=args)
    plot.generate_ctf(filename=args.output, image_size=args.image_size,
                      pixel_size=args.pixel_size,
                      num_defocus=args.num_defocus,
                      step_defocus=args.step_defocus,
                      recentre=args.recentre, **_envelope_kwargs(args))


def correct_parser():
    parser = argparse.ArgumentParser(description="Perform CTF correction")
    parser.add_argument(
        "-o", dest="output", default="corrected.mrc", type=str,
        help="The output file containing the corrected projections.")
    parser.add_argument(
        "-i", dest="input", default=None, required=True, type=str,
        help="The input file containing a stack of projection images.")
    _device_argument(parser)
    parser.add_argument(
        "--centre", dest="centre", default=None, type=float,
        help="The rotation centre in pixels.")
    for name in ("energy", "defocus", "num_defocus", "step_defocus",
                 "spherical_aberration", "astigmatism", "astigmatism_angle",
                 "phase_shift"):
        add_argument(parser, name)
    return parser


def correct(args=None):
    """The ``guanaco.correct`` command (command_line.py:497-566)."""
    args = correct_parser().parse_args(args=args)
    correct_file(
        input_filename=args.input,
        output_filename=args.output,
        centre=args.centre,
        energy=args.energy,
        defocus=args.defocus,
        num_defocus=args.num_defocus,
        step_defocus=args.step_defocus,
        spherical_aberration=args.spherical_aberration,
        astigmatism=args.astigmatism,
        astigmatism_angle=args.astigmatism_angle,
        phase_shift=args.phase_shift,
        device=args.device,
    )
