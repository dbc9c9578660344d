"""guanaco_amd: guanaco's CTF-corrected FBP hot path on AMD MI355X (gfx950).

The public names mirror the reference package ``guanaco``
(src/guanaco/__init__.py:22-28): ``reconstruct``, ``correct_projections``
and the ``guanaco_ext`` binding surface (``recon``, ``corr``, ``eDevice``,
``CTF``, ``get_electron_wavelength``). The GPU device (``device="gpu"``) is
this package's gfx950 path; the CPU device belongs to the reference.
"""
from guanaco_amd.detail.correct import *  # noqa: F401,F403
from guanaco_amd.detail.reconstruct import *  # noqa: F401,F403
from guanaco_amd.guanaco_ext import *  # noqa: F401,F403

__version__ = "0.1.0"
