"""Binding surface, as the reference's guanaco.detail (detail/__init__.py:21)."""
from guanaco_amd.guanaco_ext import *  # noqa: F401,F403
