"""nanotracer_amd — MI355X (gfx950) differentiable Gaussian-splat rasteriser.

The compute path is libnts.so (hand-written HIP kernels behind the C ABI in
include/nts.h).  This package is the host-side mirror of the reference's
nano-optimize API over that ABI.
"""
from . import synth  # noqa: F401
from ._lib import LIB_PATH, NtsError, lib  # noqa: F401

__all__ = ["synth", "lib", "LIB_PATH", "NtsError", "api"]


def __getattr__(name):
    if name == "api":
        from . import api

        return api
    raise AttributeError(name)
