"""core_amd -- MI355X (gfx950) Superconvergent Patch Recovery.

A from-scratch HIP implementation of SCOREC/core's SPR error-estimation path
(spr::getGradIPField, spr::recoverField, spr::getSPRSizeField,
spr::getTargetSPRSizeField) behind a C ABI (include/sprgpu.h), with a Python
mirror of the reference interface in core_amd.spr.
"""
from . import _lib
from ._lib import SPRError

__all__ = ["SPRError", "lib"]


def lib():
    """The loaded libsprgpu.so (raises ImportError when it is not built)."""
    return _lib.lib()
