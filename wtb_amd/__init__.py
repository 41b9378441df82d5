"""MI355X (gfx950) implementation of WTB's exciton / Bethe-Salpeter hot path.

``wtb_amd.hip`` binds the C-ABI library (``libwtb_hip.so``, include/wtb_hip.h);
``wtb_amd.tbmodel`` holds the TB model container, reference-format TB file I/O
and the seeded synthetic model generators.
"""
from . import tbmodel  # noqa: F401
from . import hip  # noqa: F401

__all__ = ["tbmodel", "hip"]
