"""MI355X-native (gfx950) electrostatic boundary-element hot path of KEMField.

Host orchestration (C++) and HIP kernels live in ``csrc/`` and are reached through the C-ABI
declared in ``include/kemhip.h``; ``kemfield`` mirrors the reference's class interface over
it, ``geometry`` builds element records and the benchmark meshes.
"""
from . import geometry  # noqa: F401

__all__ = ["geometry", "kemfield", "build"]
