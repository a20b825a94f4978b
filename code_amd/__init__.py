"""MI355X-native rebuild of GETM's 3-D baroclinic transport step.

The compute path is ``libgetm_transport.so`` (hand-written HIP for gfx950,
C-ABI in include/getm_transport.h).  ``code_amd.transport`` is its Python
binding; ``code_amd.inputs`` builds the seeded workloads.
"""
from . import inputs  # noqa: F401

__all__ = ["inputs", "transport"]
