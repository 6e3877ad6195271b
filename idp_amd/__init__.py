"""MI355X (gfx950) Newton-with-barrier inner loop of IDP.

Host side mirrors the reference's FEMShell animation-fix driver for tet
meshes; all hot-path work runs in libidp_amd.so (HIP) behind the C ABI of
``include/idp_amd.h``.
"""
from . import scenes  # noqa: F401

__all__ = ["scenes", "capi", "driver"]
