"""MI355X-native (gfx950) implementation of WoodenPBREngine's ray-tracing hot path.

The compute path is ``libwpbr_hip.so`` (hand-written HIP kernels behind a C ABI,
``include/wpbr.h``).  This package holds the Python binding of that ABI
(:mod:`woodenpbrengine_amd.lib`), seeded synthetic scene generators
(:mod:`woodenpbrengine_amd.scenes`) and the C++ host mirror of the reference's
Job/ECS surface (``host/``).
"""
from pathlib import Path

PACKAGE_DIR = Path(__file__).resolve().parent
LIB_PATH = PACKAGE_DIR / "libwpbr_hip.so"

__all__ = ["PACKAGE_DIR", "LIB_PATH"]
