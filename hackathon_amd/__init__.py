"""MI355X-native differential dynamic microscopy engine.

The compute lives in libddm.so (hand-written HIP for gfx950, C ABI in
include/ddm.h).  `hackathon_amd.engine.DdmEngine` is the Python host mirror of
the reference's engine interface; `hackathon_amd/js/ddm-engine.js` is the
TypeScript-facing facade over the N-API addon.
"""
from .engine import DdmEngine, DdmError, Radial, load_library  # noqa: F401
