"""MI355X-native (gfx950) MuhRec reconstruction hot path.

Hand-written HIP kernels behind a C ABI (``include/muhrec_hip.h``,
``imagingsuite_amd/lib/libmuhrec_hip.so``): the ProjectionFilterSingle ramp
filter, the parallel-beam back-projectors (MultiProjBPparallel, MultiProjBP,
NearestNeighborBP) and the FDKbp_single cone-beam back-projector.  MuhRec's
plugin API (BackProjectorModuleBase / PreprocModuleBase) sits on top in C++
(``imagingsuite_amd/csrc/modules``); ``recon`` mirrors it for Python callers.
"""
from ._lib import LIB_PATH, MrError, load  # noqa: F401

__all__ = ["LIB_PATH", "MrError", "load"]
