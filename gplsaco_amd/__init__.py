"""MI355X-native FACO batch engine (HIP/CDNA4) behind the gp_faco_ext boundary.

``gplsaco_amd.gp_faco_ext`` is the drop-in Python surface; the engine itself
is the C ABI library ``libfaco_hip.so`` (include/faco_engine.h).
"""
__all__ = ["gp_faco_ext"]
