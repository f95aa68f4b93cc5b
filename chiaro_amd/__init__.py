"""MI355X-native (gfx950) depth hot path of chiaro-fusion: multi-view
plane-sweep ZNCC cost field, 8-path SGM, WTA with sub-label fit, direct
fine-grid remeasurement, consistency filters and per-target cross-camera
decisions, behind the C ABI in include/chiaro_depth.h."""
from .depth import (Context, DenseDepthMap, DepthOptions, Session, TargetResult, luminance_half,
                    refine_multiview_depth)

__all__ = ["Context", "DenseDepthMap", "DepthOptions", "Session", "TargetResult", "luminance_half",
           "refine_multiview_depth"]
