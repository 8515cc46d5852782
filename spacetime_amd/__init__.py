"""spacetime_amd — MI355X-native (gfx950) Kerr–Schild null-geodesic tracer.

The product is libspacetime_hip.so (C ABI, include/spacetime_hip.h); this
package is its host-side mirror of SpaceTime.jl's tracer API.
"""
from .api import (AccretionDisc, Blackbody, Camera, DiscVolume, HipPreviewContext, Kerr, Lens,
                  MetalPreviewContext, Schwarzschild, StepRule, horizon_radius, ks_cam_params, ks_camera_tetrad,
                  make_sky, make_volume, photon_orbit_min, render, render_draft_hip, render_preview_hip,
                  render_preview_hip_, spacetime_params, spin, trace_args)
from ._lib import ST_PREC_FP32, ST_PREC_FP64, ST_PREC_MIXED, StError

__all__ = [
    "AccretionDisc", "Blackbody", "Camera", "DiscVolume", "HipPreviewContext", "Kerr", "Lens",
    "MetalPreviewContext", "Schwarzschild", "StepRule", "horizon_radius", "ks_cam_params", "ks_camera_tetrad",
    "make_sky", "make_volume", "photon_orbit_min", "render", "render_draft_hip", "render_preview_hip",
    "render_preview_hip_", "spacetime_params", "spin", "trace_args", "ST_PREC_FP32", "ST_PREC_FP64",
    "ST_PREC_MIXED", "StError",
]
