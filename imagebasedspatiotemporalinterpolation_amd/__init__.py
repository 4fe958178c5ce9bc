"""MI355X (gfx950) rebuild of the per-frame interpolation path of
DerThomy/ImageBasedSpatioTemporalInterpolation (`Renderer::render`,
src/renderer.cu): bidirectional reprojection of two keyframes, depth-sorted
64-bit atomicMin splatting, fixed-point flow search and a fused resolve
(decision, second chance blend, soft shadows, aux-camera fallback, hole
retention), as hand-written HIP kernels behind the C ABI of include/isti.h.

Python entry points:
  renderer.Renderer / RenderInfo / FrameData  - the reference's interface
  scene                                       - seeded synthetic keyframes
  build                                       - in-tree native build
"""
from .renderer import (  # noqa: F401
    AdditionalRenderInfo,
    AuxCamData,
    Batch,
    FrameData,
    KeyframeView,
    RenderInfo,
    RenderMode,
    Renderer,
)

__version__ = "0.1.0"
