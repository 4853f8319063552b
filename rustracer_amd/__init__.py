"""rustracer_amd — MI355X-native per-sample radiance path of rustracer.

The product is librtamd.so (HIP, gfx950) behind the C ABI of
include/rt_api.h.  This package holds the host-side mirror of the
reference's scene and render interfaces (scene.py, renderer.py) and the
synthetic workload builders (scenes.py).  Importing it does not load the
library; `Renderer()` does, and raises if it is missing or no GPU is present.
"""
from .scene import (Background, Box, Camera, Checker, ConstantMedium, Convex, Diamond, Dielectric,  # noqa: F401
                    DiffuseLight, GridMedium, Lambertian, Mesh, Metal, PbrMaterial, Quad, RoughDielectric,
                    Scene, SpectralDielectric, SpectralMetal, Sphere, StepCut)
from .renderer import Renderer, RtError  # noqa: F401

__all__ = ["Renderer", "RtError", "Scene", "Camera", "Background"]
