"""Synthetic workloads of the measured configurations.

Every value the configuration text leaves open is fixed here, once, and
recorded in the docstring of its builder.
"""
import math

import numpy as np

from .scene import (Background, Camera, Dielectric, DiffuseLight, Lambertian, Metal, Quad, Scene, Sphere)


def box_room(width=256, height=256):
    """Plain path tracing of a box room (the CPU-runnable reference case).

    Five quads span [0,1]^3 (floor y=0, ceiling y=1, back z=1, left x=0,
    right x=1; the z=0 side is open toward the camera).  Albedo 0.73 except
    left 0.65/0.05/0.05 and right 0.12/0.45/0.15.  A 0.25x0.25 DiffuseLight
    of radiance 15 sits at y = 0.999 centred in the ceiling, facing down, and
    is listed both in the world and in the NEE light list as `[[lights]]`
    entries are (src/scene_file.rs:1054-1066).  Two r=0.18 spheres: a GGX
    metal (PbrMaterial metallic 1, roughness 0.3, albedo fixed at
    0.80/0.85/0.88) at (0.30, 0.18, 0.55) and a Dielectric of IOR 1.5 at
    (0.70, 0.18, 0.35).  Solid black background.  Pinhole camera at
    (0.5, 0.5, -1.4) looking at (0.5, 0.5, 0.5), vfov 40, aspect W/H.
    Render: 16 spp, max depth 8, no photons.
    """
    white = Lambertian((0.73, 0.73, 0.73))
    red = Lambertian((0.65, 0.05, 0.05))
    green = Lambertian((0.12, 0.45, 0.15))
    light = DiffuseLight((15.0, 15.0, 15.0))
    objs = [
        Quad((0, 0, 0), (1, 0, 0), (0, 0, 1), white),       # floor
        Quad((0, 1, 0), (1, 0, 0), (0, 0, 1), white),       # ceiling
        Quad((0, 0, 1), (1, 0, 0), (0, 1, 0), white),       # back
        Quad((0, 0, 0), (0, 0, 1), (0, 1, 0), red),         # left
        Quad((1, 0, 0), (0, 0, 1), (0, 1, 0), green),       # right
        Sphere((0.30, 0.18, 0.55), 0.18, Metal((0.80, 0.85, 0.88), 0.3)),
        Sphere((0.70, 0.18, 0.35), 0.18, Dielectric(1.5)),
    ]
    lq = dict(q=(0.375, 0.999, 0.375), u=(0.25, 0, 0), v=(0, 0, 0.25))
    objs.append(Quad(lq["q"], lq["u"], lq["v"], light))
    lights = [Quad(lq["q"], lq["u"], lq["v"], light)]
    scene = Scene(objs, lights, Background.solid((0, 0, 0)), name="box_room")
    cam = Camera((0.5, 0.5, -1.4), (0.5, 0.5, 0.5), (0, 1, 0), 40.0, width / height)
    return scene, cam, dict(width=width, height=height, spp=16, max_depth=8, light_quad=lq)
