"""Shared test inputs: seeded ray batches and small scenes."""
import numpy as np

from rustracer_amd import _ffi as F
from rustracer_amd.scene import (Box, Dielectric, Lambertian, Mesh, Quad, Scene, Sphere, Background, DiffuseLight)


def rays_from(origins, dirs, t_min=0.001, t_max=np.inf, time=0.0):
    r = np.zeros(len(origins), dtype=F.ray_dtype())
    r["origin"] = origins
    r["dir"] = dirs
    r["t_min"] = t_min
    r["t_max"] = t_max
    r["time"] = time
    return r


def random_rays(n, seed, lo=(-1, -1, -1), hi=(2, 2, 2)):
    """Origins uniform in a box, directions uniform on the sphere (float32)."""
    g = np.random.default_rng(seed)
    o = g.uniform(lo, hi, size=(n, 3)).astype(np.float32)
    d = g.normal(size=(n, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return rays_from(o, d.astype(np.float32))


def icosphere(level, radius=1.0, center=(0, 0, 0)):
    """Subdivided icosahedron: 20 * 4**level triangles, as (T,3,3) float32."""
    t = (1.0 + 5 ** 0.5) / 2.0
    V = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    Fc = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
          (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
          (8, 6, 7), (9, 8, 1)]
    verts = np.array(V, dtype=np.float64)
    verts /= np.linalg.norm(verts, axis=1, keepdims=True)
    faces = np.array(Fc, dtype=np.int64)
    for _ in range(level):
        a, b, c = faces[:, 0], faces[:, 1], faces[:, 2]
        edges = np.concatenate([np.stack([a, b], 1), np.stack([b, c], 1), np.stack([c, a], 1)])
        key = np.sort(edges, axis=1)
        uniq, inv = np.unique(key, axis=0, return_inverse=True)
        mids = verts[uniq[:, 0]] + verts[uniq[:, 1]]
        mids /= np.linalg.norm(mids, axis=1, keepdims=True)
        base = len(verts)
        verts = np.concatenate([verts, mids])
        m = inv.reshape(3, -1) + base
        ab, bc, ca = m[0], m[1], m[2]
        faces = np.concatenate([np.stack([a, ab, ca], 1), np.stack([b, bc, ab], 1), np.stack([c, ca, bc], 1),
                                np.stack([ab, bc, ca], 1)])
    tri = (verts[faces] * radius + np.asarray(center)).astype(np.float32)
    return tri, verts, faces


def mixed_scene():
    """Spheres, quads, a box and a small mesh with transforms — every shape kind
    the trace kernels handle, for hit-parity checks."""
    from rustracer_amd.scene import Diamond
    w = Lambertian((0.7, 0.7, 0.7))
    tri, _, _ = icosphere(2, 0.4, (0.0, 0.0, 0.0))
    objs = [
        Quad((-2, -1, -2), (4, 0, 0), (0, 0, 4), w),
        Sphere((0.5, 0.0, 0.3), 0.3, Dielectric(1.5)),
        Sphere((-0.6, 0.2, -0.4), 0.25, w),
        Box((-0.2, -1.0, 0.6), (0.3, -0.4, 1.1), w).rotated((0, 25, 0), pivot=(0.05, -0.7, 0.85)),
        Mesh(tri, w).scaled(1.2).translated((0.9, 0.6, -0.5)),
        Diamond((0, 0, 0), 0.3, Dielectric(2.4), girdle_facets=56, girdle_pct=0.0).rotated((0, 15, 0)).translated((-0.5, 0.5, 0.6)),
        Quad((-0.3, 1.2, -0.3), (0.6, 0, 0), (0, 0, 0.6), DiffuseLight((5, 5, 5))),
    ]
    return Scene(objs, [], Background.solid((0, 0, 0)), name="mixed")
