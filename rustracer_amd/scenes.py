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


# --------------------------------------------------------------------- helpers
class Perlin:
    """Gradient noise with the structure of rustracer's Perlin (src/perlin.rs):
    256 random unit vectors, three permutation tables, Hermite-smoothed
    trilinear blend of the eight corner dot products; turbulence sums octaves
    at doubling frequency and halving weight and takes |.|.  Tables come from
    numpy's seeded generator (the reference seeds from thread_rng)."""

    def __init__(self, seed):
        g = np.random.default_rng(seed)
        v = g.uniform(-1, 1, size=(256, 3))
        self.ranvec = (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)
        self.perm = [g.permutation(256) for _ in range(3)]

    def noise(self, p):
        p = np.asarray(p, dtype=np.float32)
        f = np.floor(p)
        uvw = p - f
        s = uvw * uvw * (3.0 - 2.0 * uvw)
        ijk = f.astype(np.int64)
        acc = np.zeros(len(p), np.float32)
        for di in (0, 1):
            for dj in (0, 1):
                for dk in (0, 1):
                    idx = (self.perm[0][(ijk[:, 0] + di) & 255] ^ self.perm[1][(ijk[:, 1] + dj) & 255]
                           ^ self.perm[2][(ijk[:, 2] + dk) & 255])
                    c = self.ranvec[idx]
                    w = uvw - np.array([di, dj, dk], np.float32)
                    wt = ((di * s[:, 0] + (1 - di) * (1 - s[:, 0])) * (dj * s[:, 1] + (1 - dj) * (1 - s[:, 1]))
                          * (dk * s[:, 2] + (1 - dk) * (1 - s[:, 2])))
                    acc += wt * (c * w).sum(1)
        return acc

    def turb(self, p, octaves):
        p = np.asarray(p, dtype=np.float32)
        acc = np.zeros(len(p), np.float32)
        w = 1.0
        for _ in range(octaves):
            acc += w * self.noise(p)
            w *= 0.5
            p = p * 2.0
        return np.abs(acc)


def icosphere(level, radius=1.0):
    """Subdivided icosahedron: 20 * 4**level faces.  Returns (verts, faces)."""
    t = (1.0 + 5 ** 0.5) / 2.0
    V = np.array([(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
                  (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)], dtype=np.float64)
    Fc = np.array([(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
                   (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5),
                   (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)], dtype=np.int64)
    verts = V / np.linalg.norm(V, axis=1, keepdims=True)
    faces = Fc
    for _ in range(level):
        a, b, c = faces[:, 0], faces[:, 1], faces[:, 2]
        e = np.concatenate([np.stack([a, b], 1), np.stack([b, c], 1), np.stack([c, a], 1)])
        e.sort(axis=1)
        key = e[:, 0] * (len(verts) + 1) + e[:, 1]
        uniq, inv = np.unique(key, return_inverse=True)
        ua, ub = uniq // (len(verts) + 1), uniq % (len(verts) + 1)
        mids = verts[ua] + verts[ub]
        mids /= np.linalg.norm(mids, axis=1, keepdims=True)
        base = len(verts)
        verts = np.concatenate([verts, mids])
        m = inv.reshape(3, -1) + base
        ab, bc, ca = m[0], m[1], m[2]
        faces = np.concatenate([np.stack([a, ab, ca], 1), np.stack([b, bc, ab], 1), np.stack([c, ca, bc], 1),
                                np.stack([ab, bc, ca], 1)])
    return verts * radius, faces


def displaced_icosphere(level, radius=1.0, amplitude=0.05, octaves=3, frequency=4.0, seed=1):
    """Icosphere displaced along its normals by `amplitude` x Perlin turbulence
    (signed: 2*turb - mean) and smooth per-vertex normals.  Returns triangle
    soup arrays (T,3,3) vertices and normals."""
    verts, faces = icosphere(level, radius)
    n = verts / np.linalg.norm(verts, axis=1, keepdims=True)
    noise = Perlin(seed).turb((verts * frequency).astype(np.float32), octaves).astype(np.float64)
    noise = noise - noise.mean()
    pv = verts + n * (amplitude * noise / max(np.abs(noise).max(), 1e-9))[:, None]
    f = pv[faces]
    fn = np.cross(f[:, 1] - f[:, 0], f[:, 2] - f[:, 0])
    vn = np.zeros_like(pv)
    for k in range(3):
        np.add.at(vn, faces[:, k], fn)
    vn /= np.linalg.norm(vn, axis=1, keepdims=True)
    return pv[faces].astype(np.float32), vn[faces].astype(np.float32)


def sky_envmap(width=2048, height=1024, sun_elev_deg=35.0, sun_azim_deg=30.0, sun_sigma_deg=1.0, sun_peak=5.0e4):
    """Equirect float32 HDR (row 0 = +Y): a vertical sky gradient (deep blue
    zenith to pale horizon, dim brown below) plus a Gaussian sun lobe of
    angular sigma `sun_sigma_deg` and peak radiance `sun_peak`.  Texel
    directions follow EnvMapData's mapping (u = 0.5 + atan2(x,z)/2pi,
    v = acos(y)/pi, src/renderer.rs:97-121)."""
    j = (np.arange(height) + 0.5) / height
    i = (np.arange(width) + 0.5) / width
    theta = np.pi * j[:, None]
    phi = 2 * np.pi * (i[None, :] - 0.5)
    d = np.stack([np.sin(theta) * np.sin(phi), np.cos(theta) * np.ones_like(phi), np.sin(theta) * np.cos(phi)], -1)
    y = d[..., 1]
    zen = np.array([0.10, 0.25, 0.65])
    hor = np.array([0.85, 0.90, 1.00])
    gnd = np.array([0.12, 0.10, 0.08])
    up = np.clip(y, 0, 1)[..., None]
    img = np.where(y[..., None] >= 0, hor * (1 - up) + zen * up, gnd * np.ones_like(up))
    el, az = np.radians(sun_elev_deg), np.radians(sun_azim_deg)
    s = np.array([np.cos(el) * np.sin(az), np.sin(el), np.cos(el) * np.cos(az)])
    ang = np.arccos(np.clip(d @ s, -1, 1))
    lobe = sun_peak * np.exp(-0.5 * (ang / np.radians(sun_sigma_deg)) ** 2)
    img = img + lobe[..., None] * np.array([1.0, 0.95, 0.88])
    return img.astype(np.float32)


def mesh_envmap(width=1280, height=720, level=8):
    """QBVH-bound triangle mesh under an HDR environment.

    An icosphere of 20*4**level triangles (level 8: 1,310,720), radius 1 at the
    origin, displaced along its normals by 0.05 x 3-octave Perlin turbulence
    (frequency 4, permutation seed 1), smooth vertex normals, rough GGX
    dielectric (IOR 1.5, roughness 0.15).  A 20x20 diffuse 0.8 ground quad at
    y = -1.1.  Lighting: sky_envmap() 2048x1024 (sun lobe sigma 1 deg, peak
    5e4, elevation 35 deg, azimuth 30 deg), importance sampled through an
    alias table over its texels.  Camera at
    (0, 0.6, -3.6) looking at (0, -0.1, 0), vfov 40.  Render: 64 spp, max
    depth 12.
    """
    from .scene import Background, Camera, Lambertian, Mesh, Quad, RoughDielectric, Scene
    v, n = displaced_icosphere(level)
    objs = [Mesh(v, RoughDielectric(1.5, 0.15), normals=n),
            Quad((-10, -1.1, -10), (20, 0, 0), (0, 0, 20), Lambertian((0.8, 0.8, 0.8)))]
    scene = Scene(objs, [], Background.envmap(sky_envmap(), sampling="alias"), name=f"mesh_envmap_L{level}")
    cam = Camera((0.0, 0.6, -3.6), (0.0, -0.1, 0.0), (0, 1, 0), 40.0, width / height)
    return scene, cam, dict(width=width, height=height, spp=64, max_depth=12)


def gem_caustics(width=1280, height=720, num_photons=2_000_000, dispersion="cauchy", spin_deg=0.0, knn=0,
                 gather_radius=0.05):
    """Dispersive gem with caustics.

    A round brilliant (Diamond, table 57 %, crown 34.5 deg, pavilion 40.75 deg,
    girdle_facets 56 with no girdle band: 1 + 28 + 28 = 57 half-spaces) of
    radius 0.5, girdle at y = 0.55, rotated 12 deg about y.  SpectralDielectric
    diamond: the reference preset Cauchy B = 2.351, C = 0.0230
    (src/scene_file.rs:686; n(589) ~ 2.417) or, with dispersion="sellmeier",
    Peter's Sellmeier law (B 0.3306, 4.3356; C 0.175^2, 0.106^2 um^2).  Floor:
    diffuse 0.8 quad 8x8 at y = 0.  Light: a 0.3x0.3 DiffuseLight quad of
    radiance 50 at y = 2.2 above the stone, facing down, in world and light
    list.  Caustic photon map from that quad (build_from_quad), fixed gather
    radius 0.05, or with knn > 0 the k nearest photons within gather_radius.
    Camera at (0, 1.6, -2.6) looking at (0, 0.35, 0), vfov 35, black
    background.  Render: 256 spp, max depth 24.
    """
    from .scene import Background, Camera, Diamond, DiffuseLight, Lambertian, Quad, Scene, SpectralDielectric
    if dispersion == "sellmeier":
        mat = SpectralDielectric(sellmeier=(0.3306, 4.3356, 0.0, 0.175 ** 2, 0.106 ** 2, 0.0))
    else:
        mat = SpectralDielectric(2.351, 0.0230)
    gem = Diamond((0, 0, 0), 0.5, mat, table_pct=57.0, crown_angle=34.5, pavilion_angle=40.75,
                  girdle_facets=56, girdle_pct=0.0).rotated((0, 12 + spin_deg, 0)).translated((0, 0.55, 0))
    light = DiffuseLight((50.0, 50.0, 50.0))
    lq = dict(q=(-0.15, 2.2, -0.15), u=(0.3, 0, 0), v=(0, 0, 0.3))
    objs = [Quad((-4, 0, -4), (8, 0, 0), (0, 0, 8), Lambertian((0.8, 0.8, 0.8))), gem,
            Quad(lq["q"], lq["u"], lq["v"], light)]
    lights = [Quad(lq["q"], lq["u"], lq["v"], light)]
    scene = Scene(objs, lights, Background.solid((0, 0, 0)), name="gem_caustics" if not spin_deg else f"gem_spin{spin_deg:g}")
    cam = Camera((0.0, 1.6, -2.6), (0.0, 0.35, 0.0), (0, 1, 0), 35.0, width / height)
    return scene, cam, dict(width=width, height=height, spp=256, max_depth=24, light_quad=lq, light_color=(50.0, 50.0, 50.0),
                            num_photons=num_photons, gather_radius=gather_radius, knn=knn)


# The gem of the headline caustic workload: Sellmeier diamond, k = 64 nearest
# photons within the reference's default gather radius 0.15 (src/scene_file.rs:499).
GEM_BENCH = dict(dispersion="sellmeier", knn=64, gather_radius=0.15)


ADJOINT_PARAM_NAMES = (["n_d", "sellmeier_B1", "alpha_table", "floor_albedo"] +
                       [f"alpha_group{k}" for k in range(16)])


def fog_adjoint(width=64, height=64):
    """A constant-density fog ball (sigma_t 2.0, albedo 0.8, HG g 0.3) over a
    grey Lambertian floor, beside a diffuse sphere, under a quad light and a
    sky gradient: camera rays, bounces, shadow rays and misses all cross the
    fog.  cfg["params"]: the fog's sigma_t (rt_param with the medium flag) and
    the floor's grey albedo; cfg["theta"] their values."""
    from .scene import Background, Camera, ConstantMedium, DiffuseLight, Lambertian, Quad, Scene, Sphere
    from . import _ffi as F
    floor_mat = Lambertian((0.7, 0.7, 0.7))
    fog = Sphere((0.0, 0.55, 0.0), 0.5, None)
    ConstantMedium(fog, 2.0, (0.8, 0.8, 0.8), g=0.3)
    lq = Quad((-0.35, 1.9, -0.35), (0.7, 0, 0), (0, 0, 0.7), DiffuseLight((6.0, 6.0, 6.0)))
    objs = [Quad((-4, 0, -4), (8, 0, 0), (0, 0, 8), floor_mat), fog,
            Sphere((0.75, 0.3, 0.4), 0.3, Lambertian((0.6, 0.3, 0.2))), lq]
    scene = Scene(objs, [lq], Background.gradient((0.4, 0.5, 0.7), (0.9, 0.9, 0.9)), name="fog_adjoint")
    cam = Camera((0.0, 1.0, -2.4), (0.0, 0.45, 0.0), (0, 1, 0), 40.0, width / height)
    params = [(F.PARAM_MEDIUM | 0, 0, 1), (scene.material_index(floor_mat), 0, 3)]
    return scene, cam, dict(width=width, height=height, spp=16, max_depth=12, params=params,
                            theta=np.array([2.0, 0.7], np.float32))


def gem_adjoint(width=512, height=512, seed=4, rel_sigma=0.05):
    """Adjoint material gradients on the gem scene.

    The gem of gem_caustics() with a rough dispersive dielectric
    (RoughSpectralDielectric: n_d 2.417, Peter's diamond Sellmeier law with
    B1 0.3306, table alpha 0.02, 16 facet-group alphas drawn uniform in
    [0.005, 0.05] with `seed`) on the diffuse 0.8 floor, same light and camera,
    64 spp, depth 24, no photon map (caustics by path tracing through the rough
    facets, so every term of the image is differentiated).  20 parameters:
    n_d, B1, table alpha, grey floor albedo, 16 group alphas.  Returns the scene
    at the target parameters plus cfg["params"] (rt_param triples),
    cfg["theta_target"] and cfg["theta_start"] (the target perturbed by
    Gaussian noise of `rel_sigma` relative, seeded).
    """
    from .scene import Camera, Diamond, DiffuseLight, Lambertian, Quad, RoughSpectralDielectric, Scene, Background
    g = np.random.default_rng(seed)
    groups = g.uniform(0.005, 0.05, size=16).astype(np.float32)
    gem_mat = RoughSpectralDielectric(nd=2.417, alpha_table=0.02, alpha_groups=list(groups))
    floor_mat = Lambertian((0.8, 0.8, 0.8))
    gem = Diamond((0, 0, 0), 0.5, gem_mat, table_pct=57.0, crown_angle=34.5, pavilion_angle=40.75,
                  girdle_facets=56, girdle_pct=0.0).rotated((0, 12, 0)).translated((0, 0.55, 0))
    light = DiffuseLight((50.0, 50.0, 50.0))
    lq = dict(q=(-0.15, 2.2, -0.15), u=(0.3, 0, 0), v=(0, 0, 0.3))
    objs = [Quad((-4, 0, -4), (8, 0, 0), (0, 0, 8), floor_mat), gem, Quad(lq["q"], lq["u"], lq["v"], light)]
    lights = [Quad(lq["q"], lq["u"], lq["v"], light)]
    scene = Scene(objs, lights, Background.solid((0, 0, 0)), name="gem_adjoint")
    cam = Camera((0.0, 1.6, -2.6), (0.0, 0.35, 0.0), (0, 1, 0), 35.0, width / height)
    mg, mf = scene.material_index(gem_mat), scene.material_index(floor_mat)
    R = RoughSpectralDielectric
    params = [(mg, R.FIELD_ND, 1), (mg, R.FIELD_B1, 1), (mg, R.FIELD_ALPHA_TABLE, 1), (mf, 0, 3)]
    params += [(mg, R.FIELD_ALPHA_GROUP0 + k, 1) for k in range(16)]
    target = np.array([2.417, 0.3306, 0.02, 0.8] + list(groups), dtype=np.float32)
    start = (target * (1.0 + rel_sigma * g.standard_normal(target.shape))).astype(np.float32)
    return scene, cam, dict(width=width, height=height, spp=64, max_depth=24, params=params, theta_target=target,
                            theta_start=start, names=list(ADJOINT_PARAM_NAMES))


def perlin_grid(n=128, octaves=4, frequency=3.0, seed=2):
    """n^3 float32 density grid in [0,1]: `octaves`-octave Perlin turbulence
    over the unit cube, normalised by its maximum (x fastest, then y, z)."""
    ax = (np.arange(n, dtype=np.float32) + 0.5) / n
    z, y, x = np.meshgrid(ax, ax, ax, indexing="ij")
    p = np.stack([x.ravel(), y.ravel(), z.ravel()], 1) * np.float32(frequency)
    t = Perlin(seed).turb(p, octaves).reshape(n, n, n)
    return (t / max(float(t.max()), 1e-9)).astype(np.float32)


def volume_glass_box(width=640, height=480, grid_n=128):
    """Heterogeneous medium in a glass box under the HDR environment.

    A Dielectric box (IOR 1.33) spanning [-0.6,0.6]x[0,1.2]x[-0.6,0.6] on a
    diffuse 0.8 ground, filled (inset 0.02) with a delta-tracked GridMedium:
    density = perlin_grid(128, 4 octaves, frequency 3, seed 2) x sigma_t 8,
    albedo 0.9, Henyey-Greenstein g 0.6.  Lit by sky_envmap().  Camera at
    (0, 1.3, -3.0) looking at (0, 0.55, 0), vfov 35.  64 spp, depth 24.
    """
    from .scene import Background, Box, Camera, Dielectric, GridMedium, Lambertian, Quad, Scene
    inner = Box((-0.58, 0.02, -0.58), (0.58, 1.18, 0.58), None)
    GridMedium(inner, perlin_grid(grid_n), 8.0, (0.9, 0.9, 0.9), 0.6)
    objs = [Quad((-10, -0.001, -10), (20, 0, 0), (0, 0, 20), Lambertian((0.8, 0.8, 0.8))),
            Box((-0.6, 0.0, -0.6), (0.6, 1.2, 0.6), Dielectric(1.33)), inner]
    scene = Scene(objs, [], Background.envmap(sky_envmap(), sampling="alias"), name="volume_glass_box")
    cam = Camera((0.0, 1.3, -3.0), (0.0, 0.55, 0.0), (0, 1, 0), 35.0, width / height)
    return scene, cam, dict(width=width, height=height, spp=64, max_depth=24)


SPIN_RESOLUTIONS = [(256, 256), (512, 384), (960, 540)]


def spin_batch(n_stills=36, step_deg=10.0, spp=64, seed=3, num_photons=2_000_000, with_volume=True):
    """Batched renders of uneven resolution: `n_stills` spin stills of the gem
    scene (spin k * step_deg about the stone's centre, as scene_file's `spin`
    key, src/scene_file.rs:1032-1038) at resolutions drawn uniformly with
    `seed` from SPIN_RESOLUTIONS, each with its own caustic photon map,
    `spp` spp and depth 24; plus, when `with_volume`, volume_glass_box()."""
    g = np.random.default_rng(seed)
    picks = g.integers(0, len(SPIN_RESOLUTIONS), size=n_stills)
    views = []
    for k in range(n_stills):
        W, H = SPIN_RESOLUTIONS[picks[k]]
        sc, cam, cfg = gem_caustics(W, H, num_photons=num_photons, spin_deg=k * step_deg, **GEM_BENCH)
        cfg["spp"] = spp
        views.append((sc, cam, cfg))
    if with_volume:
        views.append(volume_glass_box())
    return views
