"""Host-side mirror of rustracer's scene-building surface.

The reference assembles a scene from `dyn Hittable` shapes (src/sphere.rs,
src/quad.rs, src/triangle.rs via src/mesh.rs, src/diamond.rs), SRT wrappers
(src/transform.rs), media (src/volume.rs), `dyn Material`s (src/material.rs),
a light `HittableList` and a `Background` (src/renderer.rs:183-223), the way
`scene_file::build` does (src/scene_file.rs:899-1132).  This module keeps the
same names and argument meanings and flattens the result into the
`rt_scene_desc` of include/rt_api.h — the single structure that crosses the
C ABI.
"""
import ctypes as C
import math
import numpy as np

from . import _ffi as F

f32 = np.float32


def v3(x):
    a = np.asarray(x, dtype=np.float32).reshape(3)
    return a


def unit(a):
    a = np.asarray(a, dtype=np.float32)
    return (a / np.float32(np.sqrt(np.float32(np.dot(a, a))))).astype(np.float32)


def cross(a, b):
    return np.array([a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]], dtype=np.float32)


# ---------------------------------------------------------------- textures
def perlin_tables(seed):
    """Perlin::new (src/perlin.rs:11-33) with numpy draws: 256 unit vectors
    from the cube [-1, 1]^3, then three permutations of 0..255; returned as
    the 1536 floats the descriptor takes (ranvec, perm_x, perm_y, perm_z).
    The reference draws from thread_rng (NoiseMedium) or StdRng(0x5EED)
    (Mottle); a host binding passes its own Perlin's tables instead."""
    rng = np.random.default_rng(seed)
    v = rng.uniform(-1.0, 1.0, size=(256, 3)).astype(np.float32)
    v = v / np.sqrt((v * v).sum(1, keepdims=True))
    perms = [rng.permutation(256).astype(np.float32) for _ in range(3)]
    return np.ascontiguousarray(np.concatenate([v.astype(np.float32).ravel()] + perms), dtype=np.float32)


class Texture:
    """Texture (src/texture.rs:7-18): `tex` kind and the rt_material.p slots
    it fills; `image` / `perlin` are filled with descriptor indices."""
    tex = F.TEX_SOLID
    image = None
    perlin = None

    def slots(self, image_index=0, perlin_index=0):
        return {}


class SolidTexture(Texture):
    def __init__(self, color):
        self.color = v3(color)

    def slots(self, image_index=0, perlin_index=0):
        return {0: self.color[0], 1: self.color[1], 2: self.color[2]}


class CheckerTexture(Texture):
    """Texture::Checker: parity of floor(scale * p) (src/texture.rs:29-34)."""
    tex = F.TEX_CHECKER

    def __init__(self, even, odd, scale):
        self.even, self.odd, self.scale = v3(even), v3(odd), float(scale)

    def slots(self, image_index=0, perlin_index=0):
        e, o = self.even, self.odd
        return {0: e[0], 1: e[1], 2: e[2], 3: self.scale, 4: o[0], 5: o[1], 6: o[2]}


class ImageTexture(Texture):
    """Texture::Image (src/texture.rs:35-44): nearest texel of an 8-bit sRGB
    image (H x W x 3, row 0 at the top), decoded to linear."""
    tex = F.TEX_IMAGE

    def __init__(self, rgb):
        a = np.ascontiguousarray(rgb, dtype=np.uint8)
        if a.ndim != 3 or a.shape[2] != 3 or a.shape[0] == 0 or a.shape[1] == 0:
            raise ValueError("image texture must be H x W x 3 uint8")
        self.image = a

    def slots(self, image_index=0, perlin_index=0):
        return {7: float(image_index)}


class NoiseTexture(Texture):
    """Texture::Noise (src/texture.rs:45-48): grey
    0.5 (1 + sin(scale p.z + 10 turb(p, 7)))."""
    tex = F.TEX_NOISE

    def __init__(self, scale, seed=0, tables=None):
        self.scale = float(scale)
        self.perlin = perlin_tables(seed) if tables is None else np.ascontiguousarray(tables, dtype=np.float32)

    def slots(self, image_index=0, perlin_index=0):
        return {3: self.scale, 8: float(perlin_index)}


class MottleTexture(Texture):
    """Texture::Mottle (src/texture.rs:49-52): lerp(lo, hi,
    clamp(2 turb(scale p, 7), 0, 1))."""
    tex = F.TEX_MOTTLE

    def __init__(self, scale, lo, hi, seed=0, tables=None):
        self.scale, self.lo, self.hi = float(scale), v3(lo), v3(hi)
        self.perlin = perlin_tables(seed) if tables is None else np.ascontiguousarray(tables, dtype=np.float32)

    def slots(self, image_index=0, perlin_index=0):
        lo, hi = self.lo, self.hi
        return {0: lo[0], 1: lo[1], 2: lo[2], 3: self.scale, 4: hi[0], 5: hi[1], 6: hi[2], 8: float(perlin_index)}


def as_texture(t):
    return t if isinstance(t, Texture) else SolidTexture(t)


# ---------------------------------------------------------------- materials
class Material:
    kind = 0
    tex = 0
    texture = None

    def params(self):
        return []

    def to_c(self, image_index=0, perlin_index=0):
        m = F.rt_material()
        m.kind = self.kind
        m.tex = self.texture.tex if self.texture is not None else self.tex
        for i, v in enumerate(self.params()):
            m.p[i] = float(v)
        if self.texture is not None:
            for i, v in self.texture.slots(image_index, perlin_index).items():
                m.p[i] = float(v)
        return m


class Lambertian(Material):
    """src/material.rs:85-102: albedo from a texture (a colour is a Solid one)."""
    kind = F.MAT_LAMBERTIAN

    def __init__(self, color):
        self.texture = as_texture(color)

    @property
    def color(self):
        return getattr(self.texture, "color", None)


class Checker(Lambertian):
    """Lambertian over Texture::Checker (src/texture.rs:28-33, src/scene_file.rs:1146-1147)."""

    def __init__(self, even, odd, scale):
        self.texture = CheckerTexture(even, odd, scale)


class Dielectric(Material):
    """src/material.rs:323-332."""
    kind = F.MAT_DIELECTRIC

    def __init__(self, ior):
        self.ior = float(ior)

    def params(self):
        return [self.ior]


class SpectralDielectric(Material):
    """src/material.rs:357-403.  `sellmeier=(B1,B2,B3,C1,C2,C3)` switches the
    Cauchy n(lambda) for a Sellmeier law (extension)."""
    kind = F.MAT_SPECTRAL_DIELECTRIC

    def __init__(self, cauchy_b=1.5, cauchy_c=0.0, absorption=(0, 0, 0), sellmeier=None):
        self.cauchy_b, self.cauchy_c = float(cauchy_b), float(cauchy_c)
        self.absorption = v3(absorption)
        self.sellmeier = sellmeier

    @staticmethod
    def from_ior_dispersion(ior, dispersion, absorption=(0, 0, 0)):
        """MaterialConfig::SpectralDielectric -> Cauchy (src/scene_file.rs:1166-1180)."""
        c = f32(dispersion) / f32(3.250)
        b = f32(ior) - c / (f32(0.58756) * f32(0.58756))
        return SpectralDielectric(float(b), float(c), absorption)

    def params(self):
        p = [self.cauchy_b, self.cauchy_c] + list(self.absorption)
        if self.sellmeier is None:
            return p + [0.0]
        return p + [1.0] + [float(x) for x in self.sellmeier]


class PbrMaterial(Material):
    """PbrMaterial (src/material.rs:695-1050) with its thin-film clearcoat tint
    and optional image maps (H x W x 3 uint8, row 0 at the top): albedo (sRGB),
    roughness / metallic / ambient occlusion (linear, red channel) and a
    tangent-space normal map (used on meshes, which carry tangents)."""
    kind = F.MAT_PBR
    MAP_NAMES = ("albedo_map", "roughness_map", "metallic_map", "ao_map", "normal_map")

    def __init__(self, albedo=(0.8, 0.8, 0.8), roughness=0.5, metallic=0.0, anisotropy=0.0,
                 anisotropy_angle=0.0, clearcoat=0.0, clearcoat_roughness=0.03, emission=(0, 0, 0),
                 emission_strength=0.0, sheen=0.0, sheen_tint=0.5, film_thickness=0.0, film_ior=1.5,
                 albedo_map=None, roughness_map=None, metallic_map=None, ao_map=None, normal_map=None):
        self.albedo = v3(albedo)
        self.roughness, self.metallic = float(roughness), float(metallic)
        self.anisotropy, self.anisotropy_angle = float(anisotropy), float(anisotropy_angle)
        self.clearcoat, self.clearcoat_roughness = float(clearcoat), float(clearcoat_roughness)
        self.emission, self.emission_strength = v3(emission), float(emission_strength)
        self.sheen, self.sheen_tint = float(sheen), float(sheen_tint)
        self.film_thickness, self.film_ior = float(film_thickness), float(film_ior)
        self.maps = [None if a is None else ImageTexture(a).image
                     for a in (albedo_map, roughness_map, metallic_map, ao_map, normal_map)]
        self.map_index = [0] * 5      # descriptor image index + 1, set by Scene.descriptor

    def params(self):
        return (list(self.albedo) + [self.roughness, self.metallic, self.anisotropy, self.anisotropy_angle,
                                     self.clearcoat, self.clearcoat_roughness] + list(self.emission)
                + [self.emission_strength, self.sheen, self.sheen_tint, self.film_thickness, self.film_ior]
                + [float(k) for k in self.map_index])


def Metal(color, fuzz):
    """MaterialConfig::Metal = PbrMaterial metallic 1 (src/scene_file.rs:1155-1156)."""
    return PbrMaterial(albedo=color, metallic=1.0, roughness=fuzz)


class DiffuseLight(Material):
    """src/material.rs:14-25: emission from a texture (a colour is a Solid one)."""
    kind = F.MAT_DIFFUSE_LIGHT

    def __init__(self, emit):
        self.texture = as_texture(emit)

    @property
    def emit(self):
        return getattr(self.texture, "color", None)


class SpectralMetal(Material):
    """src/material.rs:427-554; variant 'gold' | 'copper' | 'silver'."""
    kind = F.MAT_SPECTRAL_METAL

    def __init__(self, variant="gold", roughness=0.0):
        self.variant = {"gold": 0, "copper": 1, "silver": 2}[variant]
        self.roughness = float(roughness)

    def params(self):
        return [self.variant, self.roughness]


class RoughDielectric(Material):
    """GGX microfacet dielectric (extension; the reference's dielectrics are smooth)."""
    kind = F.MAT_ROUGH_DIELECTRIC

    def __init__(self, ior=1.5, roughness=0.15):
        self.ior, self.roughness = float(ior), float(roughness)

    def params(self):
        return [self.ior, self.roughness]


class BlackbodyLight(Material):
    """BlackbodyLight::new (src/material.rs:49-61): Planck emitter at
    `temp_k`, scaled by `intensity`."""
    kind = F.MAT_BLACKBODY

    def __init__(self, temp_k, intensity=1.0):
        self.temp_k, self.intensity = float(temp_k), float(intensity)

    def params(self):
        return [self.temp_k, self.intensity]


class SSSMaterial(Material):
    """SSSMaterial (src/material.rs:576-631): dielectric boundary of `ior`
    around a scattering interior (exponential free flight of `density`, HG
    phase `g`, single-scatter `albedo`, Beer-Lambert `sigma_a`)."""
    kind = F.MAT_SSS

    def __init__(self, albedo, sigma_a, ior=1.3, density=1.0, g=0.0):
        self.albedo, self.sigma_a = v3(albedo), v3(sigma_a)
        self.ior, self.density, self.g = float(ior), float(density), float(g)

    def params(self):
        return list(self.albedo) + list(self.sigma_a) + [self.ior, self.density, self.g]


class PearlMaterial(Material):
    """PearlMaterial (src/material.rs:1149-1287): GGX luster tinted by
    thin-film interference over a Lambertian body; defaults of the scene
    file (src/scene_file.rs:460-465)."""
    kind = F.MAT_PEARL

    def __init__(self, base_color=(0.98, 0.93, 0.88), ior=1.56, film_thickness=450.0, orient_strength=0.30,
                 film_scale=3.0, luster_roughness=0.05):
        self.base_color = v3(base_color)
        self.ior, self.film_thickness, self.orient_strength = float(ior), float(film_thickness), float(orient_strength)
        self.film_scale, self.luster_roughness = float(film_scale), float(luster_roughness)

    def params(self):
        return list(self.base_color) + [self.ior, self.film_thickness, self.orient_strength, self.film_scale,
                                        self.luster_roughness]


class RoughSpectralDielectric(Material):
    """Rough dispersive dielectric (extension): n(l) = n_d + s(l) - s(589.3 nm)
    with s the Sellmeier law (B, C in um^2); GGX alpha `alpha_table` on
    primitive 0 (a Diamond's table facet) and `alpha_groups[(k-1) % 16]` on
    primitive k >= 1; the spectral dielectric's fire tint on the first
    refraction.  Differentiable in n_d, B and every alpha."""
    kind = F.MAT_ROUGH_SPECTRAL
    FIELD_ND, FIELD_B1, FIELD_ALPHA_TABLE, FIELD_ALPHA_GROUP0 = 0, 1, 7, 8

    def __init__(self, nd=2.417, sellmeier=(0.3306, 4.3356, 0.0, 0.175 ** 2, 0.106 ** 2, 0.0), alpha_table=0.02,
                 alpha_groups=None):
        self.nd = float(nd)
        self.sellmeier = [float(x) for x in sellmeier]
        self.alpha_table = float(alpha_table)
        self.alpha_groups = [float(x) for x in (alpha_groups if alpha_groups is not None else [alpha_table] * 16)]
        assert len(self.alpha_groups) == 16

    def params(self):
        return [self.nd] + self.sellmeier + [self.alpha_table] + self.alpha_groups


# ------------------------------------------------------------------- shapes
class Hittable:
    shape = 0

    def __init__(self, mat):
        self.mat = mat
        self.xforms = []      # innermost first
        self.medium = None

    # SRT wrappers in scene_file order: scale -> rotate -> translate
    def scaled(self, factor):
        assert factor > 0, "UniformScale factor must be positive"
        self.xforms.append((F.XFORM_SCALE, [factor]))
        return self

    def rotated(self, angles_deg, pivot=(0.0, 0.0, 0.0)):
        """Rotate::euler_about (src/transform.rs:196-224): fwd = Ry*Rx*Rz."""
        rx, ry, rz = [f32(a) for a in angles_deg]
        if rx == 0 and ry == 0 and rz == 0:
            return self
        pv = v3(pivot)
        m = euler_matrix(rx, ry, rz)
        self.xforms.append((F.XFORM_TRANSLATE, list(-pv)))
        self.xforms.append((F.XFORM_ROTATE, list(m.reshape(9))))
        self.xforms.append((F.XFORM_TRANSLATE, list(pv)))
        return self

    def translated(self, offset):
        self.xforms.append((F.XFORM_TRANSLATE, list(v3(offset))))
        return self

    def inline(self):
        return []


def euler_matrix(rx_deg, ry_deg, rz_deg):
    def sc(d):
        r = f32(d) * f32(math.pi / 180.0)
        return f32(math.sin(r)), f32(math.cos(r))
    sx, cx = sc(rx_deg)
    sy, cy = sc(ry_deg)
    sz, cz = sc(rz_deg)
    Rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]], dtype=np.float32)
    Ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]], dtype=np.float32)
    Rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]], dtype=np.float32)

    def mm(a, b):
        out = np.zeros((3, 3), dtype=np.float32)
        for i in range(3):
            for j in range(3):
                out[i, j] = f32(f32(a[i, 0] * b[0, j]) + f32(a[i, 1] * b[1, j])) + f32(a[i, 2] * b[2, j])
        return out
    return mm(Ry, mm(Rx, Rz))


class Sphere(Hittable):
    shape = F.SHAPE_SPHERE

    def __init__(self, center, radius, mat, center1=None):
        super().__init__(mat)
        self.center, self.radius = v3(center), float(radius)
        self.center1 = None if center1 is None else v3(center1)

    def inline(self):
        c1 = self.center1 if self.center1 is not None else self.center
        return list(self.center) + [self.radius] + list(c1) + [1.0 if self.center1 is not None else 0.0]


class Quad(Hittable):
    shape = F.SHAPE_QUAD

    def __init__(self, q, u, v, mat):
        super().__init__(mat)
        self.q, self.u, self.v = v3(q), v3(u), v3(v)
        n = cross(self.u, self.v)
        area = float(np.sqrt(np.dot(n, n)))
        if not area > 1e-6 * float(np.linalg.norm(self.u)) * float(np.linalg.norm(self.v)):
            raise ValueError(f"Quad::new: degenerate quad — u and v are parallel or zero (cross length {area})")

    def inline(self):
        return list(self.q) + list(self.u) + list(self.v)


class Cylinder(Hittable):
    """Cylinder (src/cylinder.rs:124-130): y axis, closed caps, `center` the
    centre of the bottom cap."""
    shape = F.SHAPE_CYLINDER

    def __init__(self, center, radius, height, mat):
        super().__init__(mat)
        self.center, self.radius, self.height = v3(center), float(radius), float(height)

    def inline(self):
        return list(self.center) + [self.radius, self.height]


class Cone(Hittable):
    """Cone (src/cone.rs:11-16): base disk at `center`, apex `height` above."""
    shape = F.SHAPE_CONE

    def __init__(self, center, radius, height, mat):
        super().__init__(mat)
        self.center, self.radius, self.height = v3(center), float(radius), float(height)

    def inline(self):
        return list(self.center) + [self.radius, self.height]


class Disk(Hittable):
    """Disk::new (src/disk.rs:115-124); the normal is normalised on upload."""
    shape = F.SHAPE_DISK

    def __init__(self, center, normal, radius, mat):
        super().__init__(mat)
        self.center, self.normal, self.radius = v3(center), v3(normal), float(radius)
        if not float(np.dot(self.normal, self.normal)) > 1e-12:
            raise ValueError("Disk::new: zero normal would give NaN hits")

    def inline(self):
        return list(self.center) + list(self.normal) + [self.radius]


class InfinitePlane(Hittable):
    """InfinitePlane::new (src/plane.rs:24-37): point, normal (normalised on
    upload), optional ripple (wave amplitude, wave scale)."""
    shape = F.SHAPE_PLANE

    def __init__(self, point, normal, mat, wave_amplitude=0.0, wave_scale=1.0):
        super().__init__(mat)
        self.point, self.normal = v3(point), v3(normal)
        self.wave_amplitude, self.wave_scale = float(wave_amplitude), float(wave_scale)
        if not float(np.dot(self.normal, self.normal)) > 1e-12:
            raise ValueError("InfinitePlane::new: zero normal would give NaN hits")

    def inline(self):
        return list(self.point) + list(self.normal) + [self.wave_amplitude, self.wave_scale]


class Box(Hittable):
    """make_box (src/quad.rs:201-215): six quads in a nested BVH."""
    shape = F.SHAPE_BOX

    def __init__(self, p0, p1, mat):
        super().__init__(mat)
        self.p0, self.p1 = v3(p0), v3(p1)

    def inline(self):
        return list(self.p0) + list(self.p1)


class Mesh(Hittable):
    """A triangle list under its own BVH, as mesh::load_obj builds it
    (src/mesh.rs:13-92).  vertices: (T,3,3); normals/tangents (T,3,3) or None;
    uvs (T,3,2) or None."""
    shape = F.SHAPE_MESH

    def __init__(self, vertices, mat, normals=None, tangents=None, uvs=None):
        super().__init__(mat)
        self.v = np.ascontiguousarray(vertices, dtype=np.float32).reshape(-1, 9)
        self.n = None if normals is None else np.ascontiguousarray(normals, dtype=np.float32).reshape(-1, 9)
        self.t = None if tangents is None else np.ascontiguousarray(tangents, dtype=np.float32).reshape(-1, 9)
        self.uv = None if uvs is None else np.ascontiguousarray(uvs, dtype=np.float32).reshape(-1, 6)

    @property
    def num_tris(self):
        return self.v.shape[0]


class Convex(Hittable):
    """Convex polyhedron as half-spaces n.x <= d (convex_hit, src/diamond.rs:203-253)."""
    shape = F.SHAPE_CONVEX

    def __init__(self, planes, mat, center=(0, 0, 0), crown_h=1.0, pavilion_h=1.0, girdle_h=0.0,
                 polish=0.0, bbox=None):
        super().__init__(mat)
        self.planes = np.ascontiguousarray(planes, dtype=np.float32).reshape(-1, 4)
        self.center = v3(center)
        self.crown_h, self.pavilion_h, self.girdle_h, self.polish = float(crown_h), float(pavilion_h), float(girdle_h), float(polish)
        self.bbox = bbox if bbox is not None else aabb_from_planes(self.planes)

    def inline(self):
        return (list(self.center) + [self.crown_h, self.pavilion_h, self.girdle_h, self.polish]
                + list(self.bbox[0]) + list(self.bbox[1]))


def aabb_from_planes(planes):
    """aabb_from_planes (src/hittable.rs:231-255), vectorised in float64 and
    rounded outward to float32."""
    P = np.asarray(planes, dtype=np.float64)
    n, d = P[:, :3], P[:, 3]
    eps = 1e-5 * (1.0 + np.abs(d).max())
    lo = np.full(3, np.inf)
    hi = np.full(3, -np.inf)
    m = len(P)
    for i in range(m):
        for j in range(i + 1, m):
            bc = np.cross(n[j], n[j + 1:])
            det = bc @ n[i]
            ok = np.abs(det) > 1e-12
            if not ok.any():
                continue
            k_idx = np.arange(j + 1, m)[ok]
            A = np.stack([np.broadcast_to(n[i], (len(k_idx), 3)), np.broadcast_to(n[j], (len(k_idx), 3)), n[k_idx]], axis=1)
            rhs = np.stack([np.full(len(k_idx), d[i]), np.full(len(k_idx), d[j]), d[k_idx]], axis=1)
            pts = np.linalg.solve(A, rhs[..., None])[..., 0]
            inside = (pts @ n.T <= d + eps).all(axis=1)
            if inside.any():
                lo = np.minimum(lo, pts[inside].min(axis=0))
                hi = np.maximum(hi, pts[inside].max(axis=0))
    lo32 = np.nextafter(lo.astype(np.float32), np.float32(-np.inf))
    hi32 = np.nextafter(hi.astype(np.float32), np.float32(np.inf))
    return lo32, hi32


def Diamond(center, radius, mat, table_pct=53.0, crown_angle=34.5, pavilion_angle=40.75, oval_ratio=1.0,
            girdle_facets=16, girdle_pct=2.0, faceted=False, polish=0.0):
    """Round / oval brilliant as half-spaces, after Diamond::new (src/diamond.rs:93-195)."""
    c = v3(center)
    r = f32(radius)
    q = max(f32(oval_ratio), f32(0.05))
    table_r = f32(table_pct / 100.0) * r
    crown_h = (r - table_r) * f32(math.tan(float(f32(crown_angle) * f32(math.pi / 180))))
    pavilion_h = r * f32(math.tan(float(f32(pavilion_angle) * f32(math.pi / 180))))
    girdle_h = min(max(f32(girdle_pct / 100.0) * f32(2.0) * r, f32(0.0)), f32(0.3) * r)
    n = (max(int(girdle_facets), 4) // 2) * 2
    m = n // 2
    band_m = n if faceted else 64
    planes = []

    def push(nrm, g):
        nw = unit(np.array([nrm[0] / q, nrm[1], nrm[2]], dtype=np.float32))
        gw = np.array([q * g[0], g[1], g[2]], dtype=np.float32)
        planes.append(list(nw) + [float(np.dot(nw, gw))])

    push(np.array([0, 1, 0], np.float32), np.array([c[0], c[1] + f32(0.5) * girdle_h + crown_h, c[2]], np.float32))
    wavy = girdle_h > 0 and faceted

    def band_h(k):
        return f32(0.0) if (wavy and k % 2 == 1) else girdle_h
    crown_n = n if girdle_h > 0 else m
    for k in range(crown_n):
        th = f32(k) * f32(2.0 * math.pi) / f32(crown_n)
        s, co = f32(math.sin(th)), f32(math.cos(th))
        push(unit(np.array([co * crown_h, r - table_r, s * crown_h], np.float32)),
             np.array([c[0] + co * r, c[1] + f32(0.5) * band_h(k), c[2] + s * r], np.float32))
    pav_n = n if girdle_h > 0 else m
    for k in range(pav_n):
        if girdle_h > 0:
            ph = f32(k) * f32(2.0 * math.pi) / f32(n)
        else:
            ph = f32(k) * f32(2.0 * math.pi) / f32(m) + f32(math.pi) / f32(m)
        s, co = f32(math.sin(ph)), f32(math.cos(ph))
        push(unit(np.array([co * pavilion_h, -r, s * pavilion_h], np.float32)),
             np.array([c[0] + co * r, c[1] - f32(0.5) * band_h(k), c[2] + s * r], np.float32))
    if girdle_h > 0:
        for j in range(band_m):
            a = f32(j) * f32(2.0 * math.pi) / f32(band_m)
            s, co = f32(math.sin(a)), f32(math.cos(a))
            push(np.array([co, 0, s], np.float32), np.array([c[0] + co * r, c[1], c[2] + s * r], np.float32))
    P = np.array(planes, dtype=np.float32)
    if girdle_h > 0:
        circ_r = r / f32(math.cos(math.pi / band_m))
        bb = (c - np.array([circ_r * q, f32(0.5) * girdle_h + pavilion_h, circ_r], np.float32),
              c + np.array([circ_r * q, f32(0.5) * girdle_h + crown_h, circ_r], np.float32))
    else:
        bb = aabb_from_planes(P)
    pol = min(max(float(polish), 0.0), 0.25) if math.isfinite(polish) else 0.0
    return Convex(P, mat, center=c, crown_h=crown_h, pavilion_h=pavilion_h, girdle_h=girdle_h, polish=pol,
                  bbox=(np.asarray(bb[0], np.float32), np.asarray(bb[1], np.float32)))


def step_outline(kind, a, b, corner_pct=12.0, segs=4, pear=False, points=None):
    """StepOutline::vertices (src/step_cut.rs:37-108): the girdle outline in
    XZ relative to the centre, from the semi-extents a (X) and b (Z), sorted
    by angle about the vertex centroid (convex outlines come out CCW).
    kind: "rect" (45-degree corner chamfers of corner_pct % of min(a, b); 0:
    sharp), "rounded_rect" (quarter arcs of `segs` segments), "triangle",
    "hexagon", "pointed_oval" (|cos|^p pinched ellipse, 2 * max(segs, 4)
    vertices; `pear` keeps the -X end round) or "polygon" (`points`)."""
    a, b = f32(a), f32(b)
    mn = min(a, b)
    if kind == "rect":
        c = min(f32(corner_pct) / f32(100.0) * mn, f32(0.45) * mn)
        if c <= f32(1e-6):
            v = [(a, b), (a, -b), (-a, -b), (-a, b)]
        else:
            v = [(a - c, b), (a, b - c), (a, -(b - c)), (a - c, -b),
                 (-(a - c), -b), (-a, -(b - c)), (-a, b - c), (-(a - c), b)]
    elif kind == "rounded_rect":
        r = min(f32(corner_pct) / f32(100.0) * mn, f32(0.45) * mn)
        sg = max(int(segs), 1)
        v = []
        for sx in (f32(1.0), f32(-1.0)):
            for sz in (f32(1.0), f32(-1.0)):
                cx, cz = sx * (a - r), sz * (b - r)
                for j in range(sg + 1):
                    ang = f32(j) / f32(sg) * f32(math.pi / 2)
                    v.append((cx + sx * r * f32(math.cos(ang)), cz + sz * r * f32(math.sin(ang))))
    elif kind == "triangle":
        v = []
        for k in range(3):
            th = f32(math.pi / 2) + f32(k) * f32(2.0) * f32(math.pi) / f32(3.0)
            v.append((a * f32(math.cos(th)), b * f32(math.sin(th))))
    elif kind == "hexagon":
        v = []
        for k in range(6):
            th = f32(k) * f32(math.pi) / f32(3.0)
            v.append((a * f32(math.cos(th)), b * f32(math.sin(th))))
    elif kind == "pointed_oval":
        n = max(int(segs), 4) * 2
        v = []
        for i in range(n):
            th = f32(2.0) * f32(math.pi) * f32(i) / f32(n)
            c, sn = f32(math.cos(th)), f32(math.sin(th))
            pw = (f32(0.6) if c >= 0 else f32(1.0)) if pear else f32(0.65)
            sgn = f32(0.0) if c == 0 else f32(math.copysign(1.0, c))
            v.append((a * sgn * f32(abs(c) ** pw), b * sn))
    elif kind == "polygon":
        v = [(f32(x), f32(z)) for x, z in points]
    else:
        raise ValueError(f"unknown step outline {kind!r}")
    cx = f32(sum(float(q[0]) for q in v) / len(v))
    cz = f32(sum(float(q[1]) for q in v) / len(v))
    return sorted(v, key=lambda q: math.atan2(float(q[1] - cz), float(q[0] - cx)))


def _check_outline(v, scale):
    """validate_outline (src/step_cut.rs:192-232): drop repeated vertices
    (wrap-around included), then require three or more vertices, no short
    edge, non-zero area and a convex turn at every vertex."""
    eps = 1e-6 * float(scale)
    out = []
    for q in v:
        if out and math.hypot(float(q[0] - out[-1][0]), float(q[1] - out[-1][1])) < eps:
            continue
        out.append(q)
    if len(out) > 1 and math.hypot(float(out[0][0] - out[-1][0]), float(out[0][1] - out[-1][1])) < eps:
        out.pop()
    n = len(out)
    if n < 3:
        raise ValueError("StepCut: outline needs at least 3 distinct vertices")
    for i in range(n):
        p, q = out[i], out[(i + 1) % n]
        if math.hypot(float(p[0] - q[0]), float(p[1] - q[1])) < eps:
            raise ValueError("StepCut: outline has a zero-length edge")
    area2 = sum(float(out[i][0]) * float(out[(i + 1) % n][1]) - float(out[(i + 1) % n][0]) * float(out[i][1])
                for i in range(n))
    if abs(area2) < eps * eps:
        raise ValueError("StepCut: outline has zero area")
    for i in range(n):
        p, q, r = out[i], out[(i + 1) % n], out[(i + 2) % n]
        turn = float(q[0] - p[0]) * float(r[1] - q[1]) - float(q[1] - p[1]) * float(r[0] - q[0])
        if turn * area2 < 0:
            raise ValueError("StepCut: polygon outline must be convex")
    return out


def StepCut(center, half_a, half_b, mat, table_pct=65.0, crown_angle=35.0, pavilion_angle=42.0,
            outline=("rect", {"corner_pct": 12.0}), crown_steps=3, pavilion_steps=3, culet_pct=8.0,
            step_spread=0.25, chevrons=0, girdle_pct=0.0, polish=0.0):
    """Step-family cut (emerald, princess, cushion, trillion, marquise ...) as
    half-spaces, after StepCut::new (src/step_cut.rs:236-384); its hit is the
    convex hit with the same cylindrical uv as Diamond (src/step_cut.rs:
    388-405).  Per outline edge (outward normal m, offset h): an optional
    vertical girdle band face, `crown_steps` crown rows and `pavilion_steps`
    pavilion rows of equal rises whose runs follow row angles spread from
    (1 + step_spread) x to (1 - step_spread) x the nominal tangent, scaled so
    the runs add up to the table (culet) inset; `chevrons` diagonal rows per
    outline vertex toward table and culet; then the table and culet caps.
    The box is exact from the outline and the cap heights.  `outline` is
    (kind, keywords of step_outline)."""
    a, b = f32(half_a), f32(half_b)
    if not (a > 0 and b > 0 and math.isfinite(a) and math.isfinite(b)):
        raise ValueError("StepCut: half extents must be positive and finite")
    c = v3(center)
    deg = f32(math.pi) / f32(180.0)
    k = min(max(f32(table_pct) / f32(100.0), f32(0.05)), f32(0.95))
    kc = min(max(f32(culet_pct) / f32(100.0), f32(0.0)), f32(0.5))
    s_in = b * (f32(1.0) - k)
    sp_in = b * (f32(1.0) - kc)
    tan_c = f32(math.tan(float(f32(crown_angle) * deg)))
    tan_p = f32(math.tan(float(f32(pavilion_angle) * deg)))
    crown_h = s_in * tan_c
    pavilion_h = sp_in * tan_p
    girdle_h = min(max(f32(girdle_pct) / f32(100.0) * f32(2.0) * b, f32(0.0)), f32(0.3) * b)
    if not (math.isfinite(girdle_pct) and 0.0 <= girdle_pct < 100.0):
        raise ValueError(f"StepCut: girdle_pct must be in [0, 100), got {girdle_pct}")
    if not (crown_h > 0 and pavilion_h > 0):
        raise ValueError(f"StepCut: crown/pavilion angles must be in (0, 90) degrees, got {crown_angle} / {pavilion_angle}")
    kind, kw = outline if isinstance(outline, tuple) else (outline, {})
    verts = _check_outline(step_outline(kind, a, b, **kw), max(a, b))
    nv = len(verts)
    gx = f32(sum(float(q[0]) for q in verts) / nv)
    gz = f32(sum(float(q[1]) for q in verts) / nv)

    def flat_unit(x, z):
        ln = f32(math.sqrt(float(x * x + z * z)))
        return (x / ln, z / ln) if ln > 1e-12 else (f32(0.0), f32(0.0))
    edges = []
    for i in range(nv):
        p, q = verts[i], verts[(i + 1) % nv]
        mx, mz = flat_unit(q[1] - p[1], -(q[0] - p[0]))
        if mx * (p[0] - gx) + mz * (p[1] - gz) < 0:
            mx, mz = -mx, -mz
        edges.append((mx, mz, mx * p[0] + mz * p[1]))
    cs, ps = max(int(crown_steps), 1), max(int(pavilion_steps), 1)
    spread = min(max(f32(step_spread), f32(0.0)), f32(0.9))

    def factor(i, rows):
        return f32(1.0) if rows <= 1 else f32(1.0) + spread * (f32(1.0) - f32(2.0) * f32(i) / f32(rows - 1))
    dv, dvp = crown_h / f32(cs), pavilion_h / f32(ps)
    norm_c = s_in / f32(sum(float(dv / (tan_c * factor(i, cs))) for i in range(cs)))
    norm_p = sp_in / f32(sum(float(dvp / (tan_p * factor(i, ps))) for i in range(ps)))
    planes = []

    def add(n, pt):
        n = np.asarray(n, np.float32)
        planes.append([n[0], n[1], n[2], f32(np.dot(n, np.asarray(pt, np.float32)))])
    for mx, mz, h in edges:
        if girdle_h > 0:
            planes.append([mx, f32(0.0), mz, mx * c[0] + mz * c[2] + h])
        for rows, tanv, dvv, nrm, sgn in ((cs, tan_c, dv, norm_c, f32(1.0)), (ps, tan_p, dvp, norm_p, f32(-1.0))):
            dist, y = h, sgn * f32(0.5) * girdle_h
            for i in range(rows):
                dh = dvv / (tanv * factor(i, rows)) * nrm
                add(unit([mx * dvv, sgn * dh, mz * dvv]), [c[0] + mx * dist, c[1] + y, c[2] + mz * dist])
                dist = dist - dh
                y = y + sgn * dvv
    for i in range(nv):
        v, vp, vn = verts[i], verts[(i + nv - 1) % nv], verts[(i + 1) % nv]
        u1, u2 = flat_unit(vp[0] - v[0], vp[1] - v[1]), flat_unit(vn[0] - v[0], vn[1] - v[1])
        bis = (u1[0] + u2[0], u1[1] + u2[1])
        if bis[0] * bis[0] + bis[1] * bis[1] < 1e-12:
            continue
        c3 = np.array([c[0] + v[0], c[1], c[2] + v[1]], np.float32)
        d3 = np.array([-bis[1], 0.0, bis[0]], np.float32)
        tc = np.array([c[0] + v[0] * k, c[1] + f32(0.5) * girdle_h + crown_h, c[2] + v[1] * k], np.float32)
        cc = np.array([c[0] + v[0] * kc, c[1] - f32(0.5) * girdle_h - pavilion_h, c[2] + v[1] * kc], np.float32)
        for j in range(int(chevrons)):
            fj = f32(j + 1) / f32(int(chevrons))
            for pt in (c3 + (tc - c3) * fj, c3 + (cc - c3) * fj):
                n = unit(cross(d3, pt - c3))
                if np.dot(n, c) > np.dot(n, c3):
                    n = -n
                add(n, c3)
    planes.append([f32(0.0), f32(1.0), f32(0.0), c[1] + f32(0.5) * girdle_h + crown_h])
    planes.append([f32(0.0), f32(-1.0), f32(0.0), -(c[1] - f32(0.5) * girdle_h - pavilion_h)])
    top = f32(0.5) * girdle_h + s_in * tan_c
    bottom = f32(0.5) * girdle_h + sp_in * tan_p
    xs = [q[0] for q in verts]
    zs = [q[1] for q in verts]
    bb = (np.array([c[0] + min(xs), c[1] - bottom, c[2] + min(zs)], np.float32),
          np.array([c[0] + max(xs), c[1] + top, c[2] + max(zs)], np.float32))
    pol = min(max(float(polish), 0.0), 0.25) if math.isfinite(polish) else 0.0
    return Convex(np.array(planes, np.float32), mat, center=c, crown_h=crown_h, pavilion_h=pavilion_h,
                  girdle_h=girdle_h, polish=pol, bbox=bb)


class ConstantMedium:
    """src/volume.rs:78-156: wraps a boundary shape (its material is ignored)."""

    def __init__(self, boundary, density, color, g=0.0):
        self.boundary = boundary
        boundary.medium = F.rt_medium(kind=F.MEDIUM_CONSTANT, density=float(density), g=float(g))
        for k in range(3):
            boundary.medium.albedo[k] = float(v3(color)[k])
        boundary.medium_grid = None


class GridMedium:
    """Delta-tracked heterogeneous medium over a dense density grid in [0,1]
    spanning the boundary's bounding box (NoiseMedium with a grid in place of
    per-lookup Perlin turbulence, src/volume.rs:160-263)."""

    def __init__(self, boundary, grid, density, color, g=0.0):
        self.boundary = boundary
        grid = np.ascontiguousarray(grid, dtype=np.float32)
        nz, ny, nx = grid.shape
        boundary.medium = F.rt_medium(kind=F.MEDIUM_GRID, density=float(density), g=float(g))
        for k in range(3):
            boundary.medium.albedo[k] = float(v3(color)[k])
        boundary.medium.grid_n[0], boundary.medium.grid_n[1], boundary.medium.grid_n[2] = nx, ny, nz
        boundary.medium_grid = grid


class NoiseMedium:
    """NoiseMedium::new (src/volume.rs:170-188): delta tracking with density
    max(turb(p * noise_scale, 7) / 2 - threshold, 0) * density.  The reference
    draws its Perlin tables (256 unit vectors, three permutations) from an
    unseeded thread_rng; here they come from `seed` (numpy), so a scene is
    reproducible and the checker sees the same tables."""

    def __init__(self, boundary, color, density, noise_scale=1.0, threshold=0.0, g=0.0, seed=0):
        self.boundary = boundary
        boundary.medium = F.rt_medium(kind=F.MEDIUM_NOISE, density=float(density), g=float(g),
                                      noise_scale=float(noise_scale), threshold=float(threshold))
        for k in range(3):
            boundary.medium.albedo[k] = float(v3(color)[k])
        boundary.medium_grid = None
        boundary.medium_perlin = perlin_tables(seed)


# --------------------------------------------------------------- background
class Background:
    """Background::Solid / Gradient / Physical / EnvMap / EnvMapSplit."""

    def __init__(self, kind=F.BG_SOLID, color=(0, 0, 0), top=(1, 1, 1), bottom=(0, 0, 0), sun_dir=(0, 1, 0),
                 turbidity=3.0, env=None, env_rot_y_deg=0.0, visible=None, sampling="cdf"):
        self.kind, self.color, self.top, self.bottom = kind, v3(color), v3(top), v3(bottom)
        self.sun_dir, self.turbidity = v3(sun_dir), float(turbidity)
        self.env = None if env is None else np.ascontiguousarray(env, dtype=np.float32)
        self.env_rot = float(f32(env_rot_y_deg) * f32(math.pi / 180))
        self.visible = visible
        assert sampling in ("cdf", "alias")
        self.sampling = sampling

    @staticmethod
    def solid(c):
        return Background(F.BG_SOLID, color=c)

    @staticmethod
    def gradient(top, bottom):
        return Background(F.BG_GRADIENT, top=top, bottom=bottom)

    @staticmethod
    def physical(sun_azimuth_deg, sun_elevation_deg, turbidity):
        el = f32(sun_elevation_deg) * f32(math.pi / 180)
        az = f32(sun_azimuth_deg) * f32(math.pi / 180)
        d = unit(np.array([f32(math.cos(el)) * f32(math.sin(az)), f32(math.sin(el)),
                           f32(math.cos(el)) * f32(math.cos(az))], np.float32))
        return Background(F.BG_PHYSICAL, sun_dir=d, turbidity=turbidity)

    @staticmethod
    def envmap(image_hw3, rotate_y_deg=0.0, visible=None, sampling="cdf"):
        """visible: None or Background.solid / Background.gradient (EnvMapSplit).
        sampling: "cdf" (the reference's row/column CDF searches) or "alias"
        (alias table over texels, same density; extension)."""
        kind = F.BG_ENVMAP if visible is None else F.BG_ENVMAP_SPLIT
        return Background(kind, env=image_hw3, env_rot_y_deg=rotate_y_deg, visible=visible, sampling=sampling)

    def to_c(self):
        b = F.rt_background()
        b.kind = self.kind
        src = self.visible if self.visible is not None else self
        b.visible_kind = src.kind if self.visible is not None else 0
        for k in range(3):
            b.color[k] = float(src.color[k])
            b.top[k] = float(src.top[k])
            b.bottom[k] = float(src.bottom[k])
            b.sun_dir[k] = float(self.sun_dir[k])
        b.turbidity = self.turbidity
        b.env_sampling = F.ENV_ALIAS if self.sampling == "alias" else F.ENV_CDF
        if self.env is not None:
            b.env_h, b.env_w = self.env.shape[0], self.env.shape[1]
            b.env_rgb = self.env.ctypes.data_as(F.F32P)
            b.env_rot_y = self.env_rot
        return b


# -------------------------------------------------------------------- camera
class Camera:
    """Camera::new (src/camera.rs:24-45) in f32."""

    def __init__(self, lookfrom, lookat, vup, vfov_deg, aspect, aperture=0.0, focus_dist=None, blades=0):
        lf, la, up = v3(lookfrom), v3(lookat), v3(vup)
        if focus_dist is None:
            focus_dist = max(float(np.linalg.norm(lf - la)), 0.01)
        fd = f32(focus_dist)
        h = f32(math.tan(float(f32(vfov_deg) * f32(math.pi / 180)) / 2.0))
        vh = f32(2.0) * h
        vw = f32(aspect) * vh
        w = unit(lf - la)
        u = unit(cross(up, w))
        v = cross(w, u)
        self.origin = lf
        self.horizontal = (fd * vw * u).astype(np.float32)
        self.vertical = (fd * vh * v).astype(np.float32)
        self.lower_left = (lf - self.horizontal / f32(2) - self.vertical / f32(2) - fd * w).astype(np.float32)
        self.u, self.v = u, v
        self.lens_radius = f32(aperture) / f32(2)
        self.blades = int(blades)

    def to_c(self):
        c = F.rt_camera()
        for k in range(3):
            c.origin[k] = float(self.origin[k]); c.lower_left[k] = float(self.lower_left[k])
            c.horizontal[k] = float(self.horizontal[k]); c.vertical[k] = float(self.vertical[k])
            c.u[k] = float(self.u[k]); c.v[k] = float(self.v[k])
        c.lens_radius = float(self.lens_radius)
        c.aperture_blades = self.blades
        return c


# --------------------------------------------------------------------- scene
class Scene:
    """World objects (input order = obj_id), the NEE light list and the
    background: what `render_tiles` receives as world / lights / background."""

    def __init__(self, objects, lights=(), background=None, name="scene"):
        self.objects = list(objects)
        self.lights = list(lights)
        self.background = background if background is not None else Background.solid((0, 0, 0))
        self.name = name
        self._desc = None

    def descriptor(self):
        """Flatten into rt_scene_desc; the returned object keeps every array alive."""
        mats, mat_index = [], {}
        xforms, media, pool, grids = [], [], [], []
        tri_v, tri_n, tri_t, tri_uv = [], [], [], []
        ntris = 0
        any_n = any(isinstance(o, Mesh) and o.n is not None for o in self.objects + self.lights)
        any_t = any(isinstance(o, Mesh) and o.t is not None for o in self.objects + self.lights)
        any_uv = any(isinstance(o, Mesh) and o.uv is not None for o in self.objects + self.lights)

        def mat_id(m):
            if m is None:
                m = Lambertian((0.5, 0.5, 0.5))
            k = id(m)
            if k not in mat_index:
                mat_index[k] = len(mats)
                mats.append(m)
            return mat_index[k]

        def flat(o):
            nonlocal ntris
            r = F.rt_object()
            r.shape = o.shape
            r.material = mat_id(o.mat)
            r.xform_first = len(xforms)
            r.xform_count = len(o.xforms)
            for kind, vals in o.xforms:
                x = F.rt_xform(kind=kind)
                for i, val in enumerate(vals):
                    x.v[i] = float(val)
                xforms.append(x)
            for i, val in enumerate(o.inline()):
                r.p[i] = float(val)
            if isinstance(o, Mesh):
                r.data_first, r.data_count = ntris, o.num_tris
                tri_v.append(o.v)
                if any_n:
                    tri_n.append(o.n if o.n is not None else _face_normals(o.v))
                if any_t:
                    tri_t.append(o.t if o.t is not None else _edge_tangents(o.v))
                if any_uv:
                    tri_uv.append(o.uv if o.uv is not None else np.tile(np.array([0, 0, 1, 0, 0.5, 1], np.float32), (o.num_tris, 1)))
                ntris += o.num_tris
            elif isinstance(o, Convex):
                r.data_first = sum(len(p) for p in pool) // 4
                r.data_count = o.planes.shape[0]
                pool.append(o.planes.reshape(-1))
            med = getattr(o, "medium", None)
            if med is not None:
                m = F.rt_medium()
                C.pointer(m)[0] = med
                g = getattr(o, "medium_grid", None)
                if g is not None:
                    grids.append(g)
                    m.grid = g.ctypes.data_as(F.F32P)
                pt = getattr(o, "medium_perlin", None)
                if pt is not None:
                    grids.append(pt)
                    m.perlin = pt.ctypes.data_as(F.F32P)
                media.append(m)
                r.medium = len(media)
            return r

        objs = [flat(o) for o in self.objects]
        lights = [flat(o) for o in self.lights]
        keep = {}
        d = F.rt_scene_desc()
        keep["objs"] = (F.rt_object * max(len(objs), 1))(*objs)
        d.objects, d.num_objects = C.cast(keep["objs"], C.POINTER(F.rt_object)), len(objs)
        keep["lights"] = (F.rt_object * max(len(lights), 1))(*lights)
        d.lights, d.num_lights = C.cast(keep["lights"], C.POINTER(F.rt_object)), len(lights)
        images, image_index, perlins, perlin_index = [], {}, [], {}
        cm = []
        for m in mats:
            for k, img in enumerate(getattr(m, "maps", ())):
                if img is not None:
                    if id(img) not in image_index:
                        image_index[id(img)] = len(images)
                        images.append(img)
                    m.map_index[k] = image_index[id(img)] + 1
            t = m.texture
            ii = pi = 0
            if t is not None and t.image is not None:
                if id(t) not in image_index:
                    image_index[id(t)] = len(images)
                    images.append(t.image)
                ii = image_index[id(t)]
            if t is not None and t.perlin is not None:
                if id(t) not in perlin_index:
                    perlin_index[id(t)] = len(perlins)
                    perlins.append(t.perlin)
                pi = perlin_index[id(t)]
            cm.append(m.to_c(ii, pi))
        if images:
            ims = [F.rt_image(width=a.shape[1], height=a.shape[0], rgb=a.ctypes.data_as(C.POINTER(C.c_uint8))) for a in images]
            keep["images"] = (F.rt_image * len(ims))(*ims)
            keep["image_data"] = images
            d.images, d.num_images = C.cast(keep["images"], C.POINTER(F.rt_image)), len(ims)
        if perlins:
            PT = np.ascontiguousarray(np.concatenate(perlins), dtype=np.float32)
            keep["perlin"] = PT
            d.perlin, d.num_perlin = PT.ctypes.data_as(F.F32P), len(perlins)
        keep["mats"] = (F.rt_material * max(len(cm), 1))(*cm)
        d.materials, d.num_materials = C.cast(keep["mats"], C.POINTER(F.rt_material)), len(cm)
        keep["xf"] = (F.rt_xform * max(len(xforms), 1))(*xforms)
        d.xforms, d.num_xforms = C.cast(keep["xf"], C.POINTER(F.rt_xform)), len(xforms)
        keep["media"] = (F.rt_medium * max(len(media), 1))(*media)
        d.media, d.num_media = C.cast(keep["media"], C.POINTER(F.rt_medium)), len(media)
        keep["grids"] = grids
        if pool:
            P = np.ascontiguousarray(np.concatenate(pool), dtype=np.float32)
            keep["pool"] = P
            d.pool, d.pool_len = P.ctypes.data_as(F.F32P), P.size
        if ntris:
            for name, lst, attr in (("tv", tri_v, "tri_v"), ("tn", tri_n, "tri_n"), ("tt", tri_t, "tri_t"), ("tuv", tri_uv, "tri_uv")):
                if lst:
                    A = np.ascontiguousarray(np.concatenate(lst), dtype=np.float32)
                    keep[name] = A
                    setattr(d, attr, A.ctypes.data_as(F.F32P))
            d.num_tris = ntris
        d.background = self.background.to_c()
        keep["bg_env"] = self.background.env
        self._desc = (d, keep)
        self._mat_index = dict(mat_index)
        return d

    def material_index(self, mat):
        """Index of `mat` in the descriptor's material table (for rt_param)."""
        if not hasattr(self, "_mat_index"):
            self.descriptor()
        return self._mat_index[id(mat)]


def _face_normals(v):
    v = v.reshape(-1, 3, 3).astype(np.float32)
    n = np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0])
    n = n / np.sqrt((n * n).sum(1, keepdims=True))
    return np.repeat(n[:, None, :], 3, axis=1).reshape(-1, 9).astype(np.float32)


def _edge_tangents(v):
    v = v.reshape(-1, 3, 3).astype(np.float32)
    t = v[:, 1] - v[:, 0]
    t = t / np.sqrt((t * t).sum(1, keepdims=True))
    return np.repeat(t[:, None, :], 3, axis=1).reshape(-1, 9).astype(np.float32)
