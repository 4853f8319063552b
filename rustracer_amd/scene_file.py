"""TOML scene files: the reference's `scene_file::build` (src/scene_file.rs:
899-1216) onto this package's scene classes, so a user of the reference can
render its `scenes/*.toml` through the device path.

`load(path)` reads the file with tomli and returns (Scene, camera, info):
`camera(aspect)` builds the Camera for an image aspect ratio (the reference
builds it from the window or output size), and `info` holds max_samples,
the caustics settings and the first light quad (the photon emitter of
`SceneData::rebuild_caustics`, src/scene.rs:35-74).

Covered: camera; physical / solid / env_map backgrounds (Radiance .hdr, with
an optional visible solid or gradient backdrop); every [[objects]] shape
(sphere, quad, box, cylinder, cone, disk, diamond, step_cut, infinite_plane,
mesh from OBJ, constant / noise volume sphere / box) with scale / rotate /
rotate_y / spin / translate; every material (lambertian, checker, noise,
image, metal, dielectric, spectral_dielectric, diffuse_light, pbr,
textured_pbr, spectral_metal, pearl, sss); [[gems]] grading reports
(`build_gem`, src/scene_file.rs:793-897: brilliant and every step-family cut
preset, diamond / sapphire / ruby / emerald and the checker / opaque debug
materials, colour-grade tint) with the `--spin` override of `load_with`
(src/scene_file.rs:625-648); [[lights]] quads (temp_k -> blackbody);
[caustics].  Emissive objects join the light list as the reference decides
(light_area() > 0, i.e. Material::is_emissive, src/scene_file.rs:1040-1045).
Not covered: camera animation ([animation] is only flagged in `info`).
Images load with Pillow.
"""
import math
import os

import numpy as np

from . import scene as S

f32 = np.float32


def _p3(v):
    return tuple(float(x) for x in v)


def _asset(base, path):
    return path if os.path.isabs(path) else os.path.join(base, path)


def load_hdr(path):
    """Radiance RGBE (.hdr) to float32 (H, W, 3), row 0 at the top: flat or
    new-style run-length scanlines, the exponent table of the format."""
    with open(path, "rb") as fh:
        data = fh.read()
    pos = 0
    while True:                                   # header lines until an empty line
        end = data.index(b"\n", pos)
        line = data[pos:end]
        pos = end + 1
        if not line.strip():
            break
        if line.startswith(b"FORMAT=") and b"32-bit_rle_rgbe" not in line:
            raise ValueError(f"{path}: unsupported HDR format {line!r}")
    end = data.index(b"\n", pos)
    dims = data[pos:end].split()
    pos = end + 1
    if len(dims) != 4 or dims[0] != b"-Y" or dims[2] != b"+X":
        raise ValueError(f"{path}: unsupported HDR orientation {dims!r}")
    h, w = int(dims[1]), int(dims[3])
    buf = np.frombuffer(data, np.uint8, offset=pos)
    out = np.zeros((h, w, 4), np.uint8)
    i = 0
    for y in range(h):
        if 8 <= w < 32768 and buf[i] == 2 and buf[i + 1] == 2 and (int(buf[i + 2]) << 8 | int(buf[i + 3])) == w:
            i += 4
            for c in range(4):
                x = 0
                while x < w:
                    n = int(buf[i])
                    i += 1
                    if n > 128:
                        n -= 128
                        out[y, x:x + n, c] = buf[i]
                        i += 1
                    else:
                        out[y, x:x + n, c] = buf[i:i + n]
                        i += n
                    x += n
        else:
            out[y] = buf[i:i + 4 * w].reshape(w, 4)
            i += 4 * w
    e = out[..., 3].astype(np.int32)
    scale = np.where(e > 0, np.ldexp(1.0, e - 136), 0.0)
    return (out[..., :3].astype(np.float64) * scale[..., None]).astype(np.float32)


def load_image(path):
    """An 8-bit RGB image (H, W, 3) uint8 via Pillow."""
    from PIL import Image
    with Image.open(path) as im:
        return np.asarray(im.convert("RGB"), dtype=np.uint8).copy()


def load_obj(path, mat):
    """crate::mesh::load_obj (src/mesh.rs:13-92): triangles of the OBJ's faces
    (polygons fanned from their first vertex), with the file's vertex normals
    and texture coordinates when every corner of a face has them."""
    v, vt, vn = [], [], []
    tris, nrm, uvs = [], [], []
    with open(path) as fh:
        for line in fh:
            t = line.split()
            if not t:
                continue
            if t[0] == "v":
                v.append([float(x) for x in t[1:4]])
            elif t[0] == "vt":
                vt.append([float(x) for x in t[1:3]] + [0.0] * max(0, 2 - len(t[1:3])))
            elif t[0] == "vn":
                vn.append([float(x) for x in t[1:4]])
            elif t[0] == "f":
                corners = []
                for c in t[1:]:
                    parts = (c.split("/") + ["", ""])[:3]
                    idx = [int(p) if p else 0 for p in parts]
                    fix = lambda k, n: k - 1 if k > 0 else (n + k if k < 0 else -1)
                    corners.append((fix(idx[0], len(v)), fix(idx[1], len(vt)), fix(idx[2], len(vn))))
                for k in range(1, len(corners) - 1):
                    tri = [corners[0], corners[k], corners[k + 1]]
                    tris.append([v[a] for a, _, _ in tri])
                    nrm.append([vn[c] for _, _, c in tri] if all(c >= 0 for _, _, c in tri) else None)
                    uvs.append([vt[b] for _, b, _ in tri] if all(b >= 0 for _, b, _ in tri) else None)
    if not tris:
        raise ValueError(f"{path}: no faces")
    T = np.asarray(tris, np.float32)
    N = np.asarray([n if n is not None else [[0, 0, 0]] * 3 for n in nrm], np.float32) if any(n is not None for n in nrm) else None
    U = np.asarray([u if u is not None else [[0, 0]] * 3 for u in uvs], np.float32) if any(u is not None for u in uvs) else None
    return S.Mesh(T, mat, normals=N, uvs=U)


def _material(m, base):
    kind = m["type"]
    g = m.get
    if kind == "lambertian":
        return S.Lambertian(_p3(m["color"]))
    if kind == "checker":
        return S.Checker(_p3(m["even"]), _p3(m["odd"]), g("scale", 1.0))
    if kind == "noise":
        return S.Lambertian(S.MottleTexture(g("scale", 3.0), _p3(m["lo"]), _p3(m["hi"]), seed=0x5EED))
    if kind == "image":
        return S.Lambertian(S.ImageTexture(load_image(_asset(base, m["path"]))))
    if kind == "metal":
        return S.Metal(_p3(m["color"]), g("fuzz", 0.0))
    if kind == "dielectric":
        return S.Dielectric(m["ior"])
    if kind == "spectral_dielectric":
        return S.SpectralDielectric.from_ior_dispersion(m["ior"], g("dispersion", 0.0), _p3(g("absorption", [0, 0, 0])))
    if kind == "diffuse_light":
        return S.DiffuseLight(_p3(m["color"]))
    if kind == "pbr":
        return S.PbrMaterial(albedo=_p3(m["albedo"]), metallic=g("metallic", 0.0), roughness=g("roughness", 0.5),
                             anisotropy=g("anisotropy", 0.0), anisotropy_angle=g("anisotropy_angle", 0.0),
                             clearcoat=g("clearcoat", 0.0), clearcoat_roughness=g("clearcoat_roughness", 0.03),
                             film_thickness=g("film_thickness", 0.0), film_ior=g("film_ior", 1.5),
                             sheen=g("sheen", 0.0), sheen_tint=g("sheen_tint", 0.5),
                             emission=_p3(g("emission", [0, 0, 0])), emission_strength=g("emission_strength", 0.0))
    if kind == "textured_pbr":
        opt = lambda key: load_image(_asset(base, m[key])) if m.get(key) else None
        return S.PbrMaterial(albedo=(1.0, 1.0, 1.0), roughness=g("roughness", 0.5), metallic=g("metallic", 0.0),
                             albedo_map=load_image(_asset(base, m["albedo_path"])), roughness_map=opt("roughness_path"),
                             metallic_map=opt("metallic_path"), ao_map=opt("ao_path"), normal_map=opt("normal_path"))
    if kind == "spectral_metal":
        variant = str(m["variant"]).lower()
        return S.SpectralMetal(variant if variant in ("copper", "silver") else "gold", g("roughness", 0.0))
    if kind == "pearl":
        return S.PearlMaterial(base_color=_p3(g("base_color", [0.98, 0.93, 0.88])), ior=g("ior", 1.56),
                               film_thickness=g("film_thickness", 450.0), orient_strength=g("orient_strength", 0.30),
                               film_scale=g("film_scale", 3.0), luster_roughness=g("luster_roughness", 0.05))
    if kind == "sss":
        return S.SSSMaterial(_p3(m["albedo"]), _p3(m["sigma_a"]), ior=g("ior", 1.3), density=g("density", 1.0), g=g("g", 0.0))
    raise ValueError(f"unknown material type {kind!r}")


_DUMMY = S.Lambertian((0.5, 0.5, 0.5))
_OUTLINES = {"rect": "rect", "rounded_rect": "rounded_rect", "cushion": "rounded_rect", "triangle": "triangle",
             "trillion": "triangle", "hexagon": "hexagon", "marquise": "pointed_oval", "pear": "pointed_oval",
             "polygon": "polygon"}


def _shape(sh, mat, base, where):
    kind = sh["type"]
    g = sh.get
    if kind in ("constant_volume_sphere", "constant_volume_box", "noise_volume_sphere", "noise_volume_box"):
        bnd = (S.Sphere(_p3(sh["center"]), sh["radius"], _DUMMY) if kind.endswith("sphere")
               else S.Box(_p3(sh["p_min"]), _p3(sh["p_max"]), _DUMMY))
        if kind.startswith("constant"):
            S.ConstantMedium(bnd, sh["density"], _p3(sh["color"]), g=g("g", 0.0))
        else:
            S.NoiseMedium(bnd, _p3(sh["color"]), sh["density"], noise_scale=g("noise_scale", 1.0),
                          threshold=g("threshold", 0.0), g=g("g", 0.0))
        return bnd
    if mat is None:
        raise ValueError(f"{where}: material is required for this shape type")
    if kind == "sphere":
        return S.Sphere(_p3(sh["center"]), sh["radius"], mat)
    if kind == "quad":
        return S.Quad(_p3(sh["corner"]), _p3(sh["u"]), _p3(sh["v"]), mat)
    if kind == "box":
        return S.Box(_p3(sh["p_min"]), _p3(sh["p_max"]), mat)
    if kind == "cylinder":
        return S.Cylinder(_p3(sh["center"]), sh["radius"], sh["height"], mat)
    if kind == "cone":
        return S.Cone(_p3(sh["center"]), sh["radius"], sh["height"], mat)
    if kind == "disk":
        return S.Disk(_p3(sh["center"]), _p3(sh["normal"]), sh["radius"], mat)
    if kind == "infinite_plane":
        return S.InfinitePlane(_p3(sh["point"]), _p3(sh["normal"]), mat, wave_amplitude=g("wave_amplitude", 0.0),
                               wave_scale=g("wave_scale", 1.0))
    if kind == "diamond":
        finish = str(g("girdle_finish") or "").lower()
        return S.Diamond(_p3(sh["center"]), sh["radius"], mat, table_pct=g("table_pct", 53.0),
                         crown_angle=g("crown_angle", 34.5), pavilion_angle=g("pavilion_angle", 40.75),
                         oval_ratio=g("oval_ratio", 1.0), girdle_facets=g("girdle_facets", 16),
                         girdle_pct=g("girdle_pct", 2.0), faceted=finish == "faceted", polish=g("polish", 0.02))
    if kind == "step_cut":
        name = str(g("outline") or "rect").lower()
        if name not in _OUTLINES:
            raise ValueError(f"{where}: unknown step-cut outline {name!r}")
        o = _OUTLINES[name]
        kw = {}
        if o in ("rect", "rounded_rect"):
            kw = {"corner_pct": g("corner_pct", 12.0), "segs": g("corner_segs", 4)}
        elif o == "pointed_oval":
            kw = {"segs": g("oval_segs", 24), "pear": name == "pear" or bool(g("pear", False))}
        elif o == "polygon":
            if not g("points"):
                raise ValueError(f"{where}: outline = \"polygon\" needs points")
            kw = {"points": g("points")}
        return S.StepCut(_p3(sh["center"]), 0.5 * sh["length"], 0.5 * sh["width"], mat,
                         table_pct=g("table_pct", 53.0), crown_angle=g("crown_angle", 34.5),
                         pavilion_angle=g("pavilion_angle", 40.75), outline=(o, kw),
                         crown_steps=g("crown_steps", 3), pavilion_steps=g("pavilion_steps", 3),
                         culet_pct=g("culet_pct", 8.0), step_spread=g("step_spread", 0.25), chevrons=g("chevrons", 0),
                         girdle_pct=g("girdle_pct", 2.0), polish=g("polish", 0.02))
    if kind == "mesh":
        return load_obj(_asset(base, sh["path"]), mat)
    raise ValueError(f"{where}: unknown shape type {kind!r}")


def _background(b, base):
    kind = b.get("type", "physical")
    if kind == "physical":
        return S.Background.physical(b.get("sun_azimuth", 0.0), b.get("sun_elevation", 30.0), b.get("turbidity", 3.0))
    if kind == "solid":
        return S.Background.solid(_p3(b["color"]))
    if kind == "env_map":
        vis = b.get("visible")
        visible = None
        if vis is not None:
            visible = (S.Background.solid(_p3(vis["color"])) if vis["type"] == "solid"
                       else S.Background.gradient(_p3(vis["top"]), _p3(vis["bottom"])))
        return S.Background.envmap(load_hdr(_asset(base, b["path"])), rotate_y_deg=b.get("rotate_y") or 0.0,
                                   visible=visible)
    raise ValueError(f"unknown background type {kind!r}")


def _lum(c):
    """Color::luminance (src/vec3.rs:33), in f32."""
    return f32(f32(f32(0.2126) * f32(c[0])) + f32(f32(0.7152) * f32(c[1]))) + f32(f32(0.0722) * f32(c[2]))


def _srgb_decode(v):
    x = f32(v) / f32(255.0)
    return x / f32(12.92) if x <= f32(0.04045) else f32(((x + f32(0.055)) / f32(1.055)) ** f32(2.4))


def is_emissive(mat):
    """Material::is_emissive (src/material.rs:22-24, 67, 926-928): a diffuse
    light whose texture at (u, v) = (0.5, 0.5), p = 0 has positive luminance, a
    blackbody light of positive intensity, a PBR material with positive
    emission strength and emission luminance; nothing else."""
    if isinstance(mat, S.BlackbodyLight):
        return mat.intensity > 0.0
    if isinstance(mat, S.PbrMaterial):
        return mat.emission_strength > 0.0 and _lum(mat.emission) > 0.0
    if isinstance(mat, S.DiffuseLight):
        t = mat.texture
        if isinstance(t, S.SolidTexture):
            c = t.color
        elif isinstance(t, S.CheckerTexture):
            c = t.even                                     # floor(scale * 0) sums to 0: even
        elif isinstance(t, S.ImageTexture):
            h, w = t.image.shape[:2]
            x = min(int(f32(0.5) * f32(w)), w - 1)
            y = min(int(f32(0.5) * f32(h)), h - 1)
            c = [_srgb_decode(k) for k in t.image[y, x]]
        elif isinstance(t, S.MottleTexture):
            c = t.lo                                      # turb(0) = 0: the low colour
        else:                                             # noise: 0.5 (1 + sin 0) at p = 0
            c = (0.5, 0.5, 0.5)
        return _lum(c) > 0.0
    return False


# Gem material presets (src/scene_file.rs:681-697): Cauchy B, C (um^2) and a
# body-colour Beer-Lambert sigma per scene unit at scale_mm = 0.001.
GEM_PRESETS = {
    "diamond": (2.351, 0.0230, (0.0, 0.0, 0.0)),
    "sapphire": (1.741, 0.0094, (650.0, 320.0, 25.0)),
    "ruby": (1.741, 0.0094, (25.0, 550.0, 450.0)),
    "emerald": (1.556, 0.0073, (480.0, 25.0, 380.0)),
}

# Step-family cut presets (src/scene_file.rs:741-768): outline, culet %,
# crown steps, pavilion steps, step spread, default chevron rows.
STEP_CUT_PRESETS = {
    "step": (("rect", {"corner_pct": 12.0}), 8.0, 3, 3, 0.25, 0),
    "princess": (("rect", {"corner_pct": 0.0}), 0.0, 4, 4, 0.15, 3),
    "asscher": (("rect", {"corner_pct": 22.0}), 8.0, 3, 3, 0.20, 0),
    "radiant": (("rect", {"corner_pct": 12.0}), 0.0, 4, 4, 0.15, 2),
    "baguette": (("rect", {"corner_pct": 4.0}), 8.0, 2, 2, 0.10, 0),
    "cushion": (("rounded_rect", {"corner_pct": 30.0, "segs": 4}), 8.0, 3, 3, 0.15, 0),
    "trillion": (("triangle", {}), 0.0, 4, 4, 0.15, 2),
    "marquise": (("pointed_oval", {"segs": 24, "pear": False}), 0.0, 4, 4, 0.15, 2),
    "pear": (("pointed_oval", {"segs": 24, "pear": True}), 0.0, 4, 4, 0.15, 2),
}

_DEG = f32(57.2957795130823208767981548141051703)     # f32::to_degrees


def gem_absorption(g, preset_abs):
    """gem_absorption (src/scene_file.rs:717-735): preset body colour plus the
    D..Z grade tint, rescaled by 0.001 / scale_mm, plus an explicit
    `absorption` (per scene unit, not rescaled)."""
    grade = str(g.get("color_grade", "D")).strip().upper()[:1]
    sigma = {"D": 0.0, "E": 0.0, "F": 0.0, "G": 50.0, "H": 80.0, "I": 120.0, "J": 170.0, "K": 230.0, "L": 300.0,
             "M": 300.0}.get(grade, 400.0)
    tint = [f32(c) * f32(sigma) for c in (1.0, 0.75, 0.35)]
    fix = f32(0.001) / f32(g.get("scale_mm", 0.001))
    extra = g.get("absorption") or (0.0, 0.0, 0.0)
    return tuple(float(f32(f32(f32(a) + t) * fix) + f32(e)) for a, t, e in zip(preset_abs, tint, extra))


def build_gem(g, where="gem"):
    """build_gem (src/scene_file.rs:793-897): a [[gems]] grading report ->
    a Diamond (cut "brilliant") or StepCut (step-family presets) with its
    spectral-dielectric preset, placed by scale, Euler rotation about the
    gem centre (rotate + rotate_y + spin) and translate."""
    name = f"{where} '{g.get('report', '')}'"
    scale_mm = f32(g.get("scale_mm", 0.001))
    if g.get("measurements_mm") is not None:
        m = [f32(x) for x in g["measurements_mm"]]
        if m[0] <= 0 or m[1] <= 0:
            raise ValueError(f"{name}: measurements_mm must be positive")
        l, w = (m[0], m[1]) if m[0] >= m[1] else (m[1], m[0])
        oval = l / w if (l - w) / w > f32(0.02) else f32(1.0)
        w_mm = w
    elif g.get("carat") is not None:
        ct = f32(g["carat"])
        if ct <= 0:
            raise ValueError(f"{name}: carat must be positive")
        w_mm, oval = f32((ct / f32(0.00355)) ** (1.0 / 3.0)), f32(1.0)
    else:
        raise ValueError(f"{name}: provide either carat or measurements_mm")
    if scale_mm <= 0:
        raise ValueError(f"{name}: scale_mm must be positive")
    radius = f32(0.5) * w_mm * scale_mm
    table_pct = f32(g.get("table_pct", 53.0))
    if not (0.0 < table_pct < 100.0):
        raise ValueError(f"{name}: table_pct must be in (0, 100), got {table_pct}")
    girdle_pct = f32(g.get("girdle_pct", 2.0))
    if not (math.isfinite(girdle_pct) and 0.0 <= girdle_pct < 100.0):
        raise ValueError(f"{name}: girdle_pct must be in [0, 100), got {girdle_pct}")
    table_r = (table_pct / f32(100.0)) * radius
    crown = f32(g.get("crown_angle", 34.5))
    if g.get("crown_pct") is not None:
        crown = f32(math.atan(f32(f32(f32(g["crown_pct"]) / f32(100.0)) * w_mm * scale_mm) / (radius - table_r))) * _DEG
    pav = f32(g.get("pavilion_angle", 40.75))
    if g.get("pavilion_pct") is not None:
        pav = f32(math.atan(f32(f32(f32(g["pavilion_pct"]) / f32(100.0)) * w_mm * scale_mm) / radius)) * _DEG
    if not (0.0 < crown < 90.0) or not (0.0 < pav < 90.0):
        raise ValueError(f"{name}: crown / pavilion angle must be in (0, 90) deg, got {crown}, {pav}")
    finish = str(g.get("girdle_finish") or "").lower()
    if finish not in ("", "faceted", "polished", "bruted"):
        raise ValueError(f"{name}: girdle_finish must be faceted|polished|bruted, got {finish!r}")
    facets = g.get("girdle_facets") or (64 if oval > f32(1.05) else 16)
    polish = float(g.get("polish", 0.02))
    mname = g.get("material") or "diamond"
    if mname == "checker":
        mat = S.Checker((0.9, 0.9, 0.9), (0.05, 0.05, 0.05), 12.0)
    elif mname == "opaque":
        mat = S.Lambertian((0.8, 0.8, 0.8))
    else:
        if mname not in GEM_PRESETS:
            raise ValueError(f"{name}: unknown gem material '{mname}' (expected diamond, sapphire, ruby or emerald)")
        b, c, pabs = GEM_PRESETS[mname]
        if g.get("cauchy_b") is not None:
            b = g["cauchy_b"]
        if g.get("cauchy_c") is not None:
            c = g["cauchy_c"]
        mat = S.SpectralDielectric(b, c, gem_absorption(g, pabs))
    center = _p3(g["center"])
    cut = g.get("cut") or "brilliant"
    if cut == "brilliant":
        h = S.Diamond(center, float(radius), mat, table_pct=float(table_pct), crown_angle=float(crown),
                      pavilion_angle=float(pav), oval_ratio=float(oval), girdle_facets=int(facets),
                      girdle_pct=float(girdle_pct), faceted=finish == "faceted", polish=polish)
    else:
        if cut not in STEP_CUT_PRESETS:
            raise ValueError(f"{name}: unknown gem cut '{cut}' (expected brilliant, step, princess, asscher, radiant, "
                             "baguette, cushion, trillion, marquise or pear)")
        outline, culet, cs, ps, spread, chev = STEP_CUT_PRESETS[cut]
        chev = g.get("chevrons") if g.get("chevrons") is not None else chev
        h = S.StepCut(center, float(radius * oval), float(radius), mat, table_pct=float(table_pct),
                      crown_angle=float(crown), pavilion_angle=float(pav), outline=outline, crown_steps=cs,
                      pavilion_steps=ps, culet_pct=culet, step_spread=spread, chevrons=int(chev),
                      girdle_pct=float(girdle_pct), polish=polish)
    if g.get("scale") is not None:
        h = h.scaled(g["scale"])
    ang = [float(f32(a)) for a in (g.get("rotate") or [0.0, 0.0, 0.0])]
    ang[1] = float(f32(f32(ang[1]) + f32(g.get("rotate_y") or 0.0)) + f32(g.get("spin") or 0.0))
    if any(a != 0.0 for a in ang):
        h = h.rotated(tuple(ang), pivot=center)
    if g.get("translate") is not None:
        h = h.translated(_p3(g["translate"]))
    return h


def _light_shape(h):
    """NEE light sampling exists on the device for quads, spheres and disks."""
    return isinstance(h, (S.Quad, S.Sphere, S.Disk))


def load(path, gem_spin=None):
    """Scene, camera(aspect) and info of a reference TOML scene file;
    `gem_spin` overrides every [[gems]] spin (load_with, src/scene_file.rs:
    625-648)."""
    import tomli
    with open(path, "rb") as fh:
        doc = tomli.load(fh)
    base = os.path.dirname(os.path.abspath(path))
    if gem_spin is not None:
        for gm in doc.get("gems", []):
            gm["spin"] = gem_spin
    objects, lights = [], []
    for i, o in enumerate(doc.get("objects", [])):
        mat = _material(o["material"], base) if o.get("material") else None
        h = _shape(o["shape"], mat, base, f"objects[{i}]")
        if o.get("scale") is not None:
            h = h.scaled(o["scale"])
        ang = list(o.get("rotate") or [0.0, 0.0, 0.0])
        ang[1] += (o.get("rotate_y") or 0.0) + (o.get("spin") or 0.0)
        if any(a != 0.0 for a in ang):
            h = h.rotated(tuple(ang))
        if o.get("translate") is not None:
            h = h.translated(_p3(o["translate"]))
        objects.append(h)
        if h.mat is not None and is_emissive(h.mat):     # light_area() > 0 joins the NEE list
            if not _light_shape(h):
                raise ValueError(f"objects[{i}]: an emissive {type(h).__name__} would be an NEE light, which the "
                                 "device samples for quads, spheres and disks only")
            lights.append(h)
    for i, gm in enumerate(doc.get("gems", [])):
        objects.append(build_gem(gm, f"gems[{i}]"))
    for lc in doc.get("lights", []):
        mat = (S.BlackbodyLight(lc["temp_k"], lc.get("intensity", 1.0)) if lc.get("temp_k") is not None
               else S.DiffuseLight(_p3(lc["color"])))
        q = S.Quad(_p3(lc["corner"]), _p3(lc["u"]), _p3(lc["v"]), mat)
        objects.append(q)
        lights.append(q)
    bg = _background(doc.get("background", {"type": "physical"}), base)
    scene = S.Scene(objects, lights, bg, name=doc.get("name", "Custom Scene"))
    cam = doc["camera"]
    frm, at = _p3(cam["look_from"]), _p3(cam["look_at"])
    focus = cam.get("focus_dist")
    if focus is None:
        focus = max(float(np.linalg.norm(np.subtract(frm, at).astype(np.float32))), 0.01)

    def camera(aspect):
        return S.Camera(frm, at, (0.0, 1.0, 0.0), cam.get("vfov", 40.0), aspect, aperture=cam.get("aperture", 0.0),
                        focus_dist=focus, blades=cam.get("aperture_blades", 0))
    c = doc.get("caustics") or {}
    first = doc.get("lights", [None])[0] if doc.get("lights") else None
    info = {"max_samples": doc.get("max_samples", 2000), "caustics": bool(c.get("enabled", True)) if c else False,
            "gather_radius": c.get("gather_radius", 0.15), "num_photons": c.get("num_photons", 500_000),
            "light_quad": None if first is None else {"q": _p3(first["corner"]), "u": _p3(first["u"]),
                                                      "v": _p3(first["v"]), "color": _p3(first.get("color", [1, 1, 1]))},
            "animation": "animation" in doc}
    return scene, camera, info
