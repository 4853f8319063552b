"""The oracle's photon emitters: photon_bounds, the sun-disk sizing and the
emitter choice of SceneData::rebuild_caustics.

* photon_bounds: the reference's own test (src/photon.rs:296-322) restated on
  whole worlds: an infinite plane has none, a list unions its finite children,
  Translate and UniformScale forward theirs;
* the disk hugs the real geometry: src/photon.rs:345-364 (glass sphere over an
  infinite plane, 20,000 photons from a sun straight above: more than 2,000
  stored);
* rebuild_caustics (src/scene.rs:35-74): the Physical branch emits from the
  sun with pi times the sky just off the sun axis; the env branch's direction
  is pinned by an env map whose energy sits in one texel (every draw lands in
  it, so the direction is the normalised (0,1,0) + 8 x texel direction) and
  its colour by a constant map (pi x the constant); no sun and no env map
  falls back to the caustic quad, and with no quad there is no emitter.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from rustracer_amd import _ffi as F
from rustracer_amd.scene import (Background, Dielectric, InfinitePlane, Lambertian, Quad, Scene, Sphere, DiffuseLight)


def diffuse():
    return Lambertian((0.7, 0.7, 0.7))


def world(objs, bg=None, lights=()):
    return O.OracleScene(Scene(objs, list(lights), bg or Background.solid((0, 0, 0))))


def test_photon_bounds_skip_proxy_and_forward_through_wrappers():
    # lone plane: no finite geometry
    assert world([InfinitePlane((0, 0, 0), (0, 1, 0), diffuse())]).photon_bounds() is None
    # list: union of the finite children only
    mn, mx = world([InfinitePlane((0, 0, 0), (0, 1, 0), diffuse()), Sphere((3.0, 1.0, 0.0), 0.5, diffuse())]).photon_bounds()
    assert abs(mn[0] - 2.5) < 1e-4 and abs(mx[0] - 3.5) < 1e-4
    # wrappers forward the finite bounds
    mn, mx = world([Sphere((0.0, 1.0, 0.0), 0.5, diffuse()).translated((10.0, 0.0, 0.0))]).photon_bounds()
    assert abs(mn[0] - 9.5) < 1e-4 and abs(mx[0] - 10.5) < 1e-4
    mn, mx = world([Sphere((0.0, 1.0, 0.0), 0.5, diffuse()).scaled(2.0)]).photon_bounds()
    assert abs(mx[0] - 1.0) < 1e-4 and abs(mn[1] - 1.0) < 1e-4
    # a plane under a transform still has none; rotation forwards the rotated box
    assert world([InfinitePlane((0, 0, 0), (0, 1, 0), diffuse()).translated((0, 1, 0))]).photon_bounds() is None
    mn, mx = world([Sphere((1.0, 0.0, 0.0), 0.25, diffuse()).rotated((0, 90, 0))]).photon_bounds()
    assert abs(mn[2] + 1.25) < 1e-4 and abs(mx[2] + 0.75) < 1e-4


def sun_params(n, radius, sun=(0.0, 1.0, 0.0), color=(10.0, 10.0, 10.0)):
    pp = F.rt_photon_params()
    pp.emitter = F.EMIT_SUN_DISK
    for k in range(3):
        pp.sun_dir[k], pp.color[k] = sun[k], color[k]
    pp.num_photons, pp.gather_radius = n, radius
    return pp


def test_photon_disk_hugs_real_geometry_in_plane_scenes():
    orc = world([InfinitePlane((0, 0, 0), (0, 1, 0), diffuse()), Sphere((0.0, 1.0, 0.0), 0.5, Dielectric(1.5))])
    mn, mx = orc.photon_bounds()
    assert mx[0] - mn[0] < 2.0
    stored = orc.photon_build(sun_params(20_000, 0.5))
    assert stored > 2_000, stored
    ph = orc.photon_dump(stored)
    assert np.isfinite(ph).all()
    # every photon lies on the y = 0 floor; the refracted ones focus under the
    # sphere (grazing reflections land far out)
    assert np.abs(ph[:, 1]).max() < 1e-3 and np.median(np.hypot(ph[:, 0], ph[:, 2])) < 1.0


def test_sun_disk_without_finite_geometry_uses_the_fallback_disk():
    # Only infinite planes: no photon bounds, so the 20-unit disk 30 units up
    # the sun direction (src/photon.rs:72).  A glass plane at y = 1 over a
    # diffuse floor passes the vertical photons straight down: the stored
    # photons are the disk itself, uniform over radius 20 (E[r^2] = 200), each
    # carrying color * disk_area / (n * pi) = color * 400 / n before Fresnel.
    orc = world([InfinitePlane((0, 0, 0), (0, 1, 0), diffuse()), InfinitePlane((0, 1, 0), (0, 1, 0), Dielectric(1.5))])
    assert orc.photon_bounds() is None
    n = 20_000
    stored = orc.photon_build(sun_params(n, 0.5, color=(1.0, 1.0, 1.0)))
    ph = orc.photon_dump(stored)
    r = np.hypot(ph[:, 0], ph[:, 2])
    assert stored > 0.9 * n and np.abs(ph[:, 1]).max() < 1e-3
    assert 19.9 < r.max() <= 20.0 + 1e-3 and abs((r ** 2).mean() - 200.0) < 6.0
    # a plain Dielectric passes full power on the branch it picks
    assert np.allclose(ph[:, 3:6], 400.0 / n, rtol=1e-5), ph[:3, 3:6]


def _angle(a, b):
    a = np.asarray(a, np.float64); b = np.asarray(b, np.float64)
    return math.acos(max(-1.0, min(1.0, a @ b / np.linalg.norm(a) / np.linalg.norm(b))))


@pytest.mark.parametrize("az,el", [(30.0, 40.0), (90.0, 20.0), (200.0, 65.0)])
def test_rebuild_caustics_physical_sun(az, el):
    bg = Background.physical(az, el, 3.0)
    orc = world([Sphere((0, 1, 0), 0.5, Dielectric(1.5))], bg)
    pp = F.rt_photon_params()
    pp.emitter = F.EMIT_SCENE
    e = orc.photon_emitter(pp)
    assert e.emitter == F.EMIT_SUN_DISK
    sd = np.array(bg.sun_dir, np.float32)
    assert np.array_equal(np.array(e.sun_dir[:], np.float32), sd)
    # the colour is pi x the sky seen 0.04 rad (tan) off the sun axis: some
    # direction at that angle from the sun gives exactly that colour
    c = np.array(e.color[:], np.float32)
    assert np.isfinite(c).all() and (c > 0).all()
    ang = math.atan(0.04)
    best = np.inf
    a = np.cross(sd, [0.0, 0.0, 1.0] if abs(sd[0]) >= 0.9 else [1.0, 0.0, 0.0])
    a /= np.linalg.norm(a)
    b = np.cross(sd, a)
    for phi in np.linspace(0.0, 2 * math.pi, 721):
        d = (math.cos(ang) * sd + math.sin(ang) * (math.cos(phi) * a + math.sin(phi) * b)).astype(np.float32)
        sky = O.eval_fn("preetham", list(d) + list(sd) + [3.0], nout=3) * np.float32(math.pi)
        best = min(best, float(np.abs(sky - c).max() / c.max()))
    assert best < 2e-3, best


def _sph_dir(u, v):
    phi = 2.0 * math.pi * (u - 0.5)
    th = math.pi * v
    return np.array([math.sin(th) * math.sin(phi), math.cos(th), math.sin(th) * math.cos(phi)])


def test_rebuild_caustics_env_direction_one_bright_texel():
    W, H = 64, 32
    img = np.zeros((H, W, 3), np.float32)
    i, j = 41, 9                     # all the map's energy in one texel
    img[j, i] = 100.0
    orc = world([Sphere((0, 1, 0), 0.5, Dielectric(1.5))], Background.envmap(img))
    pp = F.rt_photon_params()
    pp.emitter = F.EMIT_SCENE
    e = orc.photon_emitter(pp)
    assert e.emitter == F.EMIT_SUN_DISK
    d = np.array(e.sun_dir[:], np.float64)
    assert abs(np.linalg.norm(d) - 1.0) < 1e-5
    want = np.array([0.0, 1.0, 0.0]) + 8.0 * _sph_dir((i + 0.5) / W, (j + 0.5) / H)
    # eight draws jitter inside the texel: within its angular size of the centre
    assert _angle(d, want) < 2.0 * math.pi / W, (_angle(d, want), d, want / np.linalg.norm(want))
    # the same map under the alias sampler: rebuild_caustics still takes the CDF draws
    orc2 = world([Sphere((0, 1, 0), 0.5, Dielectric(1.5))], Background.envmap(img, sampling="alias"))
    e2 = orc2.photon_emitter(pp)
    assert np.array_equal(np.array(e2.sun_dir[:]), np.array(e.sun_dir[:]))


def test_rebuild_caustics_env_colour_constant_map():
    img = np.full((16, 32, 3), (0.5, 1.0, 2.0), np.float32)
    orc = world([Sphere((0, 1, 0), 0.5, Dielectric(1.5))], Background.envmap(img, rotate_y_deg=30.0))
    pp = F.rt_photon_params()
    pp.emitter = F.EMIT_SCENE
    e = orc.photon_emitter(pp)
    assert np.allclose(np.array(e.color[:]), np.array([0.5, 1.0, 2.0]) * np.float32(math.pi), rtol=1e-6)


def test_rebuild_caustics_quad_fallback_and_none():
    light = DiffuseLight((15, 15, 15))
    q = Quad((-0.1, 2.0, -0.1), (0.2, 0, 0), (0, 0, 0.2), light)
    orc = world([Sphere((0, 1, 0), 0.5, Dielectric(1.5)), q], lights=[q])
    pp = F.rt_photon_params()
    pp.emitter = F.EMIT_SCENE
    assert orc.photon_emitter(pp) is None          # no sun, no env map, no quad given
    for k in range(3):
        pp.origin[k], pp.u[k], pp.v[k], pp.color[k] = q.q[k], q.u[k], q.v[k], 15.0
    e = orc.photon_emitter(pp)
    assert e.emitter == F.EMIT_QUAD and list(e.origin) == list(pp.origin)


def test_gem_scene_photon_map_is_populated_and_finite():
    """src/photon.rs:324-343 on the reference's own bench scene (a fixture of
    tests/golden): rebuild_caustics under the dusk sun stores more than 1,000
    finite photons, and the floor right under the diamond (girdle centre
    (0, 1.30, 2.5)) receives positive irradiance."""
    import os
    from rustracer_amd import Renderer
    from rustracer_amd import scene_file as SF
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "bench_scene", "bench.toml")
    sc, _, info = SF.load(path)
    orc = O.OracleScene(sc)
    pp = Renderer.caustic_params(info)
    e = orc.photon_emitter(pp)
    assert e.emitter == F.EMIT_SUN_DISK                # Physical background: the sun branch
    n = orc.photon_build(pp)
    assert n > 1_000, n
    ph = orc.photon_dump(n)
    assert np.isfinite(ph).all()
    irr = orc.photon_irradiance(np.array([[0.0, 0.0, 2.5, 0.0, 1.0, 0.0]], np.float32))
    assert np.isfinite(irr).all() and irr[0, 0] > 0.0, irr
