"""Device photon emitters against the oracle: the sun disk sized from the
world's photon_bounds (PhotonMap::build, src/photon.rs:41-94) and the emitter
choice of SceneData::rebuild_caustics (src/scene.rs:35-74), through the C ABI.

Both sides trace one stream per photon index, so the stored photon sets agree
up to float rounding (a transcendental's last ulp can flip a Fresnel choice
for a few photons).
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from rustracer_amd import _ffi as F
from rustracer_amd.scene import (Background, Dielectric, InfinitePlane, Lambertian, Scene, Sphere)
from rustracer_amd.scenes import gem_caustics

pytestmark = pytest.mark.gpu


def same_photon_sets(ph, pc, frac=0.95, tol=1e-3):
    a = np.sort(ph[:, 0] * 1e3 + ph[:, 2])
    b = np.sort(pc[:, 0] * 1e3 + pc[:, 2])
    m = min(len(a), len(b))
    return np.mean(np.isclose(a[:m], b[:m], atol=tol)) > frac


def plane_and_glass_sphere(bg=None):
    return Scene([InfinitePlane((0, 0, 0), (0, 1, 0), Lambertian((0.7, 0.7, 0.7))),
                  Sphere((0.0, 1.0, 0.0), 0.5, Dielectric(1.5))], [], bg or Background.solid((0, 0, 0)))


def build_both(r, sc, pp):
    r.upload(sc)
    orc = O.OracleScene(sc)
    ng = r.photon_build(pp)
    nc = orc.photon_build(pp)
    return r.photon_export(), orc.photon_dump(nc), ng, nc, orc


def test_sun_disk_sized_from_photon_bounds(gpu_renderer):
    sc = plane_and_glass_sphere()
    pp = F.rt_photon_params()
    pp.emitter = F.EMIT_SUN_DISK
    for k in range(3):
        pp.sun_dir[k], pp.color[k] = (0.0, 1.0, 0.0)[k], 10.0
    pp.num_photons, pp.gather_radius = 20_000, 0.5
    ph, pc, ng, nc, orc = build_both(gpu_renderer, sc, pp)
    # the reference's regression bound (src/photon.rs:360-363) on the device
    assert ng > 2_000 and abs(ng - nc) <= 0.01 * nc + 5, (ng, nc)
    bd, bc = gpu_renderer.photon_bounds(), orc.photon_bounds()
    assert np.array_equal(bd[0], bc[0]) and np.array_equal(bd[1], bc[1])
    assert bd[1][0] - bd[0][0] < 2.0
    assert same_photon_sets(ph, pc)
    # total stored power agrees (the same photons carry the same power)
    assert abs(ph[:, 3:6].sum() - pc[:, 3:6].sum()) <= 0.01 * pc[:, 3:6].sum()


def test_sun_disk_fallback_without_finite_geometry(gpu_renderer):
    sc = Scene([InfinitePlane((0, 0, 0), (0, 1, 0), Lambertian((0.7, 0.7, 0.7))),
                InfinitePlane((0, 1, 0), (0, 1, 0), Dielectric(1.5))], [], Background.solid((0, 0, 0)))
    pp = F.rt_photon_params()
    pp.emitter = F.EMIT_SUN_DISK
    for k in range(3):
        pp.sun_dir[k], pp.color[k] = (0.0, 1.0, 0.0)[k], 1.0
    pp.num_photons, pp.gather_radius = 20_000, 0.5
    ph, pc, ng, nc, _ = build_both(gpu_renderer, sc, pp)
    assert gpu_renderer.photon_bounds() is None
    r = np.hypot(ph[:, 0], ph[:, 2])
    assert ng > 18_000 and abs(ng - nc) <= 5
    assert 19.9 < r.max() <= 20.0 + 1e-3 and abs((r ** 2).mean() - 200.0) < 6.0
    assert same_photon_sets(ph, pc)


def scene_params(knn=0):
    pp = F.rt_photon_params()
    pp.emitter = F.EMIT_SCENE
    pp.num_photons, pp.gather_radius, pp.knn = 20_000, 0.5, knn
    return pp


def check_emitter(gpu_renderer, sc):
    gpu_renderer.upload(sc)
    orc = O.OracleScene(sc)
    pp = scene_params()
    ed, ec = gpu_renderer.photon_emitter(pp), orc.photon_emitter(pp)
    assert ed.emitter == ec.emitter == F.EMIT_SUN_DISK
    assert np.allclose(ed.sun_dir[:], ec.sun_dir[:], rtol=1e-5, atol=1e-6), (ed.sun_dir[:], ec.sun_dir[:])
    assert np.allclose(ed.color[:], ec.color[:], rtol=2e-5), (ed.color[:], ec.color[:])
    ng = gpu_renderer.photon_build(pp)
    nc = orc.photon_build(pp)
    assert ng > 500 and abs(ng - nc) <= 0.01 * nc + 5, (ng, nc)
    assert same_photon_sets(gpu_renderer.photon_export(), orc.photon_dump(nc))


@pytest.mark.parametrize("az,el", [(30.0, 40.0), (200.0, 65.0)])
def test_rebuild_caustics_physical_sun(gpu_renderer, az, el):
    check_emitter(gpu_renderer, plane_and_glass_sphere(Background.physical(az, el, 3.0)))


@pytest.mark.parametrize("sampling", ["cdf", "alias"])
def test_rebuild_caustics_env_map(gpu_renderer, sampling):
    g = np.random.default_rng(7)
    img = g.uniform(0.0, 1.0, size=(32, 64, 3)).astype(np.float32)
    img[6:9, 40:43] += 200.0               # a bright key region the draws favour
    check_emitter(gpu_renderer, plane_and_glass_sphere(Background.envmap(img, rotate_y_deg=20.0, sampling=sampling)))


def test_rebuild_caustics_quad_fallback(gpu_renderer):
    sc, cam, cfg = gem_caustics(16, 16, num_photons=50_000)
    gpu_renderer.upload(sc)
    orc = O.OracleScene(sc)
    pp = scene_params()
    assert pytest.raises(Exception, gpu_renderer.photon_emitter, pp)
    assert orc.photon_emitter(pp) is None
    lq = cfg["light_quad"]
    for k in range(3):
        pp.origin[k], pp.u[k], pp.v[k], pp.color[k] = lq["q"][k], lq["u"][k], lq["v"][k], cfg["light_color"][k]
    pp.num_photons = 200_000
    e = gpu_renderer.photon_emitter(pp)
    assert e.emitter == F.EMIT_QUAD
    ng = gpu_renderer.photon_build(pp)
    nc = orc.photon_build(pp)
    assert ng > 500 and abs(ng - nc) <= 0.02 * nc + 5, (ng, nc)
    assert same_photon_sets(gpu_renderer.photon_export(), orc.photon_dump(nc))
