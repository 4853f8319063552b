"""Parity of the HIP trace kernels with the CPU oracle on seeded ray batches:
hit/miss, object and primitive ids bit-exact; t, point and normal bit-exact
(tolerance 1e-5 relative stated for the record, the kernels and the oracle
evaluate the same unfused f32 expressions)."""
import numpy as np
import pytest

from oracle import oracle as O
from rustracer_amd import scene as S
from rustracer_amd import scenes
from tests.helpers import icosphere, mixed_scene, random_rays, rays_from

pytestmark = pytest.mark.gpu

TOL = 1e-5


def compare_hits(hg, hc):
    assert np.array_equal(hg["obj_id"], hc["obj_id"]), np.nonzero(hg["obj_id"] != hc["obj_id"])[0][:10]
    hit = hc["obj_id"] >= 0
    assert np.array_equal(hg["prim_id"][hit], hc["prim_id"][hit])
    assert np.array_equal(hg["front_face"][hit], hc["front_face"][hit])
    for f in ("t", "p", "n", "u", "v"):
        a, b = hg[f][hit], hc[f][hit]
        assert np.allclose(a, b, rtol=TOL, atol=TOL), (f, np.abs(a - b).max())
    return float(np.mean(hg[f][hit] == hc[f][hit]))


def run_scene(gpu_renderer, scene, rays):
    gpu_renderer.upload(scene)
    orc = O.OracleScene(scene)
    hg = gpu_renderer.trace_closest(rays)
    hc = orc.trace_closest(rays)
    compare_hits(hg, hc)
    assert np.array_equal(gpu_renderer.trace_any(rays), orc.trace_any(rays))
    assert np.allclose(gpu_renderer.transmittance(rays), orc.transmittance(rays), rtol=TOL, atol=TOL)
    return hg


def test_box_room_rays(gpu_renderer):
    scene, _, _ = scenes.box_room()
    hg = run_scene(gpu_renderer, scene, random_rays(200000, 1, lo=(0.02, 0.02, -1.0), hi=(0.98, 0.98, 0.98)))
    assert (hg["obj_id"] >= 0).mean() > 0.5


def test_mixed_shapes_and_transforms(gpu_renderer):
    hg = run_scene(gpu_renderer, mixed_scene(), random_rays(200000, 2, lo=(-2, -1.5, -2), hi=(2, 2, 2)))
    ids = set(np.unique(hg["obj_id"]))
    assert {0, 1, 2, 3, 4, 5} <= ids, ids


def test_mesh_bvh_bit_exact_with_prim_ids(gpu_renderer):
    tri, _, _ = icosphere(5, 1.0)   # 20480 triangles
    g = np.random.default_rng(9)
    tri = tri + g.normal(scale=0.003, size=tri.shape).astype(np.float32)
    scene = S.Scene([S.Mesh(tri, S.Lambertian((0.5, 0.5, 0.5))),
                     S.Quad((-3, -1.2, -3), (6, 0, 0), (0, 0, 6), S.Lambertian((0.8, 0.8, 0.8)))])
    hg = run_scene(gpu_renderer, scene, random_rays(200000, 3, lo=(-2, -2, -2), hi=(2, 2, 2)))
    assert len(np.unique(hg["prim_id"][hg["obj_id"] == 0])) > 10000


def test_t_interval_edges_and_degenerate_rays(gpu_renderer):
    scene = S.Scene([S.Sphere((0, 0, 0), 1.0, S.Lambertian((0.5, 0.5, 0.5)))])
    gpu_renderer.upload(scene)
    orc = O.OracleScene(scene)
    o = np.array([[0, 0, -5]] * 6, np.float32)
    d = np.array([[0, 0, 1], [0, 0, 1], [0, 0, 0], [0, 0, 1], [1, 0, 0], [np.nan, 0, 1]], np.float32)
    r = rays_from(o, d)
    r["t_max"] = [4.0, 3.999, np.inf, np.inf, np.inf, np.inf]
    r["t_min"] = [0.001, 0.001, 0.001, 4.5, 0.001, 0.001]
    hg, hc = gpu_renderer.trace_closest(r), orc.trace_closest(r)
    compare_hits(hg, hc)
    assert list(hg["obj_id"]) == [0, -1, -1, 0, -1, -1]


def test_empty_batch(gpu_renderer):
    scene, _, _ = scenes.box_room()
    gpu_renderer.upload(scene)
    assert len(gpu_renderer.trace_closest(random_rays(0, 0))) == 0


def analytic_scene():
    """Cylinders, cones, disks and infinite planes (one rippled), bare and
    under scale / rotation / translation."""
    w = S.Lambertian((0.7, 0.7, 0.7))
    objs = [
        S.InfinitePlane((0, -1.2, 0), (0, 1, 0), w),
        S.InfinitePlane((0, 0, 3.0), (0.1, 0.2, -1.0), w, wave_amplitude=0.05, wave_scale=4.0),
        S.Cylinder((0.5, -1.0, 0.2), 0.3, 0.9, w),
        S.Cylinder((0, 0, 0), 0.2, 0.6, w).rotated((30, 0, 20)).translated((-0.8, 0.2, -0.3)),
        S.Cone((-0.3, -1.0, 0.8), 0.35, 0.8, w),
        S.Cone((0, 0, 0), 0.25, 0.5, w).scaled(1.3).rotated((0, 0, 90)).translated((0.9, 0.6, 0.9)),
        S.Disk((0, 1.3, 0), (0, -1, 0), 0.4, S.DiffuseLight((4, 4, 4))),
        S.Disk((0, 0, 0), (1, 1, 0), 0.3, w).translated((-0.6, -0.5, 1.0)),
    ]
    return S.Scene(objs, [], S.Background.solid((0, 0, 0)), name="analytic")


def test_cylinder_cone_disk_plane_bit_exact(gpu_renderer):
    hg = run_scene(gpu_renderer, analytic_scene(), random_rays(200000, 12, lo=(-2, -1.1, -2), hi=(2, 2, 2)))
    ids = set(np.unique(hg["obj_id"]))
    assert set(range(8)) <= ids, ids
    # lateral surfaces and caps of the capped shapes are both hit
    for obj in (2, 4):
        ny = np.abs(hg["n"][hg["obj_id"] == obj, 1])
        assert (ny > 0.99).any() and (ny < 0.9).any()


def test_noise_medium_events_and_transmittance(gpu_renderer):
    """NoiseMedium: the tracking stream is seeded from the ray's bits, so the
    GPU and the checker draw the same free paths; an event's t differs only by
    the ulps of log, which can move a rare accept / reject decision."""
    b = S.Sphere((0, 0.2, 0), 0.9, None)
    S.NoiseMedium(b, (0.9, 0.9, 0.9), 5.0, noise_scale=2.5, threshold=0.05, g=0.3, seed=3)
    w = S.Lambertian((0.6, 0.6, 0.6))
    scene = S.Scene([S.Quad((-3, -1, -3), (6, 0, 0), (0, 0, 6), w), b], [])
    rays = random_rays(100000, 21, lo=(-1.5, -0.9, -1.5), hi=(1.5, 1.5, 1.5))
    gpu_renderer.upload(scene)
    orc = O.OracleScene(scene)
    hg, hc = gpu_renderer.trace_closest(rays), orc.trace_closest(rays)
    same = hg["obj_id"] == hc["obj_id"]
    assert same.mean() > 0.995, same.mean()
    both = same & (hc["obj_id"] == 1)
    assert both.sum() > 2000
    assert np.allclose(hg["t"][both], hc["t"][both], rtol=1e-3, atol=1e-4)
    assert np.allclose(gpu_renderer.transmittance(rays), orc.transmittance(rays), rtol=1e-4, atol=1e-6)


def test_step_cuts_bit_exact(gpu_renderer):
    """StepCut family (src/step_cut.rs) as convex half-spaces: emerald with a
    girdle band, princess with chevrons, cushion, trillion, marquise with
    polish, under transforms."""
    w = S.Lambertian((0.7, 0.7, 0.7))
    g = S.Dielectric(1.6)
    objs = [
        S.Quad((-3, -1, -3), (6, 0, 0), (0, 0, 6), w),
        S.StepCut((-1.2, 0.0, -1.0), 0.6, 0.4, g, girdle_pct=3.0),
        S.StepCut((0.0, 0.0, -1.0), 0.45, 0.45, g, outline=("rect", {"corner_pct": 0.0}), chevrons=2).rotated((0, 30, 0)),
        S.StepCut((1.2, 0.0, -1.0), 0.5, 0.4, g, outline=("rounded_rect", {"corner_pct": 30.0, "segs": 4})),
        S.StepCut((-1.0, 0.0, 0.8), 0.5, 0.5, g, outline="triangle", table_pct=55.0),
        S.StepCut((0.6, 0.2, 0.8), 0.7, 0.3, g, outline=("pointed_oval", {"segs": 12, "pear": True}), polish=0.05,
                  step_spread=0.4).rotated((20, 0, 10)),
    ]
    sc = S.Scene(objs, [], S.Background.solid((0, 0, 0)), name="step_cuts")
    hg = run_scene(gpu_renderer, sc, random_rays(200000, 21, lo=(-2, -1.0, -2), hi=(2, 1.5, 2)))
    assert set(range(6)) <= set(np.unique(hg["obj_id"]))
