"""Pins the CPU oracle (oracle/) against the reference's own known-answer
tests, published test vectors, closed forms and brute force.

Each test names the reference test or fact it restates.  These run on the CPU
only (no GPU), and are what makes the oracle trustworthy as the GPU checker.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from rustracer_amd import scene as S
from tests.helpers import rays_from, random_rays, icosphere

DUMMY = S.Lambertian((0.5, 0.5, 0.5))


def one(shape, lights=()):
    return O.OracleScene(S.Scene([shape], list(lights)))


def shoot(osc, o, d, t_min=0.001, t_max=np.inf, direct=True, sub=-1):
    """direct: Hittable::hit of object 0 itself (sub >= 0: of that mesh
    triangle), as the reference's unit tests call it; otherwise through the
    world BvhTree."""
    r = rays_from(np.array([o], np.float32), np.array([d], np.float32), t_min, t_max)
    h = (osc.trace_object(0, r, sub) if direct else osc.trace_closest(r))[0]
    return None if h["obj_id"] < 0 else h


# ---------------------------------------------------------------- RNG
def test_splitmix_seeding_matches_published_vector():
    # SplitMix64 from state 0 yields e220a8397b1dcdaf, 6e789e6aa1b965f4,
    # 06c45d188009454f, f88bb8a8724c81ec (Vigna's splitmix64.c); SmallRng's
    # seed_from_u64(0) therefore starts xoshiro256++ from that state.
    s = [0xe220a8397b1dcdaf, 0x6e789e6aa1b965f4, 0x06c45d188009454f, 0xf88bb8a8724c81ec]
    assert np.array_equal(O.rng_u64(0, 4), O.xoshiro_raw(np.array(s, np.uint64), 4))


def test_xoshiro256pp_published_vector():
    # xoshiro256++ reference implementation from state {1, 2, 3, 4}:
    # first output rotl(1 + 4, 23) + 1 = 41943041, second 58720359.
    out = O.xoshiro_raw(np.array([1, 2, 3, 4], np.uint64), 2)
    assert int(out[0]) == 41943041
    assert int(out[1]) == 58720359


def test_f32_draws_are_24_bit_grid_in_unit_interval():
    x = O.rng_f32(42, 100000)
    assert x.min() >= 0.0 and x.max() < 1.0
    assert np.all((x * 16777216.0) == np.floor(x * 16777216.0))
    assert abs(x.mean() - 0.5) < 0.005


def test_gen_range_stays_in_half_open_range():
    x = O.rng_range(7, 380.0, 700.0, 100000)
    assert x.min() >= 380.0 and x.max() < 700.0
    assert abs(x.mean() - 540.0) < 2.0


# ---------------------------------------------------------------- quad (src/quad.rs:129-198)
def unit_quad():
    return one(S.Quad((0, 0, 0), (1, 0, 0), (0, 1, 0), DUMMY))


def test_quad_hit_returns_correct_t():
    h = shoot(unit_quad(), (0.5, 0.5, 3.0), (0, 0, -1))
    assert abs(h["t"] - 3.0) < 1e-5


def test_quad_hit_returns_correct_uv():
    h = shoot(unit_quad(), (0.25, 0.75, 1.0), (0, 0, -1))
    assert abs(h["u"] - 0.25) < 1e-5 and abs(h["v"] - 0.75) < 1e-5


def test_quad_hit_corners():
    q = unit_quad()
    for ox, oy in [(0, 0), (1, 0), (0, 1), (1, 1)]:
        assert shoot(q, (ox, oy, 1.0), (0, 0, -1)) is not None


def test_bvh_slab_nan_semantics_cull_face_grazing_ray():
    # AVX2 minps/maxps return the second operand on NaN and the <= compare is
    # ordered, so a ray on a box face with a zero direction component misses
    # the node (src/bvh.rs:445-456 documents this); the quad itself is hit.
    q = unit_quad()
    assert shoot(q, (0, 0, 1.0), (0, 0, -1), direct=True) is not None
    assert shoot(q, (0, 0, 1.0), (0, 0, -1), direct=False) is None
    assert shoot(q, (0.5, 0.5, 1.0), (0, 0, -1), direct=False) is not None


def test_quad_misses_and_parallel():
    q = unit_quad()
    assert shoot(q, (1.01, 0.5, 1.0), (0, 0, -1)) is None
    assert shoot(q, (0.5, -0.01, 1.0), (0, 0, -1)) is None
    assert shoot(q, (0.5, 0.5, 0.0), (1, 0, 0)) is None


def test_quad_front_face_orientation():
    q = unit_quad()
    h = shoot(q, (0.5, 0.5, 1.0), (0, 0, -1))
    assert h["front_face"] == 1 and h["n"][2] > 0
    h2 = shoot(q, (0.5, 0.5, -1.0), (0, 0, 1))
    assert h2["front_face"] == 0


def test_degenerate_quad_rejected():
    with pytest.raises(ValueError, match="degenerate quad"):
        S.Quad((0, 0, 0), (1, 0, 0), (2, 0, 0), DUMMY)


# ---------------------------------------------------------------- triangle (src/triangle.rs:160-345)
def tri_mesh(v, n=None, uv=None):
    v = np.array(v, np.float32).reshape(1, 3, 3)
    return one(S.Mesh(v, DUMMY, normals=None if n is None else np.array(n, np.float32).reshape(1, 3, 3),
                      uvs=None if uv is None else np.array(uv, np.float32).reshape(1, 3, 2)))


XY = [(0, 0, 0), (2, 0, 0), (0, 2, 0)]


def test_triangle_centroid_t_and_vertex():
    m = tri_mesh(XY)
    assert abs(shoot(m, (2 / 3, 2 / 3, -5), (0, 0, 1), sub=0)["t"] - 5.0) < 1e-5
    assert abs(shoot(m, (0, 0, -3), (0, 0, 1), sub=0)["t"] - 3.0) < 1e-5


def test_triangle_misses():
    m = tri_mesh(XY)
    assert shoot(m, (3, 0, -1), (0, 0, 1), sub=0) is None
    assert shoot(m, (0, -1, -1), (0, 0, 1), sub=0) is None
    assert shoot(m, (1.5, 1.5, -1), (0, 0, 1), sub=0) is None
    assert shoot(m, (0.5, 0.5, 0), (1, 0, 0), sub=0) is None


def test_triangle_uv_interpolated_at_centroid():
    m = tri_mesh(XY, n=[(0, 0, 1)] * 3, uv=[(0, 0), (1, 0), (0, 1)])
    h = shoot(m, (2 / 3, 2 / 3, -1), (0, 0, 1), sub=0)
    assert abs(h["u"] - 1 / 3) < 1e-5 and abs(h["v"] - 1 / 3) < 1e-5


def test_triangle_shading_normal_interpolated():
    n0 = np.array([0, 0, 1.0]); n1 = np.array([1, 0, 1.0]) / math.sqrt(2); n2 = np.array([-1, 0, 1.0]) / math.sqrt(2)
    m = tri_mesh(XY, n=[n0, n1, n2], uv=[(0, 0), (1, 0), (0, 1)])
    h = shoot(m, (2 / 3, 2 / 3, -1), (0, 0, 1), sub=0)
    avg = (n0 + n1 + n2) / np.linalg.norm(n0 + n1 + n2)
    got = h["n"] if h["front_face"] else -h["n"]
    assert abs(float(np.dot(got, avg)) - 1.0) < 1e-4


def test_triangle_zero_interpolated_normal_falls_back_to_face_normal():
    m = tri_mesh([(0, 0, 0), (1, 0, 0), (0, 1, 0)], n=[(0, 0, 1), (0, 0, -1), (0, 0, 0)], uv=[(0, 0), (1, 0), (0.5, 1)])
    h = shoot(m, (0.5, 0.0, -1.0), (0, 0, 1), sub=0)
    assert np.all(np.isfinite(h["n"]))
    assert np.linalg.norm(h["n"] - np.array([0, 0, -1])) < 1e-5


# ---------------------------------------------------------------- sphere (src/sphere.rs:133-138)
def test_sphere_zero_direction_misses():
    s = one(S.Sphere((0, 0, 0), 1.0, DUMMY))
    assert shoot(s, (0, 0, -5), (0, 0, 0)) is None
    r = rays_from(np.array([[0, 0, -5]], np.float32), np.array([[0, 0, 0]], np.float32))
    assert s.trace_any(r)[0] == 0


def test_sphere_closed_interval_at_t_max():
    # a hit at exactly t == t_max is accepted (src/hittable.rs:97-101)
    s = one(S.Sphere((0, 0, 0), 1.0, DUMMY))
    assert shoot(s, (0, 0, -5), (0, 0, 1), t_max=4.0) is not None
    assert shoot(s, (0, 0, -5), (0, 0, 1), t_max=3.999) is None


# ---------------------------------------------------------------- BVH (src/bvh.rs:684-745)
def unit_spheres(centers):
    return O.OracleScene(S.Scene([S.Sphere(c, 1.0, DUMMY) for c in centers]))


def test_bvh_hits_and_misses_one_sphere():
    b = unit_spheres([(0, 0, 0)])
    assert shoot(b, (0, 0, -5), (0, 0, 1), direct=False) is not None
    assert shoot(b, (2, 0, -5), (0, 0, 1), direct=False) is None


def test_bvh_finds_nearest_of_two():
    b = unit_spheres([(0, 0, -3), (0, 0, 3)])
    h = shoot(b, (0, 0, -10), (0, 0, 1), direct=False)
    assert h["t"] < 8.0 and h["obj_id"] == 0


def test_bvh_all_spheres_in_grid_reachable():
    pos = [(x * 10.0, y * 10.0, 0.0) for x in (-1, 0, 1) for y in (-1, 0, 1)]
    b = unit_spheres(pos)
    for i, p in enumerate(pos):
        h = shoot(b, (p[0], p[1], -20), (0, 0, 1), direct=False)
        assert h is not None and h["obj_id"] == i


def test_bvh_closest_equals_linear_scan():
    """Brute force: the SAH QBVH's closest hit is the minimum over a linear
    scan of every primitive (HittableList), ids by input order."""
    g = np.random.default_rng(3)
    spheres = [S.Sphere(tuple(g.uniform(-3, 3, 3)), float(g.uniform(0.1, 0.6)), DUMMY) for _ in range(60)]
    quads = [S.Quad(tuple(g.uniform(-3, 3, 3)), tuple(g.uniform(-1, 1, 3)), tuple(g.uniform(-1, 1, 3)), DUMMY) for _ in range(40)]
    objs = spheres + quads
    bvh = O.OracleScene(S.Scene(objs))
    rays = random_rays(3000, 11, lo=(-4, -4, -4), hi=(4, 4, 4))
    hb = bvh.trace_closest(rays, nthreads=4)
    best_t = np.full(len(rays), np.inf, np.float32)
    best_id = np.full(len(rays), -1)
    for i, o in enumerate(objs):
        h = O.OracleScene(S.Scene([o])).trace_closest(rays, nthreads=4)
        hit = h["obj_id"] >= 0
        better = hit & (h["t"] <= best_t)
        best_t[better] = h["t"][better]
        best_id[better] = i
    assert np.array_equal(hb["obj_id"], best_id)
    m = best_id >= 0
    assert np.array_equal(hb["t"][m], best_t[m])


def test_mesh_bvh_closest_equals_brute_force():
    tri, _, _ = icosphere(2, 1.0)
    g = np.random.default_rng(5)
    tri = tri + g.normal(scale=0.02, size=tri.shape).astype(np.float32)
    mesh = O.OracleScene(S.Scene([S.Mesh(tri, DUMMY)]))
    rays = random_rays(2000, 12, lo=(-2, -2, -2), hi=(2, 2, 2))
    hm = mesh.trace_closest(rays, nthreads=4)
    # brute force Moller-Trumbore in float32 numpy, same operation order
    o = rays["origin"].astype(np.float32)[:, None, :]
    d = rays["dir"].astype(np.float32)[:, None, :]
    v0, v1, v2 = tri[None, :, 0], tri[None, :, 1], tri[None, :, 2]
    e1, e2 = v1 - v0, v2 - v0

    def cr(a, b):
        return np.stack([a[..., 1] * b[..., 2] - a[..., 2] * b[..., 1], a[..., 2] * b[..., 0] - a[..., 0] * b[..., 2],
                         a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0]], -1)

    def dt(a, b):
        return (a[..., 0] * b[..., 0] + a[..., 1] * b[..., 1]) + a[..., 2] * b[..., 2]
    with np.errstate(all="ignore"):
        h = cr(d, e2)
        a = dt(e1, h)
        f = np.float32(1.0) / a
        s = o - v0
        b1 = f * dt(s, h)
        q = cr(s, e1)
        b2 = f * dt(d, q)
        t = f * dt(e2, q)
        ok = (np.abs(a) >= 1e-8) & (b1 >= 0) & (b1 <= 1) & (b2 >= 0) & (b1 + b2 <= 1) & (t >= 0.001)
        t = np.where(ok, t, np.inf)
    best = t.min(axis=1)
    hit = np.isfinite(best)
    assert np.array_equal(hm["obj_id"] >= 0, hit)
    assert np.array_equal(hm["t"][hit], best[hit])
    assert np.array_equal(hm["prim_id"][hit], t.argmin(axis=1)[hit])


# ---------------------------------------------------------------- convex / diamond (src/diamond.rs:290-519)
def test_convex_cube_matches_box_and_interior_exit():
    planes = [(1, 0, 0, 1), (-1, 0, 0, 1), (0, 1, 0, 1), (0, -1, 0, 1), (0, 0, 1, 1), (0, 0, -1, 1)]
    c = one(S.Convex(planes, DUMMY))
    h = shoot(c, (0.2, 0.3, -5), (0, 0, 1))
    assert abs(h["t"] - 4.0) < 1e-6 and h["n"][2] < 0
    h = shoot(c, (0.2, 0.3, 0.0), (0, 0, 1))       # interior ray exits
    assert abs(h["t"] - 1.0) < 1e-6 and h["front_face"] == 0


def test_diamond_knife_edge_has_57_planes_and_hits_table_from_above():
    d = S.Diamond((0, 0, 0), 1.0, DUMMY, table_pct=57.0, crown_angle=34.5, pavilion_angle=40.75,
                  girdle_facets=56, girdle_pct=0.0)
    assert d.planes.shape[0] == 57
    h = shoot(one(d), (0, 5, 0), (0, -1, 0))
    crown_h = (1 - 0.57) * math.tan(math.radians(34.5))
    assert abs(h["t"] - (5 - crown_h)) < 1e-4 and h["n"][1] > 0.99


def test_diamond_default_banded_plane_count():
    d = S.Diamond((0, 0, 0), 1.0, DUMMY)      # 16 girdle facets + 2% band -> 1 + 16 + 16 + 64
    assert d.planes.shape[0] == 97


# ---------------------------------------------------------------- transforms (src/transform.rs:301-419)
def test_translate_and_rotate_move_hits():
    s = one(S.Sphere((0, 0, 0), 1.0, DUMMY).translated((5, 0, 0)))
    assert shoot(s, (0, 0, -5), (0, 0, 1)) is None
    assert abs(shoot(s, (5, 0, -5), (0, 0, 1))["t"] - 4.0) < 1e-5
    q = one(S.Quad((-1, -1, 0), (2, 0, 0), (0, 2, 0), DUMMY).rotated((0, 90, 0)))
    h = shoot(q, (5, 0, 0), (-1, 0, 0))
    assert h is not None and abs(h["t"] - 5.0) < 1e-5 and abs(abs(h["n"][0]) - 1.0) < 1e-6


def test_uniform_scale_scales_hit_point():
    s = one(S.Sphere((0, 0, 0), 1.0, DUMMY).scaled(2.0))
    h = shoot(s, (0, 0, -5), (0, 0, 1))
    assert abs(h["t"] - 3.0) < 1e-5 and abs(h["p"][2] + 2.0) < 1e-5


# ---------------------------------------------------------------- media (src/volume.rs:280-339)
def test_constant_medium_beer_lambert():
    b = S.Sphere((0, 0, 0), 1.0, None)
    S.ConstantMedium(b, 1.0, (1, 1, 1))
    m = one(b)
    r = rays_from(np.array([[0, 0, -5]], np.float32), np.array([[0, 0, 1]], np.float32))
    assert abs(m.transmittance(r)[0] - math.exp(-2.0)) < 1e-5
    assert m.trace_any(r)[0] == 0          # media never report any_hit


def test_hg_sample_mean_cosine_is_g():
    # closed form: E[cos theta] of Henyey-Greenstein = g
    for g in (0.0, 0.6, -0.3):
        c = np.mean([O.eval_fn("hg_sample_seeded", [0, 0, 1, g, float(s)])[2] for s in range(4000)])
        assert abs(c - g) < 0.03


# ---------------------------------------------------------------- materials (src/material.rs:1299-1639)
def test_ggx_ndf_integrates_to_one():
    for alpha in (0.1, 0.3, 0.7, 1.0):
        n = 1000
        dth = (math.pi / 2) / n
        s = 0.0
        for i in range(n):
            th = (i + 0.5) * dth
            s += O.eval_fn("ggx_ndf", [math.cos(th), alpha], 1) * math.cos(th) * math.sin(th) * dth * 2 * math.pi
        assert abs(s - 1.0) < 0.01


def test_smith_g1_range_smooth_limit_and_lambda():
    for a in (0.01, 0.1, 0.3, 0.5, 0.8, 1.0):
        for c in (0.01, 0.1, 0.3, 0.5, 0.7, 0.9, 1.0):
            g = O.eval_fn("smith_g1", [c, a], 1)
            assert 0.0 <= g <= 1.0
    for c in (0.1, 0.5, 0.9, 1.0):
        assert abs(O.eval_fn("smith_g1", [c, 1e-6], 1) - 1.0) < 1e-4
    for a in (0.01, 0.1, 0.5, 1.0):
        for c in (0.05, 0.3, 0.7, 1.0):
            lam = O.eval_fn("smith_lambda", [c, a], 1)
            g1 = O.eval_fn("smith_g1", [c, a], 1)
            assert abs(lam - (1 / g1 - 1)) < 1e-5


def test_e_lut_bounds_and_smooth_limit():
    for mu in (0.02, 0.25, 0.5, 0.9, 1.0):
        for a in (0.01, 0.25, 0.7, 1.0):
            e = O.eval_fn("e_ss", [mu, a], 1)
            assert 0 < e <= 1 + 1e-4
            c = O.eval_fn("energy_comp", [mu, a], 1)
            assert 1 - 1e-6 <= c <= 10
    assert O.eval_fn("e_ss", [1.0, 0.001], 1) > 0.99


def furnace(mat, n, seed):
    out = O.scatter(mat, (0, 0, -1), (0, 0, 1), 550.0, (0, 0, 0), (0, 0, -1), seed=seed, n=n)
    a = out[:, :3]
    lum = 0.2126 * a[:, 0] + 0.7152 * a[:, 1] + 0.0722 * a[:, 2]
    return float(np.where(out[:, 7] > 0, lum, 0).mean())


def test_white_furnace_metallic():
    for r in (0.1, 0.5, 0.9):
        m = furnace(S.PbrMaterial(albedo=(1, 1, 1), metallic=1.0, roughness=r), 8000, 42)
        assert 0.85 < m <= 1.05
    assert furnace(S.PbrMaterial(albedo=(1, 1, 1), metallic=1.0, roughness=0.9), 20000, 7) > 0.9
    m = furnace(S.PbrMaterial(albedo=(1, 1, 1), metallic=1.0, roughness=0.05), 8000, 123)
    assert 0.95 < m <= 1.05


def test_fresnel_r0_grazing_and_tir():
    assert abs(O.eval_fn("fresnel_dielectric", [1.0, 1 / 1.5], 1) - 0.04) < 1e-4
    assert abs(O.eval_fn("fresnel_dielectric", [0.0, 1 / 1.5], 1) - 1.0) < 1e-6
    # closed form r0 = ((n-1)/(n+1))^2
    for n in (1.33, 2.417):
        assert abs(O.eval_fn("fresnel_dielectric", [1.0, 1 / n], 1) - ((n - 1) / (n + 1)) ** 2) < 1e-6
    # TIR from inside glass beyond the critical angle
    assert O.eval_fn("fresnel_dielectric", [math.cos(math.radians(60)), 1.5], 1) == 1.0


def test_dielectric_transmits_most_light_at_normal_incidence():
    out = O.scatter(S.Dielectric(1.5), (0, 0, -1), (0, 0, 1), 550.0, (0, 0, 0), (0, 0, -1), seed=1, n=4000)
    refracted = out[:, 5] > 0
    assert refracted.mean() >= 0.9


def test_dielectric_tir_always_reflects():
    d = np.array([math.sin(math.radians(60)), 0, math.cos(math.radians(60))], np.float32)
    out = O.scatter(S.Dielectric(1.5), (0, 0, 0), d, 550.0, (0, 0, 1), (0, 0, -1), front_face=False, seed=2, n=500)
    assert np.all(out[:, 5] < 0)


# ---------------------------------------------------------------- spectrum (src/spectrum.rs:283-452)
def test_cmf_and_fire_tint_means_are_white():
    lam = 380.0 + np.arange(65) * 5.0
    s = np.mean([O.eval_fn("spectral_to_rgb", [x]) for x in lam], axis=0)
    f = np.mean([O.eval_fn("spectral_fire_rgb", [x]) for x in lam], axis=0)
    assert np.allclose(s, 1.0, atol=1e-4) and np.allclose(f, 1.0, atol=1e-4)


def test_cauchy_crown_glass_and_diamond_dispersion():
    assert abs(O.eval_fn("cauchy", [589.3, 1.507, 0.00375], 1) - 1.52) < 0.005
    # diamond preset B=2.351, C=0.0230 (src/scene_file.rs:686): nF - nC = 0.044 +- 0.002
    nf = O.eval_fn("cauchy", [486.1, 2.351, 0.0230], 1)
    nc = O.eval_fn("cauchy", [656.3, 2.351, 0.0230], 1)
    assert abs((nf - nc) - 0.044) < 0.002


def test_sellmeier_diamond_closed_form():
    # Peter (1923): n^2 = 1 + 0.3306 l^2/(l^2 - 0.175^2) + 4.3356 l^2/(l^2 - 0.106^2) -> n(589.3) = 2.4175
    n = O.eval_fn("sellmeier", [589.3, 0.3306, 4.3356, 0.0, 0.175 ** 2, 0.106 ** 2, 0.0], 1)
    assert abs(n - 2.4175) < 0.001


def test_conductor_normal_incidence_formula():
    for n, k in ((0.2, 3.5), (1.1, 2.6), (0.15, 4.0)):
        want = ((n - 1) ** 2 + k * k) / ((n + 1) ** 2 + k * k)
        assert abs(O.eval_fn("fresnel_conductor", [1.0, n, k], 1) - want) < 1e-5


def test_gold_and_silver_reflectance():
    n, k = O.eval_fn("metal_ior", [0, 650.0], 2)
    assert O.eval_fn("fresnel_conductor", [1.0, n, k], 1) > 0.9       # gold reflects red
    n, k = O.eval_fn("metal_ior", [0, 430.0], 2)
    assert O.eval_fn("fresnel_conductor", [1.0, n, k], 1) < 0.5       # and absorbs blue
    for lam in (400.0, 550.0, 650.0):
        n, k = O.eval_fn("metal_ior", [2, lam], 2)
        assert O.eval_fn("fresnel_conductor", [1.0, n, k], 1) > 0.85  # silver is bright


# ---------------------------------------------------------------- env map (src/renderer.rs:737-815)
def test_env_sample_is_bilinear_and_wraps_seam():
    img = np.zeros((2, 4, 3), np.float32)
    img[:, 2:] = 1.0
    s4 = math.sin(math.pi / 4), math.cos(math.pi / 4)
    s34 = math.sin(3 * math.pi / 4), math.cos(3 * math.pi / 4)
    dirs = [(0, 0, 1), (s4[0], 0, s4[1]), (0, 0, -1), (s34[0], 0, s34[1])]
    out = O.env_eval(img, 0.0, dirs)
    assert np.allclose(out[:, 0], [0.5, 1.0, 0.5, 1.0], atol=1e-6)


def test_env_rotation_shifts_azimuth_and_sampling():
    img = np.zeros((4, 8, 3), np.float32)
    img[:, 1:3] = 5.0
    rot = math.radians(90.0)
    out = O.env_eval(img, rot, [(0, 0, 1), (-1, 0, 0)])
    assert out[0, 0] > 4.0 and out[1, 0] < 1.0
    s = O.env_sample(img, rot, 1, 1)[0]
    d = s[:3] / np.linalg.norm(s[:3])
    assert s[3] > 0 and d[2] > abs(d[0])


def test_env_pdf_integrates_to_one():
    g = np.random.default_rng(0)
    img = g.uniform(0, 2, size=(16, 32, 3)).astype(np.float32)
    img[3, 5] = 50.0
    # integrate the pdf over the sphere on a fine equal-angle grid
    nt, npf = 256, 512
    th = (np.arange(nt) + 0.5) / nt * math.pi
    ph = (np.arange(npf) + 0.5) / npf * 2 * math.pi - math.pi
    T, P = np.meshgrid(th, ph, indexing="ij")
    dirs = np.stack([np.sin(T) * np.sin(P), np.cos(T), np.sin(T) * np.cos(P)], -1).reshape(-1, 3)
    pdf = O.env_eval(img, 0.0, dirs)[:, 3].reshape(nt, npf)
    integral = float((pdf * np.sin(T)).sum() * (math.pi / nt) * (2 * math.pi / npf))
    assert abs(integral - 1.0) < 0.01


# ---------------------------------------------------------------- cosine pdf (src/pdf.rs:73-128)
def test_lambertian_scatter_is_diffuse_placeholder():
    out = O.scatter(S.Lambertian((0.5, 0.25, 1.0)), (0, 0, -1), (0, 0, 1), 550.0, (0, 0, 0), (0, 0, -1), n=1)
    assert out[0, 6] == 0 and np.allclose(out[0, :3], [0.5, 0.25, 1.0])


# ---------------------------------------------------------------- BlackbodyLight (src/material.rs:27-83, src/spectrum.rs:204-208)
def test_planck_reference_kats_and_closed_form():
    p = lambda lam, T: O.eval_fn("planck_raw", [lam, T], nout=1)
    # planck_hotter_blackbody_is_bluer, planck_monotone_in_temperature (src/spectrum.rs:362-376)
    assert p(450, 6500) / p(650, 6500) > p(450, 3000) / p(650, 3000)
    for lam in (400.0, 550.0, 700.0):
        assert p(lam, 6000) > p(lam, 3000) and p(lam, 10000) > p(lam, 6000)
    for lam, T in [(380.0, 2700.0), (555.0, 5500.0), (700.0, 10000.0)]:
        x = 14388000.0 / (lam * T)
        assert p(lam, T) == pytest.approx(1.0 / (lam ** 5 * math.expm1(x)), rel=2e-5)


def test_blackbody_white_point_and_spectral_mean():
    """avg_color is the mean over the 65 CIE samples of spectral_to_rgb times
    the normalised Planck weight: 6500 K is near neutral, 3000 K warm; the
    normalised weights average to 1 over the samples."""
    def bb(T, I=1.0):
        return O.eval_fn("blackbody", [T, I], nout=4)
    c65 = np.array(bb(6500.0)[:3])
    c30 = np.array(bb(3000.0)[:3])
    assert c30[0] > c30[2] * 2 and abs(c65[0] / c65[2] - 1) < 0.35
    lam = 380.0 + 5.0 * np.arange(65)
    T = 4200.0
    out = bb(T, 2.5)
    w = np.array([O.eval_fn("planck_raw", [l, T], nout=1) for l in lam]) * out[3]
    assert w.mean() == pytest.approx(1.0, rel=1e-5)
    rgb = np.array([O.eval_fn("spectral_to_rgb", [l], nout=3) for l in lam])
    assert np.allclose((rgb * w[:, None]).mean(0) * 2.5, out[:3], rtol=1e-5)


# ---------------------------------------------------------------- SSSMaterial (src/material.rs:576-631)
def test_sss_free_flight_and_exit_absorption():
    """From inside (back face at distance L along the ray), a scatter event
    happens with probability 1 - exp(-density L), weighted albedo *
    exp(-sigma_a t); otherwise the boundary is crossed with weight
    exp(-sigma_a L)."""
    albedo, sa, dens, L = (0.9, 0.7, 0.5), (0.2, 0.4, 0.8), 1.3, 0.8
    m = S.SSSMaterial(albedo, sa, ior=1.4, density=dens, g=0.3)
    out = O.scatter(m, (0, 0, -L), (0, 0, 1), 550.0, (0, 0, 0), (0, 0, -1), t=L, front_face=False, n=200000, seed=3)
    assert (out[:, 7] == 1).all() and (out[:, 6] == 1).all()
    exitw = np.exp(-np.array(sa) * L)
    is_exit = np.isclose(out[:, :3], exitw, rtol=1e-5).all(1)
    p_scat = 1 - math.exp(-dens * L)
    frac = 1 - is_exit.mean()
    assert abs(frac - p_scat) < 4 * math.sqrt(p_scat * (1 - p_scat) / len(out))
    sc = out[~is_exit]
    t = -np.log(sc[:, 0] / albedo[0]) / sa[0]                 # recover the free-flight length
    assert (t > 0).all() and (t < L).all()
    assert np.allclose(sc[:, 1], albedo[1] * np.exp(-sa[1] * t), rtol=1e-4)


def test_sss_without_density_is_the_dielectric():
    d = O.scatter(S.Dielectric(1.4), (0, 1, -1), (0, -0.6, 0.8), 550.0, (0, 0, 0), (0, 1, 0), n=64, seed=9)
    s = O.scatter(S.SSSMaterial((1, 1, 1), (0, 0, 0), ior=1.4, density=0.0), (0, 1, -1), (0, -0.6, 0.8), 550.0,
                  (0, 0, 0), (0, 1, 0), n=64, seed=9)
    assert np.array_equal(d, s)
