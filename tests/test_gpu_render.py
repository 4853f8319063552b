"""Parity of the plain megakernel with the CPU oracle's render_tiles.

Two checks per scene:
  * same streams: the GPU draws one stream per (pixel, sample); the oracle's
    per-pixel mode draws the same numbers, so most pixels agree to float
    rounding (transcendentals differ by ulps between libm and the device
    library, which occasionally flips a Russian-roulette or Fresnel decision);
  * independent streams: per-pixel means of the GPU and of the oracle's
    reference row streams agree within 3 sigma (<= 0.5 % of channels beyond
    3 sigma and none beyond 6 sigma), as the north-star image test requires:
    128 frames of 32 passes per side (4,096 samples per pixel), z-scores of
    the frame means (tests/imgstats.py).  The wavefront renderer is gated the
    same way, directly against the oracle.
"""
import numpy as np
import pytest

from oracle import oracle as O
from rustracer_amd import Renderer, scenes

import imgstats as IS

pytestmark = pytest.mark.gpu

IS_K, IS_S = 128, 32      # frames per side x passes per frame of the 3-sigma gate


def passes_gpu(r, cam, W, H, K, depth, seed=0, photons=False):
    out = np.zeros((K, H, W, 3), np.float32)
    for k in range(K):
        r.render_tiles(out[k], None, k, 0, W, H, cam, max_depth=depth, seed=seed, photon_map=photons)
    return out


def passes_cpu(orc, cam, W, H, K, depth, mode, seed=0, photons=False):
    out = np.zeros((K, H, W, 3), np.float32)
    for k in range(K):
        p = Renderer.params(W, H, sample_idx=k, max_depth=depth, seed=seed, use_photons=photons)
        out[k] = orc.render_sample(cam, p, mode=mode)
    return out


def three_sigma(a, b):
    """Welch z-scores of two stacks of single-pass frames (K, H, W, 3)."""
    K = a.shape[0]
    ma, mb = a.mean(0), b.mean(0)
    va, vb = a.var(0, ddof=1), b.var(0, ddof=1)
    se = np.sqrt(va / K + vb / K)
    z = np.abs(ma - mb) / np.maximum(se, 1e-12)
    both_const = (va == 0) & (vb == 0)
    z = np.where(both_const, np.where(np.abs(ma - mb) <= 1e-6 * np.maximum(np.abs(ma), 1), 0.0, np.inf), z)
    return z


def gpu_frames(r, cam, W, H, depth, seed, photons=False, K=IS_K, S=IS_S, wavefront=False, **kw):
    """K frames of S passes each through the megakernel (or the wavefront)."""
    def render(first, s):
        out = np.zeros((H, W, 3), np.float32)
        if wavefront:
            r.render_wavefront(out, cam, W, H, s, sample_idx=first, max_depth=depth, seed=seed, photon_map=photons, **kw)
        else:
            r.render_accumulate(out, cam, W, H, s, sample_idx=first, max_depth=depth, seed=seed, photon_map=photons)
        return out
    return IS.frames(render, K, S)


def cpu_frames(orc, cam, W, H, depth, seed, photons=False, K=IS_K, S=IS_S, mode=0):
    """K frames of S passes each of the oracle (mode 0: the reference's row streams)."""
    def render(first, s):
        p = Renderer.params(W, H, sample_idx=first, max_depth=depth, seed=seed, use_photons=photons)
        return orc.render_accumulate(cam, p, s, mode=mode)
    return IS.frames(render, K, S)


def three_sigma_gate(r, orc, cam, W, H, depth, seed=2, photons=False, wavefront=False, **kw):
    """Per-pixel means of the device and of the oracle's reference row
    streams (independent samples, K x S per side) within 3 sigma: at most
    0.5 % of channels beyond 3 sigma, none beyond 6."""
    g = gpu_frames(r, cam, W, H, depth, seed, photons, wavefront=wavefront, **kw)
    c = cpu_frames(orc, cam, W, H, depth, seed + 1, photons)
    assert np.isfinite(g.mean).all()
    return IS.gate(g, c)


def test_box_room_same_streams(gpu_renderer):
    scene, cam, cfg = scenes.box_room(64, 64)
    gpu_renderer.upload(scene)
    orc = O.OracleScene(scene)
    g = passes_gpu(gpu_renderer, cam, 64, 64, 4, cfg["max_depth"])
    c = passes_cpu(orc, cam, 64, 64, 4, cfg["max_depth"], mode=1)
    close = np.isclose(g, c, rtol=1e-3, atol=1e-5).all(axis=-1)
    assert close.mean() > 0.97, close.mean()
    assert np.isfinite(g).all()


def test_box_room_three_sigma_vs_reference_streams(gpu_renderer):
    scene, cam, cfg = scenes.box_room(48, 48)
    gpu_renderer.upload(scene)
    orc = O.OracleScene(scene)
    three_sigma_gate(gpu_renderer, orc, cam, 48, 48, cfg["max_depth"], seed=1)


@pytest.mark.parametrize("sort,streams", [(1, 1), (2, 3)])
def test_box_room_wavefront_three_sigma_vs_reference_streams(gpu_renderer, sort, streams):
    """The wavefront renderer itself against the oracle's reference streams."""
    scene, cam, cfg = scenes.box_room(48, 48)
    gpu_renderer.upload(scene)
    orc = O.OracleScene(scene)
    three_sigma_gate(gpu_renderer, orc, cam, 48, 48, cfg["max_depth"], seed=4, wavefront=True, sort=sort, streams=streams,
                     capacity=16384)


def test_accumulate_equals_sum_of_passes(gpu_renderer):
    scene, cam, cfg = scenes.box_room(32, 32)
    gpu_renderer.upload(scene)
    acc = np.zeros((32, 32, 3), np.float32)
    gpu_renderer.render_accumulate(acc, cam, 32, 32, 6, max_depth=cfg["max_depth"])
    g = passes_gpu(gpu_renderer, cam, 32, 32, 6, cfg["max_depth"])
    assert np.allclose(acc, g.sum(0), rtol=1e-5, atol=1e-6)


def test_converged_pixels_are_black(gpu_renderer):
    scene, cam, cfg = scenes.box_room(16, 16)
    gpu_renderer.upload(scene)
    conv = np.zeros((16, 16), np.uint8)
    conv[3:7, 2:9] = 1
    out = np.full((16, 16, 3), 7.0, np.float32)
    gpu_renderer.render_tiles(out, conv, 0, 0, 16, 16, cam, max_depth=cfg["max_depth"])
    assert np.all(out[conv == 1] == 0.0)


def test_device_buffer_output_matches_host(gpu_renderer):
    torch = pytest.importorskip("torch")
    scene, cam, cfg = scenes.box_room(32, 32)
    gpu_renderer.upload(scene)
    host = np.zeros((32, 32, 3), np.float32)
    gpu_renderer.render_tiles(host, None, 3, 0, 32, 32, cam, max_depth=cfg["max_depth"])
    dev = torch.zeros((32, 32, 3), dtype=torch.float32, device="cuda")
    gpu_renderer.render_tiles(dev, None, 3, 0, 32, 32, cam, max_depth=cfg["max_depth"])
    assert np.array_equal(dev.cpu().numpy(), host)


def small_env_scene(W=48, H=48, sampling="cdf"):
    from rustracer_amd import scene as S
    v, n = scenes.displaced_icosphere(3, amplitude=0.05)
    env = scenes.sky_envmap(128, 64, sun_sigma_deg=4.0, sun_peak=200.0)
    sc = S.Scene([S.Mesh(v, S.RoughDielectric(1.5, 0.15), normals=n),
                  S.Quad((-10, -1.1, -10), (20, 0, 0), (0, 0, 20), S.Lambertian((0.8, 0.8, 0.8)))],
                 [], S.Background.envmap(env, sampling=sampling), name="small_env")
    cam = S.Camera((0.0, 0.6, -3.6), (0.0, -0.1, 0.0), (0, 1, 0), 40.0, W / H)
    return sc, cam


@pytest.mark.parametrize("sampling", ["cdf", "alias"])
def test_env_mesh_rough_dielectric_same_streams_and_three_sigma(gpu_renderer, sampling):
    sc, cam = small_env_scene(sampling=sampling)
    gpu_renderer.upload(sc)
    orc = O.OracleScene(sc)
    g = passes_gpu(gpu_renderer, cam, 48, 48, 4, 12)
    c = passes_cpu(orc, cam, 48, 48, 4, 12, mode=1)
    close = np.isclose(g, c, rtol=1e-3, atol=1e-4).all(axis=-1)
    assert close.mean() > 0.9, close.mean()
    three_sigma_gate(gpu_renderer, orc, cam, 32, 32, 12, seed=2)


def gem_small(W, H, photons):
    sc, cam, cfg = scenes.gem_caustics(W, H, num_photons=photons)
    return sc, cam, cfg


def test_photon_map_matches_checker_and_brute_force(gpu_renderer):
    from rustracer_amd import _ffi as F
    sc, cam, cfg = gem_small(32, 32, 200000)
    gpu_renderer.upload(sc)
    lq = cfg["light_quad"]
    ng = gpu_renderer.photon_build_from_quad(lq["q"], lq["u"], lq["v"], cfg["light_color"], 200000, cfg["gather_radius"])
    orc = O.OracleScene(sc)
    pp = F.rt_photon_params()
    pp.emitter = F.EMIT_QUAD
    for k in range(3):
        pp.origin[k], pp.u[k], pp.v[k], pp.color[k] = lq["q"][k], lq["u"][k], lq["v"][k], cfg["light_color"][k]
    pp.num_photons, pp.gather_radius = 200000, cfg["gather_radius"]
    nc = orc.photon_build(pp)
    assert ng > 500 and abs(ng - nc) <= 0.02 * nc + 5, (ng, nc)
    ph = gpu_renderer.photon_export()
    # same photon positions as the checker (same per-photon streams), up to ulps
    pc = orc.photon_dump(nc)
    a = np.sort(ph[:, 0] * 1e3 + ph[:, 2])
    b = np.sort(pc[:, 0] * 1e3 + pc[:, 2])
    m = min(len(a), len(b))
    assert np.mean(np.isclose(a[:m], b[:m], atol=1e-3)) > 0.95
    # kd gather equals the O(N) brute-force Epanechnikov sum over the same photons
    g = np.random.default_rng(4)
    q = np.zeros((400, 6), np.float32)
    q[:, [0, 2]] = g.uniform(-0.6, 0.6, size=(400, 2))
    q[:, 4] = 1.0
    irr = gpu_renderer.photon_irradiance(q)
    r2 = cfg["gather_radius"] ** 2
    d2 = ((ph[None, :, :3] - q[:, None, :3]) ** 2).sum(-1)
    w = np.where((d2 < r2) & ((ph[None, :, 6:9] * q[:, None, 3:6]).sum(-1) > 0), 1 - d2 / r2, 0)
    brute = (w @ ph[:, 3:6]) * (2.0 / (np.pi ** 2 * r2))
    assert np.allclose(irr, brute, rtol=1e-4, atol=1e-6)
    assert (irr.sum(1) > 0).mean() > 0.2


def test_gem_with_photons_three_sigma(gpu_renderer):
    from rustracer_amd import _ffi as F
    sc, cam, cfg = gem_small(32, 32, 100000)
    gpu_renderer.upload(sc)
    lq = cfg["light_quad"]
    gpu_renderer.photon_build_from_quad(lq["q"], lq["u"], lq["v"], cfg["light_color"], 100000, cfg["gather_radius"])
    orc = O.OracleScene(sc)
    pp = F.rt_photon_params()
    pp.emitter = F.EMIT_QUAD
    for k in range(3):
        pp.origin[k], pp.u[k], pp.v[k], pp.color[k] = lq["q"][k], lq["u"][k], lq["v"][k], cfg["light_color"][k]
    pp.num_photons, pp.gather_radius = 100000, cfg["gather_radius"]
    orc.photon_build(pp)
    three_sigma_gate(gpu_renderer, orc, cam, 32, 32, 24, seed=3, photons=True)


@pytest.mark.parametrize("sort,graph,cap,occ,streams,block", [
    (1, False, 0, 0, 1, 0), (0, False, 4096, 4, 1, 64), (1, True, 8192, 5, 1, 256), (1, True, 2048, 0, 1, 0),
    (2, False, 4096, 0, 1, 0), (2, True, 2048, 4, 1, 64), (1, False, 2048, 0, 2, 256), (2, True, 2048, 0, 2, 0),
    (1, True, 4096, 0, 2, 0), (1, True, 2048, 0, 3, 64), (0, False, 2048, 0, 4, 0), (1, True, 4096, 5, 4, 256)])
def test_wavefront_equals_megakernel(gpu_renderer, sort, graph, cap, occ, streams, block):
    """Same per-pixel streams, same bounce code: the wavefront frame equals the
    megakernel frame up to the order of float atomic additions.  Capacity 2048
    with the graph replays one captured chunk seven times; with n streams the
    chunks go round-robin over n captured chunks (fewer when there are fewer
    full chunks than streams) and the remainder follows on the next lane."""
    scene, cam, cfg = scenes.box_room(48, 40)
    gpu_renderer.upload(scene)
    a = np.zeros((40, 48, 3), np.float32)
    b = np.zeros((40, 48, 3), np.float32)
    gpu_renderer.render_accumulate(a, cam, 48, 40, 8, max_depth=cfg["max_depth"], seed=5)
    gpu_renderer.render_wavefront(b, cam, 48, 40, 8, max_depth=cfg["max_depth"], seed=5, sort=sort, use_graph=graph,
                                  capacity=cap, occupancy=occ, streams=streams, block=block)
    assert np.allclose(a, b, rtol=1e-4, atol=1e-5), np.abs(a - b).max()


@pytest.mark.parametrize("knn_wave_max", [1, 1 << 30])
def test_wavefront_knn_search_regimes_equal_megakernel(gpu_renderer, knn_wave_max):
    """The dense k-NN queries searched one per lane (heap in LDS) and one per
    wave (sorted set across the lanes) give the megakernel's frame."""
    sc, cam, cfg = scenes.gem_caustics(40, 32, num_photons=100_000, knn=64, gather_radius=0.15)
    gpu_renderer.upload(sc)
    lq = cfg["light_quad"]
    gpu_renderer.photon_build_from_quad(lq["q"], lq["u"], lq["v"], cfg["light_color"], 100_000, 0.15, knn=64)
    a = np.zeros((32, 40, 3), np.float32)
    b = np.zeros((32, 40, 3), np.float32)
    gpu_renderer.render_accumulate(a, cam, 40, 32, 8, max_depth=24, seed=6, photon_map=True)
    gpu_renderer.render_wavefront(b, cam, 40, 32, 8, max_depth=24, seed=6, photon_map=True, capacity=4096,
                                  knn_wave_max=knn_wave_max, block=64)
    assert (a.sum(-1) > 0).mean() > 0.3
    assert np.allclose(a, b, rtol=1e-4, atol=1e-5), np.abs(a - b).max()


@pytest.mark.parametrize("sort", [0, 2])
def test_wavefront_gem_with_photons_equals_megakernel(gpu_renderer, sort):
    sc, cam, cfg = gem_small(40, 32, 50000)
    gpu_renderer.upload(sc)
    lq = cfg["light_quad"]
    gpu_renderer.photon_build_from_quad(lq["q"], lq["u"], lq["v"], cfg["light_color"], 50000, cfg["gather_radius"])
    a = np.zeros((32, 40, 3), np.float32)
    b = np.zeros((32, 40, 3), np.float32)
    gpu_renderer.render_accumulate(a, cam, 40, 32, 16, max_depth=24, seed=6, photon_map=True)
    gpu_renderer.render_wavefront(b, cam, 40, 32, 16, max_depth=24, seed=6, photon_map=True, capacity=8192, sort=sort)
    assert np.allclose(a, b, rtol=1e-4, atol=1e-5), np.abs(a - b).max()


@pytest.mark.parametrize("which,sort", [("fog", 1), ("fog", 2), ("volume", 0)])
def test_wavefront_with_media_equals_megakernel(gpu_renderer, which, sort):
    """Media draw their free-flight distance from a stream seeded by the ray
    (origin, direction, wavelength), so the wavefront's extend stage must
    rebuild the ray whole to take the megakernel's medium events."""
    if which == "fog":
        sc, cam, cfg = scenes.fog_adjoint(32, 24)
        D = cfg["max_depth"]
    else:
        sc, cam, cfg = scenes.volume_glass_box(32, 24, grid_n=16)
        D = 24
    gpu_renderer.upload(sc)
    a = np.zeros((24, 32, 3), np.float32)
    b = np.zeros((24, 32, 3), np.float32)
    gpu_renderer.render_accumulate(a, cam, 32, 24, 8, max_depth=D, seed=4)
    gpu_renderer.render_wavefront(b, cam, 32, 24, 8, max_depth=D, seed=4, capacity=4096, sort=sort)
    assert np.allclose(a, b, rtol=1e-4, atol=1e-5), np.abs(a - b).max()


def test_volume_glass_box_same_streams_and_three_sigma(gpu_renderer):
    """Dielectric box around a delta-tracked grid medium under the HDR sky."""
    sc, cam, cfg = scenes.volume_glass_box(32, 24, grid_n=16)
    gpu_renderer.upload(sc)
    orc = O.OracleScene(sc)
    g = passes_gpu(gpu_renderer, cam, 32, 24, 4, 24, seed=1)
    c = passes_cpu(orc, cam, 32, 24, 4, 24, mode=1, seed=1)
    # the medium's tracking stream is seeded from the ray's bits, so a 1-ulp
    # difference at the box entry point reseeds a whole volume path
    close = np.isclose(g, c, rtol=1e-3, atol=1e-4).all(axis=-1)
    assert close.mean() > 0.75, close.mean()
    three_sigma_gate(gpu_renderer, orc, cam, 32, 24, 24, seed=2)


@pytest.mark.parametrize("graph,cap,streams", [(False, 4096, 1), (True, 2048, 1), (True, 2048, 2), (False, 2048, 2),
                                               (True, 2048, 4)])
def test_batch_equals_individual_wavefront_renders(gpu_renderer, graph, cap, streams):
    """One batched launch over views of uneven resolution, spp, depth, seed and
    scene (gem spin stills with their own photon maps, plus the volume box)
    gives each view the frame its own wavefront render gives."""
    r = gpu_renderer
    views = []
    try:
        for k, (W, H) in enumerate([(24, 16), (40, 32), (16, 16)]):
            sc, cam, cfg = scenes.gem_caustics(W, H, num_photons=20000, spin_deg=30.0 * k)
            h = r.upload_extra(sc)
            lq = cfg["light_quad"]
            r.photon_build_from_quad(lq["q"], lq["u"], lq["v"], cfg["light_color"], 20000, cfg["gather_radius"], scene=h)
            views.append(dict(scene=h, camera=cam, width=W, height=H, spp=3 + k, max_depth=24, photon_map=True,
                              seed=7 + k))
        sc, cam, cfg = scenes.volume_glass_box(32, 24, grid_n=16)
        views.append(dict(scene=r.upload_extra(sc), camera=cam, width=32, height=24, spp=5, max_depth=12, seed=11))
        for v in views:
            v["out"] = np.zeros((v["height"], v["width"], 3), np.float32)
        r.render_batch(views, capacity=cap, use_graph=graph, streams=streams)
        for v in views:
            ref = np.zeros_like(v["out"])
            r.render_wavefront(ref, v["camera"], v["width"], v["height"], v["spp"], max_depth=v["max_depth"],
                               photon_map=v.get("photon_map", False), seed=v["seed"], capacity=cap, scene=v["scene"])
            assert np.isfinite(v["out"]).all()
            assert np.allclose(v["out"], ref, rtol=1e-4, atol=1e-4), np.abs(v["out"] - ref).max()
            assert (v["out"].sum(-1) > 0).mean() > 0.5
    finally:
        for v in views:
            r.destroy(v["scene"])


def test_batch_rejects_bad_views(gpu_renderer):
    from rustracer_amd.renderer import RtError
    sc, cam, cfg = scenes.box_room(16, 16)
    r = gpu_renderer.upload(sc)
    out = np.zeros((16, 16, 3), np.float32)
    with pytest.raises(RtError):
        r.render_batch([dict(scene=r.scene, camera=cam, width=16, height=16, spp=0, out=out)])
    with pytest.raises(RtError):
        r.render_batch([dict(scene=r.scene, camera=cam, width=16, height=16, spp=2, out=out, photon_map=True)])


def test_disk_light_and_analytic_shapes_same_streams_and_three_sigma(gpu_renderer):
    """NEE through a disk light (Disk::pdf_value / pdf_generate) over a floor,
    a cylinder, a cone and a rippled plane."""
    from rustracer_amd import scene as S
    light = S.DiffuseLight((6.0, 6.0, 6.0))
    w = S.Lambertian((0.7, 0.6, 0.5))
    objs = [S.InfinitePlane((0, 0, 0), (0, 1, 0), S.Lambertian((0.8, 0.8, 0.8)), wave_amplitude=0.03, wave_scale=5.0),
            S.Cylinder((-0.5, 0, 0.2), 0.25, 0.6, w), S.Cone((0.4, 0, 0.1), 0.3, 0.7, S.Metal((0.9, 0.8, 0.7), 0.2)),
            S.Disk((0, 1.5, 0), (0, -1, 0), 0.4, light)]
    sc = S.Scene(objs, [S.Disk((0, 1.5, 0), (0, -1, 0), 0.4, light)], S.Background.solid((0.05, 0.05, 0.08)))
    cam = S.Camera((0.0, 0.9, -2.5), (0.0, 0.3, 0.0), (0, 1, 0), 40.0, 1.0)
    gpu_renderer.upload(sc)
    orc = O.OracleScene(sc)
    g = passes_gpu(gpu_renderer, cam, 40, 40, 4, 8)
    c = passes_cpu(orc, cam, 40, 40, 4, 8, mode=1)
    close = np.isclose(g, c, rtol=1e-3, atol=1e-5).all(axis=-1)
    assert close.mean() > 0.95, close.mean()
    three_sigma_gate(gpu_renderer, orc, cam, 32, 32, 8, seed=2)


def test_noise_medium_three_sigma(gpu_renderer):
    from rustracer_amd import scene as S
    b = S.Sphere((0, 0.6, 0), 0.6, None)
    S.NoiseMedium(b, (0.9, 0.8, 0.7), 4.0, noise_scale=3.0, threshold=0.1, g=0.3, seed=5)
    sc = S.Scene([S.Quad((-3, 0, -3), (6, 0, 0), (0, 0, 6), S.Lambertian((0.8, 0.8, 0.8))), b],
                 [], S.Background.gradient((1, 1, 1), (0.3, 0.4, 0.6)))
    cam = S.Camera((0, 1.0, -3.0), (0, 0.5, 0), (0, 1, 0), 35.0, 1.0)
    gpu_renderer.upload(sc)
    orc = O.OracleScene(sc)
    three_sigma_gate(gpu_renderer, orc, cam, 32, 32, 12, seed=2)


def test_imported_host_photon_map_gives_the_checkers_frame(gpu_renderer):
    """With the checker's own photons imported, the photon gather and the
    same-stream frames match the checker's (the map is no longer a separate
    random draw)."""
    from rustracer_amd import _ffi as F
    sc, cam, cfg = gem_small(32, 32, 100000)
    gpu_renderer.upload(sc)
    orc = O.OracleScene(sc)
    lq = cfg["light_quad"]
    pp = F.rt_photon_params()
    pp.emitter = F.EMIT_QUAD
    for k in range(3):
        pp.origin[k], pp.u[k], pp.v[k], pp.color[k] = lq["q"][k], lq["u"][k], lq["v"][k], cfg["light_color"][k]
    pp.num_photons, pp.gather_radius = 100000, cfg["gather_radius"]
    n = orc.photon_build(pp)
    ph = orc.photon_dump(n)
    assert gpu_renderer.photon_import(ph, cfg["gather_radius"]) == n
    back = gpu_renderer.photon_export()
    assert np.array_equal(np.sort(back, axis=0), np.sort(ph, axis=0))
    g = np.random.default_rng(5)
    q = np.zeros((500, 6), np.float32)
    q[:, [0, 2]] = g.uniform(-0.6, 0.6, size=(500, 2))
    q[:, 4] = 1.0
    assert np.allclose(gpu_renderer.photon_irradiance(q), orc.photon_irradiance(q), rtol=1e-4, atol=1e-6)
    gp = passes_gpu(gpu_renderer, cam, 32, 32, 4, 24, seed=4, photons=True)
    cp = passes_cpu(orc, cam, 32, 32, 4, 24, mode=1, seed=4, photons=True)
    close = np.isclose(gp, cp, rtol=1e-3, atol=1e-4).all(axis=-1)
    assert close.mean() > 0.95, close.mean()


def test_blackbody_light_and_sss_same_streams_and_three_sigma(gpu_renderer):
    """A blackbody area light (spectral paths through a dispersive sphere see
    its Planck weight, others its white-point average) over an SSS sphere."""
    from rustracer_amd import scene as S
    bb = S.BlackbodyLight(3200.0, 8.0)
    lq = S.Quad((-0.3, 1.6, -0.3), (0.6, 0, 0), (0, 0, 0.6), bb)
    objs = [S.Quad((-4, 0, -4), (8, 0, 0), (0, 0, 8), S.Lambertian((0.8, 0.8, 0.8))),
            S.Sphere((-0.45, 0.35, 0.1), 0.35, S.SSSMaterial((0.95, 0.8, 0.7), (0.3, 0.8, 1.2), 1.35, 3.0, 0.2)),
            S.Sphere((0.45, 0.35, 0.0), 0.35, S.SpectralDielectric(1.6, 0.02)), lq]
    sc = S.Scene(objs, [S.Quad((-0.3, 1.6, -0.3), (0.6, 0, 0), (0, 0, 0.6), bb)], S.Background.solid((0.02, 0.02, 0.03)))
    cam = S.Camera((0.0, 1.0, -2.6), (0.0, 0.4, 0.0), (0, 1, 0), 40.0, 1.0)
    gpu_renderer.upload(sc)
    orc = O.OracleScene(sc)
    g = passes_gpu(gpu_renderer, cam, 40, 40, 4, 16)
    c = passes_cpu(orc, cam, 40, 40, 4, 16, mode=1)
    close = np.isclose(g, c, rtol=1e-3, atol=1e-5).all(axis=-1)
    assert close.mean() > 0.95, close.mean()
    three_sigma_gate(gpu_renderer, orc, cam, 32, 32, 16, seed=2)


def test_textures_same_streams_and_three_sigma(gpu_renderer):
    """Image, noise, mottle and checker albedo textures and a checker-textured
    area light (emission and NEE both read the light's (u, v, p))."""
    from rustracer_amd import scene as S
    img = ((np.arange(6 * 8 * 3).reshape(6, 8, 3) * 37) % 256).astype(np.uint8)
    emit = S.DiffuseLight(S.CheckerTexture((6.0, 5.0, 4.0), (2.0, 3.0, 6.0), 5.0))
    lq = S.Quad((-0.4, 1.8, -0.4), (0.8, 0, 0), (0, 0, 0.8), emit)
    objs = [S.Quad((-4, 0, -4), (8, 0, 0), (0, 0, 8), S.Lambertian(S.MottleTexture(2.0, (0.3, 0.3, 0.32), (0.7, 0.6, 0.5), seed=4))),
            S.Quad((-1.5, 0, 1.2), (3, 0, 0), (0, 2, 0), S.Lambertian(S.ImageTexture(img))),
            S.Sphere((-0.5, 0.4, 0.2), 0.4, S.Lambertian(S.NoiseTexture(3.0, seed=6))),
            S.Sphere((0.5, 0.35, 0.0), 0.35, S.Checker((0.9, 0.9, 0.9), (0.1, 0.2, 0.6), 6.0)), lq]
    sc = S.Scene(objs, [lq], S.Background.solid((0.05, 0.05, 0.06)))
    cam = S.Camera((0.0, 1.0, -2.6), (0.0, 0.6, 0.0), (0, 1, 0), 45.0, 1.0)
    gpu_renderer.upload(sc)
    orc = O.OracleScene(sc)
    g = passes_gpu(gpu_renderer, cam, 40, 40, 4, 8)
    c = passes_cpu(orc, cam, 40, 40, 4, 8, mode=1)
    close = np.isclose(g, c, rtol=1e-3, atol=1e-5).all(axis=-1)
    assert close.mean() > 0.95, close.mean()
    three_sigma_gate(gpu_renderer, orc, cam, 32, 32, 8, seed=2)


def test_texture_upload_rejections(gpu_renderer):
    from rustracer_amd import scene as S
    bad = S.Dielectric(1.5)
    bad.texture = S.NoiseTexture(1.0)
    with pytest.raises(RuntimeError):
        gpu_renderer.upload(S.Scene([S.Sphere((0, 0, 0), 1.0, bad)]))
    t = S.perlin_tables(0)
    t[800] = 300.0
    with pytest.raises(RuntimeError):
        gpu_renderer.upload(S.Scene([S.Sphere((0, 0, 0), 1.0, S.Lambertian(S.NoiseTexture(1.0, tables=t)))]))


def test_knn_photon_gather_matches_checker_and_megakernel(gpu_renderer):
    """k-NN gather (extension) over the checker's imported photons: the
    irradiance queries, the megakernel frame (LDS heap in the kernel) and the
    wavefront frame (radii from the k_wf_knn stage) against the checker."""
    from rustracer_amd import _ffi as F
    sc, cam, cfg = gem_small(32, 32, 200000)
    gpu_renderer.upload(sc)
    orc = O.OracleScene(sc)
    lq = cfg["light_quad"]
    pp = F.rt_photon_params()
    pp.emitter = F.EMIT_QUAD
    for k in range(3):
        pp.origin[k], pp.u[k], pp.v[k], pp.color[k] = lq["q"][k], lq["u"][k], lq["v"][k], cfg["light_color"][k]
    pp.num_photons, pp.gather_radius = 200000, cfg["gather_radius"]
    n = orc.photon_build(pp)
    ph = orc.photon_dump(n)
    gpu_renderer.photon_import(ph, cfg["gather_radius"])
    g = np.random.default_rng(6)
    q = np.zeros((600, 6), np.float32)
    near = ph[g.integers(0, n, 300), :3]
    q[:300, [0, 2]] = near[:, [0, 2]] + g.normal(0, 0.05, size=(300, 2))
    q[300:, [0, 2]] = g.uniform(-1.0, 1.0, size=(300, 2))
    q[:, 4] = 1.0
    for k in (1, 8, 64):
        gpu_renderer.photon_set_gather(0.25, k)
        orc.photon_set_gather(0.25, k)
        a, b = gpu_renderer.photon_irradiance(q), orc.photon_irradiance(q)
        assert np.allclose(a, b, rtol=1e-4, atol=1e-6), (k, np.abs(a - b).max())
        assert k == 1 or (a.sum(1) > 0).mean() > 0.3
    gpu_renderer.photon_set_gather(0.25, 64)
    orc.photon_set_gather(0.25, 64)
    gp = passes_gpu(gpu_renderer, cam, 32, 32, 4, 24, seed=4, photons=True)
    cp = passes_cpu(orc, cam, 32, 32, 4, 24, mode=1, seed=4, photons=True)
    close = np.isclose(gp, cp, rtol=1e-3, atol=1e-4).all(axis=-1)
    assert close.mean() > 0.95, close.mean()
    a = np.zeros((32, 32, 3), np.float32)
    b = np.zeros((32, 32, 3), np.float32)
    gpu_renderer.render_accumulate(a, cam, 32, 32, 16, max_depth=24, seed=7, photon_map=True)
    for sort in (0, 1, 2):
        b[:] = 0
        gpu_renderer.render_wavefront(b, cam, 32, 32, 16, max_depth=24, seed=7, photon_map=True, capacity=4096, sort=sort)
        assert np.allclose(a, b, rtol=1e-4, atol=1e-5), (sort, np.abs(a - b).max())
    from rustracer_amd.renderer import RtError
    with pytest.raises(RtError, match="fixed radius"):
        gpu_renderer.backward(np.zeros((32, 32, 3), np.float32), cam, 32, 32, 4, [(0, 0, 1)], max_depth=8, photon_map=True)


@pytest.mark.parametrize("sky", [True, False])
def test_pearl_same_streams_and_three_sigma(gpu_renderer, sky):
    """PearlMaterial: under a Physical sky the body glow follows the pearl
    sun (and the luster gets sun NEE); under a quad light it follows the view."""
    from rustracer_amd import scene as S
    pearl = S.PearlMaterial(film_scale=4.0)
    objs = [S.Quad((-4, 0, -4), (8, 0, 0), (0, 0, 8), S.Lambertian((0.6, 0.6, 0.6))),
            S.Sphere((0.0, 0.4, 0.0), 0.4, pearl),
            S.Sphere((0.7, 0.25, -0.3), 0.25, S.PearlMaterial(base_color=(0.2, 0.2, 0.25), luster_roughness=0.2))]
    lights = []
    if sky:
        bg = S.Background.physical(30.0, 35.0, 3.0)
    else:
        lq = S.Quad((-0.3, 1.8, -0.3), (0.6, 0, 0), (0, 0, 0.6), S.DiffuseLight((8.0, 8.0, 8.0)))
        objs.append(lq)
        lights = [lq]
        bg = S.Background.solid((0.05, 0.05, 0.06))
    sc = S.Scene(objs, lights, bg)
    cam = S.Camera((0.0, 1.0, -2.4), (0.0, 0.35, 0.0), (0, 1, 0), 40.0, 1.0)
    gpu_renderer.upload(sc)
    orc = O.OracleScene(sc)
    g = passes_gpu(gpu_renderer, cam, 40, 40, 4, 12)
    c = passes_cpu(orc, cam, 40, 40, 4, 12, mode=1)
    close = np.isclose(g, c, rtol=1e-3, atol=1e-4).all(axis=-1)
    assert close.mean() > 0.95, close.mean()
    three_sigma_gate(gpu_renderer, orc, cam, 32, 32, 12, seed=2)


def test_pbr_maps_and_film_same_streams_and_three_sigma(gpu_renderer):
    """PbrMaterial maps (albedo, roughness, metallic, ambient occlusion,
    normal on a tangent-carrying mesh) and a thin-film clearcoat."""
    from rustracer_amd import scene as S
    g = np.random.default_rng(11)
    alb = g.integers(0, 256, size=(8, 8, 3), dtype=np.uint8)
    rough = g.integers(20, 240, size=(8, 8, 3), dtype=np.uint8)
    metal = np.zeros((4, 4, 3), np.uint8)
    metal[::2, ::2] = 255
    ao = g.integers(120, 256, size=(4, 4, 3), dtype=np.uint8)
    nmap = np.zeros((6, 6, 3), np.uint8)
    nmap[..., 0] = g.integers(90, 170, size=(6, 6))
    nmap[..., 1] = g.integers(90, 170, size=(6, 6))
    nmap[..., 2] = 230
    wall = S.PbrMaterial(albedo_map=alb, roughness_map=rough, metallic_map=metal, ao_map=ao, normal_map=nmap)
    # a wall of two triangles with uvs and tangents along +x
    v = np.array([[[-1.2, 0, 0.9], [1.2, 0, 0.9], [1.2, 1.6, 0.9]], [[-1.2, 0, 0.9], [1.2, 1.6, 0.9], [-1.2, 1.6, 0.9]]], np.float32)
    uv = np.array([[[0, 0], [1, 0], [1, 1]], [[0, 0], [1, 1], [0, 1]]], np.float32)
    tang = np.tile(np.array([1, 0, 0], np.float32), (2, 3, 1))
    nrm = np.tile(np.array([0, 0, -1], np.float32), (2, 3, 1))
    objs = [S.Quad((-4, 0, -4), (8, 0, 0), (0, 0, 8), S.PbrMaterial(albedo_map=alb, roughness=0.6)),
            S.Mesh(v, wall, normals=nrm, tangents=tang, uvs=uv),
            S.Sphere((0.0, 0.35, 0.0), 0.35, S.PbrMaterial(albedo=(0.1, 0.1, 0.12), roughness=0.4, clearcoat=1.0,
                                                            clearcoat_roughness=0.05, film_thickness=520.0, film_ior=1.45))]
    lq = S.Quad((-0.3, 1.8, -0.6), (0.6, 0, 0), (0, 0, 0.6), S.DiffuseLight((10.0, 10.0, 10.0)))
    sc = S.Scene(objs + [lq], [lq], S.Background.gradient((0.3, 0.4, 0.6), (0.9, 0.9, 0.9)))
    cam = S.Camera((0.0, 0.9, -2.4), (0.0, 0.5, 0.5), (0, 1, 0), 45.0, 1.0)
    gpu_renderer.upload(sc)
    orc = O.OracleScene(sc)
    gp = passes_gpu(gpu_renderer, cam, 40, 40, 4, 12)
    cp = passes_cpu(orc, cam, 40, 40, 4, 12, mode=1)
    close = np.isclose(gp, cp, rtol=1e-3, atol=1e-4).all(axis=-1)
    assert close.mean() > 0.95, close.mean()
    three_sigma_gate(gpu_renderer, orc, cam, 32, 32, 12, seed=2)


def test_toml_scene_renders_like_the_checker(gpu_renderer, tmp_path):
    """A scene file of the reference's TOML format with every shape, material,
    light and volume kind (rustracer_amd.scene_file) uploads and renders on
    the device like the checker on the same per-pixel streams (media reseed
    their tracking stream from the ray's bits, so a few pixels diverge)."""
    from rustracer_amd import scene_file as SF
    from tests.test_scene_file import write_scene
    sc, camera, info = SF.load(write_scene(tmp_path))
    cam = camera(48 / 32)
    gpu_renderer.upload(sc)
    orc = O.OracleScene(sc)
    g = passes_gpu(gpu_renderer, cam, 48, 32, 2, 12, seed=3)
    c = passes_cpu(orc, cam, 48, 32, 2, 12, mode=1, seed=3)
    assert np.isfinite(g).all()
    close = np.isclose(g, c, rtol=1e-3, atol=1e-4).all(axis=-1)
    assert close.mean() > 0.85, close.mean()
