"""Path-replay adjoint on the device (rt_backward, rt_render_detached).

* the rough spectral gem renders like the checker (same streams, 3 sigma);
* the detached renderer at theta = theta_s is the plain frame; away from
  theta_s it estimates the image at theta (3 sigma against the plain frame
  rendered after set_params(theta)), and equals the oracle's detached
  renderer sample by sample;
* the adjoint equals the checker's host gradient (forward score accumulation,
  numeric dlog f) on the same per-pixel streams;
* the adjoint equals finite differences of the detached renderer with common
  random numbers within 2 %, no absolute slack: Richardson-extrapolated
  central differences (h = 1e-3 and 5e-4 relative) for the gem's 20
  parameters, central differences (h = 1e-3) for a constant medium's sigma_t;
* the wavefront replay gives the megakernel replay's gradient.
"""
import numpy as np
import pytest

from oracle import oracle as O

import imgstats as IS
from rustracer_amd import Renderer, scenes

pytestmark = pytest.mark.gpu


def setup(r, W, H):
    sc, cam, cfg = scenes.gem_adjoint(W, H)
    r.upload(sc)
    return sc, cam, cfg


def test_rough_spectral_gem_matches_checker(gpu_renderer):
    W, H = 32, 24
    sc, cam, cfg = setup(gpu_renderer, W, H)
    orc = O.OracleScene(sc)
    g = np.zeros((4, H, W, 3), np.float32)
    c = np.zeros((4, H, W, 3), np.float32)
    for k in range(4):
        gpu_renderer.render_tiles(g[k], None, k, 0, W, H, cam, max_depth=24, seed=1)
        c[k] = orc.render_sample(cam, Renderer.params(W, H, sample_idx=k, max_depth=24, seed=1), mode=1)
    close = np.isclose(g, c, rtol=1e-3, atol=1e-4).all(axis=-1)
    assert close.mean() > 0.95, close.mean()
    def gpu(first, spp):
        out = np.zeros((H, W, 3), np.float32)
        gpu_renderer.render_accumulate(out, cam, W, H, spp, sample_idx=first, max_depth=24, seed=2)
        return out

    def cpu(first, spp):
        return orc.render_accumulate(cam, Renderer.params(W, H, sample_idx=first, max_depth=24, seed=3), spp, mode=0)
    IS.gate(IS.frames(gpu, 128, 32), IS.frames(cpu, 128, 32))


def detached_thetas(cfg):
    """Parameter vectors away from theta_s whose BSDF ratios f_theta / f_theta_s
    stay bounded: the 16 group alphas, the table alpha and the floor albedo 5 %
    off (a GGX lobe's ratio is bounded by (alpha / alpha_s)^+-2), and every
    parameter 0.5 % off.  A 5 % change of n_d or B1 is not one of them: the
    refracted half vector then moves by several lobe widths of the 0.005
    alpha facets, the ratio's distribution is heavy-tailed (the largest
    frame-mean difference grows with the frame count), and the oracle's own
    detached renderer fails the normal-theory gate against its plain frame
    just as the device's does; that case is checked sample by sample against
    the oracle instead."""
    ts, tt = cfg["theta_target"], cfg["theta_start"]
    rough = ts.copy()
    rough[2:] = tt[2:]
    return {"alphas_albedo_5pct": rough, "all_0p5pct": (ts + 0.1 * (tt - ts)).astype(np.float32)}


@pytest.mark.parametrize("which", ["alphas_albedo_5pct", "all_0p5pct"])
def test_detached_away_from_the_scene_parameters_renders_that_image(gpu_renderer, which):
    """render_detached(theta) with decisions at theta_s estimates the image at
    theta: its per-pixel means agree within the 3-sigma gate with
    render_accumulate after set_params(theta)."""
    W, H = 32, 24
    sc, cam, cfg = setup(gpu_renderer, W, H)
    P = cfg["params"]
    th_s, th = cfg["theta_target"], detached_thetas(cfg)[which]

    def detached(first, spp):
        out = np.zeros((H, W, 3), np.float32)
        gpu_renderer.render_detached(out, cam, W, H, spp, P, th, sample_idx=first, max_depth=24, seed=2)
        return out

    def plain(first, spp):
        out = np.zeros((H, W, 3), np.float32)
        gpu_renderer.render_accumulate(out, cam, W, H, spp, sample_idx=first, max_depth=24, seed=3)
        return out
    gpu_renderer.set_params(P, th_s)
    a = IS.frames(detached, 128, 32)
    gpu_renderer.set_params(P, th)
    b = IS.frames(plain, 128, 32)
    gpu_renderer.set_params(P, th_s)
    assert np.abs(a.mean - b.mean).max() > 0 and (b.mean > 0).mean() > 0.3
    IS.gate(a, b)


def test_detached_far_from_the_scene_parameters_matches_the_oracle_per_sample(gpu_renderer):
    """Every parameter 5 % off (n_d and B1 included): the device's detached
    frame equals the oracle's detached frame on the same per-pixel streams,
    sample by sample."""
    W, H = 32, 24
    sc, cam, cfg = setup(gpu_renderer, W, H)
    P = cfg["params"]
    orc = O.OracleScene(sc)
    g = np.zeros((4, H, W, 3), np.float32)
    c = np.zeros((4, H, W, 3), np.float32)
    for k in range(4):
        gpu_renderer.render_detached(g[k], cam, W, H, 1, P, cfg["theta_start"], sample_idx=k, max_depth=24, seed=5)
        c[k] = orc.render_detached(cam, Renderer.params(W, H, sample_idx=k, max_depth=24, seed=5), 1, P,
                                   cfg["theta_start"], mode=1)
    assert (g.sum(axis=0) > 0).mean() > 0.3
    close = np.isclose(g, c, rtol=1e-3, atol=1e-4).all(axis=-1)
    assert close.mean() > 0.95, close.mean()


def test_detached_at_current_parameters_is_the_plain_frame(gpu_renderer):
    W, H = 40, 32
    sc, cam, cfg = setup(gpu_renderer, W, H)
    a = np.zeros((H, W, 3), np.float32)
    b = np.zeros((H, W, 3), np.float32)
    gpu_renderer.render_accumulate(a, cam, W, H, 8, max_depth=24, seed=3)
    gpu_renderer.render_detached(b, cam, W, H, 8, cfg["params"], cfg["theta_target"], max_depth=24, seed=3)
    assert np.allclose(a, b, rtol=1e-5, atol=1e-6), np.abs(a - b).max()
    # parameters round-trip through the scene
    got = gpu_renderer.get_params(cfg["params"])
    assert np.array_equal(got, cfg["theta_target"])
    gpu_renderer.set_params(cfg["params"], cfg["theta_start"])
    assert np.array_equal(gpu_renderer.get_params(cfg["params"]), cfg["theta_start"])


@pytest.mark.parametrize("wavefront", [False, True])
def test_adjoint_matches_host_replay(gpu_renderer, wavefront):
    """Megakernel and wavefront replay against the oracle's host gradient
    (forward score accumulation, a different algorithm) on the same streams."""
    W, H, spp = 24, 24, 8
    sc, cam, cfg = setup(gpu_renderer, W, H)
    P = cfg["params"]
    w = np.random.default_rng(5).standard_normal((H, W, 3)).astype(np.float32)
    kw = dict(wavefront=True, capacity=2048, streams=2) if wavefront else {}
    g = gpu_renderer.backward(w, cam, W, H, spp, P, max_depth=24, seed=4, **kw)
    orc = O.OracleScene(sc)
    h = orc.backward(cam, Renderer.params(W, H, max_depth=24, seed=4), spp, P, w, mode=1)
    scale = np.abs(h).max()
    bad = np.abs(g - h) > 2e-2 * np.abs(h) + 2e-3 * scale
    assert not bad.any(), np.c_[np.arange(len(P)), g, h][bad]


def test_adjoint_matches_finite_differences_with_common_random_numbers(gpu_renderer):
    W, H, spp = 64, 64, 16
    sc, cam, cfg = setup(gpu_renderer, W, H)
    P = cfg["params"]
    theta = cfg["theta_start"]
    gpu_renderer.set_params(P, theta)
    target = np.zeros((H, W, 3), np.float32)
    gpu_renderer.render_detached(target, cam, W, H, spp, P, cfg["theta_target"], max_depth=24, seed=99)
    target /= spp
    img = np.zeros((H, W, 3), np.float32)
    gpu_renderer.render_accumulate(img, cam, W, H, spp, max_depth=24, seed=7)
    img /= spp
    dl = (2.0 * (img - target) / img.size).astype(np.float32)      # L = mean squared error
    g = gpu_renderer.backward(dl, cam, W, H, spp, P, max_depth=24, seed=7)

    def loss(th):
        out = np.zeros((H, W, 3), np.float32)
        gpu_renderer.render_detached(out, cam, W, H, spp, P, th, max_depth=24, seed=7)
        return float((((out / spp).astype(np.float64) - target) ** 2).mean())
    def central(j, hrel):
        h = hrel * abs(float(theta[j]))
        tp, tm = theta.astype(np.float64).copy(), theta.astype(np.float64).copy()
        tp[j] += h
        tm[j] -= h
        return (loss(tp) - loss(tm)) / (float(np.float32(tp[j])) - float(np.float32(tm[j])))
    # Richardson over h = 1e-3 and 5e-4 relative cancels the central
    # difference's O(h^2) term (the nearly specular facets curve the BSDF
    # ratio strongly in n_d); no absolute slack
    fd = np.array([(4.0 * central(j, 5e-4) - central(j, 1e-3)) / 3.0 for j in range(len(P))])
    bad = np.abs(g - fd) > 2e-2 * np.abs(fd)
    assert not bad.any(), np.c_[np.arange(len(P)), g, fd][bad]
    assert np.count_nonzero(np.abs(fd) > 1e-2 * np.abs(fd).max()) >= 10


@pytest.mark.parametrize("cap,sort,streams", [(4096, 1, 1), (1 << 20, 0, 1), (8192, 2, 1), (1024, 1, 3), (2048, 0, 4)])
def test_wavefront_adjoint_equals_megakernel_adjoint(gpu_renderer, cap, sort, streams):
    """The replay as wavefronts (primal radiance per sample, then the replay
    stages) gives the megakernel replay's gradient; capacity 4096 splits the
    frame into two chunks, capacity 1024 into nine over three streams."""
    W, H, spp = 24, 24, 8
    sc, cam, cfg = setup(gpu_renderer, W, H)
    P = cfg["params"]
    gpu_renderer.set_params(P, cfg["theta_start"])
    w = np.random.default_rng(6).standard_normal((H, W, 3)).astype(np.float32)
    a = gpu_renderer.backward(w, cam, W, H, spp, P, max_depth=24, seed=8)
    b = gpu_renderer.backward(w, cam, W, H, spp, P, max_depth=24, seed=8, wavefront=True, capacity=cap, sort=sort,
                              streams=streams)
    scale = np.abs(a).max()
    assert np.allclose(a, b, rtol=2e-3, atol=2e-4 * scale), np.c_[a, b]


def test_medium_sigma_adjoint_matches_fd_and_wavefront(gpu_renderer):
    """A constant medium's sigma_t (and a floor albedo beside it): the device
    replay (megakernel and wavefront) equals central differences of the
    device's detached renderer with common random numbers (h = 1e-3 relative)
    within 2 %.  (The checker's own score gradient equals its own differences
    in the CPU suite; the two do not share streams here, because a medium's
    free-flight stream is seeded from the ray's bits and an ulp of difference
    at the boundary reseeds the rest of the path.)"""
    W, H, spp = 24, 24, 16
    sc, cam, cfg = scenes.fog_adjoint(W, H)
    gpu_renderer.upload(sc)
    P, theta = cfg["params"], cfg["theta"]
    D = cfg["max_depth"]
    assert np.array_equal(gpu_renderer.get_params(P), theta)
    a = np.zeros((H, W, 3), np.float32)
    b = np.zeros((H, W, 3), np.float32)
    gpu_renderer.render_accumulate(a, cam, W, H, spp, max_depth=D, seed=3)
    gpu_renderer.render_detached(b, cam, W, H, spp, P, theta, max_depth=D, seed=3)
    assert np.allclose(a, b, rtol=1e-5, atol=1e-6), np.abs(a - b).max()
    w = np.random.default_rng(7).standard_normal((H, W, 3)).astype(np.float32)
    g = gpu_renderer.backward(w, cam, W, H, spp, P, max_depth=D, seed=4)
    for cap, sort in ((4096, 2), (1 << 20, 1)):
        gw = gpu_renderer.backward(w, cam, W, H, spp, P, max_depth=D, seed=4, wavefront=True, capacity=cap, sort=sort)
        assert np.allclose(gw, g, rtol=1e-3, atol=1e-4 * np.abs(g).max()), np.c_[gw, g]

    def img(th):
        out = np.zeros((H, W, 3), np.float32)
        gpu_renderer.render_detached(out, cam, W, H, spp, P, th, max_depth=D, seed=4)
        return out.astype(np.float64) / spp
    fd = np.zeros(len(P))
    for j in range(len(P)):
        hh = 1e-3 * abs(float(theta[j]))
        tp, tm = theta.astype(np.float64).copy(), theta.astype(np.float64).copy()
        tp[j] += hh
        tm[j] -= hh
        fd[j] = ((img(tp) - img(tm)) * w).sum() / (float(np.float32(tp[j])) - float(np.float32(tm[j])))
    assert np.all(np.abs(fd) > 1e-3)
    assert np.allclose(g, fd, rtol=2e-2), np.c_[g, fd]
    # the density is a scene parameter: set, read back, and it changes the frame
    gpu_renderer.set_params(P, [2.5, 0.7])
    assert np.allclose(gpu_renderer.get_params(P), [2.5, 0.7])
    c = np.zeros((H, W, 3), np.float32)
    gpu_renderer.render_accumulate(c, cam, W, H, spp, max_depth=D, seed=3)
    assert not np.allclose(a, c)
