"""Measurement and tuning surface: the megakernel's register budget, the
per-stage wavefront timing and the counting pass's photon-gather bytes."""
import numpy as np
import pytest
import torch

from rustracer_amd import scenes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("knn", [0, 64])
def test_megakernel_occupancy_two_gives_the_same_frame(gpu_renderer, knn):
    """k_render at two waves per SIMD (at most 128 VGPRs) runs the same code
    in the same order as at the compiler's budget."""
    r = gpu_renderer
    if knn:
        sc, cam, cfg = scenes.gem_caustics(40, 32, num_photons=50_000, knn=64, gather_radius=0.15)
        r.upload(sc)
        lq = cfg["light_quad"]
        r.photon_build_from_quad(lq["q"], lq["u"], lq["v"], cfg["light_color"], 50_000, 0.15, knn=64)
        W, H, D, ph = 40, 32, 24, True
    else:
        sc, cam, cfg = scenes.box_room(48, 40)
        r.upload(sc)
        W, H, D, ph = 48, 40, cfg["max_depth"], False
    a = np.zeros((H, W, 3), np.float32)
    b = np.zeros((H, W, 3), np.float32)
    try:
        r.render_accumulate(a, cam, W, H, 4, max_depth=D, seed=3, photon_map=ph)
        r.set_megakernel_occupancy(2)
        r.render_accumulate(b, cam, W, H, 4, max_depth=D, seed=3, photon_map=ph)
    finally:
        r.set_megakernel_occupancy(0)
    assert (a.sum(-1) > 0).mean() > 0.3
    assert np.allclose(a, b, rtol=1e-5, atol=1e-7), np.abs(a - b).max()


def test_megakernel_occupancy_rejects_other_budgets(gpu_renderer):
    from rustracer_amd.renderer import RtError
    with pytest.raises(RtError):
        gpu_renderer.set_megakernel_occupancy(3)


def test_stage_times_cover_every_stage(gpu_renderer):
    r = gpu_renderer
    sc, cam, cfg = scenes.gem_caustics(40, 32, num_photons=50_000, knn=64, gather_radius=0.15)
    r.upload(sc)
    lq = cfg["light_quad"]
    r.photon_build_from_quad(lq["q"], lq["u"], lq["v"], cfg["light_color"], 50_000, 0.15, knn=64)
    acc = torch.zeros((32, 40, 3), dtype=torch.float32, device="cuda")
    r.set_timing(True)
    r.render_wavefront(acc, cam, 40, 32, 4, max_depth=6, seed=1, photon_map=True, capacity=2048)
    torch.cuda.synchronize()
    st = r.stage_times()
    ems, en = r.stage_timing()
    r.set_timing(False)
    chunks = (40 * 32 * 4 + 2047) // 2048
    for k in ("extend", "knn", "shade", "shadow"):
        ms, n = st[k]
        assert n == 6 * chunks and ms > 0.0, (k, ms, n)
    assert (ems, en) == st["extend"]
    # no k-NN map: the k-NN stage is not timed
    sc, cam, cfg = scenes.box_room(16, 16)
    r.upload(sc)
    acc = torch.zeros((16, 16, 3), dtype=torch.float32, device="cuda")
    r.set_timing(True)
    r.render_wavefront(acc, cam, 16, 16, 2, max_depth=4, seed=1)
    torch.cuda.synchronize()
    st = r.stage_times()
    r.set_timing(False)
    assert st["knn"] == (0.0, 0) and st["extend"][1] == 4


@pytest.mark.parametrize("knn", [0, 64])
def test_counting_pass_reports_photon_gather_bytes(gpu_renderer, knn):
    r = gpu_renderer
    sc, cam, cfg = scenes.gem_caustics(32, 24, num_photons=50_000, knn=knn, gather_radius=0.15)
    r.upload(sc)
    lq = cfg["light_quad"]
    r.photon_build_from_quad(lq["q"], lq["u"], lq["v"], cfg["light_color"], 50_000, 0.15, knn=knn)
    a = np.zeros((24, 32, 3), np.float32)
    r.set_counting(True)
    r.render_accumulate(a, cam, 32, 24, 2, max_depth=8, seed=1, photon_map=True)
    c = r.counters()
    r.set_counting(False)
    assert c["samples"] == 32 * 24 * 2
    # records are 32 B, powers 16 B, planes 8 B: a multiple of 8, at least one record per gather
    assert c["photon_bytes"] > 32 * c["samples"] // 4 and c["photon_bytes"] % 8 == 0
    r.set_counting(True)
    r.render_accumulate(a, cam, 32, 24, 2, max_depth=8, seed=1, photon_map=False)
    c = r.counters()
    r.set_counting(False)
    assert c["photon_bytes"] == 0
