"""Camera and time features on the device against the oracle.

* thin lens: Camera::get_ray's lens offset (src/camera.rs:47-63) with
  lens_radius > 0, drawn from the unit disk (blades 0) or from a 6-blade
  polygon (sample_polygon, src/camera.rs:66-79);
* shutter time: render_tiles' frame_time draw (src/renderer.rs:721-724) and
  the default uniform time over [0, 1) (Camera::get_ray's time draw);
* motion: a moving sphere (Sphere::center_at, src/sphere.rs:30-34), one of
  them an emissive sphere in the NEE light list, so the light's pdf_value /
  pdf_generate (src/sphere.rs:87-110) are evaluated at the ray's time.

Each case is checked twice, as the other render tests are: per-pixel streams
shared with the oracle (most pixels agree to float rounding), and per-pixel
means against the oracle's reference row streams within the 3-sigma gate.
"""
import numpy as np
import pytest

from oracle import oracle as O
from rustracer_amd import Renderer
from rustracer_amd.scene import (Background, Camera, Dielectric, DiffuseLight, Lambertian, Metal, Quad, Scene,
                                 Sphere)

import imgstats as IS

pytestmark = pytest.mark.gpu

DEPTH = 8


def motion_room(W, H, aperture, blades):
    """The box room with a sphere moving 0.2 up over t in [0, 1), a small
    emissive sphere moving sideways and listed as a light, and a lens focused
    on the glass sphere."""
    white = Lambertian((0.73, 0.73, 0.73))
    light = DiffuseLight((15.0, 15.0, 15.0))
    glow = DiffuseLight((4.0, 3.0, 1.5))
    lq = Quad((0.375, 0.999, 0.375), (0.25, 0, 0), (0, 0, 0.25), light)
    lamp = Sphere((0.20, 0.75, 0.60), 0.06, glow, center1=(0.45, 0.75, 0.60))
    objs = [
        Quad((0, 0, 0), (1, 0, 0), (0, 0, 1), white),
        Quad((0, 1, 0), (1, 0, 0), (0, 0, 1), white),
        Quad((0, 0, 1), (1, 0, 0), (0, 1, 0), white),
        Quad((0, 0, 0), (0, 0, 1), (0, 1, 0), Lambertian((0.65, 0.05, 0.05))),
        Quad((1, 0, 0), (0, 0, 1), (0, 1, 0), Lambertian((0.12, 0.45, 0.15))),
        Sphere((0.30, 0.18, 0.55), 0.18, Lambertian((0.2, 0.3, 0.8)), center1=(0.30, 0.38, 0.55)),
        Sphere((0.70, 0.18, 0.35), 0.18, Dielectric(1.5)),
        Sphere((0.55, 0.10, 0.15), 0.10, Metal((0.9, 0.8, 0.6), 0.05)),
        lq, lamp,
    ]
    lights = [Quad((0.375, 0.999, 0.375), (0.25, 0, 0), (0, 0, 0.25), light),
              Sphere((0.20, 0.75, 0.60), 0.06, glow, center1=(0.45, 0.75, 0.60))]
    sc = Scene(objs, lights, Background.solid((0, 0, 0)), name="motion_room")
    lookfrom = np.array([0.5, 0.5, -1.4])
    focus = float(np.linalg.norm(lookfrom - np.array([0.70, 0.18, 0.35])))
    cam = Camera(tuple(lookfrom), (0.5, 0.5, 0.5), (0, 1, 0), 40.0, W / H, aperture=aperture, focus_dist=focus,
                 blades=blades)
    return sc, cam


CASES = {
    "disk_lens": dict(aperture=0.12, blades=0, frame_time=None),
    "hexagon_lens_frame_time": dict(aperture=0.12, blades=6, frame_time=(0.25, 0.75)),
    "pinhole_default_time": dict(aperture=0.0, blades=0, frame_time=None),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_camera_and_motion_same_streams(gpu_renderer, case):
    c = CASES[case]
    W, H = 48, 36
    sc, cam = motion_room(W, H, c["aperture"], c["blades"])
    gpu_renderer.upload(sc)
    orc = O.OracleScene(sc)
    g = np.zeros((4, H, W, 3), np.float32)
    o = np.zeros((4, H, W, 3), np.float32)
    for k in range(4):
        gpu_renderer.render_tiles(g[k], None, k, 0, W, H, cam, frame_time=c["frame_time"], max_depth=DEPTH, seed=3)
        p = Renderer.params(W, H, sample_idx=k, max_depth=DEPTH, frame_time=c["frame_time"], seed=3)
        o[k] = orc.render_sample(cam, p, mode=1)
    assert np.isfinite(g).all() and (g.sum(0) > 0).mean() > 0.5
    close = np.isclose(g, o, rtol=1e-3, atol=1e-5).all(axis=-1)
    assert close.mean() > 0.97, close.mean()


@pytest.mark.parametrize("case", sorted(CASES))
def test_camera_and_motion_three_sigma_vs_reference_streams(gpu_renderer, case):
    c = CASES[case]
    W, H = 32, 24
    sc, cam = motion_room(W, H, c["aperture"], c["blades"])
    gpu_renderer.upload(sc)
    orc = O.OracleScene(sc)

    def gpu(first, s):
        out = np.zeros((H, W, 3), np.float32)
        gpu_renderer.render_accumulate(out, cam, W, H, s, sample_idx=first, max_depth=DEPTH, seed=5,
                                       frame_time=c["frame_time"])
        return out

    def cpu(first, s):
        p = Renderer.params(W, H, sample_idx=first, max_depth=DEPTH, frame_time=c["frame_time"], seed=6)
        return orc.render_accumulate(cam, p, s, mode=0)
    IS.gate(IS.frames(gpu, 128, 32), IS.frames(cpu, 128, 32))


def test_lens_and_shutter_change_the_frame(gpu_renderer):
    """The features are live: a lens blurs away from the focus plane, the
    shutter window moves the sphere, and the wavefront renderer draws the same
    camera rays as the megakernel."""
    W, H = 48, 36
    sc, pin = motion_room(W, H, 0.0, 0)
    _, lens = motion_room(W, H, 0.12, 6)
    gpu_renderer.upload(sc)

    def frame(cam, ft, wavefront=False):
        out = np.zeros((H, W, 3), np.float32)
        if wavefront:
            gpu_renderer.render_wavefront(out, cam, W, H, 16, max_depth=DEPTH, seed=9, frame_time=ft, capacity=8192)
        else:
            gpu_renderer.render_accumulate(out, cam, W, H, 16, max_depth=DEPTH, seed=9, frame_time=ft)
        return out
    a, b = frame(pin, (0.0, 0.1)), frame(lens, (0.0, 0.1))
    assert np.abs(a - b).max() > 0.1
    c = frame(pin, (0.9, 1.0))
    assert np.abs(a - c).max() > 0.1
    w = frame(lens, (0.25, 0.75), wavefront=True)
    m = frame(lens, (0.25, 0.75))
    assert np.allclose(w, m, rtol=1e-4, atol=1e-5), np.abs(w - m).max()
