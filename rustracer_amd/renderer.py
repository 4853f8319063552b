"""Host mirror of rustracer's render entry points over the C ABI.

`Renderer.render_tiles` keeps the argument meaning of the reference's
`render_tiles` (src/renderer.rs:655-669): one sample pass into a caller-owned
scratch buffer (row 0 = image top, 3 floats per pixel), an optional
converged-pixel mask, the pass index, the strata count, the camera, the scene
(world + lights + background), `bg_scale`, the photon map switch and the
frame-time window.  The reference hard-codes MAX_DEPTH = 50
(src/renderer.rs:16); here it is an argument with that default.
Errors from the library raise RuntimeError with the library's message.
"""
import ctypes as C
import numpy as np

from . import _ffi as F


class RtError(RuntimeError):
    pass


def _ptr(a):
    if a is None:
        return None
    if hasattr(a, "data_ptr"):          # torch tensor (host or device)
        return C.c_void_p(a.data_ptr())
    return C.c_void_p(a.ctypes.data)


class Renderer:
    """One rt_ctx on one device plus the uploaded scene."""

    def __init__(self, device=0):
        self.L = F.lib()
        ctx = C.c_void_p()
        rc = self.L.rt_ctx_create(device, C.byref(ctx))
        if rc != F.RT_OK:
            raise RtError(f"rt_ctx_create({device}) failed with {rc}: no usable HIP device")
        self.ctx = ctx
        self.scene = None
        self.num_photons = 0

    def _check(self, rc, what):
        if rc != F.RT_OK:
            msg = self.L.rt_last_error(self.ctx)
            raise RtError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    def close(self):
        for h in getattr(self, "_extra", []):
            self.L.rt_scene_destroy(h)
        self._extra = []
        if self.scene is not None:
            self.L.rt_scene_destroy(self.scene)
            self.scene = None
        if self.ctx:
            self.L.rt_ctx_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---------------------------------------------------------------- scene
    def upload(self, scene):
        """scene: rustracer_amd.scene.Scene; becomes the current scene."""
        if self.scene is not None:
            self.L.rt_scene_destroy(self.scene)
            self.scene = None
        self.scene = self.upload_extra(scene)
        self._extra.remove(self.scene)
        self.num_photons = 0
        return self

    def upload_extra(self, scene):
        """Upload another scene (for batches) without replacing the current
        one; the returned handle lives until close() or destroy(handle)."""
        d = scene.descriptor()
        s = C.c_void_p()
        self._check(self.L.rt_scene_upload(self.ctx, C.byref(d), C.byref(s)), "rt_scene_upload")
        self._extra = getattr(self, "_extra", [])
        self._extra.append(s)
        return s

    def destroy(self, handle):
        if handle is not None:
            self.L.rt_scene_destroy(handle)
            if handle in getattr(self, "_extra", []):
                self._extra.remove(handle)

    def stats(self):
        st = (C.c_uint64 * 8)()
        self._check(self.L.rt_scene_stats(self.scene, st), "rt_scene_stats")
        return {"nodes": st[0], "triangles": st[1], "node_bytes": st[2], "stack_bound": st[3],
                "objects": st[4], "triangle_bytes": st[5]}

    # ---------------------------------------------------------------- queries
    def trace_closest(self, rays):
        rays = np.ascontiguousarray(rays)
        hits = np.zeros(len(rays), dtype=F.hit_dtype())
        self._check(self.L.rt_trace_closest(self.ctx, self.scene, _ptr(rays), len(rays), _ptr(hits)), "rt_trace_closest")
        return hits

    def trace_any(self, rays):
        rays = np.ascontiguousarray(rays)
        out = np.zeros(len(rays), dtype=np.uint8)
        self._check(self.L.rt_trace_any(self.ctx, self.scene, _ptr(rays), len(rays), _ptr(out)), "rt_trace_any")
        return out

    def transmittance(self, rays):
        rays = np.ascontiguousarray(rays)
        out = np.zeros(len(rays), dtype=np.float32)
        self._check(self.L.rt_transmittance(self.ctx, self.scene, _ptr(rays), len(rays), _ptr(out)), "rt_transmittance")
        return out

    # ---------------------------------------------------------------- render
    @staticmethod
    def params(width, height, sample_idx=0, strata=0, bg_scale=1.0, max_depth=50, frame_time=None,
               use_photons=False, seed=0, wavelengths=0):
        """rt_sample_params; wavelengths: 0 the reference's one hero wavelength
        over [380, 700) nm, 1 one over [380, 780), 4 four hero wavelengths over
        [380, 780) (extension)."""
        p = F.rt_sample_params()
        p.sample_idx, p.strata, p.width, p.height = sample_idx, strata, width, height
        p.bg_scale, p.max_depth = bg_scale, max_depth
        if frame_time is not None:
            p.use_frame_time, p.frame_t0, p.frame_t1 = 1, frame_time[0], frame_time[1]
        p.use_photons = 1 if use_photons else 0
        p.seed = seed
        p.wavelengths = wavelengths
        return p

    def render_tiles(self, scratch, converged, sample_idx, strata, width, height, camera, bg_scale=1.0,
                     photon_map=False, frame_time=None, max_depth=50, seed=0, wavelengths=0):
        """render_tiles: one pass into `scratch` (H*W*3 float32, host array or
        device tensor)."""
        p = self.params(width, height, sample_idx, strata, bg_scale, max_depth, frame_time, photon_map, seed, wavelengths)
        cam = camera.to_c()
        conv = None if converged is None else np.ascontiguousarray(converged, dtype=np.uint8)
        self._check(self.L.rt_render_sample(self.ctx, self.scene, C.byref(cam), C.byref(p), _ptr(conv), _ptr(scratch)),
                    "rt_render_sample")
        return scratch

    def render_accumulate(self, out, camera, width, height, spp, sample_idx=0, strata=0, bg_scale=1.0,
                          photon_map=False, max_depth=50, seed=0, scene=None, frame_time=None, wavelengths=0):
        """Sum of `spp` passes (sample_idx .. sample_idx+spp-1) into `out`;
        `scene` is a handle from upload_extra (default: the current scene);
        frame_time = (t0, t1) draws each camera ray's time in [t0, t1)."""
        p = self.params(width, height, sample_idx, strata, bg_scale, max_depth, frame_time, photon_map, seed, wavelengths)
        cam = camera.to_c()
        self._check(self.L.rt_render_accumulate(self.ctx, self.scene if scene is None else scene, C.byref(cam), C.byref(p), spp, _ptr(out)),
                    "rt_render_accumulate")
        return out

    @staticmethod
    def accum_opts(firefly_clamp=8.0, firefly_min_samples=4, firefly_abs_floor=5.0, adaptive=False,
                   min_samples=128, convergence_pct=99.0, threshold=0.05):
        """rt_accum_opts; the defaults are the reference's (FireflyConfig::default,
        src/main.rs:124-133, 176, 696-718)."""
        return F.rt_accum_opts(firefly_clamp=firefly_clamp, firefly_min_samples=firefly_min_samples,
                               firefly_abs_floor=firefly_abs_floor, adaptive=1 if adaptive else 0,
                               min_samples=min_samples, convergence_pct=convergence_pct, threshold=threshold)

    @staticmethod
    def pixel_state(state):
        """rt_pixel_state over a dict of arrays (numpy or torch, host or
        device): accum (n,3) float32, samples (n,) uint32 (torch: int32),
        var_m2 (n,3) float32, converged (n,) uint8."""
        return F.rt_pixel_state(accum=_ptr(state["accum"]), samples=_ptr(state["samples"]),
                                var_m2=_ptr(state["var_m2"]), converged=_ptr(state["converged"]))

    @staticmethod
    def new_pixel_state(npix):
        """A fresh host state for npix pixels (numpy)."""
        return dict(accum=np.zeros((npix, 3), np.float32), samples=np.zeros(npix, np.uint32),
                    var_m2=np.zeros((npix, 3), np.float32), converged=np.zeros(npix, np.uint8))

    def accumulate_sample(self, scratch, state, opts=None):
        """accumulate_sample (src/main.rs:603-644) of one pass into `state`."""
        o = opts if opts is not None else self.accum_opts()
        st = self.pixel_state(state)
        npix = int(np.prod(scratch.shape[:-1]))
        self._check(self.L.rt_accumulate_sample(self.ctx, _ptr(scratch), npix, C.byref(o), C.byref(st)),
                    "rt_accumulate_sample")
        return state

    def mark_converged(self, state, opts=None):
        """mark_converged (src/main.rs:646-677); returns the converged count."""
        o = opts if opts is not None else self.accum_opts()
        st = self.pixel_state(state)
        n = C.c_uint32()
        self._check(self.L.rt_mark_converged(self.ctx, len(state["converged"]), C.byref(o), C.byref(st), C.byref(n)),
                    "rt_mark_converged")
        return n.value

    def render_progressive(self, state, camera, width, height, samples_max, opts=None, sample_idx=0, strata=0,
                           bg_scale=1.0, photon_map=False, max_depth=50, seed=0, wavefront=False, capacity=0, grid=0,
                           sort=1, use_graph=False, occupancy=0, streams=1, block=0, knn_wave_max=0, wavelengths=0):
        """The reference's render loop (src/main.rs:463-506) on the device;
        returns (passes run, converged pixels).  wavefront=True renders the
        passes with the wavefront stages (rt_render_progressive_wavefront)."""
        o = opts if opts is not None else self.accum_opts()
        p = self.params(width, height, sample_idx, strata, bg_scale, max_depth, None, photon_map, seed, wavelengths)
        cam = camera.to_c()
        st = self.pixel_state(state)
        done, nconv = C.c_uint32(), C.c_uint32()
        if wavefront:
            w = F.rt_wavefront_opts(capacity=capacity, grid=grid, sort=int(sort), use_graph=1 if use_graph else 0,
                                    occupancy=occupancy, streams=streams, block=block, knn_wave_max=knn_wave_max)
            self._check(self.L.rt_render_progressive_wavefront(self.ctx, self.scene, C.byref(cam), C.byref(p), samples_max,
                                                               C.byref(o), C.byref(st), C.byref(w), C.byref(done),
                                                               C.byref(nconv)), "rt_render_progressive_wavefront")
        else:
            self._check(self.L.rt_render_progressive(self.ctx, self.scene, C.byref(cam), C.byref(p), samples_max, C.byref(o),
                                                     C.byref(st), C.byref(done), C.byref(nconv)), "rt_render_progressive")
        return done.value, nconv.value

    def render_aux(self, camera, width, height, albedo=None, normal=None):
        """render_aux_pass (src/renderer.rs:597-645): first-hit albedo and
        normal buffers for a denoiser; returns (albedo, normal), each
        (height, width, 3) float32 (host arrays unless given)."""
        if albedo is None:
            albedo = np.zeros((height, width, 3), np.float32)
        if normal is None:
            normal = np.zeros((height, width, 3), np.float32)
        cam = camera.to_c()
        self._check(self.L.rt_render_aux(self.ctx, self.scene, C.byref(cam), width, height, _ptr(albedo), _ptr(normal)),
                    "rt_render_aux")
        return albedo, normal

    def resolve_rgb8(self, accum, width, height, samples=None, spp=1, exposure=1.0, tonemapper="aces",
                     vignette=0.0, distortion=0.0, out=None):
        """The saved 8-bit image (src/main.rs:571-582, src/output.rs): mean
        radiance (accum / samples, or / spp), lens optics, ACES or AgX tone
        map, sRGB encode; returns (height, width, 3) uint8."""
        o = F.rt_resolve_opts(exposure=exposure, tonemapper=F.TONE_AGX if tonemapper == "agx" else F.TONE_ACES,
                              vignette=vignette, distortion=distortion)
        if out is None:
            out = np.zeros((height, width, 3), np.uint8)
        self._check(self.L.rt_resolve_rgb8(self.ctx, _ptr(accum), _ptr(samples), spp, width, height, C.byref(o), _ptr(out)),
                    "rt_resolve_rgb8")
        return out

    def render_wavefront(self, out, camera, width, height, spp, sample_idx=0, strata=0, bg_scale=1.0,
                         photon_map=False, max_depth=50, seed=0, capacity=0, grid=0, sort=True, use_graph=False,
                         occupancy=0, streams=1, scene=None, block=0, knn_wave_max=0, frame_time=None, wavelengths=0):
        """Same frame as render_accumulate, through the wavefront stage kernels."""
        p = self.params(width, height, sample_idx, strata, bg_scale, max_depth, frame_time, photon_map, seed, wavelengths)
        cam = camera.to_c()
        o = F.rt_wavefront_opts(capacity=capacity, grid=grid, sort=int(sort), use_graph=1 if use_graph else 0,
                                occupancy=occupancy, streams=streams, block=block, knn_wave_max=knn_wave_max)
        self._check(self.L.rt_render_wavefront(self.ctx, self.scene if scene is None else scene, C.byref(cam), C.byref(p), spp, _ptr(out), C.byref(o)),
                    "rt_render_wavefront")
        return out

    def render_wavefront_noalloc(self, out, camera, width, height, spp, sample_idx=0, strata=0, bg_scale=1.0,
                                 photon_map=False, max_depth=50, seed=0, capacity=0, grid=0, sort=True, occupancy=0,
                                 scene=None, block=0, knn_wave_max=0, wavelengths=0):
        """render_wavefront without device allocations (rt_render_wavefront_noalloc):
        out must be device memory (e.g. a torch cuda tensor); capacity at most
        NOALLOC_CAPACITY paths, one stream, no graphs."""
        p = self.params(width, height, sample_idx, strata, bg_scale, max_depth, None, photon_map, seed, wavelengths)
        cam = camera.to_c()
        o = F.rt_wavefront_opts(capacity=capacity, grid=grid, sort=int(sort), use_graph=0, occupancy=occupancy, streams=1,
                                block=block, knn_wave_max=knn_wave_max)
        self._check(self.L.rt_render_wavefront_noalloc(self.ctx, self.scene if scene is None else scene, C.byref(cam),
                                                       C.byref(p), spp, _ptr(out), C.byref(o)), "rt_render_wavefront_noalloc")
        return out

    def render_batch(self, views, capacity=0, grid=0, sort=True, use_graph=False, occupancy=0, streams=1, block=0,
                     knn_wave_max=0):
        """views: list of dicts with keys scene (handle), camera, width, height,
        spp, out, and optionally max_depth, photon_map, seed, bg_scale, wavelengths."""
        arr = (F.rt_batch_view * len(views))()
        for i, v in enumerate(views):
            arr[i].scene = v["scene"]
            arr[i].camera = v["camera"].to_c()
            arr[i].params = self.params(v["width"], v["height"], 0, v.get("strata", 0), v.get("bg_scale", 1.0),
                                        v.get("max_depth", 50), None, v.get("photon_map", False), v.get("seed", 0),
                                        v.get("wavelengths", 0))
            arr[i].spp = v["spp"]
            arr[i].accum_rgb = _ptr(v["out"])
        o = F.rt_wavefront_opts(capacity=capacity, grid=grid, sort=int(sort), use_graph=1 if use_graph else 0,
                                occupancy=occupancy, streams=streams, block=block, knn_wave_max=knn_wave_max)
        self._check(self.L.rt_render_batch(self.ctx, arr, len(views), C.byref(o)), "rt_render_batch")
        return [v["out"] for v in views]

    # ---------------------------------------------------------------- differentiable parameters
    @staticmethod
    def _params(params):
        """params: list of (material index, field, count) -> rt_param array."""
        arr = (F.rt_param * len(params))()
        for i, (m, f, c) in enumerate(params):
            arr[i].material, arr[i].field, arr[i].count = int(m), int(f), int(c)
        return arr

    def set_params(self, params, values, scene=None):
        vals = np.ascontiguousarray(values, dtype=np.float32)
        assert len(vals) == len(params)
        self._check(self.L.rt_scene_set_params(self.ctx, self.scene if scene is None else scene, self._params(params),
                                               len(params), _ptr(vals)), "rt_scene_set_params")

    def get_params(self, params, scene=None):
        out = np.zeros(len(params), np.float32)
        self._check(self.L.rt_scene_get_params(self.ctx, self.scene if scene is None else scene, self._params(params),
                                               len(params), _ptr(out)), "rt_scene_get_params")
        return out

    def render_detached(self, out, camera, width, height, spp, params, eval_values, sample_idx=0, strata=0,
                        bg_scale=1.0, photon_map=False, max_depth=50, seed=0, scene=None, wavelengths=0):
        """The frame of render_accumulate with decisions at the scene's current
        parameters and BSDF values at `eval_values` (finite differences with
        common random numbers); wavelengths 0 or 1."""
        p = self.params(width, height, sample_idx, strata, bg_scale, max_depth, None, photon_map, seed, wavelengths)
        cam = camera.to_c()
        vals = np.ascontiguousarray(eval_values, dtype=np.float32)
        self._check(self.L.rt_render_detached(self.ctx, self.scene if scene is None else scene, C.byref(cam), C.byref(p),
                                              spp, self._params(params), len(params), _ptr(vals), _ptr(out)),
                    "rt_render_detached")
        return out

    def backward(self, dloss_dimage, camera, width, height, spp, params, sample_idx=0, strata=0, bg_scale=1.0,
                 photon_map=False, max_depth=50, seed=0, scene=None, grad=None, wavefront=False, capacity=0, grid=0,
                 sort=True, occupancy=0, streams=1, block=0, wavelengths=0):
        """Path-replay adjoint: dLoss/dtheta for dloss_dimage = dLoss/d(mean
        image) (H x W x 3, host array or device tensor); returns grad (float32
        numpy array, or fills the given device tensor).  wavefront=True runs
        rt_backward_wavefront with the given launch options; wavelengths 0 or 1."""
        p = self.params(width, height, sample_idx, strata, bg_scale, max_depth, None, photon_map, seed, wavelengths)
        cam = camera.to_c()
        if hasattr(dloss_dimage, "data_ptr"):
            d = dloss_dimage
        else:
            d = np.ascontiguousarray(dloss_dimage, dtype=np.float32)
        g = np.zeros(len(params), np.float32) if grad is None else grad
        sc = self.scene if scene is None else scene
        if wavefront:
            o = F.rt_wavefront_opts(capacity=capacity, grid=grid, sort=int(sort), use_graph=0, occupancy=occupancy,
                                    streams=streams, block=block)
            self._check(self.L.rt_backward_wavefront(self.ctx, sc, C.byref(cam), C.byref(p), spp, self._params(params),
                                                     len(params), _ptr(d), _ptr(g), C.byref(o)), "rt_backward_wavefront")
        else:
            self._check(self.L.rt_backward(self.ctx, sc, C.byref(cam), C.byref(p), spp, self._params(params), len(params),
                                           _ptr(d), _ptr(g)), "rt_backward")
        return g

    # ---------------------------------------------------------------- photons
    def photon_build_from_quad(self, origin, u, v, color, num_photons, gather_radius, scene=None, knn=0, wavelengths=0):
        """PhotonMap::build_from_quad (src/photon.rs:97-135); knn > 0 gathers
        the k nearest photons within gather_radius (extension); wavelengths 1
        or 4 draws the photon wavelengths over [380, 780) (extension)."""
        pp = F.rt_photon_params()
        pp.emitter = F.EMIT_QUAD
        for k in range(3):
            pp.origin[k], pp.u[k], pp.v[k], pp.color[k] = float(origin[k]), float(u[k]), float(v[k]), float(color[k])
        pp.num_photons, pp.gather_radius, pp.knn, pp.wavelengths = num_photons, gather_radius, knn, wavelengths
        n = C.c_uint32()
        self._check(self.L.rt_photon_build(self.ctx, scene if scene is not None else self.scene, C.byref(pp), C.byref(n)),
                    "rt_photon_build")
        self.num_photons = n.value
        return n.value

    def photon_build_sun(self, sun_dir, sun_color, num_photons, gather_radius, knn=0, wavelengths=0):
        """PhotonMap::build (src/photon.rs:41-94): directional disk toward `sun_dir`."""
        pp = F.rt_photon_params()
        pp.emitter = F.EMIT_SUN_DISK
        for k in range(3):
            pp.sun_dir[k], pp.color[k] = float(sun_dir[k]), float(sun_color[k])
        pp.num_photons, pp.gather_radius, pp.knn, pp.wavelengths = num_photons, gather_radius, knn, wavelengths
        n = C.c_uint32()
        self._check(self.L.rt_photon_build(self.ctx, self.scene, C.byref(pp), C.byref(n)), "rt_photon_build")
        self.num_photons = n.value
        return n.value

    def photon_build(self, pp, scene=None):
        """rt_photon_build with a filled rt_photon_params."""
        n = C.c_uint32()
        self._check(self.L.rt_photon_build(self.ctx, self.scene if scene is None else scene, C.byref(pp), C.byref(n)),
                    "rt_photon_build")
        self.num_photons = n.value
        return n.value

    @staticmethod
    def caustic_params(info, knn=0, wavelengths=0):
        """rt_photon_params for SceneData::rebuild_caustics (src/scene.rs:35-74):
        RT_EMIT_SCENE with the scene's caustic quad (its first light) as the
        fallback emitter.  info: the caustics settings of a loaded scene
        (num_photons, gather_radius, light_quad)."""
        pp = F.rt_photon_params()
        pp.emitter = F.EMIT_SCENE
        lq = info.get("light_quad")
        if lq is not None:
            for k in range(3):
                pp.origin[k], pp.u[k], pp.v[k], pp.color[k] = (float(lq["q"][k]), float(lq["u"][k]), float(lq["v"][k]),
                                                               float(lq["color"][k]))
        pp.num_photons, pp.gather_radius, pp.knn = int(info["num_photons"]), float(info["gather_radius"]), int(knn)
        pp.wavelengths = int(wavelengths)
        return pp

    def photon_emitter(self, pp, scene=None):
        """The emitter rt_photon_build uses for pp (RT_EMIT_SCENE resolved)."""
        out = F.rt_photon_params()
        self._check(self.L.rt_photon_emitter(self.ctx, self.scene if scene is None else scene, C.byref(pp), C.byref(out)),
                    "rt_photon_emitter")
        return out

    def rebuild_caustics(self, info, knn=0, scene=None, wavelengths=0):
        """SceneData::rebuild_caustics (src/scene.rs:35-74): nothing when the
        scene's caustics are off or no emitter exists (no sun, no env map, no
        light quad); otherwise the map of the emitter the background picks.
        Returns the number of stored photons (0 when nothing was built)."""
        if not info.get("caustics"):
            return 0
        sc = self.scene if scene is None else scene
        pp = self.caustic_params(info, knn, wavelengths)
        try:
            self.photon_emitter(pp, sc)
        except RtError:
            return 0
        return self.photon_build(pp, sc)

    def photon_bounds(self, scene=None):
        """The uploaded world's photon_bounds: (min, max) or None."""
        out = np.zeros(6, np.float32)
        has = C.c_int()
        self._check(self.L.rt_scene_photon_bounds(self.ctx, self.scene if scene is None else scene, _ptr(out), C.byref(has)),
                    "rt_scene_photon_bounds")
        return (out[:3], out[3:]) if has.value else None

    def photon_import(self, photons, gather_radius, scene=None):
        """Attach a host-built photon map: (n, 9) float32 rows of position,
        power, normal (the layout photon_export returns)."""
        ph = np.ascontiguousarray(photons, dtype=np.float32).reshape(-1, 9)
        self._check(self.L.rt_photon_import(self.ctx, self.scene if scene is None else scene, _ptr(ph), len(ph),
                                            float(gather_radius)), "rt_photon_import")
        self.num_photons = len(ph)
        return len(ph)

    def photon_set_gather(self, gather_radius, knn=0, scene=None):
        """Fixed-radius (knn 0) or k-NN gather for the attached photon map."""
        self._check(self.L.rt_photon_set_gather(self.ctx, self.scene if scene is None else scene, float(gather_radius),
                                                int(knn)), "rt_photon_set_gather")

    def photon_export(self):
        n = C.c_uint32()
        self._check(self.L.rt_photon_export(self.ctx, self.scene, None, 0, C.byref(n)), "rt_photon_export")
        out = np.zeros((n.value, 9), np.float32)
        self._check(self.L.rt_photon_export(self.ctx, self.scene, _ptr(out), n.value, C.byref(n)), "rt_photon_export")
        return out

    def photon_irradiance(self, pos_nrm):
        pos_nrm = np.ascontiguousarray(pos_nrm, dtype=np.float32).reshape(-1, 6)
        out = np.zeros((len(pos_nrm), 3), dtype=np.float32)
        self._check(self.L.rt_photon_irradiance(self.ctx, self.scene, _ptr(pos_nrm), len(pos_nrm), _ptr(out)),
                    "rt_photon_irradiance")
        return out

    # ---------------------------------------------------------------- measurement
    def set_timing(self, enable=True):
        self._check(self.L.rt_ctx_set_timing(self.ctx, 1 if enable else 0), "rt_ctx_set_timing")

    def timing(self):
        ms, n = C.c_double(), C.c_uint64()
        self._check(self.L.rt_ctx_get_timing(self.ctx, C.byref(ms), C.byref(n)), "rt_ctx_get_timing")
        return ms.value, n.value

    def set_megakernel_occupancy(self, occupancy):
        """Register budget of the plain megakernel: 0 (compiler's choice) or 2
        (at least two waves per SIMD)."""
        self._check(self.L.rt_ctx_set_megakernel_occupancy(self.ctx, int(occupancy)), "rt_ctx_set_megakernel_occupancy")

    def set_counting(self, enable=True):
        self._check(self.L.rt_ctx_set_counting(self.ctx, 1 if enable else 0), "rt_ctx_set_counting")

    def counters(self):
        out = (C.c_uint64 * 8)()
        self._check(self.L.rt_ctx_get_counters(self.ctx, out), "rt_ctx_get_counters")
        return {"node_tests": out[0], "prim_tests": out[1], "rays": out[2], "samples": out[3],
                "shadow_node_tests": out[4], "shadow_prim_tests": out[5], "shadow_rays": out[6],
                "photon_bytes": out[7]}

    def stage_times(self):
        """Per-stage wavefront timing since timing was enabled (graphs off):
        {stage: (ms, timed bounces)} for extend, knn, shade, shadow."""
        ms, n = (C.c_double * 4)(), (C.c_uint64 * 4)()
        self._check(self.L.rt_ctx_get_stage_times(self.ctx, ms, n), "rt_ctx_get_stage_times")
        return {k: (ms[i], n[i]) for i, k in enumerate(("extend", "knn", "shade", "shadow"))}

    def stage_timing(self):
        ms, n = C.c_double(), C.c_uint64()
        self._check(self.L.rt_ctx_get_stage_timing(self.ctx, C.byref(ms), C.byref(n)), "rt_ctx_get_stage_timing")
        return ms.value, n.value
