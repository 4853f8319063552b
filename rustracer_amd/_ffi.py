"""ctypes mirror of include/rt_api.h and the loader for librtamd.so.

The structures below must stay byte-identical to the C declarations; the
`not gpu` test suite checks their sizes against the C compiler's.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librtamd.so")

RT_OK = 0
RT_ERR_INVALID = -1
RT_ERR_UNSUPPORTED = -2
RT_ERR_DEVICE = -3
RT_ERR_NOMEM = -4

(SHAPE_SPHERE, SHAPE_QUAD, SHAPE_MESH, SHAPE_BOX, SHAPE_CONVEX, SHAPE_CYLINDER, SHAPE_DISK, SHAPE_CONE,
 SHAPE_PLANE) = 1, 2, 3, 4, 5, 6, 7, 8, 9
XFORM_TRANSLATE, XFORM_ROTATE, XFORM_SCALE = 1, 2, 3
MEDIUM_CONSTANT, MEDIUM_GRID, MEDIUM_NOISE = 1, 2, 3
TEX_SOLID, TEX_CHECKER, TEX_IMAGE, TEX_NOISE, TEX_MOTTLE = 0, 1, 2, 3, 4
PARAM_MEDIUM = 0x40000000    # rt_param.material | k: the density of rt_scene_desc.media[k]
(MAT_LAMBERTIAN, MAT_DIELECTRIC, MAT_SPECTRAL_DIELECTRIC, MAT_PBR, MAT_DIFFUSE_LIGHT,
 MAT_SPECTRAL_METAL, MAT_ROUGH_DIELECTRIC, MAT_ROUGH_SPECTRAL, MAT_BLACKBODY, MAT_SSS, MAT_PEARL) = \
    1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11
BG_SOLID, BG_GRADIENT, BG_PHYSICAL, BG_ENVMAP, BG_ENVMAP_SPLIT = 0, 1, 2, 3, 4
API_VERSION = 3          # include/rt_api.h RT_API_VERSION: the layouts below
EMIT_QUAD, EMIT_SUN_DISK, EMIT_SCENE = 0, 1, 2
NOALLOC_CAPACITY = 1 << 17

F32P = C.POINTER(C.c_float)


class rt_xform(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("v", C.c_float * 9)]


class rt_medium(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("density", C.c_float), ("albedo", C.c_float * 3),
                ("g", C.c_float), ("grid_n", C.c_uint32 * 3), ("grid", F32P), ("noise_scale", C.c_float),
                ("threshold", C.c_float), ("perlin", F32P)]


class rt_object(C.Structure):
    _fields_ = [("shape", C.c_uint32), ("material", C.c_uint32), ("xform_first", C.c_uint32),
                ("xform_count", C.c_uint32), ("data_first", C.c_uint32), ("data_count", C.c_uint32),
                ("medium", C.c_uint32), ("flags", C.c_uint32), ("p", C.c_float * 16)]


class rt_material(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("tex", C.c_uint32), ("p", C.c_float * 24)]


ENV_CDF, ENV_ALIAS = 0, 1


class rt_background(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("visible_kind", C.c_uint32), ("color", C.c_float * 3),
                ("top", C.c_float * 3), ("bottom", C.c_float * 3), ("sun_dir", C.c_float * 3),
                ("turbidity", C.c_float), ("env_w", C.c_uint32), ("env_h", C.c_uint32),
                ("env_rgb", F32P), ("env_rot_y", C.c_float), ("env_sampling", C.c_uint32)]


class rt_image(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("rgb", C.POINTER(C.c_uint8))]


class rt_scene_desc(C.Structure):
    _fields_ = [("objects", C.POINTER(rt_object)), ("num_objects", C.c_uint32),
                ("lights", C.POINTER(rt_object)), ("num_lights", C.c_uint32),
                ("materials", C.POINTER(rt_material)), ("num_materials", C.c_uint32),
                ("xforms", C.POINTER(rt_xform)), ("num_xforms", C.c_uint32),
                ("media", C.POINTER(rt_medium)), ("num_media", C.c_uint32),
                ("pool", F32P), ("pool_len", C.c_uint32),
                ("tri_v", F32P), ("tri_n", F32P), ("tri_t", F32P), ("tri_uv", F32P),
                ("num_tris", C.c_uint32), ("background", rt_background),
                ("images", C.POINTER(rt_image)), ("num_images", C.c_uint32),
                ("perlin", F32P), ("num_perlin", C.c_uint32)]


class rt_ray(C.Structure):
    _fields_ = [("origin", C.c_float * 3), ("dir", C.c_float * 3), ("t_min", C.c_float),
                ("t_max", C.c_float), ("time", C.c_float)]


class rt_hit(C.Structure):
    _fields_ = [("t", C.c_float), ("p", C.c_float * 3), ("n", C.c_float * 3), ("u", C.c_float),
                ("v", C.c_float), ("obj_id", C.c_int32), ("prim_id", C.c_int32), ("front_face", C.c_uint32)]


class rt_camera(C.Structure):
    _fields_ = [("origin", C.c_float * 3), ("lower_left", C.c_float * 3), ("horizontal", C.c_float * 3),
                ("vertical", C.c_float * 3), ("u", C.c_float * 3), ("v", C.c_float * 3),
                ("lens_radius", C.c_float), ("aperture_blades", C.c_uint32)]


class rt_sample_params(C.Structure):
    _fields_ = [("sample_idx", C.c_uint32), ("strata", C.c_uint32), ("width", C.c_uint32),
                ("height", C.c_uint32), ("bg_scale", C.c_float), ("max_depth", C.c_uint32),
                ("use_frame_time", C.c_uint32), ("frame_t0", C.c_float), ("frame_t1", C.c_float),
                ("use_photons", C.c_uint32), ("seed", C.c_uint64), ("wavelengths", C.c_uint32)]


class rt_wavefront_opts(C.Structure):
    _fields_ = [("capacity", C.c_uint32), ("grid", C.c_uint32), ("sort", C.c_uint32), ("use_graph", C.c_uint32),
                ("occupancy", C.c_uint32), ("streams", C.c_uint32), ("block", C.c_uint32), ("knn_wave_max", C.c_uint32)]


class rt_accum_opts(C.Structure):
    _fields_ = [("firefly_clamp", C.c_float), ("firefly_min_samples", C.c_uint32), ("firefly_abs_floor", C.c_float),
                ("adaptive", C.c_uint32), ("min_samples", C.c_uint32), ("convergence_pct", C.c_float),
                ("threshold", C.c_float), ("reserved", C.c_uint32)]


class rt_pixel_state(C.Structure):
    _fields_ = [("accum", C.c_void_p), ("samples", C.c_void_p), ("var_m2", C.c_void_p), ("converged", C.c_void_p)]


class rt_resolve_opts(C.Structure):
    _fields_ = [("exposure", C.c_float), ("tonemapper", C.c_uint32), ("vignette", C.c_float), ("distortion", C.c_float)]


TONE_ACES, TONE_AGX = 0, 1


class rt_batch_view(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("camera", rt_camera), ("params", rt_sample_params), ("spp", C.c_uint32),
                ("accum_rgb", C.c_void_p)]


class rt_param(C.Structure):
    _fields_ = [("material", C.c_uint32), ("field", C.c_uint32), ("count", C.c_uint32)]


class rt_photon_params(C.Structure):
    _fields_ = [("emitter", C.c_uint32), ("origin", C.c_float * 3), ("u", C.c_float * 3),
                ("v", C.c_float * 3), ("color", C.c_float * 3), ("sun_dir", C.c_float * 3),
                ("num_photons", C.c_uint32), ("gather_radius", C.c_float), ("knn", C.c_uint32),
                ("wavelengths", C.c_uint32)]


# Numpy dtypes with the same layout, for batches of rays / hits.
def ray_dtype():
    import numpy as np
    return np.dtype([("origin", "<f4", 3), ("dir", "<f4", 3), ("t_min", "<f4"), ("t_max", "<f4"), ("time", "<f4")])


def hit_dtype():
    import numpy as np
    return np.dtype([("t", "<f4"), ("p", "<f4", 3), ("n", "<f4", 3), ("u", "<f4"), ("v", "<f4"),
                     ("obj_id", "<i4"), ("prim_id", "<i4"), ("front_face", "<u4")])


# Every entry point include/rt_api.h declares, with its ctypes signature.
VP = C.c_void_p
SIGNATURES = {
    "rt_api_version": (C.c_int, []),
    "rt_ctx_create": (C.c_int, [C.c_int, C.POINTER(VP)]),
    "rt_ctx_destroy": (None, [VP]),
    "rt_last_error": (C.c_char_p, [VP]),
    "rt_scene_upload": (C.c_int, [VP, C.POINTER(rt_scene_desc), C.POINTER(VP)]),
    "rt_scene_destroy": (None, [VP]),
    "rt_scene_stats": (C.c_int, [VP, C.POINTER(C.c_uint64)]),
    "rt_trace_closest": (C.c_int, [VP, VP, VP, C.c_uint32, VP]),
    "rt_trace_any": (C.c_int, [VP, VP, VP, C.c_uint32, VP]),
    "rt_transmittance": (C.c_int, [VP, VP, VP, C.c_uint32, VP]),
    "rt_render_sample": (C.c_int, [VP, VP, C.POINTER(rt_camera), C.POINTER(rt_sample_params), VP, VP]),
    "rt_render_accumulate": (C.c_int, [VP, VP, C.POINTER(rt_camera), C.POINTER(rt_sample_params), C.c_uint32, VP]),
    "rt_render_wavefront": (C.c_int, [VP, VP, C.POINTER(rt_camera), C.POINTER(rt_sample_params), C.c_uint32, VP,
                                      C.POINTER(rt_wavefront_opts)]),
    "rt_render_wavefront_noalloc": (C.c_int, [VP, VP, C.POINTER(rt_camera), C.POINTER(rt_sample_params), C.c_uint32,
                                              VP, C.POINTER(rt_wavefront_opts)]),
    "rt_render_batch": (C.c_int, [VP, C.POINTER(rt_batch_view), C.c_uint32, C.POINTER(rt_wavefront_opts)]),
    "rt_scene_set_params": (C.c_int, [VP, VP, C.POINTER(rt_param), C.c_uint32, VP]),
    "rt_scene_get_params": (C.c_int, [VP, VP, C.POINTER(rt_param), C.c_uint32, VP]),
    "rt_render_detached": (C.c_int, [VP, VP, C.POINTER(rt_camera), C.POINTER(rt_sample_params), C.c_uint32,
                                     C.POINTER(rt_param), C.c_uint32, VP, VP]),
    "rt_backward": (C.c_int, [VP, VP, C.POINTER(rt_camera), C.POINTER(rt_sample_params), C.c_uint32,
                              C.POINTER(rt_param), C.c_uint32, VP, VP]),
    "rt_backward_wavefront": (C.c_int, [VP, VP, C.POINTER(rt_camera), C.POINTER(rt_sample_params), C.c_uint32,
                                        C.POINTER(rt_param), C.c_uint32, VP, VP, C.POINTER(rt_wavefront_opts)]),
    "rt_photon_build": (C.c_int, [VP, VP, C.POINTER(rt_photon_params), C.POINTER(C.c_uint32)]),
    "rt_photon_emitter": (C.c_int, [VP, VP, C.POINTER(rt_photon_params), C.POINTER(rt_photon_params)]),
    "rt_scene_photon_bounds": (C.c_int, [VP, VP, VP, C.POINTER(C.c_int)]),
    "rt_photon_import": (C.c_int, [VP, VP, VP, C.c_uint32, C.c_float]),
    "rt_photon_set_gather": (C.c_int, [VP, VP, C.c_float, C.c_uint32]),
    "rt_photon_irradiance": (C.c_int, [VP, VP, VP, C.c_uint32, VP]),
    "rt_photon_export": (C.c_int, [VP, VP, VP, C.c_uint32, C.POINTER(C.c_uint32)]),
    "rt_accumulate_sample": (C.c_int, [VP, VP, C.c_uint32, C.POINTER(rt_accum_opts), C.POINTER(rt_pixel_state)]),
    "rt_mark_converged": (C.c_int, [VP, C.c_uint32, C.POINTER(rt_accum_opts), C.POINTER(rt_pixel_state),
                                    C.POINTER(C.c_uint32)]),
    "rt_render_progressive": (C.c_int, [VP, VP, C.POINTER(rt_camera), C.POINTER(rt_sample_params), C.c_uint32,
                                        C.POINTER(rt_accum_opts), C.POINTER(rt_pixel_state), C.POINTER(C.c_uint32),
                                        C.POINTER(C.c_uint32)]),
    "rt_resolve_rgb8": (C.c_int, [VP, VP, VP, C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(rt_resolve_opts), VP]),
    "rt_render_aux": (C.c_int, [VP, VP, C.POINTER(rt_camera), C.c_uint32, C.c_uint32, VP, VP]),
    "rt_render_progressive_wavefront": (C.c_int, [VP, VP, C.POINTER(rt_camera), C.POINTER(rt_sample_params), C.c_uint32,
                                                  C.POINTER(rt_accum_opts), C.POINTER(rt_pixel_state),
                                                  C.POINTER(rt_wavefront_opts), C.POINTER(C.c_uint32),
                                                  C.POINTER(C.c_uint32)]),
    "rt_ctx_set_timing": (C.c_int, [VP, C.c_int]),
    "rt_ctx_get_timing": (C.c_int, [VP, C.POINTER(C.c_double), C.POINTER(C.c_uint64)]),
    "rt_ctx_set_counting": (C.c_int, [VP, C.c_int]),
    "rt_ctx_get_counters": (C.c_int, [VP, C.POINTER(C.c_uint64)]),
    "rt_ctx_get_stage_timing": (C.c_int, [VP, C.POINTER(C.c_double), C.POINTER(C.c_uint64)]),
    "rt_ctx_get_stage_times": (C.c_int, [VP, C.POINTER(C.c_double), C.POINTER(C.c_uint64)]),
    "rt_ctx_set_megakernel_occupancy": (C.c_int, [VP, C.c_uint32]),
}

_lib = None


def lib():
    """Load librtamd.so (built in-tree).  Raises if it is missing: there is no
    CPU fallback for the product path."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() "
                              f"(hipcc --offload-arch=gfx950)")
        # torch ships its own libamdhip64 under the same soname; load it first
        # so the process holds one HIP runtime that both torch and this library
        # use (the reverse order leaves torch without a device).
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        if L.rt_api_version() != API_VERSION:
            raise ImportError(f"{LIB_PATH} has ABI version {L.rt_api_version()}, these bindings {API_VERSION}: rebuild")
        _lib = L
    return _lib
