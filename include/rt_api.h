/*
 * rt_api.h — C ABI of the MI355X-native per-sample radiance path.
 *
 * This library stands in for rustracer's `render_tiles` / `ray_color` path
 * (src/renderer.rs:655-730 and 371-585) and for the BVH queries behind it
 * (src/bvh.rs:229-575).  The reference has no FFI of its own: its callers are
 * Rust functions (`bench_scene` src/main.rs:68-82, `run_render`
 * src/main.rs:350-594) that hand `render_tiles` a `&dyn Hittable` world, a
 * `HittableList` of lights, a `Background`, a `Camera` and an optional
 * `PhotonMap`.  Here the same information crosses the boundary once, as a
 * flattened scene descriptor (`rt_scene_desc`), and every later call refers to
 * the uploaded scene by handle.
 *
 * Conventions (all entry points):
 *   - return 0 (RT_OK) on success and a negative RT_ERR_* code on failure;
 *     `rt_last_error(ctx)` returns a message for the last failure on `ctx`.
 *     Nothing throws across the ABI.
 *   - buffers are owned by the caller.  A buffer argument may be a host pointer
 *     or a device pointer (hipMalloc'd on the context's device); the library
 *     detects which and copies only when it has to.
 *   - one host thread per context.  Calls are synchronous: when a call returns,
 *     its outputs are complete.
 *   - geometry is IEEE f32 and follows the reference's closed interval
 *     [t_min, t_max] (src/hittable.rs:97-101).
 */
#ifndef RT_API_H
#define RT_API_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RT_API_VERSION 3

/* ---- status codes ------------------------------------------------------- */
#define RT_OK               0
#define RT_ERR_INVALID     -1   /* bad argument / malformed descriptor        */
#define RT_ERR_UNSUPPORTED -2   /* a Hittable/Material kind the GPU path lacks */
#define RT_ERR_DEVICE      -3   /* HIP runtime error                          */
#define RT_ERR_NOMEM       -4   /* device allocation failed                   */

typedef struct rt_ctx   rt_ctx;
typedef struct rt_scene rt_scene;

/* ---- shapes (Hittable implementors) -------------------------------------
 * Inline parameter layout of rt_object.p[] per shape:
 *   SPHERE   p0-2 center, p3 radius, p4-6 center1, p7 moving (0/1)   src/sphere.rs:17-35
 *   QUAD     p0-2 corner q, p3-5 u, p6-8 v                            src/quad.rs:9-33
 *   MESH     triangles [data_first, data_first+data_count) of the
 *            descriptor's tri_* arrays, wrapped in a nested BVH       src/mesh.rs:13-92
 *   BOX      p0-2 p0, p3-5 p1: six quads in a nested BVH             src/quad.rs:201-215
 *   CONVEX   half-spaces (n.xyz, d) x data_count in pool[data_first*4..];
 *            p0-2 center, p3 crown_h, p4 pavilion_h, p5 girdle_h,
 *            p6 polish, p7-9 bbox min, p10-12 bbox max                src/diamond.rs:203-274
 *            (also StepCut, whose hit is the same, src/step_cut.rs:388-405)
 *   CYLINDER p0-2 center of the bottom cap, p3 radius, p4 height; y axis,
 *            closed caps                                              src/cylinder.rs:124-215
 *   DISK     p0-2 center, p3-5 normal (normalised on upload), p6 radius;
 *            a valid NEE light                                        src/disk.rs:105-170
 *   CONE     p0-2 center of the base, p3 radius, p4 height; apex on +y src/cone.rs:11-129
 *   PLANE    p0-2 point, p3-5 normal (normalised on upload), p6 wave
 *            amplitude, p7 wave scale (InfinitePlane)                 src/plane.rs:14-84
 */
enum {
    RT_SHAPE_SPHERE   = 1,
    RT_SHAPE_QUAD     = 2,
    RT_SHAPE_MESH     = 3,
    RT_SHAPE_BOX      = 4,
    RT_SHAPE_CONVEX   = 5,
    RT_SHAPE_CYLINDER = 6,
    RT_SHAPE_DISK     = 7,
    RT_SHAPE_CONE     = 8,
    RT_SHAPE_PLANE    = 9
};

/* ---- transforms ----------------------------------------------------------
 * An object's transform chain lists wrappers from the innermost (applied to
 * the shape first) to the outermost, exactly as scene_file.rs builds them
 * (scale -> rotate -> translate, src/scene_file.rs:1032-1038;
 * Rotate::euler_about = Translate(-pivot), Rotate, Translate(pivot),
 * src/transform.rs:218-224).
 *   TRANSLATE v0-2 offset                       src/transform.rs:86-136
 *   ROTATE    v0-8 forward matrix, row-major    src/transform.rs:140-277
 *   SCALE     v0 factor (> 0)                   src/transform.rs:15-81
 */
enum { RT_XFORM_TRANSLATE = 1, RT_XFORM_ROTATE = 2, RT_XFORM_SCALE = 3 };

typedef struct {
    uint32_t kind;
    float    v[9];
} rt_xform;

/* ---- participating media ---------------------------------------------------
 * An object with medium != 0 is the boundary of a medium (its material is
 * ignored).  medium = index + 1 into rt_scene_desc.media.
 *   CONSTANT  ConstantMedium (src/volume.rs:78-156)
 *   GRID      delta tracking through a dense float density grid spanning the
 *             boundary's bounding box (extension of NOISE: a table lookup in
 *             place of the per-lookup turbulence)
 *   NOISE     NoiseMedium (src/volume.rs:160-263): delta tracking with
 *             density max(turb(p * noise_scale, 7) / 2 - threshold, 0) * density
 *             (Perlin, src/perlin.rs:35-72); the Perlin tables are given (the
 *             reference draws them from an unseeded thread_rng)
 */
enum { RT_MEDIUM_CONSTANT = 1, RT_MEDIUM_GRID = 2, RT_MEDIUM_NOISE = 3 };

typedef struct {
    uint32_t     kind;
    float        density;      /* sigma_t (CONSTANT) or majorant scale (GRID, NOISE) */
    float        albedo[3];
    float        g;            /* Henyey-Greenstein asymmetry                 */
    uint32_t     grid_n[3];    /* GRID only                                   */
    const float *grid;         /* GRID only: grid_n[0]*grid_n[1]*grid_n[2], x fastest, values in [0,1] */
    float        noise_scale;  /* NOISE: feature frequency                    */
    float        threshold;    /* NOISE: clipped to [0, 0.99]                 */
    const float *perlin;       /* NOISE: 1536 floats: ranvec[256][3] (unit vectors),
                                  then perm_x, perm_y, perm_z (256 each, integers 0..255) */
} rt_medium;

typedef struct {
    uint32_t shape;
    uint32_t material;      /* index into rt_scene_desc.materials           */
    uint32_t xform_first;   /* chain in rt_scene_desc.xforms                */
    uint32_t xform_count;
    uint32_t data_first;    /* MESH: first triangle; CONVEX: first plane    */
    uint32_t data_count;    /* MESH: triangle count; CONVEX: plane count     */
    uint32_t medium;        /* 0 = surface; k = media[k-1]                   */
    uint32_t flags;         /* reserved, 0                                   */
    float    p[16];
} rt_object;

/* ---- materials -----------------------------------------------------------
 * Parameter layout of rt_material.p[] per kind:
 *   LAMBERTIAN  albedo texture (below)           src/material.rs:85-102
 *   DIELECTRIC  p0 ior                             src/material.rs:323-332
 *   SPECTRAL_DIELECTRIC p0 cauchy_b, p1 cauchy_c, p2-4 absorption,
 *               p5 model (0 Cauchy, 1 Sellmeier), p6-8 Sellmeier B1..B3,
 *               p9-11 Sellmeier C1..C3 (um^2)       src/material.rs:357-403
 *   PBR         p0-2 albedo, p3 roughness, p4 metallic, p5 anisotropy,
 *               p6 anisotropy_angle, p7 clearcoat, p8 clearcoat_roughness,
 *               p9-11 emission, p12 emission_strength, p13 sheen,
 *               p14 sheen_tint, p15 film_thickness (nm; 0 = no thin-film tint of
 *               the clearcoat), p16 film_ior, p17-21 image maps as index + 1
 *               into rt_scene_desc.images (0 = none): albedo (sRGB), roughness,
 *               metallic, ambient occlusion (linear, red channel), tangent-space
 *               normal (applies where the hit has a tangent: meshes); bilinear
 *               fetches                             src/material.rs:632-1050
 *   DIFFUSE_LIGHT emission texture (below)          src/material.rs:14-25
 *   SPECTRAL_METAL p0 variant (0 gold, 1 copper, 2 silver), p1 roughness
 *                                                  src/material.rs:427-554
 *   ROUGH_DIELECTRIC p0 ior, p1 roughness (GGX transmission; extension, no
 *               counterpart in the reference)
 *   BLACKBODY p0 temperature (K), p1 intensity; emits spectral_to_rgb(l) x
 *               normalised Planck on spectral paths, its white-point average
 *               otherwise                             src/material.rs:27-83
 *   SSS         p0-2 albedo, p3-5 sigma_a, p6 ior, p7 density, p8 g
 *               (dielectric boundary with HG scattering inside)
 *                                                  src/material.rs:576-631
 *   PEARL       p0-2 base_color, p3 ior, p4 film_thickness (nm),
 *               p5 orient_strength, p6 film_scale, p7 luster_roughness: GGX
 *               luster tinted by thin-film interference over a Lambertian body;
 *               the nacre glow follows the Physical background's sun when there
 *               is one (the renderer's pearl sun), the view angle otherwise
 *                                                  src/material.rs:1052-1287
 *   ROUGH_SPECTRAL p0 n_d (IOR at 589.3 nm), p1-3 Sellmeier B1..B3,
 *               p4-6 Sellmeier C1..C3 (um^2): n(l) = n_d + s(l) - s(589.3 nm);
 *               p7 GGX alpha of primitive 0 (the table facet of a Diamond),
 *               p8-23 GGX alpha of 16 facet groups (primitive k >= 1 uses
 *               p[8 + (k-1) % 16]).  Rough dispersive dielectric with the
 *               fire tint of SPECTRAL_DIELECTRIC on the first refraction;
 *               extension (the differentiable gem of the adjoint path).
 *
 * Texture of LAMBERTIAN (albedo) and DIFFUSE_LIGHT (emission), selected by
 * rt_material.tex (Texture::value, src/texture.rs:26-56):
 *   RT_TEX_SOLID   p0-2 colour
 *   RT_TEX_CHECKER p0-2 even, p3 scale, p4-6 odd: parity of the floors of
 *                  scale * p
 *   RT_TEX_IMAGE   p7 index into rt_scene_desc.images: nearest texel at
 *                  (clamp(u), 1 - clamp(v)), sRGB-decoded through the 256-entry
 *                  IEC 61966-2-1 table (src/output.rs:88-98)
 *   RT_TEX_NOISE   p3 scale, p8 Perlin table index: grey
 *                  0.5 (1 + sin(scale p.z + 10 turb(p, 7)))
 *   RT_TEX_MOTTLE  p0-2 lo, p3 scale, p4-6 hi, p8 Perlin table index:
 *                  lerp(lo, hi, clamp(2 turb(scale p, 7), 0, 1))
 */
enum { RT_TEX_SOLID = 0, RT_TEX_CHECKER = 1, RT_TEX_IMAGE = 2, RT_TEX_NOISE = 3, RT_TEX_MOTTLE = 4 };
enum {
    RT_MAT_LAMBERTIAN          = 1,
    RT_MAT_DIELECTRIC          = 2,
    RT_MAT_SPECTRAL_DIELECTRIC = 3,
    RT_MAT_PBR                 = 4,
    RT_MAT_DIFFUSE_LIGHT       = 5,
    RT_MAT_SPECTRAL_METAL      = 6,
    RT_MAT_ROUGH_DIELECTRIC    = 7,
    RT_MAT_ROUGH_SPECTRAL      = 8,
    RT_MAT_BLACKBODY           = 9,
    RT_MAT_SSS                 = 10,
    RT_MAT_PEARL               = 11
};

typedef struct {
    uint32_t kind;
    uint32_t tex;
    float    p[24];
} rt_material;

/* ---- background (src/renderer.rs:183-223) -------------------------------- */
enum {
    RT_BG_SOLID     = 0,
    RT_BG_GRADIENT  = 1,
    RT_BG_PHYSICAL  = 2,   /* Preetham sky + sun disc */
    RT_BG_ENVMAP    = 3,
    RT_BG_ENVMAP_SPLIT = 4
};

typedef struct {
    uint32_t     kind;
    uint32_t     visible_kind;  /* ENVMAP_SPLIT: RT_BG_SOLID or RT_BG_GRADIENT */
    float        color[3];      /* SOLID (or visible SOLID)                  */
    float        top[3];        /* GRADIENT                                  */
    float        bottom[3];
    float        sun_dir[3];    /* PHYSICAL, unit                           */
    float        turbidity;
    uint32_t     env_w, env_h;  /* ENVMAP*: equirect float RGB, row-major, row 0 = +Y pole */
    const float *env_rgb;
    float        env_rot_y;     /* radians                                   */
    uint32_t     env_sampling;  /* ENVMAP*: RT_ENV_CDF (the reference's row/column CDF
                                   searches, src/renderer.rs:126-154) or RT_ENV_ALIAS
                                   (Walker/Vose alias table over texels: same density,
                                   same four draws, O(1) instead of two binary searches) */
} rt_background;

enum { RT_ENV_CDF = 0, RT_ENV_ALIAS = 1 };

/* An 8-bit sRGB image (image::RgbImage): width x height texels of 3 bytes,
 * row-major, row 0 at the top (v = 1). */
typedef struct {
    uint32_t       width, height;
    const uint8_t *rgb;
} rt_image;

typedef struct {
    const rt_object   *objects;    uint32_t num_objects;   /* world, input order */
    const rt_object   *lights;     uint32_t num_lights;    /* NEE light list     */
    const rt_material *materials;  uint32_t num_materials;
    const rt_xform    *xforms;     uint32_t num_xforms;
    const rt_medium   *media;      uint32_t num_media;
    const float       *pool;       uint32_t pool_len;      /* CONVEX planes      */
    /* triangle soup: 9 floats of vertices per triangle; normals (9), tangents (9)
     * and uvs (6) per triangle; normals/tangents/uvs may be NULL (face normal,
     * tangent = unit(v1-v0), uv (0,0),(1,0),(0.5,1) as Triangle::flat)         */
    const float *tri_v; const float *tri_n; const float *tri_t; const float *tri_uv;
    uint32_t     num_tris;
    rt_background background;
    /* textures: images for RT_TEX_IMAGE; Perlin tables for RT_TEX_NOISE /
     * RT_TEX_MOTTLE, num_perlin blocks of 1536 floats laid out as rt_medium.perlin
     * (the host's Perlin::new draws, src/perlin.rs:11-33) */
    const rt_image *images;     uint32_t num_images;
    const float    *perlin;     uint32_t num_perlin;
} rt_scene_desc;

/* ---- rays and hits (parity surface) --------------------------------------- */
typedef struct {
    float origin[3];
    float dir[3];
    float t_min, t_max;
    float time;
} rt_ray;                          /* 36 bytes */

typedef struct {
    float    t;
    float    p[3];
    float    n[3];                 /* face-forwarded normal, as HitRecord.normal */
    float    u, v;
    int32_t  obj_id;               /* world object index in input order, -1 = miss */
    int32_t  prim_id;              /* triangle / box-face index inside the object  */
    uint32_t front_face;
} rt_hit;                          /* 48 bytes */

/* ---- camera (src/camera.rs:10-59) ----------------------------------------- */
typedef struct {
    float    origin[3];
    float    lower_left[3];
    float    horizontal[3];
    float    vertical[3];
    float    u[3];
    float    v[3];
    float    lens_radius;
    uint32_t aperture_blades;
} rt_camera;

/* ---- per-pass render parameters (render_tiles arguments) ------------------ */
typedef struct {
    uint32_t sample_idx;     /* render_tiles sample_idx                       */
    uint32_t strata;         /* floor(sqrt(spp_max)), src/main.rs:64          */
    uint32_t width, height;
    float    bg_scale;
    uint32_t max_depth;      /* the reference hard-codes 50 (src/renderer.rs:16) */
    uint32_t use_frame_time; /* frame_time: Some((t0, t1))                    */
    float    frame_t0, frame_t1;
    uint32_t use_photons;    /* photon_map: Some(..) when a map was built     */
    uint64_t seed;           /* mixed into every per-pixel stream             */
    uint32_t wavelengths;    /* 0: one hero wavelength uniform over [380, 700)
                                nm, the reference's (src/renderer.rs:726);
                                extension: 1 one wavelength over [380, 780)
                                with the CIE table extended to 780 nm; 4 four
                                hero wavelengths over [380, 780) (lambda_0 and
                                lambda_0 + 100 k mod 400 nm) combined by the
                                balance heuristic (spectral MIS); the adjoint
                                entry points take 0 or 1                     */
} rt_sample_params;

/* ---- wavefront renderer options ----------------------------------------------
 * capacity: paths in flight per chunk (0: 4M); grid: blocks of the persistent
 * stage kernels (0: 2048); sort: 0 shades paths in extension order, 1 files
 * them by material kind first, 2 splits them into two shade classes (misses,
 * media and the specular materials; everything else), each shaded by a kernel
 * compiled for its materials only;
 * use_graph: 1 captures one chunk's launches in a hipGraph and replays it;
 * occupancy: register budget of the traversal stages — 0 compiler's choice,
 * 4 or 5: at least that many waves per SIMD (fewer VGPRs, some spills);
 * streams: n in 2..RT_WF_MAX_STREAMS runs consecutive chunks round-robin on
 * n HIP streams, each with queues of its own, so a chunk's last bounces (few
 * paths, latency-bound) overlap the next chunks' first ones; 0 or 1: one
 * stream;
 * block: lanes per block of the traversal stages (closest hit, shadow): 64,
 * 128 or 256 (0: 128);
 * knn_wave_max: k-NN gathers that still need the nearest-photon search after
 * the first pass are searched one query per lane (each with a heap in LDS)
 * when more than this many wait, else one query per wave (0: 65536).        */
#define RT_WF_MAX_STREAMS 4
typedef struct {
    uint32_t capacity;
    uint32_t grid;
    uint32_t sort;
    uint32_t use_graph;
    uint32_t occupancy;
    uint32_t streams;
    uint32_t block;
    uint32_t knn_wave_max;
} rt_wavefront_opts;

/* ---- caustic photon map (src/photon.rs, src/scene.rs:35-74) --------------- */
/* RT_EMIT_QUAD: PhotonMap::build_from_quad (src/photon.rs:97-135).
 * RT_EMIT_SUN_DISK: PhotonMap::build (src/photon.rs:41-94): a disk facing
 *   sun_dir, sized to the world's photon_bounds (finite geometry only: an
 *   infinite plane has none, src/plane.rs:102-106; transforms forward theirs),
 *   or a 20-unit disk 30 units up sun_dir when no finite geometry exists.
 * RT_EMIT_SCENE: SceneData::rebuild_caustics (src/scene.rs:35-74): the scene's
 *   background picks the emitter.  Physical: SUN_DISK toward its sun_dir with
 *   colour pi * background(power_dir), power_dir = unit(sun_dir + 0.04 *
 *   perp); env map (EnvMap / EnvMapSplit): SUN_DISK toward the normalised sum
 *   of (0, 1, 0) and eight CDF direction draws of SmallRng(0xC0FFEE), colour
 *   pi * env(dir); otherwise QUAD with origin / u / v / color (the scene's
 *   caustic_quad, its first light).  rt_photon_emitter reports the choice.   */
enum { RT_EMIT_QUAD = 0, RT_EMIT_SUN_DISK = 1, RT_EMIT_SCENE = 2 };

typedef struct {
    uint32_t emitter;          /* RT_EMIT_*                                      */
    float    origin[3], u[3], v[3];   /* QUAD (and SCENE's fallback quad)         */
    float    color[3];         /* QUAD: light colour; SUN_DISK: sun_color        */
    float    sun_dir[3];       /* SUN_DISK                                       */
    uint32_t num_photons;
    float    gather_radius;
    uint32_t knn;              /* 0: fixed-radius gather (PhotonMap::irradiance,
                                  src/photon.rs:148-157); 1..RT_KNN_MAX: k-NN
                                  gather (extension): the squared radius is the
                                  k-th smallest squared distance among the
                                  hemisphere's photons closer than gather_radius
                                  (gather_radius when fewer), normalised by
                                  2 / (pi^2 R^2)                               */
    uint32_t wavelengths;      /* 0: photon wavelength over [380, 700), the
                                  reference's (src/photon.rs:244); 1 or 4
                                  (extension): over [380, 780) with the
                                  extended CIE table, for renders with
                                  wavelengths 1 or 4                           */
} rt_photon_params;

#define RT_KNN_MAX 64

/* ---- per-pixel accumulation (src/main.rs:124-133, 603-677, 696-718) --------
 * The firefly clamp and adaptive-sampling settings of the reference's render
 * loop.  Defaults there: firefly_clamp 8, firefly_min_samples 4,
 * firefly_abs_floor 5 (FireflyConfig::default), min_samples 128, threshold
 * 0.05 (ADAPTIVE_THRESHOLD).                                                 */
typedef struct {
    float    firefly_clamp;        /* 0 disables the clamp                      */
    uint32_t firefly_min_samples;  /* clamp once a pixel holds this many samples */
    float    firefly_abs_floor;    /* threshold = max(clamp * mean luminance, floor) */
    uint32_t adaptive;             /* 1: converged pixels stop being rendered   */
    uint32_t min_samples;          /* adaptive: mark_converged from this many passes */
    float    convergence_pct;      /* adaptive: stop once this % of pixels converged */
    float    threshold;            /* relative standard error of every non-dark channel */
    uint32_t reserved;
} rt_accum_opts;

/* The caller-owned per-pixel state of one image (width*height pixels): the
 * accumulator (3 floats per pixel), the count of finite samples, the Welford
 * M2 per channel (3 floats) and the convergence flags.  Host or device
 * memory; the four may not mix host and device pointers.                     */
typedef struct {
    float    *accum;
    uint32_t *samples;
    float    *var_m2;
    uint8_t  *converged;
} rt_pixel_state;

/* ---- context ---------------------------------------------------------------- */
int         rt_ctx_create(int device, rt_ctx **out);
void        rt_ctx_destroy(rt_ctx *ctx);
const char *rt_last_error(rt_ctx *ctx);
int         rt_api_version(void);

/* Upload a scene: builds the SAH BVH (src/bvh.rs:199-323) and the device
 * layout.  Replaces the world BvhTree + lights HittableList + Background. */
int  rt_scene_upload(rt_ctx *ctx, const rt_scene_desc *desc, rt_scene **out);
void rt_scene_destroy(rt_scene *scene);

/* Scene statistics for measurement.  stats[0] BVH nodes (world + nested),
 * [1] triangles, [2] node bytes, [3] worst-case traversal stack depth,
 * [4] world objects, [5] triangle bytes; [6], [7] reserved.                 */
int  rt_scene_stats(rt_scene *scene, uint64_t stats[8]);

/* Closest hit per ray: BvhTree::hit_index (src/bvh.rs:229-288). */
int  rt_trace_closest(rt_ctx *ctx, rt_scene *scene, const rt_ray *rays, uint32_t n, rt_hit *hits);
/* Occlusion per ray: BvhTree::any_hit (src/bvh.rs:487-530). */
int  rt_trace_any(rt_ctx *ctx, rt_scene *scene, const rt_ray *rays, uint32_t n, uint8_t *occluded);
/* Shadow-ray transmittance: BvhTree::transmittance (src/bvh.rs:532-575). */
int  rt_transmittance(rt_ctx *ctx, rt_scene *scene, const rt_ray *rays, uint32_t n, float *tr);

/* One sample pass over the image: render_tiles (src/renderer.rs:655-730).
 * scratch_rgb: width*height*3 floats, row 0 = image top.  converged may be
 * NULL; nonzero entries are written black.                                   */
int  rt_render_sample(rt_ctx *ctx, rt_scene *scene, const rt_camera *cam,
                      const rt_sample_params *params, const uint8_t *converged,
                      float *scratch_rgb);

/* spp consecutive passes (sample_idx .. sample_idx+spp-1) summed on the device
 * into accum_rgb (width*height*3, overwritten).  Plain megakernel.           */
int  rt_render_accumulate(rt_ctx *ctx, rt_scene *scene, const rt_camera *cam,
                          const rt_sample_params *params, uint32_t spp,
                          float *accum_rgb);

/* The same frame as rt_render_accumulate (same per-pixel streams, same sums up
 * to float atomics' order), rendered as a wavefront: persistent stage kernels
 * (extend / shade / shadow) over device-resident path queues with compaction
 * and material sorting.  opts may be NULL.  No counterpart in the reference,
 * which renders each pixel's path in one loop (src/renderer.rs:379-582).     */
int  rt_render_wavefront(rt_ctx *ctx, rt_scene *scene, const rt_camera *cam,
                         const rt_sample_params *params, uint32_t spp,
                         float *accum_rgb, const rt_wavefront_opts *opts);

/* rt_render_wavefront for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  The path queues live in a
 * fixed device array of RT_NOALLOC_CAPACITY paths (53.5 MiB) that every GPU
 * the library runs on holds until the process ends; the call allocates and
 * frees nothing (no hipMalloc, hipFree, streams, graphs or events) and clears
 * with kernels, not hipMemsetAsync.  accum_rgb must be device memory.  opts
 * (may be NULL): capacity 0 = RT_NOALLOC_CAPACITY, larger is RT_ERR_INVALID;
 * streams must be <= 1 and use_graph 0.  Same frame as rt_render_wavefront
 * with the same options.  Calls that overlap on one device share the array
 * and race; each call leaves its queues in it.                               */
#define RT_NOALLOC_CAPACITY (1u << 17)
int  rt_render_wavefront_noalloc(rt_ctx *ctx, rt_scene *scene, const rt_camera *cam,
                                 const rt_sample_params *params, uint32_t spp,
                                 float *accum_rgb, const rt_wavefront_opts *opts);

/* One render of a batch: the scene it draws (each view may differ, e.g. the
 * same gem spun to another angle), its camera and pass parameters (width,
 * height, max_depth, bg_scale, use_photons, seed, sample_idx, strata), its
 * samples per pixel and its accumulator (width*height*3 floats, host or
 * device, overwritten with the sum of its spp passes).                       */
typedef struct {
    rt_scene        *scene;
    rt_camera        camera;
    rt_sample_params params;
    uint32_t         spp;
    float           *accum_rgb;
} rt_batch_view;

/* Every view of a batch in one wavefront: renders of uneven resolution, depth
 * and scene share the path queues, so short views never leave the device
 * idle.  Replaces one process per still (spin_stills.py:64-101).           */
int  rt_render_batch(rt_ctx *ctx, const rt_batch_view *views, uint32_t nviews,
                     const rt_wavefront_opts *opts);

/* ---- differentiable material parameters (extension: no counterpart in the
 * reference, which has no derivatives) ----------------------------------------
 * Parameter i of a vector theta sets rt_material.p[field .. field+count-1] of
 * material `material` to theta[i] (count 3 on a LAMBERTIAN albedo makes one
 * grey albedo parameter).  Differentiable fields: LAMBERTIAN (solid, tex 0)
 * p0-2; ROUGH_SPECTRAL p0 (n_d), p1-3 (Sellmeier B), p7-23 (GGX alpha).
 * material = RT_PARAM_MEDIUM | k addresses rt_scene_desc.media[k] instead:
 * field 0 (count 1) is a CONSTANT medium's density sigma_t (its boundary
 * objects must carry no transform).  Others return RT_ERR_UNSUPPORTED.      */
typedef struct {
    uint32_t material;
    uint32_t field;
    uint32_t count;
} rt_param;
#define RT_PARAM_MEDIUM 0x40000000u

/* Write theta into the scene's materials (device and host copies); the
 * geometry, BVH and photon map are untouched.                                */
int  rt_scene_set_params(rt_ctx *ctx, rt_scene *scene, const rt_param *params, uint32_t n,
                         const float *values);
int  rt_scene_get_params(rt_ctx *ctx, rt_scene *scene, const rt_param *params, uint32_t n,
                         float *values);

/* The frame of rt_render_accumulate with every sampling decision (directions,
 * lobe and Fresnel choices, Russian roulette) drawn with the scene's current
 * parameters and every BSDF value evaluated at eval_values: the estimator the
 * path-replay adjoint differentiates.  With eval_values equal to the current
 * parameters it is rt_render_accumulate's frame; at theta +- h with the same
 * seed it gives central finite differences with common random numbers.       */
int  rt_render_detached(rt_ctx *ctx, rt_scene *scene, const rt_camera *cam,
                        const rt_sample_params *params, uint32_t spp,
                        const rt_param *dparams, uint32_t n, const float *eval_values,
                        float *accum_rgb);

/* Path-replay adjoint.  For a loss of the mean image I = accum / spp, with
 * dloss_dimage = dLoss/dI (width*height*3 floats, host or device), writes
 * grad[n] = dLoss/dtheta at the scene's current parameters (host or device,
 * overwritten).  Each sample is traced once for its radiance and once more
 * from its stored stream seed to distribute that radiance over the path's
 * vertices; no tape is kept.                                                  */
int  rt_backward(rt_ctx *ctx, rt_scene *scene, const rt_camera *cam,
                 const rt_sample_params *params, uint32_t spp,
                 const rt_param *dparams, uint32_t n, const float *dloss_dimage,
                 float *grad);

/* rt_backward as wavefronts: per chunk of paths, the primal wavefront writes
 * each sample's radiance, then the replay wavefront (the same extend / shade
 * / shadow stages with per-path adjoint state) distributes it.  Same gradient
 * as rt_backward up to the order of float additions; opts as for
 * rt_render_wavefront (may be NULL).                                          */
int  rt_backward_wavefront(rt_ctx *ctx, rt_scene *scene, const rt_camera *cam,
                           const rt_sample_params *params, uint32_t spp,
                           const rt_param *dparams, uint32_t n, const float *dloss_dimage,
                           float *grad, const rt_wavefront_opts *opts);

/* Trace photons and build the kd-tree on the device: PhotonMap::build /
 * build_from_quad (src/photon.rs:41-135, 165-191).  The map stays attached to
 * the scene; rt_sample_params.use_photons enables the gather.                */
int  rt_photon_build(rt_ctx *ctx, rt_scene *scene, const rt_photon_params *pp,
                     uint32_t *num_stored);
/* The emitter rt_photon_build would use for pp: RT_EMIT_SCENE resolved into
 * the SUN_DISK (sun_dir, color) or QUAD rebuild_caustics picks
 * (src/scene.rs:35-74); other kinds are copied.  RT_ERR_INVALID when
 * RT_EMIT_SCENE finds neither a sun, an env map nor a quad (u x v = 0): the
 * reference then builds no map.                                              */
int  rt_photon_emitter(rt_ctx *ctx, rt_scene *scene, const rt_photon_params *pp,
                       rt_photon_params *out);
/* The world's photon_bounds (src/bvh.rs:206-208, src/hittable.rs:199-211):
 * min[3], max[3] into out6; *has = 0 when only proxy geometry (infinite
 * planes) exists.  Host-side data of the uploaded scene.                     */
int  rt_scene_photon_bounds(rt_ctx *ctx, rt_scene *scene, float *out6, int *has);
/* Attach a photon map built elsewhere (e.g. by the host's own
 * rebuild_caustics): n photons of 9 floats (position, power, normal), the
 * layout rt_photon_export writes; the device orders them for its gather.    */
int  rt_photon_import(rt_ctx *ctx, rt_scene *scene, const float *photons, uint32_t n,
                      float gather_radius);
/* Change the gather of the attached map: fixed radius (knn 0) or k-NN with
 * gather_radius as the search bound (knn 1..RT_KNN_MAX; see rt_photon_params).
 * No counterpart in the reference, which gathers in a fixed radius.          */
int  rt_photon_set_gather(rt_ctx *ctx, rt_scene *scene, float gather_radius, uint32_t knn);
/* Irradiance queries against the attached map: PhotonMap::irradiance
 * (src/photon.rs:148-157).  pos_nrm: n * 6 floats (p.xyz, n.xyz).            */
int  rt_photon_irradiance(rt_ctx *ctx, rt_scene *scene, const float *pos_nrm,
                          uint32_t n, float *irr_rgb);
/* Copy the attached map's photons (kd order) out: up to cap x 9 floats
 * (position, power, normal); *count receives the map size.                   */
int  rt_photon_export(rt_ctx *ctx, rt_scene *scene, float *out, uint32_t cap,
                      uint32_t *count);

/* accumulate_sample (src/main.rs:603-644): add one sample pass (scratch_rgb,
 * npix*3 floats) into the state.  Per pixel: converged pixels and samples
 * with a non-finite channel are skipped; once the pixel holds
 * firefly_min_samples samples a sample whose luminance exceeds
 * max(firefly_clamp * mean luminance, firefly_abs_floor) is scaled down to
 * it; then the sum, the count and the per-channel Welford M2 are updated.    */
int  rt_accumulate_sample(rt_ctx *ctx, const float *scratch_rgb, uint32_t npix,
                          const rt_accum_opts *opts, const rt_pixel_state *state);

/* mark_converged (src/main.rs:646-677): flag every pixel with samples whose
 * non-dark channels (mean >= 1e-4) all have a relative standard error below
 * opts->threshold; *n_converged receives the number of flagged pixels.       */
int  rt_mark_converged(rt_ctx *ctx, uint32_t npix, const rt_accum_opts *opts,
                       const rt_pixel_state *state, uint32_t *n_converged);

/* The reference's render loop (src/main.rs:463-506) on the device: for
 * s = 0 .. samples_max-1 one pass of sample index k = params->sample_idx + s
 * (render_tiles, converged pixels skipped when adaptive), accumulated as
 * rt_accumulate_sample does, then, when adaptive and k + 1 >= min_samples,
 * rt_mark_converged (k + 1 is the frame's pass count, so a call continuing a
 * frame at sample_idx = k marks as the reference's loop does at pass k); the
 * loop stops after a pass when adaptive and the converged count reaches
 * convergence_pct % of the pixels.  Render, accumulation and the convergence
 * test run fused per pixel, one launch per pass, with no host round trip: the
 * stop test is read by the next launch.  *samples_done receives the passes
 * run, *n_converged the flagged pixels (either may be NULL).  The state is
 * read and updated in place, so a later call continues it.                  */
int  rt_render_progressive(rt_ctx *ctx, rt_scene *scene, const rt_camera *cam,
                           const rt_sample_params *params, uint32_t samples_max,
                           const rt_accum_opts *opts, const rt_pixel_state *state,
                           uint32_t *samples_done, uint32_t *n_converged);

/* rt_render_progressive's loop with the passes rendered by the wavefront
 * renderer (opts as rt_render_wavefront's; NULL: its defaults): chunks of
 * whole passes (capacity / pixels passes each, one when adaptive) render
 * their per-sample radiance, then fold into the state pass by pass in order.
 * Same per-pixel streams as the megakernel loop; each sample's radiance sums
 * its contributions in the wavefront's stage order, so states agree to float
 * rounding.  No counterpart in the reference.                               */
int  rt_render_progressive_wavefront(rt_ctx *ctx, rt_scene *scene, const rt_camera *cam,
                                     const rt_sample_params *params, uint32_t samples_max,
                                     const rt_accum_opts *opts, const rt_pixel_state *state,
                                     const rt_wavefront_opts *wf_opts, uint32_t *samples_done,
                                     uint32_t *n_converged);

/* render_aux_pass (src/renderer.rs:597-645): the denoiser's auxiliary
 * buffers, width*height*3 floats each, row 0 = image top.  One ray per pixel
 * through the pixel centre; a hit stores the material's albedo hint clamped
 * to [0, 1] and the hit's normal, a miss the background colour and a zero
 * normal.                                                                    */
int  rt_render_aux(rt_ctx *ctx, rt_scene *scene, const rt_camera *cam, uint32_t width, uint32_t height,
                   float *albedo, float *normal);

/* The saved image (src/main.rs:571-582, src/output.rs:103-190): per pixel the
 * mean radiance accum / max(n, 1) (n = samples[i], or spp when samples is
 * NULL), then apply_optics (vignette 1 - vignette * r^2, barrel distortion
 * sampling c + (p - c) * (1 - distortion * r^2) bilinearly, r = 1 at the
 * corners; both 0: untouched), then tone_map with scale = exposure (ACES or
 * AgX "Punchy"), the IEC sRGB encode and the 8-bit store
 * (min(encode, 0.999) * 256, truncated).  rgb8: width*height*3 bytes, row 0 =
 * image top.                                                                */
enum { RT_TONE_ACES = 0, RT_TONE_AGX = 1 };
typedef struct {
    float    exposure;
    uint32_t tonemapper;           /* RT_TONE_ACES / RT_TONE_AGX */
    float    vignette;
    float    distortion;
} rt_resolve_opts;
int  rt_resolve_rgb8(rt_ctx *ctx, const float *accum, const uint32_t *samples, uint32_t spp,
                     uint32_t width, uint32_t height, const rt_resolve_opts *opts, uint8_t *rgb8);

/* ---- measurement (no counterpart in the reference: its --bench mode times
 * render_tiles with Instant, src/main.rs:68-117) ------------------------------ */
/* Bracket every render call's kernels (and those of rt_accumulate_sample,
 * rt_mark_converged, rt_resolve_rgb8 and rt_render_aux) with HIP events on
 * the context's stream; enabling resets the totals.                           */
int  rt_ctx_set_timing(rt_ctx *ctx, int enable);
/* Summed kernel time (ms) and kernel launch count since timing was enabled. */
int  rt_ctx_get_timing(rt_ctx *ctx, double *kernel_ms, uint64_t *launches);
/* Counting pass: renders add their traversal work to device counters.
 * Enabling zeroes them.                                                       */
int  rt_ctx_set_counting(rt_ctx *ctx, int enable);
/* Megakernel renders since counting was enabled: out[0] BVH node tests,
 * [1] primitive tests, [2] rays of closest-hit queries, [3] samples,
 * [4] node tests, [5] primitive tests, [6] rays of shadow queries, [7] bytes
 * of photon-map records the photon gathers read (kd photon: 32 B position +
 * normal, 16 B power when summed; bucket tree: 8 B per split plane, 32 B
 * record per photon of a visited leaf, 16 B power when summed).            */
int  rt_ctx_get_counters(rt_ctx *ctx, uint64_t out[8]);
/* Wavefront renders with timing on and graphs off: summed duration (ms) and
 * count of the k_wf_extend launches (closest-hit traversal stage).           */
int  rt_ctx_get_stage_timing(rt_ctx *ctx, double *extend_ms, uint64_t *extend_launches);
/* The same for every stage of a bounce: [0] closest hit (k_wf_extend), [1]
 * k-NN photon gather (k_wf_knn + k_wf_knn_dense + k_wf_knn_wave, bounces of
 * renders with a k-NN photon map only), [2] shade (k_wf_shade), [3] shadow
 * (k_wf_shadow): summed duration (ms) and number of timed bounces.           */
int  rt_ctx_get_stage_times(rt_ctx *ctx, double ms[4], uint64_t launches[4]);
/* Register budget of the plain megakernel (rt_render_sample,
 * rt_render_accumulate, rt_render_tiles): 0 the compiler's choice (all 256
 * VGPRs, one wave per SIMD), 2 at least two waves per SIMD (at most 128
 * VGPRs, the rest spilled).  Same frames either way; a tuning knob, as the
 * wavefront's occupancy option is.                                           */
int  rt_ctx_set_megakernel_occupancy(rt_ctx *ctx, uint32_t occupancy);

#ifdef __cplusplus
}
#endif
#endif /* RT_API_H */
