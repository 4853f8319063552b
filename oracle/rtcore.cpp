// ORACLE — test infrastructure only.  The product library never links this
// file; tests/, __graft_entry__.smoke() and bench.py's cpu_baseline load the
// built liboracle_rtcore.so through ctypes, as the checker.
//
// Builds rustracer's object graph (dyn Hittable / dyn Material, SAH BvhTree)
// from the same flattened rt_scene_desc the product consumes, the way
// scene_file::build does (src/scene_file.rs:899-1132): shape, then the medium
// wrapper, then the SRT chain; emissive world objects are NOT shared into the
// light list here — the descriptor lists lights explicitly.
#include "rc_render.h"
#include "rc_adjoint.h"
#include "../include/rt_api.h"
#include <string>
#include <cstdio>

using namespace rc;

namespace {

struct OScene {
    PearlSun pearl_sun;                      // set from a Physical background (render_tiles, src/renderer.rs:676-679)
    std::vector<MediumSet> media;            // per descriptor medium: its constant-medium objects (sigma_t parameters)
    std::vector<std::unique_ptr<Material>> mats;
    std::vector<rt_material> desc_mats;     // the descriptor's table, kept for the adjoint checker
    std::vector<HPtr> world_objs;
    BvhTree world;
    HittableList lights;
    Background bg;
    PhotonMap photons;
    bool have_photons = false;
    std::string err;
};

static bool perlin_perms_ok(const float* t) {
    for (int k = 768; k < 1536; k++)
        if (!(t[k] >= 0.0f && t[k] <= 255.0f && t[k] == std::floor(t[k]))) return false;
    return true;
}

// Texture of a LAMBERTIAN / DIFFUSE_LIGHT entry; images and Perlin tables
// come from the descriptor (null: solid and checker only).
static bool make_texture(const rt_material& m, const rt_scene_desc* d, Texture& t, std::string& err) {
    t.kind = (int)m.tex;
    t.even = Color(m.p[0], m.p[1], m.p[2]);
    t.scale = m.p[3];
    t.odd = Color(m.p[4], m.p[5], m.p[6]);
    if (m.tex == RT_TEX_IMAGE) {
        float f = m.p[7];
        if (!d || !d->images || !(f >= 0.0f && f < (float)d->num_images && f == std::floor(f))) { err = "image index"; return false; }
        const rt_image& im = d->images[(uint32_t)f];
        if (!im.rgb || !im.width || !im.height) { err = "empty image"; return false; }
        auto img = std::make_shared<RgbImage>();
        img->w = im.width; img->h = im.height;
        img->rgb.assign(im.rgb, im.rgb + (size_t)im.width * im.height * 3);
        t.img = img;
    } else if (m.tex == RT_TEX_NOISE || m.tex == RT_TEX_MOTTLE) {
        float f = m.p[8];
        if (!d || !d->perlin || !(f >= 0.0f && f < (float)d->num_perlin && f == std::floor(f))) { err = "Perlin table index"; return false; }
        const float* tab = d->perlin + (size_t)1536 * (uint32_t)f;
        if (!perlin_perms_ok(tab)) { err = "Perlin permutation entries must be integers in 0..255"; return false; }
        auto pn = std::make_shared<Perlin>();
        pn->load(tab);
        t.perlin = pn;
    } else if (m.tex > RT_TEX_MOTTLE) { err = "unsupported texture kind"; return false; }
    return true;
}

Material* make_material(const rt_material& m, const rt_scene_desc* d, std::string& err) {
    if (m.tex != RT_TEX_SOLID && m.kind != RT_MAT_LAMBERTIAN && m.kind != RT_MAT_DIFFUSE_LIGHT) {
        err = "only LAMBERTIAN and DIFFUSE_LIGHT take a texture"; return nullptr;
    }
    switch (m.kind) {
        case RT_MAT_LAMBERTIAN: {
            auto* l = new Lambertian();
            if (!make_texture(m, d, l->tex, err)) { delete l; return nullptr; }
            return l;
        }
        case RT_MAT_DIELECTRIC: { auto* d = new Dielectric(); d->ir = m.p[0]; return d; }
        case RT_MAT_SPECTRAL_DIELECTRIC: {
            auto* d = new SpectralDielectric();
            d->cauchy_b = m.p[0]; d->cauchy_c = m.p[1]; d->absorption = Color(m.p[2], m.p[3], m.p[4]);
            d->model = (int)m.p[5];
            for (int k = 0; k < 3; k++) { d->sB[k] = m.p[6 + k]; d->sC[k] = m.p[9 + k]; }
            return d;
        }
        case RT_MAT_PBR: {
            auto* p = new PbrMaterial();
            p->albedo = Color(m.p[0], m.p[1], m.p[2]); p->roughness = m.p[3]; p->metallic = m.p[4];
            p->anisotropy = m.p[5]; p->anisotropy_angle = m.p[6]; p->clearcoat = m.p[7]; p->clearcoat_roughness = m.p[8];
            p->emission = Color(m.p[9], m.p[10], m.p[11]); p->emission_strength = m.p[12];
            p->sheen = m.p[13]; p->sheen_tint = m.p[14];
            p->film_thickness = m.p[15]; p->film_ior = m.p[16];
            std::shared_ptr<const RgbImage>* maps[5] = {&p->albedo_tex, &p->roughness_tex, &p->metallic_tex, &p->ao_tex, &p->normal_tex};
            for (int k = 0; k < 5; k++) {
                float f = m.p[17 + k];
                if (f == 0.0f) continue;
                if (!d || !d->images || !(f >= 1.0f && f <= (float)d->num_images && f == std::floor(f))) { err = "PBR map image index"; delete p; return nullptr; }
                const rt_image& im = d->images[(uint32_t)f - 1];
                if (!im.rgb || !im.width || !im.height) { err = "empty image"; delete p; return nullptr; }
                auto img = std::make_shared<RgbImage>();
                img->w = im.width; img->h = im.height;
                img->rgb.assign(im.rgb, im.rgb + (size_t)im.width * im.height * 3);
                *maps[k] = img;
            }
            return p;
        }
        case RT_MAT_DIFFUSE_LIGHT: {
            auto* l = new DiffuseLight();
            if (!make_texture(m, d, l->emit, err)) { delete l; return nullptr; }
            return l;
        }
        case RT_MAT_SPECTRAL_METAL: { auto* s = new SpectralMetal(); s->variant = (int)m.p[0]; s->roughness = m.p[1]; s->init(); return s; }
        case RT_MAT_ROUGH_DIELECTRIC: { auto* r = new RoughDielectric(); r->ior = m.p[0]; r->roughness = m.p[1]; return r; }
        case RT_MAT_BLACKBODY: { auto* b = new BlackbodyLight(); b->temp_k = m.p[0]; b->intensity = m.p[1]; b->init(); return b; }
        case RT_MAT_PEARL: {
            auto* q = new PearlMaterial();
            q->base_color = Color(m.p[0], m.p[1], m.p[2]); q->ior = m.p[3]; q->film_thickness = m.p[4];
            q->orient_strength = m.p[5]; q->film_scale = m.p[6]; q->luster_roughness = m.p[7];
            return q;
        }
        case RT_MAT_SSS: {
            auto* q = new SSSMaterial();
            q->albedo = Color(m.p[0], m.p[1], m.p[2]); q->sigma_a = Color(m.p[3], m.p[4], m.p[5]);
            q->ior = m.p[6]; q->density = m.p[7]; q->g = m.p[8];
            return q;
        }
        case RT_MAT_ROUGH_SPECTRAL: {
            auto* r = new RoughSpectralDielectric();
            r->nd = m.p[0];
            for (int k = 0; k < 3; k++) { r->sB[k] = m.p[1 + k]; r->sC[k] = m.p[4 + k]; }
            for (int k = 0; k < 17; k++) r->alpha[k] = m.p[7 + k];
            return r;
        }
        default: return nullptr;
    }
}

HPtr make_box(Point3 p0, Point3 p1, const Material* mat) {
    Point3 mn(rmin(p0.x, p1.x), rmin(p0.y, p1.y), rmin(p0.z, p1.z));
    Point3 mx(rmax(p0.x, p1.x), rmax(p0.y, p1.y), rmax(p0.z, p1.z));
    Vec3 dx(mx.x - mn.x, 0, 0), dy(0, mx.y - mn.y, 0), dz(0, 0, mx.z - mn.z);
    std::vector<HPtr> sides = {
        std::make_shared<Quad>(Point3(mn.x, mn.y, mx.z), dx, dy, mat),
        std::make_shared<Quad>(Point3(mx.x, mn.y, mx.z), -dz, dy, mat),
        std::make_shared<Quad>(Point3(mx.x, mn.y, mn.z), -dx, dy, mat),
        std::make_shared<Quad>(Point3(mn.x, mn.y, mn.z), dz, dy, mat),
        std::make_shared<Quad>(Point3(mn.x, mx.y, mx.z), dx, -dz, mat),
        std::make_shared<Quad>(Point3(mn.x, mn.y, mn.z), dx, dz, mat)};
    for (int i = 0; i < 6; i++) sides[i]->tag = i;
    auto b = std::make_shared<BvhTree>();
    b->from_list(sides);
    return b;
}

HPtr make_object(OScene& S, const rt_scene_desc& d, const rt_object& o) {
    if (o.material >= d.num_materials && o.medium == 0) { S.err = "material index out of range"; return nullptr; }
    const Material* mat = o.medium ? nullptr : S.mats[o.material].get();
    HPtr h;
    const float* p = o.p;
    switch (o.shape) {
        case RT_SHAPE_SPHERE: {
            auto s = std::make_shared<Sphere>(Point3(p[0], p[1], p[2]), p[3], mat);
            if (p[7] != 0.0f) { s->moving = true; s->center1 = Point3(p[4], p[5], p[6]); }
            h = s; break;
        }
        case RT_SHAPE_QUAD: h = std::make_shared<Quad>(Point3(p[0], p[1], p[2]), Vec3(p[3], p[4], p[5]), Vec3(p[6], p[7], p[8]), mat); break;
        case RT_SHAPE_BOX: h = make_box(Point3(p[0], p[1], p[2]), Point3(p[3], p[4], p[5]), mat); break;
        case RT_SHAPE_MESH: {
            if ((uint64_t)o.data_first + o.data_count > d.num_tris || o.data_count == 0) { S.err = "mesh triangle range"; return nullptr; }
            std::vector<HPtr> tris;
            tris.reserve(o.data_count);
            for (uint32_t k = 0; k < o.data_count; k++) {
                size_t t = (size_t)o.data_first + k;
                auto tr = std::make_shared<Triangle>(d.tri_v + t * 9, d.tri_n ? d.tri_n + t * 9 : nullptr,
                                                     d.tri_t ? d.tri_t + t * 9 : nullptr, d.tri_uv ? d.tri_uv + t * 6 : nullptr, mat);
                tr->tag = (int)k;
                tris.push_back(tr);
            }
            auto b = std::make_shared<BvhTree>();
            b->from_list(tris);
            h = b; break;
        }
        case RT_SHAPE_CONVEX: {
            if ((uint64_t)(o.data_first + o.data_count) * 4 > d.pool_len) { S.err = "convex plane range"; return nullptr; }
            auto c = std::make_shared<Convex>();
            for (uint32_t k = 0; k < o.data_count; k++) {
                const float* q = d.pool + ((size_t)o.data_first + k) * 4;
                c->planes.push_back(Plane4{Vec3(q[0], q[1], q[2]), q[3]});
            }
            c->mat = mat; c->center = Point3(p[0], p[1], p[2]);
            c->crown_h = p[3]; c->pavilion_h = p[4]; c->girdle_h = p[5]; c->polish = p[6];
            c->bbox = Aabb(Point3(p[7], p[8], p[9]), Point3(p[10], p[11], p[12]));
            h = c; break;
        }
        case RT_SHAPE_CYLINDER: {
            auto c = std::make_shared<Cylinder>();
            c->center = Point3(p[0], p[1], p[2]); c->radius = p[3]; c->height = p[4]; c->mat = mat;
            h = c; break;
        }
        case RT_SHAPE_CONE: {
            auto c = std::make_shared<Cone>();
            c->center = Point3(p[0], p[1], p[2]); c->radius = p[3]; c->height = p[4]; c->mat = mat;
            h = c; break;
        }
        case RT_SHAPE_DISK: {
            auto dk = std::make_shared<Disk>();
            dk->init(Point3(p[0], p[1], p[2]), Vec3(p[3], p[4], p[5]), p[6], mat);
            h = dk; break;
        }
        case RT_SHAPE_PLANE: {
            auto pl = std::make_shared<InfinitePlane>();
            pl->init(Point3(p[0], p[1], p[2]), Vec3(p[3], p[4], p[5]), p[6], p[7], mat);
            h = pl; break;
        }
        default: S.err = "unsupported shape kind " + std::to_string(o.shape); return nullptr;
    }
    if (o.medium) {
        if (o.medium > d.num_media) { S.err = "medium index"; return nullptr; }
        const rt_medium& m = d.media[o.medium - 1];
        if (S.media.size() < d.num_media) S.media.resize(d.num_media);
        MediumSet& mset = S.media[o.medium - 1];
        if (o.xform_count) mset.transformed = true;
        if (m.kind != RT_MEDIUM_CONSTANT) mset.constant = false;
        if (m.kind == RT_MEDIUM_CONSTANT) {
            auto cm = std::make_shared<ConstantMedium>();
            cm->boundary = h; cm->density = m.density; cm->neg_inv_density = -1.0f / m.density;
            cm->phase.albedo = Color(m.albedo[0], m.albedo[1], m.albedo[2]); cm->phase.g = m.g;
            mset.objs.push_back(cm.get());
            h = cm;
        } else if (m.kind == RT_MEDIUM_GRID) {
            auto gm = std::make_shared<GridMedium>();
            gm->boundary = h; gm->density_scale = rmax(m.density, FLT_EPSILON);
            gm->phase.albedo = Color(m.albedo[0], m.albedo[1], m.albedo[2]); gm->phase.g = m.g;
            gm->nx = m.grid_n[0]; gm->ny = m.grid_n[1]; gm->nz = m.grid_n[2];
            gm->grid.assign(m.grid, m.grid + (size_t)gm->nx * gm->ny * gm->nz);
            h->bounding_box(gm->box);
            h = gm;
        } else if (m.kind == RT_MEDIUM_NOISE) {
            if (!m.perlin) { S.err = "noise medium without Perlin tables"; return nullptr; }
            auto nm = std::make_shared<NoiseMedium>();
            nm->boundary = h; nm->density_scale = rmax(m.density, FLT_EPSILON);
            nm->phase.albedo = Color(m.albedo[0], m.albedo[1], m.albedo[2]); nm->phase.g = m.g;
            nm->noise_scale = m.noise_scale;
            nm->threshold = rclamp(m.threshold, 0.0f, 0.99f);
            if (!perlin_perms_ok(m.perlin)) { S.err = "noise medium permutation entries must be integers in 0..255"; return nullptr; }
            nm->perlin.load(m.perlin);
            h->bounding_box(nm->box);
            h = nm;
        } else { S.err = "medium kind"; return nullptr; }
    }
    for (uint32_t k = 0; k < o.xform_count; k++) {
        if (o.xform_first + k >= d.num_xforms) { S.err = "xform range"; return nullptr; }
        const rt_xform& x = d.xforms[o.xform_first + k];
        if (x.kind == RT_XFORM_TRANSLATE) {
            auto t = std::make_shared<Translate>(); t->object = h; t->offset = Vec3(x.v[0], x.v[1], x.v[2]); h = t;
        } else if (x.kind == RT_XFORM_ROTATE) {
            auto r = std::make_shared<Rotate>(); r->object = h;
            for (int i = 0; i < 9; i++) r->m[i / 3][i % 3] = x.v[i];
            r->init_bbox(); h = r;
        } else if (x.kind == RT_XFORM_SCALE) {
            auto s = std::make_shared<UniformScale>(); s->object = h; s->factor = x.v[0]; h = s;
        } else { S.err = "xform kind"; return nullptr; }
    }
    return h;
}

bool build_scene(OScene& S, const rt_scene_desc& d) {
    for (uint32_t i = 0; i < d.num_materials; i++) {
        std::string err = "unsupported material kind";
        Material* m = make_material(d.materials[i], &d, err);
        if (!m) { S.err = "material " + std::to_string(i) + ": " + err; return false; }
        m->id = (int)i;
        S.mats.emplace_back(m);
        S.desc_mats.push_back(d.materials[i]);
    }
    if (d.num_objects == 0) { S.err = "empty world"; return false; }
    for (uint32_t i = 0; i < d.num_objects; i++) {
        HPtr h = make_object(S, d, d.objects[i]);
        if (!h) return false;
        h->tag = (int)i;
        S.world_objs.push_back(h);
    }
    S.world.top = true;
    S.world.from_list(S.world_objs);
    for (uint32_t i = 0; i < d.num_lights; i++) {
        HPtr h = make_object(S, d, d.lights[i]);
        if (!h) return false;
        h->tag = (int)i;
        S.lights.objects.push_back(h);
    }
    const rt_background& b = d.background;
    S.bg.kind = (int)b.kind; S.bg.visible_kind = (int)b.visible_kind;
    S.pearl_sun.active = b.kind == RT_BG_PHYSICAL;
    S.pearl_sun.dir = Vec3(b.sun_dir[0], b.sun_dir[1], b.sun_dir[2]);
    for (auto& m : S.mats) if (auto* q = dynamic_cast<PearlMaterial*>(m.get())) q->sun = &S.pearl_sun;
    S.bg.color = Color(b.color[0], b.color[1], b.color[2]);
    S.bg.top = Color(b.top[0], b.top[1], b.top[2]);
    S.bg.bottom = Color(b.bottom[0], b.bottom[1], b.bottom[2]);
    S.bg.sun_dir = Vec3(b.sun_dir[0], b.sun_dir[1], b.sun_dir[2]);
    S.bg.turbidity = b.turbidity;
    if (b.kind == RT_BG_ENVMAP || b.kind == RT_BG_ENVMAP_SPLIT) {
        if (!b.env_rgb || !b.env_w || !b.env_h) { S.err = "env map missing"; return false; }
        S.bg.env.init(b.env_rgb, b.env_w, b.env_h, b.env_rot_y);
        if (b.env_sampling == RT_ENV_ALIAS && !S.bg.env.init_alias()) { S.err = "env map has no energy"; return false; }
    }
    return true;
}

Ray to_ray(const rt_ray& r) {
    return Ray::at_time(Point3(r.origin[0], r.origin[1], r.origin[2]), Vec3(r.dir[0], r.dir[1], r.dir[2]), r.time);
}

Camera to_cam(const rt_camera& c) {
    Camera k;
    k.origin = Point3(c.origin[0], c.origin[1], c.origin[2]);
    k.lower_left = Point3(c.lower_left[0], c.lower_left[1], c.lower_left[2]);
    k.horizontal = Vec3(c.horizontal[0], c.horizontal[1], c.horizontal[2]);
    k.vertical = Vec3(c.vertical[0], c.vertical[1], c.vertical[2]);
    k.u = Vec3(c.u[0], c.u[1], c.u[2]); k.v = Vec3(c.v[0], c.v[1], c.v[2]);
    k.lens_radius = c.lens_radius; k.blades = c.aperture_blades;
    return k;
}

SampleParams to_params(const rt_sample_params& p) {
    SampleParams s;
    s.sample_idx = p.sample_idx; s.strata = p.strata; s.width = p.width; s.height = p.height;
    s.max_depth = p.max_depth; s.bg_scale = p.bg_scale; s.use_frame_time = p.use_frame_time != 0;
    s.t0 = p.frame_t0; s.t1 = p.frame_t1; s.seed = p.seed;
    s.wl = p.wavelengths == 4 ? 2 : (p.wavelengths == 1 ? 1 : 0);
    return s;
}

SceneView view(OScene& S, bool photons) {
    SceneView v;
    v.world = &S.world; v.lights = &S.lights; v.bg = &S.bg;
    v.photons = (photons && S.have_photons) ? &S.photons : nullptr;
    return v;
}

}  // namespace

extern "C" {

void* rtc_scene_create(const rt_scene_desc* d, char* err, int err_len) {
    auto* S = new OScene();
    if (!d || !build_scene(*S, *d)) {
        if (err && err_len > 0) snprintf(err, (size_t)err_len, "%s", S->err.c_str());
        delete S;
        return nullptr;
    }
    return S;
}

void rtc_scene_destroy(void* s) { delete (OScene*)s; }

int rtc_trace_closest(void* s, const rt_ray* rays, uint32_t n, rt_hit* hits, int threads) {
    OScene& S = *(OScene*)s;
    parallel_for(n, threads, [&](size_t i) {
        Ray r = to_ray(rays[i]);
        HitRecord rec;
        rt_hit& h = hits[i];
        memset(&h, 0, sizeof h);
        if (S.world.hit(r, rays[i].t_min, rays[i].t_max, rec)) {
            h.t = rec.t; h.p[0] = rec.p.x; h.p[1] = rec.p.y; h.p[2] = rec.p.z;
            h.n[0] = rec.normal.x; h.n[1] = rec.normal.y; h.n[2] = rec.normal.z;
            h.u = rec.u; h.v = rec.v; h.obj_id = rec.obj_id; h.prim_id = rec.prim_id; h.front_face = rec.front_face;
        } else { h.obj_id = -1; h.prim_id = -1; }
    });
    return 0;
}

// Hittable::hit of one world object, bypassing the world BVH (primitive KATs).
// sub >= 0 selects the nested primitive with that tag (a mesh triangle) and
// calls its own hit, bypassing the nested BVH too.
int rtc_trace_object(void* s, uint32_t obj, int sub, const rt_ray* rays, uint32_t n, rt_hit* hits) {
    OScene& S = *(OScene*)s;
    if (obj >= S.world_objs.size()) return -1;
    const Hittable* target = S.world_objs[obj].get();
    if (sub >= 0) {
        auto* b = dynamic_cast<const BvhTree*>(target);
        if (!b) return -1;
        target = nullptr;
        for (auto& o : b->objects) if (o->tag == sub) target = o.get();
        if (!target) return -1;
    }
    for (uint32_t i = 0; i < n; i++) {
        HitRecord rec;
        rt_hit& h = hits[i];
        memset(&h, 0, sizeof h);
        if (target->hit(to_ray(rays[i]), rays[i].t_min, rays[i].t_max, rec)) {
            h.t = rec.t; h.p[0] = rec.p.x; h.p[1] = rec.p.y; h.p[2] = rec.p.z;
            h.n[0] = rec.normal.x; h.n[1] = rec.normal.y; h.n[2] = rec.normal.z;
            h.u = rec.u; h.v = rec.v; h.obj_id = (int)obj; h.prim_id = rec.prim_id; h.front_face = rec.front_face;
        } else { h.obj_id = -1; h.prim_id = -1; }
    }
    return 0;
}

int rtc_trace_any(void* s, const rt_ray* rays, uint32_t n, uint8_t* out, int threads) {
    OScene& S = *(OScene*)s;
    parallel_for(n, threads, [&](size_t i) { out[i] = S.world.any_hit(to_ray(rays[i]), rays[i].t_min, rays[i].t_max) ? 1 : 0; });
    return 0;
}

int rtc_transmittance(void* s, const rt_ray* rays, uint32_t n, float* out, int threads) {
    OScene& S = *(OScene*)s;
    parallel_for(n, threads, [&](size_t i) { out[i] = S.world.transmittance(to_ray(rays[i]), rays[i].t_min, rays[i].t_max); });
    return 0;
}

int rtc_render_sample(void* s, const rt_camera* cam, const rt_sample_params* p, const uint8_t* conv, float* out, int mode, int threads) {
    OScene& S = *(OScene*)s;
    render_tiles(out, conv, view(S, p->use_photons != 0), to_cam(*cam), to_params(*p), mode, threads);
    return 0;
}

int rtc_render_accumulate(void* s, const rt_camera* cam, const rt_sample_params* p, uint32_t spp, float* accum, int mode, int threads) {
    OScene& S = *(OScene*)s;
    size_t n = (size_t)p->width * p->height * 3;
    std::vector<float> tmp(n);
    for (size_t i = 0; i < n; i++) accum[i] = 0.0f;
    SampleParams sp = to_params(*p);
    for (uint32_t k = 0; k < spp; k++) {
        SampleParams q = sp; q.sample_idx = sp.sample_idx + k;
        render_tiles(tmp.data(), nullptr, view(S, p->use_photons != 0), to_cam(*cam), q, mode, threads);
        for (size_t i = 0; i < n; i++) accum[i] += tmp[i];
    }
    return 0;
}

// SceneData::rebuild_caustics (src/scene.rs:35-74) for RT_EMIT_SCENE: a
// Physical background emits from its sun with the colour of the sky just off
// the sun axis; an env map from the normalised sum of (0, 1, 0) and eight
// luminance-sampled directions of SmallRng(0xC0FFEE) with the map's colour
// there; otherwise the caustic quad in pp.  Other kinds are copied.
// Returns -1 when RT_EMIT_SCENE finds no emitter (no quad given).
int rtc_photon_emitter(void* s, const rt_photon_params* pp, rt_photon_params* out) {
    OScene& S = *(OScene*)s;
    *out = *pp;
    if (pp->emitter != RT_EMIT_SCENE) return 0;
    Vec3 dir; Color c;
    if (S.bg.kind == RT_BG_PHYSICAL) {
        Vec3 sun_dir = S.bg.sun_dir;
        Vec3 perp = std::fabs(sun_dir.x) < 0.9f ? Vec3(0.0f, -sun_dir.z, sun_dir.y).unit()
                                                 : Vec3(-sun_dir.z, 0.0f, sun_dir.x).unit();
        Vec3 power_dir = (sun_dir + perp * 0.04f).unit();
        dir = sun_dir;
        c = S.bg.eval(power_dir) * PI;
    } else if (S.bg.has_env()) {
        SmallRng rng = SmallRng::seed_from_u64(0xC0FFEEull);
        dir = Vec3(0.0f, 1.0f, 0.0f);
        for (int i = 0; i < 8; i++) { float pdf; dir = dir + S.bg.env.sample_dir(rng, pdf, true); }
        dir = dir.unit();
        c = S.bg.env.sample(dir) * PI;
    } else {
        Vec3 u(pp->u[0], pp->u[1], pp->u[2]), v(pp->v[0], pp->v[1], pp->v[2]);
        if (!(u.cross(v).length() > 0.0f)) return -1;
        out->emitter = RT_EMIT_QUAD;
        return 0;
    }
    out->emitter = RT_EMIT_SUN_DISK;
    out->sun_dir[0] = dir.x; out->sun_dir[1] = dir.y; out->sun_dir[2] = dir.z;
    out->color[0] = c.x; out->color[1] = c.y; out->color[2] = c.z;
    return 0;
}

// The world's photon_bounds (src/bvh.rs:206-208): out[0..6) = min, max;
// returns 0 when only proxy geometry (infinite planes) is present.
int rtc_photon_bounds(void* s, float* out) {
    OScene& S = *(OScene*)s;
    Aabb bb;
    if (!S.world.photon_bounds(bb)) return 0;
    out[0] = bb.min.x; out[1] = bb.min.y; out[2] = bb.min.z; out[3] = bb.max.x; out[4] = bb.max.y; out[5] = bb.max.z;
    return 1;
}

int rtc_photon_build(void* s, const rt_photon_params* pp_in, int threads, uint32_t* stored) {
    OScene& S = *(OScene*)s;
    if (pp_in->knn > RT_KNN_MAX) return -1;
    rt_photon_params res;
    if (rtc_photon_emitter(s, pp_in, &res) != 0) return -1;
    const rt_photon_params* pp = &res;
    bool sun_was = S.pearl_sun.active;
    S.pearl_sun.active = false;              // photon tracing sees no pearl sun
    if (pp->emitter == RT_EMIT_QUAD) {
        photon_build_from_quad(S.photons, S.world, Point3(pp->origin[0], pp->origin[1], pp->origin[2]),
                               Vec3(pp->u[0], pp->u[1], pp->u[2]), Vec3(pp->v[0], pp->v[1], pp->v[2]),
                               Color(pp->color[0], pp->color[1], pp->color[2]), pp->num_photons, pp->gather_radius, threads,
                               pp->wavelengths != 0);
    } else {
        Aabb bb; bool have = S.world.photon_bounds(bb);   // PhotonMap::build sizes its disk from photon_bounds (src/photon.rs:61)
        photon_build_sun(S.photons, S.world, have, bb, Vec3(pp->sun_dir[0], pp->sun_dir[1], pp->sun_dir[2]),
                         Color(pp->color[0], pp->color[1], pp->color[2]), pp->num_photons, pp->gather_radius, threads,
                         pp->wavelengths != 0);
    }
    S.photons.knn = pp->knn;
    S.pearl_sun.active = sun_was;
    S.have_photons = true;
    if (stored) *stored = (uint32_t)S.photons.nodes.size();
    return 0;
}

// Gather parameters of the built map: radius and k (0: fixed radius).
int rtc_photon_set_gather(void* s, float radius, uint32_t knn) {
    OScene& S = *(OScene*)s;
    if (knn > RT_KNN_MAX || !(radius > 0.0f)) return -1;
    S.photons.set_radius(radius);
    S.photons.knn = knn;
    return 0;
}

// Photons in kd order: n x 9 floats (pos, power, normal).
int rtc_photon_dump(void* s, float* out, uint32_t cap) {
    OScene& S = *(OScene*)s;
    uint32_t n = (uint32_t)std::min<size_t>(cap, S.photons.nodes.size());
    for (uint32_t i = 0; i < n; i++) {
        const KdPhoton& k = S.photons.nodes[i];
        for (int j = 0; j < 3; j++) { out[i * 9 + j] = k.pos[j]; out[i * 9 + 3 + j] = k.power[j]; out[i * 9 + 6 + j] = k.nor[j]; }
    }
    return (int)n;
}

int rtc_photon_irradiance(void* s, const float* pn, uint32_t n, float* out, int threads) {
    OScene& S = *(OScene*)s;
    parallel_for(n, threads, [&](size_t i) {
        Color c = S.photons.irradiance(Point3(pn[i * 6], pn[i * 6 + 1], pn[i * 6 + 2]), Vec3(pn[i * 6 + 3], pn[i * 6 + 4], pn[i * 6 + 5]));
        out[i * 3] = c.x; out[i * 3 + 1] = c.y; out[i * 3 + 2] = c.z;
    });
    return 0;
}

// Flattened world QBVH for layout checks: nodes x (48 floats + 8 ints).
int rtc_world_nodes(void* s, float* bounds, int32_t* children, int32_t* leaf_tags, uint32_t cap) {
    OScene& S = *(OScene*)s;
    uint32_t n = (uint32_t)S.world.nodes.size();
    if (bounds && children) {
        for (uint32_t i = 0; i < n && i < cap; i++) {
            const QbvhNode& q = S.world.nodes[i];
            for (int a = 0; a < 3; a++) for (int c = 0; c < QW; c++) {
                bounds[i * 48 + a * 8 + c] = q.mn[a][c];
                bounds[i * 48 + 24 + a * 8 + c] = q.mx[a][c];
            }
            for (int c = 0; c < QW; c++) children[i * 8 + c] = q.children[c];
        }
    }
    if (leaf_tags) for (size_t i = 0; i < S.world.objects.size(); i++) leaf_tags[i] = S.world.objects[i]->tag;
    return (int)n;
}

// ---- unit-level entry points for known-answer tests --------------------------------
void rtc_rng_u64(uint64_t seed, uint32_t n, uint64_t* out) {
    SmallRng r = SmallRng::seed_from_u64(seed);
    for (uint32_t i = 0; i < n; i++) out[i] = r.next_u64();
}
// xoshiro256++ from a raw 4-word state (no SplitMix64 seeding).
void rtc_xoshiro_raw(const uint64_t* state, uint32_t n, uint64_t* out) {
    SmallRng r; for (int i = 0; i < 4; i++) r.s[i] = state[i];
    for (uint32_t i = 0; i < n; i++) out[i] = r.next_u64();
}
void rtc_rng_f32(uint64_t seed, uint32_t n, float* out) {
    SmallRng r = SmallRng::seed_from_u64(seed);
    for (uint32_t i = 0; i < n; i++) out[i] = r.gen();
}
void rtc_rng_range(uint64_t seed, float lo, float hi, uint32_t n, float* out) {
    SmallRng r = SmallRng::seed_from_u64(seed);
    for (uint32_t i = 0; i < n; i++) out[i] = r.gen_range(lo, hi);
}

// Texture::value of a LAMBERTIAN / DIFFUSE_LIGHT entry (images and Perlin
// tables from `d`) at n points (u, v, x, y, z): n x 3 colours.
int rtc_texture_eval(const rt_material* m, const rt_scene_desc* d, const float* uvp, uint32_t n, float* out) {
    Texture t;
    std::string err;
    if (!make_texture(*m, d, t, err)) return -1;
    for (uint32_t i = 0; i < n; i++) {
        const float* q = uvp + (size_t)i * 5;
        Color c = t.value(q[0], q[1], Point3(q[2], q[3], q[4]));
        out[3 * i] = c.x; out[3 * i + 1] = c.y; out[3 * i + 2] = c.z;
    }
    return 0;
}

// PbrMaterial::params for n hit records (u, v, normal xyz, tangent xyz):
// n x 8 = shading normal, albedo (maps and ambient occlusion), roughness, metallic.
int rtc_pbr_params(const rt_material* m, const rt_scene_desc* d, const float* recs, uint32_t n, float* out) {
    std::string err;
    std::unique_ptr<Material> mat(make_material(*m, d, err));
    auto* p = dynamic_cast<PbrMaterial*>(mat.get());
    if (!p) return -1;
    for (uint32_t i = 0; i < n; i++) {
        const float* q = recs + (size_t)i * 8;
        HitRecord h;
        h.u = q[0]; h.v = q[1];
        h.normal = Vec3(q[2], q[3], q[4]); h.tangent = Vec3(q[5], q[6], q[7]);
        PbrParams pp = p->params(h);
        float* o = out + (size_t)i * 8;
        o[0] = pp.normal.x; o[1] = pp.normal.y; o[2] = pp.normal.z;
        o[3] = pp.albedo.x; o[4] = pp.albedo.y; o[5] = pp.albedo.z; o[6] = pp.roughness; o[7] = pp.metallic;
    }
    return 0;
}

// Perlin::noise and Perlin::turb(depth) over 1536-float tables at n points.
int rtc_perlin(const float* tables, const float* pts, uint32_t n, int depth, float* noise, float* turb) {
    if (!perlin_perms_ok(tables)) return -1;
    Perlin P;
    P.load(tables);
    for (uint32_t i = 0; i < n; i++) {
        Point3 p(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
        noise[i] = P.noise(p);
        turb[i] = P.turb(p, depth);
    }
    return 0;
}

// Scalar functions by name: returns 0 on success.
int rtc_eval(const char* fn, const float* a, float* out) {
    std::string f(fn);
    if (f == "ggx_ndf") { out[0] = ggx_ndf(a[0], a[1]); return 0; }
    if (f == "nacre_color") { Color c = nacre_color(a[0], a[1], a[2]); out[0] = c.x; out[1] = c.y; out[2] = c.z; return 0; }
    if (f == "oil_noise") { out[0] = oil_noise(Point3(a[0], a[1], a[2])); return 0; }
    if (f == "srgb_decode") { int c = (int)a[0]; if (c < 0 || c > 255) return -1; out[0] = srgb_decode_lut()[c]; return 0; }
    if (f == "smith_g1") { out[0] = smith_g1(a[0], a[1]); return 0; }
    if (f == "smith_lambda") { out[0] = smith_lambda(a[0], a[1]); return 0; }
    if (f == "smith_g2") { out[0] = smith_g2_hc(a[0], a[1]); return 0; }
    if (f == "e_ss") { out[0] = e_ss_lut(a[0], a[1]); return 0; }
    if (f == "energy_comp") { out[0] = specular_energy_comp(a[0], a[1]); return 0; }
    if (f == "fresnel_dielectric") { out[0] = fresnel_dielectric(a[0], a[1]); return 0; }
    if (f == "schlick") { out[0] = schlick(a[0], a[1]); return 0; }
    if (f == "fresnel_conductor") { out[0] = fresnel_conductor(a[0], a[1], a[2]); return 0; }
    if (f == "cauchy") { out[0] = cauchy_ior(a[0], a[1], a[2]); return 0; }
    if (f == "sellmeier") { out[0] = sellmeier_ior(a[0], a + 1, a + 4); return 0; }
    if (f == "metal_ior") { metal_ior((int)a[0], a[1], out[0], out[1]); return 0; }
    if (f == "spectral_to_rgb") { Color c = spectral_to_rgb(a[0]); out[0] = c.x; out[1] = c.y; out[2] = c.z; return 0; }
    if (f == "spectral_fire_rgb") { Color c = spectral_fire_rgb(a[0]); out[0] = c.x; out[1] = c.y; out[2] = c.z; return 0; }
    if (f == "spectral_to_rgb_wide") { Color c = spectral_to_rgb(a[0], true); out[0] = c.x; out[1] = c.y; out[2] = c.z; return 0; }
    if (f == "spectral_fire_rgb_wide") { Color c = spectral_fire_rgb(a[0], true); out[0] = c.x; out[1] = c.y; out[2] = c.z; return 0; }
    if (f == "hero_lambda") { out[0] = hero_lambda(a[0], (int)a[1]); return 0; }
    if (f == "blackbody_wide") {
        BlackbodyLight b; b.temp_k = a[0]; b.intensity = a[1]; b.init();
        out[0] = b.avg_color_w.x; out[1] = b.avg_color_w.y; out[2] = b.avg_color_w.z; out[3] = b.norm_w; return 0;
    }
    if (f == "planck_raw") { out[0] = planck_raw(a[0], a[1]); return 0; }
    if (f == "blackbody") {
        BlackbodyLight b; b.temp_k = a[0]; b.intensity = a[1]; b.init();
        out[0] = b.avg_color.x; out[1] = b.avg_color.y; out[2] = b.avg_color.z; out[3] = b.norm; return 0;
    }
    if (f == "vndf") {
        Vec3 m = vndf_sample_aniso(Vec3(a[0], a[1], a[2]), a[3], a[4], a[5], a[6]);
        out[0] = m.x; out[1] = m.y; out[2] = m.z; return 0;
    }
    if (f == "preetham") {
        Color c = preetham_sky(Vec3(a[0], a[1], a[2]), Vec3(a[3], a[4], a[5]), a[6]);
        out[0] = c.x; out[1] = c.y; out[2] = c.z; return 0;
    }
    if (f == "hg_sample_seeded") {
        SmallRng r = SmallRng::seed_from_u64((uint64_t)a[4]);
        Vec3 d = hg_sample(Vec3(a[0], a[1], a[2]), a[3], r);
        out[0] = d.x; out[1] = d.y; out[2] = d.z; return 0;
    }
    return -1;
}

void rtc_e_lut(float* out) {
    const ELut& E = e_lut();
    for (int i = 0; i < E_N; i++) for (int j = 0; j < E_N; j++) out[i * E_N + j] = E.lut[i][j];
}

// Repeated Material::scatter on one hand-built HitRecord with one stream.
// ray: origin[3], dir[3], wavelength; rec: p[3], normal[3], t.
// out: n x 8 = attenuation rgb, direction xyz, skip_pdf, ok.
int rtc_scatter(const rt_material* m, const float* ray, int spectral_weighted, const float* rec, int front_face,
                uint64_t seed, uint32_t n, float* out) {
    std::string err;
    std::unique_ptr<Material> mat(make_material(*m, nullptr, err));
    if (!mat) return -1;
    Ray r = Ray::make(Point3(ray[0], ray[1], ray[2]), Vec3(ray[3], ray[4], ray[5]));
    r.wavelength = ray[6]; r.spectral_weighted = spectral_weighted != 0;
    HitRecord h; h.p = Point3(rec[0], rec[1], rec[2]); h.normal = Vec3(rec[3], rec[4], rec[5]); h.t = rec[6];
    h.front_face = front_face != 0; h.mat = mat.get();
    SmallRng rng = SmallRng::seed_from_u64(seed);
    for (uint32_t i = 0; i < n; i++) {
        ScatterRecord sr;
        bool ok = mat->scatter(r, h, rng, sr);
        float* o = out + (size_t)i * 8;
        if (ok) {
            o[0] = sr.attenuation.x; o[1] = sr.attenuation.y; o[2] = sr.attenuation.z;
            o[3] = sr.ray.direction.x; o[4] = sr.ray.direction.y; o[5] = sr.ray.direction.z;
            o[6] = sr.skip_pdf ? 1.0f : 0.0f; o[7] = 1.0f;
        } else { for (int k = 0; k < 8; k++) o[k] = 0.0f; }
    }
    return 0;
}

// Env-map CDF sampling against a given image: n samples -> dir xyz + pdf.
int rtc_env_sample(const float* rgb, uint32_t w, uint32_t h, float rot, uint64_t seed, uint32_t n, float* out) {
    EnvMap e; e.init(rgb, w, h, rot);
    SmallRng rng = SmallRng::seed_from_u64(seed);
    for (uint32_t i = 0; i < n; i++) {
        float pdf; Vec3 d = e.sample_dir(rng, pdf);
        out[i * 4] = d.x; out[i * 4 + 1] = d.y; out[i * 4 + 2] = d.z; out[i * 4 + 3] = pdf;
    }
    return 0;
}
int rtc_env_eval(const float* rgb, uint32_t w, uint32_t h, float rot, const float* dirs, uint32_t n, float* out) {
    EnvMap e; e.init(rgb, w, h, rot);
    for (uint32_t i = 0; i < n; i++) {
        Vec3 d(dirs[i * 3], dirs[i * 3 + 1], dirs[i * 3 + 2]);
        Color c = e.sample(d);
        out[i * 4] = c.x; out[i * 4 + 1] = c.y; out[i * 4 + 2] = c.z; out[i * 4 + 3] = e.dir_pdf(d);
    }
    return 0;
}

// ---- adjoint checker (extension; see rc_adjoint.h) ------------------------------------------
// Writes theta into the descriptor table and the live materials in place
// (the object graph keeps Material pointers).
int rtc_set_params(void* s, const rt_param* prm, uint32_t n, const float* vals) {
    OScene& S = *(OScene*)s;
    ParamMap pm;
    if (!pm.build(S.desc_mats, prm, n, &S.media)) return -1;
    for (uint32_t i = 0; i < n; i++) {
        if (prm[i].material & RT_PARAM_MEDIUM) {
            for (const ConstantMedium* cm : S.media[prm[i].material & ~RT_PARAM_MEDIUM].objs) {
                auto* w = const_cast<ConstantMedium*>(cm);
                w->density = vals[i]; w->neg_inv_density = -1.0f / vals[i];
            }
            continue;
        }
        for (uint32_t f = prm[i].field; f < prm[i].field + prm[i].count; f++) S.desc_mats[prm[i].material].p[f] = vals[i];
    }
    for (size_t i = 0; i < S.mats.size(); i++) {
        const rt_material& m = S.desc_mats[i];
        if (auto* l = dynamic_cast<Lambertian*>(S.mats[i].get())) l->tex.even = Color(m.p[0], m.p[1], m.p[2]);
        if (auto* r = dynamic_cast<RoughSpectralDielectric*>(S.mats[i].get())) {
            r->nd = m.p[0];
            for (int k = 0; k < 3; k++) { r->sB[k] = m.p[1 + k]; r->sC[k] = m.p[4 + k]; }
            for (int k = 0; k < 17; k++) r->alpha[k] = m.p[7 + k];
        }
    }
    return 0;
}

static uint64_t row_or_pixel_seed(int mode, uint32_t row, uint32_t i, const SampleParams& P) {
    return mode == 1 ? pixel_seed(i, P.sample_idx, P.seed)
                     : (((uint64_t)row * 0x9E3779B97F4A7C15ULL) ^ ((uint64_t)P.sample_idx * 6364136223846793005ULL));
}

// Detached estimator (decisions at the scene's parameters, BSDF values at
// eval_vals), summed over spp passes into accum.
int rtc_render_detached(void* s, const rt_camera* cam, const rt_sample_params* p, uint32_t spp, const rt_param* prm,
                        uint32_t n, const float* eval_vals, float* accum, int mode, int threads) {
    OScene& S = *(OScene*)s;
    ParamMap pm;
    if (!pm.build(S.desc_mats, prm, n, &S.media)) return -1;
    std::vector<rt_material> me = S.desc_mats;
    for (uint32_t i = 0; i < n; i++) {
        if (prm[i].material & RT_PARAM_MEDIUM) continue;
        for (uint32_t f = prm[i].field; f < prm[i].field + prm[i].count; f++) me[prm[i].material].p[f] = eval_vals[i];
    }
    const std::vector<MedParam> mp = med_params(pm, S.media, eval_vals);
    SceneView V = view(S, p->use_photons != 0);
    Camera C = to_cam(*cam);
    SampleParams base = to_params(*p);
    size_t npix = (size_t)p->width * p->height;
    for (size_t i = 0; i < npix * 3; i++) accum[i] = 0.0f;
    for (uint32_t k = 0; k < spp; k++) {
        SampleParams P = base; P.sample_idx = base.sample_idx + k;
        parallel_for(P.height, threads, [&](size_t row) {
            SmallRng rng = SmallRng::seed_from_u64(row_or_pixel_seed(0, (uint32_t)row, 0, P));
            for (uint32_t col = 0; col < P.width; col++) {
                size_t i = row * P.width + col;
                if (mode == 1) rng = SmallRng::seed_from_u64(row_or_pixel_seed(1, (uint32_t)row, (uint32_t)i, P));
                RatioHook hk; hk.ms = &S.desc_mats; hk.me = &me; hk.med = &mp;
                Color c = render_pixel_h(V, C, P, (uint32_t)row, col, rng, hk);
                accum[3 * i] += c.x; accum[3 * i + 1] += c.y; accum[3 * i + 2] += c.z;
            }
        });
    }
    return 0;
}

// grad[j] = sum over pixels and samples of dloss_dimage / spp . dC/dtheta_j,
// by forward score accumulation (double accumulators).
int rtc_backward(void* s, const rt_camera* cam, const rt_sample_params* p, uint32_t spp, const rt_param* prm, uint32_t n,
                 const float* dimg, double* grad, int mode, int threads) {
    OScene& S = *(OScene*)s;
    ParamMap pm;
    if (!pm.build(S.desc_mats, prm, n, &S.media)) return -1;
    const std::vector<MedParam> mp = med_params(pm, S.media, nullptr);
    SceneView V = view(S, p->use_photons != 0);
    Camera C = to_cam(*cam);
    SampleParams base = to_params(*p);
    std::vector<std::vector<double>> rows(p->height, std::vector<double>(n, 0.0));
    for (uint32_t k = 0; k < spp; k++) {
        SampleParams P = base; P.sample_idx = base.sample_idx + k;
        parallel_for(P.height, threads, [&](size_t row) {
            SmallRng rng = SmallRng::seed_from_u64(row_or_pixel_seed(0, (uint32_t)row, 0, P));
            ScoreHook hk;
            for (uint32_t col = 0; col < P.width; col++) {
                size_t i = row * P.width + col;
                if (mode == 1) rng = SmallRng::seed_from_u64(row_or_pixel_seed(1, (uint32_t)row, (uint32_t)i, P));
                Color d(dimg[3 * i] / (float)spp, dimg[3 * i + 1] / (float)spp, dimg[3 * i + 2] / (float)spp);
                hk.init(&S.desc_mats, &pm, d, rows[row].data(), &mp);
                render_pixel_h(V, C, P, (uint32_t)row, col, rng, hk);
            }
        });
    }
    for (uint32_t j = 0; j < n; j++) {
        double t = 0.0;
        for (auto& r : rows) t += r[j];
        grad[j] = t;
    }
    return 0;
}

// log f of the rough spectral dielectric for explicit directions (tests).
int rtc_rsd_logf(const float* i, const float* o, const float* nrm, double k, double alpha, int refl, double* lf) {
    return logf_rsd(D3{i[0], i[1], i[2]}, D3{o[0], o[1], o[2]}, D3{nrm[0], nrm[1], nrm[2]}, k, alpha, refl != 0, *lf) ? 1 : 0;
}

// Alias table of an env image (tests: implied probabilities).
int rtc_env_alias(const float* rgb, uint32_t w, uint32_t h, float* prob, uint32_t* alias) {
    EnvMap e; e.init(rgb, w, h, 0.0f);
    if (!e.init_alias()) return -1;
    for (size_t t = 0; t < (size_t)w * h; t++) { prob[t] = e.aprob[t]; alias[t] = e.alias[t]; }
    return 0;
}
// Env sampling through a given sampler (0 CDF, 1 alias): n draws of dir xyz + pdf.
int rtc_env_sample_mode(const float* rgb, uint32_t w, uint32_t h, float rot, int mode, uint64_t seed, uint32_t n, float* out) {
    EnvMap e; e.init(rgb, w, h, rot);
    if (mode == 1 && !e.init_alias()) return -1;
    SmallRng rng = SmallRng::seed_from_u64(seed);
    for (uint32_t i = 0; i < n; i++) {
        float pdf; Vec3 d = e.sample_dir(rng, pdf);
        out[i * 4] = d.x; out[i * 4 + 1] = d.y; out[i * 4 + 2] = d.z; out[i * 4 + 3] = pdf;
    }
    return 0;
}

// ---- per-pixel accumulation (src/main.rs:597-677) -------------------------------
// accumulate_sample: per pixel, skip converged pixels and non-finite samples;
// firefly clamp on luminance once old_n >= min_samples; sum, count and
// per-channel Welford M2.  Color / f32 divides each component.
static inline float acc_lum(float x, float y, float z) { return 0.2126f * x + 0.7152f * y + 0.0722f * z; }

int rtc_accumulate_sample(const float* sc, uint32_t npix, float ff_clamp, uint32_t ff_min, float ff_floor,
                          float* acc, uint32_t* ns, float* m2, const uint8_t* conv) {
    for (uint32_t i = 0; i < npix; i++) {
        if (conv[i]) continue;
        float s[3] = {sc[3 * i], sc[3 * i + 1], sc[3 * i + 2]};
        if (!std::isfinite(s[0]) || !std::isfinite(s[1]) || !std::isfinite(s[2])) continue;
        float* a = acc + 3 * (size_t)i;
        uint32_t old_n = ns[i];
        float om[3] = {0.0f, 0.0f, 0.0f};
        if (old_n > 0) for (int c = 0; c < 3; c++) om[c] = a[c] / (float)old_n;
        float old_lum = acc_lum(om[0], om[1], om[2]);
        if (ff_clamp > 0.0f && old_n >= ff_min) {
            float l = acc_lum(s[0], s[1], s[2]);
            float th = std::fmax(ff_clamp * old_lum, ff_floor);
            if (l > th) { float f = th / l; for (int c = 0; c < 3; c++) s[c] = s[c] * f; }
        }
        for (int c = 0; c < 3; c++) a[c] = a[c] + s[c];
        ns[i] = old_n + 1;
        for (int c = 0; c < 3; c++) {
            float nm = a[c] / (float)ns[i];
            m2[3 * (size_t)i + c] = m2[3 * (size_t)i + c] + (s[c] - om[c]) * (s[c] - nm);
        }
    }
    return 0;
}

// mark_converged: returns the number of converged pixels after the update.
int rtc_mark_converged(uint32_t npix, float threshold, const float* acc, const uint32_t* ns, const float* m2, uint8_t* conv) {
    int total = 0;
    for (uint32_t i = 0; i < npix; i++) {
        if (!conv[i] && ns[i] > 0) {
            uint32_t n = ns[i];
            bool ok = true;
            for (int c = 0; c < 3 && ok; c++) {
                float m = acc[3 * (size_t)i + c] / (float)(n > 1 ? n : 1);
                if (m < 1e-4f) continue;
                float var = m2[3 * (size_t)i + c] / (float)(n - 1 > 1 ? n - 1 : 1);
                float se = std::sqrt(var / (float)n);
                if (se / m >= threshold) ok = false;
            }
            if (ok) conv[i] = 1;
        }
        total += conv[i] ? 1 : 0;
    }
    return total;
}

// ---- the saved image (src/main.rs:571-582, src/output.rs:11-186) -----------------
static inline float o_aces_curve(float x) { x = std::fmax(x, 0.0f); return (x * (2.51f * x + 0.03f)) / (x * (2.43f * x + 0.59f) + 0.14f); }
static inline void o_aces(const float c[3], float out[3]) {
    float r = c[0] * 0.59719f + c[1] * 0.35458f + c[2] * 0.04823f;
    float g = c[0] * 0.07600f + c[1] * 0.90834f + c[2] * 0.01566f;
    float b = c[0] * 0.02840f + c[1] * 0.13383f + c[2] * 0.83777f;
    r = o_aces_curve(r); g = o_aces_curve(g); b = o_aces_curve(b);
    out[0] = r * 1.60475f + g * -0.53108f + b * -0.07367f;
    out[1] = r * -0.10208f + g * 1.10813f + b * -0.00605f;
    out[2] = r * -0.00327f + g * -0.07276f + b * 1.07602f;
}
static inline float o_clamp(float x, float lo, float hi) { if (x < lo) x = lo; if (x > hi) x = hi; return x; }
static inline float o_agx_curve(float x) {
    float x2 = x * x, x4 = x2 * x2;
    return 15.5f * x4 * x2 - 40.14f * x4 * x + 31.96f * x4 - 6.868f * x2 * x + 0.4298f * x2 + 0.1191f * x - 0.00232f;
}
static inline void o_agx(const float c[3], float out[3]) {
    float v[3] = {c[0] * 0.842479062253094f + c[1] * 0.0784335999999992f + c[2] * 0.0792237451477643f,
                  c[0] * 0.0423282422610123f + c[1] * 0.878468636469772f + c[2] * 0.0791661274605434f,
                  c[0] * 0.0423756549057051f + c[1] * 0.0784336f + c[2] * 0.879142973793104f};
    for (int k = 0; k < 3; k++) v[k] = o_agx_curve(o_clamp((std::log2(std::fmax(v[k], 1e-10f)) - -12.47393f) / 16.5f, 0.0f, 1.0f));
    float luma = acc_lum(v[0], v[1], v[2]);
    for (int k = 0; k < 3; k++) v[k] = luma + 1.4f * (std::pow(v[k], 1.35f) - luma);
    out[0] = v[0] * 1.19687900512017f + v[1] * -0.0980208811401368f + v[2] * -0.0990297440797205f;
    out[1] = v[0] * -0.0528968517574562f + v[1] * 1.15190312990417f + v[2] * -0.0989611768448433f;
    out[2] = v[0] * -0.0529716355144234f + v[1] * -0.0980434501171241f + v[2] * 1.15107567775604f;
}
static inline uint8_t o_byte(float x) {
    x = std::fmax(x, 0.0f);
    float e = x <= 0.0031308f ? 12.92f * x : 1.055f * std::pow(x, 1.0f / 2.4f) - 0.055f;
    float v = o_clamp(e, 0.0f, 0.999f) * 256.0f;
    if (!(v > 0.0f)) return 0;
    return v >= 255.0f ? 255 : (uint8_t)v;
}
// apply_optics (src/output.rs:150-186) on a linear image of w*h RGB floats.
int rtc_apply_optics(const float* raw, uint32_t w, uint32_t h, float vignette, float distortion, float* img) {
    size_t npix = (size_t)w * h;
    if (vignette == 0.0f && distortion == 0.0f) { std::copy(raw, raw + 3 * npix, img); return 0; }
    float cx = ((float)w - 1.0f) * 0.5f, cy = ((float)h - 1.0f) * 0.5f;
    float rc = std::fmax(std::sqrt(cx * cx + cy * cy), 1e-6f);
    for (uint32_t y = 0; y < h; y++)
        for (uint32_t x = 0; x < w; x++) {
            float dx = ((float)x - cx) / rc, dy = ((float)y - cy) / rc;
            float r2 = dx * dx + dy * dy;
            float s = 1.0f - distortion * r2;
            float sx = o_clamp(cx + dx * rc * s, 0.0f, (float)w - 1.0f);
            float sy = o_clamp(cy + dy * rc * s, 0.0f, (float)h - 1.0f);
            size_t x0 = (size_t)sx, y0 = (size_t)sy;
            size_t x1 = std::min<size_t>(x0 + 1, w - 1), y1 = std::min<size_t>(y0 + 1, h - 1);
            float fx = sx - (float)x0, fy = sy - (float)y0;
            float vig = std::fmax(1.0f - vignette * r2, 0.0f);
            for (int k = 0; k < 3; k++) {
                float c00 = raw[3 * (y0 * w + x0) + k], c10 = raw[3 * (y0 * w + x1) + k];
                float c01 = raw[3 * (y1 * w + x0) + k], c11 = raw[3 * (y1 * w + x1) + k];
                float c = (c00 * (1.0f - fx) + c10 * fx) * (1.0f - fy) + (c01 * (1.0f - fx) + c11 * fx) * fy;
                img[3 * ((size_t)y * w + x) + k] = c * vig;
            }
        }
    return 0;
}

// Tone-mapping pieces for tests: which 0 aces_tonemap, 1 agx_tonemap (3 -> 3
// floats per item), 2 the 8-bit sRGB store of one channel, 3 aces_curve,
// 4 agx_curve (1 -> 1).
int rtc_tone_eval(int which, const float* in, uint32_t n, float* out) {
    for (uint32_t i = 0; i < n; i++) {
        if (which == 0) o_aces(in + 3 * i, out + 3 * i);
        else if (which == 1) o_agx(in + 3 * i, out + 3 * i);
        else if (which == 2) out[i] = (float)o_byte(in[i]);
        else if (which == 3) out[i] = o_aces_curve(in[i]);
        else if (which == 4) out[i] = o_agx_curve(in[i]);
        else return -1;
    }
    return 0;
}

int rtc_resolve_rgb8(const float* acc, const uint32_t* ns, uint32_t spp, uint32_t w, uint32_t h, float exposure, int agx,
                     float vignette, float distortion, uint8_t* out) {
    size_t npix = (size_t)w * h;
    std::vector<float> raw(3 * npix), img(3 * npix);
    for (size_t i = 0; i < npix; i++) {
        uint32_t n0 = ns ? ns[i] : spp;
        float n = (float)(n0 > 1 ? n0 : 1);
        for (int k = 0; k < 3; k++) raw[3 * i + k] = acc[3 * i + k] / n;
    }
    rtc_apply_optics(raw.data(), w, h, vignette, distortion, img.data());
    for (size_t i = 0; i < npix; i++) {
        float c[3] = {img[3 * i] * exposure, img[3 * i + 1] * exposure, img[3 * i + 2] * exposure};
        float m[3];
        if (agx) o_agx(c, m); else o_aces(c, m);
        for (int k = 0; k < 3; k++) out[3 * i + k] = o_byte(m[k]);
    }
    return 0;
}

// render_aux_pass (src/renderer.rs:597-645).
int rtc_render_aux(void* s, const rt_camera* cam, uint32_t w, uint32_t h, float* alb, float* nrm) {
    OScene& S = *(OScene*)s;
    Camera c = to_cam(*cam);
    float wd = (float)std::max<uint32_t>(w - 1, 1), hd = (float)std::max<uint32_t>(h - 1, 1);
    for (uint32_t i = 0; i < w * h; i++) {
        uint32_t row = i / w, col = i % w;
        SmallRng rng = SmallRng::seed_from_u64((uint64_t)i ^ 0x9E3779B97F4A7C15ull);
        float u = ((float)col + 0.5f) / wd;
        float v = ((float)(h - 1 - row) + 0.5f) / hd;
        Ray ray = c.get_ray(u, v, rng);
        HitRecord rec;
        float a[3], n[3];
        if (!S.world.hit(ray, 0.001f, INF, rec)) {
            Color bg = S.bg.eval(ray.direction);
            a[0] = bg.x; a[1] = bg.y; a[2] = bg.z;
            n[0] = n[1] = n[2] = 0.0f;
        } else {
            Color al = rec.mat->albedo_hint(rec.u, rec.v, rec.p);
            float cl[3] = {al.x, al.y, al.z};
            for (int k = 0; k < 3; k++) { float x = cl[k]; if (x < 0.0f) x = 0.0f; if (x > 1.0f) x = 1.0f; a[k] = x; }
            n[0] = rec.normal.x; n[1] = rec.normal.y; n[2] = rec.normal.z;
        }
        for (int k = 0; k < 3; k++) { alb[3 * (size_t)i + k] = a[k]; nrm[3 * (size_t)i + k] = n[k]; }
    }
    return 0;
}

}  // extern "C"
