// ORACLE — test infrastructure only (see rc_math.h).
//
// Integrator layer restated from rustracer:
//   ConstantMedium / NoiseMedium (grid density variant)   src/volume.rs:66-263
//   EnvMapData (2D CDF, bilinear, pdf), Background         src/renderer.rs:23-223
//   preetham_sky, sun helpers                              src/renderer.rs:231-366
//   ray_color                                              src/renderer.rs:371-585 (max depth is a parameter)
//   render_tiles                                           src/renderer.rs:655-730
//   Camera::get_ray                                        src/camera.rs:47-59
//   PhotonMap build / build_from_quad / kd / gather        src/photon.rs:41-277
#pragma once
#include "rc_material.h"
#include <thread>
#include <atomic>

namespace rc {

// ---- media -----------------------------------------------------------------
static inline uint32_t fbits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline SmallRng medium_rng(const Ray& r) {
    return SmallRng::seed_from_u64(((uint64_t)fbits(r.origin.x) * 6364136223846793005ULL)
                                   ^ ((uint64_t)fbits(r.origin.y) * 0x9E3779B97F4A7C15ULL)
                                   ^ ((uint64_t)fbits(r.origin.z) * 0xBF58476D1CE4E5B9ULL)
                                   ^ ((uint64_t)fbits(r.direction.x) * 0x94D049BB133111EBULL)
                                   ^ (uint64_t)fbits(r.wavelength));
}
static inline bool inside_length(const Hittable& boundary, const Ray& r, float t_min, float t_max, float& len) {
    HitRecord r1, r2;
    if (!boundary.hit(r, -INF, INF, r1)) return false;
    if (!boundary.hit(r, r1.t + 0.0001f, INF, r2)) return false;
    float t1 = rmax(rmax(r1.t, 0.0f), t_min);
    float t2 = rmin(r2.t, t_max);
    if (t1 >= t2) return false;
    len = (t2 - t1) * r.direction.length();
    return true;
}

struct ConstantMedium : Hittable {
    HPtr boundary; float density, neg_inv_density; HenyeyGreenstein phase;
    bool hit(const Ray& r, float t_min, float t_max, HitRecord& rec) const override {
        HitRecord r1, r2;
        if (!boundary->hit(r, -INF, INF, r1)) return false;
        if (!boundary->hit(r, r1.t + 0.0001f, INF, r2)) return false;
        if (r1.t < t_min) r1.t = t_min;
        if (r2.t > t_max) r2.t = t_max;
        if (r1.t >= r2.t) return false;
        if (r1.t < 0.0f) r1.t = 0.0f;
        float ray_length = r.direction.length();
        float dist_inside = (r2.t - r1.t) * ray_length;
        SmallRng rng = medium_rng(r);
        float random = rmax(rng.gen(), FLT_EPSILON);
        float hit_dist = neg_inv_density * std::log(random);
        if (hit_dist > dist_inside) return false;
        float t = r1.t + hit_dist / ray_length;
        rec = HitRecord(); rec.p = r.at(t); rec.normal = Vec3(1, 0, 0); rec.mat = &phase; rec.t = t; rec.front_face = true;
        return true;
    }
    bool any_hit(const Ray&, float, float) const override { return false; }
    float transmittance(const Ray& r, float t_min, float t_max) const override {
        float l;
        if (inside_length(*boundary, r, t_min, t_max, l)) return std::exp(-density * l);
        return 1.0f;
    }
    bool bounding_box(Aabb& out) const override { return boundary->bounding_box(out); }
    bool photon_bounds(Aabb& out) const override { return boundary->photon_bounds(out); }   // src/volume.rs:153-155, 260-262
};

// NoiseMedium with the per-lookup Perlin turbulence replaced by a trilinear
// lookup into a dense grid over the boundary's bounding box (extension).
struct GridMedium : Hittable {
    HPtr boundary; float density_scale; HenyeyGreenstein phase;
    std::vector<float> grid; uint32_t nx, ny, nz; Aabb box;
    virtual ~GridMedium() {}
    virtual float density_at(const Point3& p) const {
        float fx = (p.x - box.min.x) / (box.max.x - box.min.x) * (float)nx - 0.5f;
        float fy = (p.y - box.min.y) / (box.max.y - box.min.y) * (float)ny - 0.5f;
        float fz = (p.z - box.min.z) / (box.max.z - box.min.z) * (float)nz - 0.5f;
        fx = rclamp(fx, 0.0f, (float)(nx - 1)); fy = rclamp(fy, 0.0f, (float)(ny - 1)); fz = rclamp(fz, 0.0f, (float)(nz - 1));
        uint32_t x0 = (uint32_t)f2usize(fx), y0 = (uint32_t)f2usize(fy), z0 = (uint32_t)f2usize(fz);
        uint32_t x1 = std::min(x0 + 1, nx - 1), y1 = std::min(y0 + 1, ny - 1), z1 = std::min(z0 + 1, nz - 1);
        float tx = fx - (float)x0, ty = fy - (float)y0, tz = fz - (float)z0;
        auto g = [&](uint32_t x, uint32_t y, uint32_t z) { return grid[((size_t)z * ny + y) * nx + x]; };
        float c00 = g(x0, y0, z0) * (1.0f - tx) + g(x1, y0, z0) * tx;
        float c10 = g(x0, y1, z0) * (1.0f - tx) + g(x1, y1, z0) * tx;
        float c01 = g(x0, y0, z1) * (1.0f - tx) + g(x1, y0, z1) * tx;
        float c11 = g(x0, y1, z1) * (1.0f - tx) + g(x1, y1, z1) * tx;
        float c0 = c00 * (1.0f - ty) + c10 * ty;
        float c1 = c01 * (1.0f - ty) + c11 * ty;
        return (c0 * (1.0f - tz) + c1 * tz) * density_scale;
    }
    bool hit(const Ray& r, float t_min, float t_max, HitRecord& rec) const override {
        HitRecord r1, r2;
        if (!boundary->hit(r, -INF, INF, r1)) return false;
        if (!boundary->hit(r, r1.t + 0.0001f, INF, r2)) return false;
        if (r1.t < t_min) r1.t = t_min;
        if (r2.t > t_max) r2.t = t_max;
        if (r1.t >= r2.t) return false;
        if (r1.t < 0.0f) r1.t = 0.0f;
        float ray_length = r.direction.length();
        float inv_maj = 1.0f / (density_scale * ray_length);
        SmallRng rng = medium_rng(r);
        float t = r1.t;
        for (;;) {
            t -= inv_maj * std::log(rmax(rng.gen(), FLT_EPSILON));
            if (t >= r2.t) return false;
            float ld = density_at(r.at(t));
            if (rng.gen() < ld / density_scale) {
                rec = HitRecord(); rec.p = r.at(t); rec.normal = Vec3(1, 0, 0); rec.mat = &phase; rec.t = t; rec.front_face = true;
                return true;
            }
        }
    }
    bool any_hit(const Ray&, float, float) const override { return false; }
    float transmittance(const Ray& r, float t_min, float t_max) const override {
        float l;
        if (!inside_length(*boundary, r, t_min, t_max, l)) return 1.0f;
        HitRecord r1;
        boundary->hit(r, -INF, INF, r1);
        float t1 = rmax(rmax(r1.t, 0.0f), t_min);
        const int K = 16;
        float seg_t = l / r.direction.length();
        float sum = 0.0f;
        for (int i = 0; i < K; i++) {
            float t = t1 + seg_t * ((float)i + 0.5f) / (float)K;
            sum += density_at(r.at(t));
        }
        return std::exp(-sum / (float)K * l);
    }
    bool bounding_box(Aabb& out) const override { return boundary->bounding_box(out); }
    bool photon_bounds(Aabb& out) const override { return boundary->photon_bounds(out); }   // src/volume.rs:153-155, 260-262
};

// NoiseMedium (src/volume.rs:160-263): the same tracking and quadrature as
// GridMedium with density max(turb(p * noise_scale, 7) / 2 - threshold, 0) *
// density_scale, Perlin::noise / turb as src/perlin.rs:35-72 over given tables.
struct NoiseMedium : GridMedium {
    Perlin perlin;
    float noise_scale = 1.0f, threshold = 0.0f;
    float density_at(const Point3& p) const override {
        float n = perlin.turb(p * noise_scale, 7) * 0.5f;
        return rmax(n - threshold, 0.0f) * density_scale;
    }
};

// ---- background ----------------------------------------------------------------
static const float COS_SUN_MAX = 0.9999892f;

struct EnvMap {
    uint32_t w = 0, h = 0;
    std::vector<float> img;   // w*h*3, row-major
    std::vector<float> marginal, cond;
    float inv_total = 0.0f, rot = 0.0f;
    // alias-table sampling (extension): acceptance probability and alias per
    // texel, built by Vose's method in double over the same weights
    bool use_alias = false;
    std::vector<float> aprob;
    std::vector<uint32_t> alias;
    bool init_alias() {
        size_t n = (size_t)w * h;
        std::vector<double> q(n);
        double total = 0.0;
        for (uint32_t j = 0; j < h; j++) {
            float sin_j = std::sin(PI * ((float)j + 0.5f) / (float)h);
            for (uint32_t i = 0; i < w; i++) {
                const float* px = &img[((size_t)j * w + i) * 3];
                float wt = Color(px[0], px[1], px[2]).luminance() * sin_j;
                q[(size_t)j * w + i] = wt > 0.0f ? (double)wt : 0.0;
                total += q[(size_t)j * w + i];
            }
        }
        if (!(total > 0.0)) return false;
        aprob.assign(n, 1.0f);
        alias.resize(n);
        std::vector<uint32_t> small, large;
        for (size_t t = 0; t < n; t++) {
            q[t] = q[t] * (double)n / total;
            alias[t] = (uint32_t)t;
            if (q[t] < 1.0) small.push_back((uint32_t)t); else large.push_back((uint32_t)t);
        }
        size_t si = 0, li = 0;
        while (si < small.size() && li < large.size()) {
            uint32_t a = small[si++], g = large[li];
            aprob[a] = (float)q[a];
            alias[a] = g;
            q[g] = (q[g] + q[a]) - 1.0;
            if (q[g] < 1.0) { li++; small.push_back(g); }
        }
        use_alias = true;
        return true;
    }
    void init(const float* rgb, uint32_t w_, uint32_t h_, float rot_y) {
        w = w_; h = h_; rot = rot_y;
        img.assign(rgb, rgb + (size_t)w * h * 3);
        std::vector<float> row(h, 0.0f);
        cond.assign((size_t)w * h, 0.0f);
        for (uint32_t j = 0; j < h; j++) {
            float sin_j = std::sin(PI * ((float)j + 0.5f) / (float)h);
            float sum = 0.0f;
            for (uint32_t i = 0; i < w; i++) {
                const float* px = &img[((size_t)j * w + i) * 3];
                sum += Color(px[0], px[1], px[2]).luminance() * sin_j;
                cond[(size_t)j * w + i] = sum;
            }
            if (sum > 0.0f) for (uint32_t i = 0; i < w; i++) cond[(size_t)j * w + i] /= sum;
            else for (uint32_t i = 0; i < w; i++) cond[(size_t)j * w + i] = (float)(i + 1) / (float)w;
            row[j] = sum;
        }
        float total = 0.0f;
        for (uint32_t j = 0; j < h; j++) total += row[j];
        marginal.assign(h, 0.0f);
        float cum = 0.0f;
        for (uint32_t j = 0; j < h; j++) {
            cum += row[j];
            marginal[j] = total > 0.0f ? cum / total : (float)(j + 1) / (float)h;
        }
        inv_total = total > 0.0f ? 1.0f / total : 0.0f;
    }
    Vec3 rot_y(const Vec3& d) const { float s = std::sin(rot), c = std::cos(rot); return Vec3(d.x * c + d.z * s, d.y, -d.x * s + d.z * c); }
    Vec3 unrot_y(const Vec3& d) const { float s = std::sin(rot), c = std::cos(rot); return Vec3(d.x * c - d.z * s, d.y, d.x * s + d.z * c); }
    Color px(int32_t ix, int32_t iy) const {
        int32_t wi = (int32_t)w, hi = (int32_t)h;
        int32_t x = ((ix % wi) + wi) % wi;
        int32_t y = iy < 0 ? 0 : (iy > hi - 1 ? hi - 1 : iy);
        const float* p = &img[((size_t)y * w + x) * 3];
        return Color(p[0], p[1], p[2]);
    }
    Color sample(const Vec3& dir) const {
        Vec3 d = unrot_y(dir.unit());
        float u = rem_euclid(0.5f + std::atan2(d.x, d.z) / (2.0f * PI), 1.0f);
        float v = std::acos(rclamp(d.y, -1.0f, 1.0f)) / PI;
        float x = u * (float)w - 0.5f, y = v * (float)h - 0.5f;
        int32_t x0 = f2i32(std::floor(x)), y0 = f2i32(std::floor(y));
        float fx = x - (float)x0, fy = y - (float)y0;
        Color c00 = px(x0, y0), c10 = px(x0 + 1, y0), c01 = px(x0, y0 + 1), c11 = px(x0 + 1, y0 + 1);
        return (c00 * (1.0f - fx) + c10 * fx) * (1.0f - fy) + (c01 * (1.0f - fx) + c11 * fx) * fy;
    }
    static size_t partition_point(const float* a, size_t n, float xi) {
        size_t lo = 0, hi = n;   // first index with a[i] >= xi
        while (lo < hi) { size_t mid = lo + (hi - lo) / 2; if (a[mid] < xi) lo = mid + 1; else hi = mid; }
        return lo;
    }
    float dir_pdf_at(size_t i, size_t j) const {
        const float* p = &img[(j * w + i) * 3];
        float lum = Color(p[0], p[1], p[2]).luminance();
        return lum * (float)((size_t)w * h) * inv_total / (2.0f * PI * PI);
    }
    // cdf: the two CDF searches even when the alias table samples the map
    // (rebuild_caustics draws through EnvMapData::sample_dir, src/scene.rs:58-62)
    Vec3 sample_dir(SmallRng& rng, float& pdf, bool cdf = false) const {
        size_t i, j;
        if (use_alias && !cdf) {
            size_t n = (size_t)w * h;
            size_t t = (size_t)(((unsigned __int128)rng.next_u64() * (unsigned __int128)n) >> 64);
            if (t >= n) t = n - 1;
            if (rng.gen() >= aprob[t]) t = alias[t];
            j = t / w;
            i = t - j * w;
        } else {
            j = std::min<size_t>(partition_point(marginal.data(), h, rng.gen()), h - 1);
            i = std::min<size_t>(partition_point(&cond[j * w], w, rng.gen()), w - 1);
        }
        float u = ((float)i + rng.gen()) / (float)w;
        float v = ((float)j + rng.gen()) / (float)h;
        float phi = 2.0f * PI * (u - 0.5f);
        float theta = PI * v;
        float sin_t = std::sin(theta);
        Vec3 dir = rot_y(Vec3(sin_t * std::sin(phi), std::cos(theta), sin_t * std::cos(phi)));
        pdf = dir_pdf_at(i, j);
        return dir;
    }
    float dir_pdf(const Vec3& dir) const {
        Vec3 d = unrot_y(dir.unit());
        float u = rem_euclid(0.5f + std::atan2(d.x, d.z) / (2.0f * PI), 1.0f);
        float v = std::acos(rclamp(d.y, -1.0f, 1.0f)) / PI;
        size_t i = std::min<size_t>(f2usize(u * (float)w), w - 1);
        size_t j = std::min<size_t>(f2usize(v * (float)h), h - 1);
        return dir_pdf_at(i, j);
    }
};

static inline Color preetham_sky(const Vec3& dir, const Vec3& sun_dir, float turbidity) {
    Vec3 d = dir.unit(), sun = sun_dir.unit();
    float cos_gamma = rclamp(d.dot(sun), -1.0f, 1.0f);
    float gamma = std::acos(cos_gamma);
    float t = rclamp(turbidity, 1.7f, 20.0f);
    float cos_theta_s = rmax(sun.y, 0.0175f);
    float theta_s = std::acos(cos_theta_s);
    auto perez = [](float a, float b, float c, float dp, float e, float ct, float g, float cg) {
        return (1.0f + a * std::exp(b / ct)) * (1.0f + c * std::exp(dp * g) + e * cg * cg);
    };
    float ay = 0.1787f * t - 1.4630f, by_ = -0.3554f * t + 0.4275f, cy = 0.0227f * t + 5.3251f, dy = 0.1206f * t - 2.5771f, ey = -0.0670f * t + 0.3703f;
    float ax = -0.0193f * t - 0.2592f, bx = -0.0665f * t + 0.0008f, cx = -0.0004f * t + 0.2125f, dx = -0.0641f * t - 0.8989f, ex = -0.0033f * t + 0.0452f;
    float acy = -0.0167f * t - 0.2608f, bcy = -0.0950f * t + 0.0092f, ccy = -0.0079f * t + 0.2102f, dcy = -0.0441f * t - 1.6537f, ecy = -0.0109f * t + 0.0529f;
    float fn_y = perez(ay, by_, cy, dy, ey, 1.0f, theta_s, cos_theta_s);
    float fn_x = perez(ax, bx, cx, dx, ex, 1.0f, theta_s, cos_theta_s);
    float fn_cy = perez(acy, bcy, ccy, dcy, ecy, 1.0f, theta_s, cos_theta_s);
    float yz = rmax((4.0453f * t - 4.9710f) * std::tan((4.0f / 9.0f - t / 120.0f) * (PI - 2.0f * theta_s)) - 0.2155f * t + 2.4192f, 0.0f);
    if (cos_gamma > COS_SUN_MAX && sun.y > 0.0f && d.y > 0.0f) {
        float disc = yz * 111.0f;
        return Color(disc, disc * 0.95f, disc * 0.85f);
    }
    float ts = theta_s, ts2 = ts * ts, ts3 = ts2 * ts;
    float xz = t * t * (0.00166f * ts3 - 0.00375f * ts2 + 0.00209f * ts)
             + t * (-0.02903f * ts3 + 0.06377f * ts2 - 0.03202f * ts + 0.00394f)
             + (0.11693f * ts3 - 0.21196f * ts2 + 0.06052f * ts + 0.25886f);
    float yzc = t * t * (0.00275f * ts3 - 0.00610f * ts2 + 0.00317f * ts)
              + t * (-0.04214f * ts3 + 0.08970f * ts2 - 0.04153f * ts + 0.00516f)
              + (0.15346f * ts3 - 0.26756f * ts2 + 0.06670f * ts + 0.26688f);
    float ct = rmax(d.y, 0.001f);
    float fy = perez(ay, by_, cy, dy, ey, ct, gamma, cos_gamma);
    float fx = perez(ax, bx, cx, dx, ex, ct, gamma, cos_gamma);
    float fcy = perez(acy, bcy, ccy, dcy, ecy, ct, gamma, cos_gamma);
    float cap_y = rmax(yz * fy / fn_y, 0.0f);
    float cap_x = rclamp(xz * fx / fn_x, 0.001f, 0.998f);
    float cap_yc = rclamp(yzc * fcy / fn_cy, 0.001f, 0.998f);
    float lum = cap_y * 0.05f;
    float big_x = cap_x / cap_yc * lum, big_y = lum;
    float big_z = rmax((1.0f - cap_x - cap_yc) / cap_yc * lum, 0.0f);
    float r = rmax(3.2406f * big_x - 1.5372f * big_y - 0.4986f * big_z, 0.0f);
    float g = rmax(-0.9689f * big_x + 1.8758f * big_y + 0.0415f * big_z, 0.0f);
    float b = rmax(0.0557f * big_x - 0.2040f * big_y + 1.0570f * big_z, 0.0f);
    Color sky(r, g, b);
    if (d.y < 0.0f) {
        float tt = rmin(-d.y * 8.0f, 1.0f);
        Color ground(0.07f, 0.06f, 0.05f);
        return sky * (1.0f - tt) + ground * tt;
    }
    return sky;
}

struct Background {
    int kind = 0, visible_kind = 0;
    Color color, top, bottom;
    Vec3 sun_dir; float turbidity = 3.0f;
    EnvMap env;
    Color eval_simple(int k, const Vec3& dir) const {
        if (k == 0) return color;
        float t = (rclamp(dir.y, -1.0f, 1.0f) + 1.0f) * 0.5f;
        return bottom * (1.0f - t) + top * t;
    }
    Color eval(const Vec3& dir) const {
        switch (kind) {
            case 0: case 1: return eval_simple(kind, dir);
            case 2: return preetham_sky(dir, sun_dir, turbidity);
            case 3: return env.sample(dir);
            default: return eval_simple(visible_kind, dir);
        }
    }
    bool has_env() const { return kind == 3 || kind == 4; }
    Color env_eval(const Vec3& dir) const { return has_env() ? env.sample(dir) : Color(); }
    bool physical() const { return kind == 2; }
};

static inline float sun_pdf_value(const Vec3& dir, const Background& bg) {
    if (bg.physical() && bg.sun_dir.y > 0.0f && dir.unit().dot(bg.sun_dir.unit()) > COS_SUN_MAX)
        return 1.0f / (2.0f * PI * (1.0f - COS_SUN_MAX));
    return 0.0f;
}
static inline Vec3 sample_sun_cone(const Vec3& axis, float cos_theta_max, SmallRng& rng) {
    float r1 = rng.gen(), r2 = rng.gen();
    float cos_theta = 1.0f - r1 * (1.0f - cos_theta_max);
    float sin_theta = std::sqrt(rmax(1.0f - cos_theta * cos_theta, 0.0f));
    float phi = 2.0f * PI * r2;
    Vec3 up = std::fabs(axis.x) < 0.999f ? Vec3(1, 0, 0) : Vec3(0, 1, 0);
    Vec3 t = axis.cross(up).unit();
    Vec3 b = axis.cross(t);
    return t * (sin_theta * std::cos(phi)) + b * (sin_theta * std::sin(phi)) + axis * cos_theta;
}
static inline float env_pdf_value(const Vec3& dir, const Background& bg) { return bg.has_env() ? bg.env.dir_pdf(dir) : 0.0f; }

// ---- camera ---------------------------------------------------------------------
struct Camera {
    Point3 origin, lower_left; Vec3 horizontal, vertical, u, v; float lens_radius = 0.0f; uint32_t blades = 0;
    Ray get_ray(float s, float t, SmallRng& rng) const {
        Vec3 rd;
        if (blades >= 3) {
            float n = (float)blades, rot = PI / n;
            float k = (float)rng.gen_range_u32(0, blades);
            float a1 = 2.0f * PI * k / n + rot, a2 = 2.0f * PI * (k + 1.0f) / n + rot;
            Vec3 p1(std::cos(a1), std::sin(a1), 0.0f), p2(std::cos(a2), std::sin(a2), 0.0f);
            float r1 = rng.gen(), r2 = rng.gen();
            if (r1 + r2 > 1.0f) { r1 = 1.0f - r1; r2 = 1.0f - r2; }
            rd = lens_radius * (p1 * r1 + p2 * r2);
        } else {
            rd = lens_radius * random_in_unit_disk(rng);
        }
        Vec3 offset = u * rd.x + v * rd.y;
        return Ray::at_time(origin + offset, lower_left + s * horizontal + t * vertical - origin - offset, rng.gen());
    }
};

// ---- photon map ---------------------------------------------------------------------
struct KdPhoton { float pos[3], power[3], nor[3]; uint8_t axis; };
static const float DISK_R = 20.0f;
static const int PHOTON_MAX_DEPTH = 12;

static inline void kd_build(KdPhoton* nodes, size_t n) {
    if (n == 0) return;
    float mn[3] = {INF, INF, INF}, mx[3] = {-INF, -INF, -INF};
    for (size_t i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) {
            if (nodes[i].pos[k] < mn[k]) mn[k] = nodes[i].pos[k];
            if (nodes[i].pos[k] > mx[k]) mx[k] = nodes[i].pos[k];
        }
    // Iterator::max_by keeps the LAST maximum on ties.
    int axis = 0;
    for (int a = 1; a < 3; a++) if (!((mx[a] - mn[a]) < (mx[axis] - mn[axis]))) axis = a;
    size_t mid = n / 2;
    std::nth_element(nodes, nodes + mid, nodes + n, [axis](const KdPhoton& a, const KdPhoton& b) { return a.pos[axis] < b.pos[axis]; });
    nodes[mid].axis = (uint8_t)axis;
    kd_build(nodes, mid);
    kd_build(nodes + mid + 1, n - mid - 1);
}
static inline void kd_gather(const KdPhoton* nodes, size_t n, const float pos[3], float r2, const float nrm[3], Color& acc) {
    if (n == 0) return;
    size_t mid = n / 2;
    const KdPhoton& node = nodes[mid];
    float dx = node.pos[0] - pos[0], dy = node.pos[1] - pos[1], dz = node.pos[2] - pos[2];
    float d2 = dx * dx + dy * dy + dz * dz;
    if (d2 < r2) {
        float dot = node.nor[0] * nrm[0] + node.nor[1] * nrm[1] + node.nor[2] * nrm[2];
        if (dot > 0.0f) acc += Color(node.power[0], node.power[1], node.power[2]) * (1.0f - d2 / r2);
    }
    int axis = node.axis;
    float sq = pos[axis] - node.pos[axis];
    if (sq <= 0.0f) {
        kd_gather(nodes, mid, pos, r2, nrm, acc);
        if (sq * sq < r2) kd_gather(nodes + mid + 1, n - mid - 1, pos, r2, nrm, acc);
    } else {
        kd_gather(nodes + mid + 1, n - mid - 1, pos, r2, nrm, acc);
        if (sq * sq < r2) kd_gather(nodes, mid, pos, r2, nrm, acc);
    }
}

// k-NN gather (extension, no reference counterpart): the same Epanechnikov
// estimate with the squared radius R2 = the k-th smallest squared distance
// among the photons on the query's hemisphere closer than the gather radius,
// and the normalisation 2 / (pi^2 R2).  With fewer than k such photons (or a
// zero k-th distance) it is the fixed-radius gather.  Photons at exactly R2
// weigh 0, so ties at the k-th distance do not matter.
// k smallest squared distances (candidates: d2 < bound, positive normal dot)
// by the usual kd-tree search with a max-heap; bound shrinks to the heap top.
static inline void kd_knn(const KdPhoton* nodes, size_t n, const float pos[3], const float nrm[3], size_t k,
                          std::vector<float>& heap, float& bound) {
    if (n == 0) return;
    size_t mid = n / 2;
    const KdPhoton& node = nodes[mid];
    float dx = node.pos[0] - pos[0], dy = node.pos[1] - pos[1], dz = node.pos[2] - pos[2];
    float d2 = dx * dx + dy * dy + dz * dz;
    if (d2 < bound && node.nor[0] * nrm[0] + node.nor[1] * nrm[1] + node.nor[2] * nrm[2] > 0.0f) {
        if (heap.size() == k) { std::pop_heap(heap.begin(), heap.end()); heap.pop_back(); }
        heap.push_back(d2);
        std::push_heap(heap.begin(), heap.end());
        if (heap.size() == k) bound = heap.front();
    }
    int axis = node.axis;
    float sq = pos[axis] - node.pos[axis];
    const KdPhoton* lo = nodes; size_t nlo = mid;
    const KdPhoton* hi = nodes + mid + 1; size_t nhi = n - mid - 1;
    if (sq <= 0.0f) {
        kd_knn(lo, nlo, pos, nrm, k, heap, bound);
        if (sq * sq < bound) kd_knn(hi, nhi, pos, nrm, k, heap, bound);
    } else {
        kd_knn(hi, nhi, pos, nrm, k, heap, bound);
        if (sq * sq < bound) kd_knn(lo, nlo, pos, nrm, k, heap, bound);
    }
}

struct PhotonMap {
    std::vector<KdPhoton> nodes; float gather_r2 = 0, norm = 0;
    uint32_t knn = 0;
    void finish(float radius) {
        kd_build(nodes.data(), nodes.size());
        set_radius(radius);
    }
    void set_radius(float radius) {
        gather_r2 = radius * radius;
        norm = 2.0f / (PI * PI * radius * radius);
    }
    float knn_r2(const Point3& pos, const Vec3& normal) const {
        float p[3] = {pos.x, pos.y, pos.z}, n[3] = {normal.x, normal.y, normal.z};
        std::vector<float> heap;
        heap.reserve(knn);
        float bound = gather_r2;
        kd_knn(nodes.data(), nodes.size(), p, n, knn, heap, bound);
        if (heap.size() < knn) return gather_r2;
        float r2 = heap.front();
        return r2 > 0.0f ? r2 : gather_r2;
    }
    Color irradiance(const Point3& pos, const Vec3& normal) const {
        if (nodes.empty()) return Color();
        float p[3] = {pos.x, pos.y, pos.z}, n[3] = {normal.x, normal.y, normal.z};
        float r2 = knn ? knn_r2(pos, normal) : gather_r2;
        float nm = r2 == gather_r2 ? norm : 2.0f / (PI * PI * r2);
        Color acc;
        kd_gather(nodes.data(), nodes.size(), p, r2, n, acc);
        return acc * nm;
    }
};

// wide (extension): the photon's wavelength over [380, 780) with the extended table
static inline bool trace_photon(const Hittable& world, Point3 pos, Vec3 dir, float lum_clamp, Color power, SmallRng& rng, KdPhoton& out,
                                bool wide = false) {
    uint32_t spec_depth = 0;
    float wavelength = rng.gen_range(380.0f, wide ? 780.0f : 700.0f);
    for (int i = 0; i < PHOTON_MAX_DEPTH; i++) {
        Ray ray = Ray::at_time(pos, dir, 0.0f);
        ray.wavelength = wavelength;
        ray.wl = wide ? 1 : 0;
        HitRecord rec;
        if (!world.hit(ray, 0.001f, INF, rec)) return false;
        ScatterRecord sr;
        if (!rec.mat->scatter(ray, rec, rng, sr)) return false;
        if (sr.skip_pdf) {
            power *= sr.attenuation;
            pos = sr.ray.origin; dir = sr.ray.direction;
            spec_depth++;
        } else {
            if (spec_depth > 0 && rec.mat->can_receive_caustics()) {
                float lum = power.luminance();
                if (lum > lum_clamp) power *= lum_clamp / lum;
                out.pos[0] = rec.p.x; out.pos[1] = rec.p.y; out.pos[2] = rec.p.z;
                out.power[0] = power.x; out.power[1] = power.y; out.power[2] = power.z;
                out.nor[0] = rec.normal.x; out.nor[1] = rec.normal.y; out.nor[2] = rec.normal.z;
                out.axis = 0;
                return true;
            }
            return false;
        }
    }
    return false;
}

template <class F>
static void parallel_for(size_t n, int threads, F f) {
    if (threads <= 1 || n < 2) { for (size_t i = 0; i < n; i++) f(i); return; }
    std::atomic<size_t> next(0);
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; t++)
        pool.emplace_back([&]() { for (;;) { size_t i = next.fetch_add(1); if (i >= n) break; f(i); } });
    for (auto& th : pool) th.join();
}

static inline uint64_t photon_seed(uint32_t i) { return ((uint64_t)i * 6364136223846793005ULL) ^ 0x9E3779B97F4A7C15ULL; }

static inline void photon_build_from_quad(PhotonMap& pm, const Hittable& world, Point3 lo, Vec3 lu, Vec3 lv, Color lc, uint32_t num, float radius, int threads,
                                          bool wide = false) {
    Vec3 cr = lu.cross(lv);
    float quad_area = cr.length();
    Vec3 normal = cr / quad_area;
    Vec3 ta, ba; vonb(normal, ta, ba);
    Color power = lc * (quad_area * PI / (float)num);
    std::vector<KdPhoton> tmp(num); std::vector<uint8_t> ok(num, 0);
    parallel_for(num, threads, [&](size_t i) {
        SmallRng rng = SmallRng::seed_from_u64(photon_seed((uint32_t)i));
        float s = rng.gen(), t = rng.gen();
        Point3 origin = lo + lu * s + lv * t;
        float cos_theta = std::sqrt(rng.gen());
        float sin_theta = std::sqrt(1.0f - cos_theta * cos_theta);
        float phi = 2.0f * PI * rng.gen();
        Vec3 dir = sin_theta * std::cos(phi) * ta + sin_theta * std::sin(phi) * ba + cos_theta * normal;
        ok[i] = trace_photon(world, origin, dir, 15.0f, power, rng, tmp[i], wide);
    });
    pm.nodes.clear();
    for (size_t i = 0; i < num; i++) if (ok[i]) pm.nodes.push_back(tmp[i]);
    pm.finish(radius);
}

static inline void photon_build_sun(PhotonMap& pm, const Hittable& world, bool have_bounds, Aabb bb, Vec3 sun_dir, Color sun_color, uint32_t num, float radius, int threads,
                                    bool wide = false) {
    Vec3 sun_down = (-sun_dir).unit();
    Vec3 up = std::fabs(sun_down.x) < 0.999f ? Vec3(1, 0, 0) : Vec3(0, 1, 0);
    Vec3 t = sun_down.cross(up).unit();
    Vec3 b = sun_down.cross(t);
    const float SCENE_CLIP = 1.0e3f;
    Point3 dc; float dr;
    if (have_bounds) {
        auto clip = [&](float v) { return rclamp(v, -SCENE_CLIP, SCENE_CLIP); };
        Vec3 mn(clip(bb.min.x), clip(bb.min.y), clip(bb.min.z)), mx(clip(bb.max.x), clip(bb.max.y), clip(bb.max.z));
        Vec3 c = (mn + mx) * 0.5f;
        float r = rmax((mx - mn).length() * 0.5f, 1.0e-3f);
        dc = c + sun_dir.unit() * r; dr = r;
    } else { dc = sun_dir.unit() * 30.0f; dr = DISK_R; }
    float disk_area = PI * dr * dr;
    Color power = sun_color * (disk_area / ((float)num * PI));
    float lum_clamp = 15.0f * disk_area / (PI * DISK_R * DISK_R);
    std::vector<KdPhoton> tmp(num); std::vector<uint8_t> ok(num, 0);
    parallel_for(num, threads, [&](size_t i) {
        SmallRng rng = SmallRng::seed_from_u64(photon_seed((uint32_t)i));
        float r = dr * std::sqrt(rng.gen());
        float phi = 2.0f * PI * rng.gen();
        Point3 origin = dc + t * (r * std::cos(phi)) + b * (r * std::sin(phi));
        ok[i] = trace_photon(world, origin, sun_down, lum_clamp, power, rng, tmp[i], wide);
    });
    pm.nodes.clear();
    for (size_t i = 0; i < num; i++) if (ok[i]) pm.nodes.push_back(tmp[i]);
    pm.finish(radius);
}

// ---- integrator ---------------------------------------------------------------------
struct SceneView {
    const Hittable* world;
    const HittableList* lights;
    const Background* bg;
    const PhotonMap* photons;  // may be null
};

// Radiance hooks: every term added to the pixel passes through H::add (terms
// fixed at this vertex) or H::add_v (terms that carry this vertex's BSDF
// value), every sampled continuation through H::scatter.  The plain render
// uses NoHook; the adjoint checker's hooks live in rc_adjoint.h.
struct NoHook {
    Color add(const Color& c) { return c; }
    Color add_v(const Color& c, const HitRecord&, const Ray&, const Vec3&) { return c; }
    void scatter(const HitRecord&, const Ray&, const Vec3&) {}
    void segment(const Ray&, bool, const HitRecord&) {}     // a traced path segment (to the hit, or a miss)
    void shadow_seg(const Ray&, float) {}                    // the shadow ray of the next contribution
};

// ---- four hero wavelengths (extension) ---------------------------------------
// Hero wavelength spectral sampling (Wilkie, Nawaz, Droske, Weidlich and
// Hanika, EGSR 2014) over the tinted-RGB estimator: the camera ray's lambda_0
// is uniform over [380, 780) and hero_lambda(lambda_0, k) are its rotations.
// Decisions and contributions are made at lambda_0 with no spectral tint;
// each contribution is then multiplied by the balance-heuristic combination
// of the four wavelengths,
//     (T(l_0) E_0 + sum_k q_k b_k E_k T(l_k)) / (E_0 (1 + sum_k r_k)),
// q_k and r_k the running products of the BSDF-value and density ratios of
// the path's sampled vertices (Material::hero_scatter), b_k the next-event
// connection's own BSDF ratio (Material::hero_brdf), E a blackbody emitter's
// Planck radiance, T the tint of the path's first spectral event per
// wavelength.  No spectral event: weight 1 (the one-wavelength estimator).
struct Hero4 {
    float q[3] = {1, 1, 1}, r[3] = {1, 1, 1};
    bool on = false;
    static Color tint(int kind, float lam) {
        if (kind == 1) return spectral_to_rgb(lam, true);
        if (kind == 2) return spectral_fire_rgb(lam, true);
        return Color(1, 1, 1);
    }
    Color weight(const Ray& ray, int kind, const float* b = nullptr, float bb_temp = 0.0f) const {
        if (ray.wl != 2) return Color(1, 1, 1);
        if (!on && kind == 0 && !b && !(bb_temp > 0.0f)) return Color(1, 1, 1);
        float l0 = ray.wavelength;
        float e0 = bb_temp > 0.0f ? planck_raw(l0, bb_temp) : 1.0f;
        Color num = tint(kind, l0);
        float den = 1.0f;
        for (int k = 0; k < 3; k++) {
            float lk = hero_lambda(l0, k + 1);
            float w = q[k] * (b ? b[k] : 1.0f) * (bb_temp > 0.0f ? planck_raw(lk, bb_temp) / e0 : 1.0f);
            if (w != 0.0f) num += tint(kind, lk) * w;
            den += r[k];
        }
        return num / den;
    }
};
// A blackbody emitter reached on a tinted path emits per wavelength.
static inline float hero_bb_temp(const HitRecord& rec, const Ray& ray) {
    if (!ray.spectral_weighted) return 0.0f;
    const BlackbodyLight* bb = dynamic_cast<const BlackbodyLight*>(rec.mat);
    return bb ? bb->temp_k : 0.0f;
}

template <class H>
static inline Color ray_color_h(const Ray& r0, const SceneView& S, float bg_scale, int max_depth, SmallRng& rng, H& hk) {
    const Hittable& world = *S.world;
    const HittableList& lights = *S.lights;
    const Background& background = *S.bg;
    Color throughput(1, 1, 1), color;
    Ray ray = r0;
    bool prev_specular = true;
    float prev_mis_w_brdf = 1.0f, prev_spec_sun_weight = 1.0f;
    Hero4 h4;
    for (int depth = 0; depth < max_depth; depth++) {
        HitRecord rec;
        bool hit = world.hit(ray, 0.001f, INF, rec);
        hk.segment(ray, hit, rec);
        const int ptint = ray.spectral_weighted ? ray.tint : 0;
        if (!hit) {
            float sun_pdf = sun_pdf_value(ray.direction, background);
            float env_pdf = env_pdf_value(ray.direction, background);
            float emit_w = (sun_pdf + env_pdf == 0.0f) ? 1.0f : (prev_specular ? prev_spec_sun_weight : prev_mis_w_brdf);
            Color bg = (depth == 0 || !background.has_env()) ? background.eval(ray.direction) : background.env_eval(ray.direction);
            color += hk.add(throughput * bg * bg_scale * emit_w * h4.weight(ray, ptint));
            break;
        }
        float emit_w = (prev_specular || lights.objects.empty()) ? 1.0f : prev_mis_w_brdf;
        Color em = rec.mat->emitted_wl(rec.u, rec.v, rec.p, ray.wavelength, ray.spectral_weighted, ray.wl);
        if (em.x != 0.0f || em.y != 0.0f || em.z != 0.0f) em = em * h4.weight(ray, ptint, nullptr, hero_bb_temp(rec, ray));
        color += hk.add(throughput * em * emit_w);
        const Color w4 = h4.weight(ray, ptint);      // this vertex's next-event weight
        ScatterRecord sr;
        if (!rec.mat->scatter(ray, rec, rng, sr)) break;
        if (sr.skip_pdf) {
            prev_spec_sun_weight = 1.0f;
            if (background.physical() && background.sun_dir.y > 0.0f) {
                float p_sun = 1.0f / (2.0f * PI * (1.0f - COS_SUN_MAX));
                Vec3 sun_u = background.sun_dir.unit();
                Vec3 sun_sample = sample_sun_cone(sun_u, COS_SUN_MAX, rng);
                Ray sun_shadow = Ray::at_time(rec.p, sun_sample, ray.time);
                float tr = world.transmittance(sun_shadow, 0.001f, INF);
                hk.shadow_seg(sun_shadow, INF);
                if (tr > 0.0f) {
                    Color brdf_cos = rec.mat->specular_brdf_cos(ray, rec, sun_sample);
                    if (brdf_cos.x > 0.0f || brdf_cos.y > 0.0f || brdf_cos.z > 0.0f) {
                        float p_mat = rec.mat->specular_sampling_pdf(ray, rec, sun_sample);
                        Color sun_color = background.eval(sun_sample);
                        float b[3];
                        bool spec = rec.mat->hero_brdf(ray, rec, sun_sample, b);
                        sun_color = sun_color * h4.weight(ray, ray.spectral_weighted ? ray.tint : (spec ? 1 : 0), spec ? b : nullptr);
                        float mis_d = p_sun * p_sun + p_mat * p_mat;
                        if (mis_d > 0.0f) color += hk.add(throughput * tr * brdf_cos * sun_color * p_sun / mis_d);
                    }
                }
                float p_mat_brdf = rec.mat->specular_sampling_pdf(ray, rec, sr.ray.direction);
                if (p_mat_brdf > 0.0f)
                    prev_spec_sun_weight = (p_mat_brdf * p_mat_brdf) / rmax(p_mat_brdf * p_mat_brdf + p_sun * p_sun, 1e-8f);
            }
            hk.scatter(rec, ray, sr.ray.direction);
            if (ray.wl == 2 && rec.mat->hero_scatter(ray, rec, sr.ray.direction, h4.q, h4.r)) h4.on = true;
            throughput *= sr.attenuation;
            ray = sr.ray;
            prev_specular = true;
            prev_mis_w_brdf = 1.0f;
        } else {
            if (S.photons) {
                Color irr = S.photons->irradiance(rec.p, rec.normal);
                if (irr.x > 0.0f || irr.y > 0.0f || irr.z > 0.0f) {
                    Color alb = rec.mat->albedo_hint(rec.u, rec.v, rec.p);
                    color += hk.add_v(throughput * alb * irr * w4, rec, ray, -ray.direction);
                }
            }
            if (!lights.objects.empty()) {
                Vec3 light_dir = lights.pdf_generate(rec.p, rng, ray.time);
                Ray shadow = Ray::at_time(rec.p, light_dir, ray.time);
                HitRecord lrec;
                if (lights.hit(shadow, 0.001f, INF, lrec)) {
                    float tr = world.transmittance(shadow, 0.001f, lrec.t * (1.0f - 1e-4f));
                    hk.shadow_seg(shadow, lrec.t * (1.0f - 1e-4f));
                    if (tr > 0.0f) {
                        float l_pdf = lights.pdf_value(rec.p, light_dir, ray.time);
                        float brdf = rec.mat->scattering_pdf(ray, rec, shadow);
                        float sun_p = sun_pdf_value(light_dir, background);
                        float env_p = env_pdf_value(light_dir, background);
                        float mis_d = l_pdf * l_pdf + brdf * brdf + sun_p * sun_p + env_p * env_p;
                        if (mis_d > 0.0f && brdf > 0.0f) {
                            Color nee_emit = lrec.mat->emitted_wl(lrec.u, lrec.v, lrec.p, ray.wavelength, ray.spectral_weighted, ray.wl);
                            nee_emit = nee_emit * h4.weight(ray, ptint, nullptr, hero_bb_temp(lrec, ray));
                            color += hk.add_v(throughput * tr * sr.attenuation * brdf * l_pdf * nee_emit / mis_d, rec, ray, light_dir);
                        }
                    }
                }
            }
            if (background.physical() && background.sun_dir.y > 0.0f) {
                Vec3 sun_u = background.sun_dir.unit();
                Vec3 sun_sample = sample_sun_cone(sun_u, COS_SUN_MAX, rng);
                Ray sun_shadow = Ray::at_time(rec.p, sun_sample, ray.time);
                float tr = world.transmittance(sun_shadow, 0.001f, INF);
                hk.shadow_seg(sun_shadow, INF);
                if (tr > 0.0f) {
                    float brdf = rec.mat->scattering_pdf(ray, rec, sun_shadow);
                    if (brdf > 0.0f) {
                        float sun_pdf = 1.0f / (2.0f * PI * (1.0f - COS_SUN_MAX));
                        Color sun_color = background.eval(sun_sample);
                        float area_pdf = lights.objects.empty() ? 0.0f : lights.pdf_value(rec.p, sun_sample, ray.time);
                        float mis_d = sun_pdf * sun_pdf + brdf * brdf + area_pdf * area_pdf;
                        color += hk.add_v(throughput * tr * sr.attenuation * brdf * sun_pdf * sun_color * w4 / mis_d, rec, ray, sun_sample);
                    }
                }
            }
            if (background.has_env()) {
                float env_pdf;
                Vec3 env_dir = background.env.sample_dir(rng, env_pdf);
                if (env_pdf > 0.0f) {
                    Ray shadow = Ray::at_time(rec.p, env_dir, ray.time);
                    float tr = world.transmittance(shadow, 0.001f, INF);
                    hk.shadow_seg(shadow, INF);
                    if (tr > 0.0f) {
                        float brdf = rec.mat->scattering_pdf(ray, rec, shadow);
                        if (brdf > 0.0f) {
                            float area_pdf = lights.objects.empty() ? 0.0f : lights.pdf_value(rec.p, env_dir, ray.time);
                            Color env_color = background.env_eval(env_dir);
                            float mis_d = env_pdf * env_pdf + brdf * brdf + area_pdf * area_pdf;
                            color += hk.add_v(throughput * tr * sr.attenuation * brdf * env_pdf * env_color * w4 / mis_d, rec, ray, env_dir);
                        }
                    }
                }
            }
            Onb uvw = Onb::from_w(rec.mat->shading_normal(rec));
            float r1 = rng.gen(), r2 = rng.gen();
            float phi = 2.0f * PI * r1;
            float sqrt_r2 = std::sqrt(r2);
            Vec3 ind_dir = uvw.local(Vec3(std::cos(phi) * sqrt_r2, std::sin(phi) * sqrt_r2, std::sqrt(1.0f - r2)));
            float pdf_val = rmax(ind_dir.unit().dot(uvw.w) / PI, 0.0f);
            if (pdf_val <= 0.0f) break;
            Ray scattered = Ray::scatter_from(rec.p, ind_dir, ray);
            float scat_pdf = rec.mat->scattering_pdf(ray, rec, scattered);
            if (scat_pdf <= 0.0f) break;
            if (ind_dir.dot(rec.normal) <= 0.0f) break;
            float area = lights.objects.empty() ? 0.0f : lights.pdf_value(rec.p, ind_dir, ray.time);
            float sun = sun_pdf_value(ind_dir, background);
            float env = env_pdf_value(ind_dir, background);
            float nee_pdf_sq = area * area + sun * sun + env * env;
            prev_mis_w_brdf = (pdf_val * pdf_val) / rmax(pdf_val * pdf_val + nee_pdf_sq, 1e-8f);
            hk.scatter(rec, ray, ind_dir);
            throughput *= sr.attenuation * (scat_pdf / pdf_val);
            ray = scattered;
            prev_specular = false;
        }
        if (depth >= 2) {
            Color eff = throughput * h4.weight(ray, ray.spectral_weighted ? ray.tint : 0);
            float survive = rmax(rmax(eff.x, eff.y), eff.z);
            float p = rmin(survive, 1.0f);
            if (survive <= 0.0f || rng.gen() >= p) break;
            throughput /= p;
        }
    }
    return color;
}

static inline Color ray_color(const Ray& r0, const SceneView& S, float bg_scale, int max_depth, SmallRng& rng) {
    NoHook nh;
    return ray_color_h(r0, S, bg_scale, max_depth, rng, nh);
}

struct SampleParams {
    uint32_t sample_idx, strata, width, height, max_depth;
    float bg_scale;
    bool use_frame_time; float t0, t1;
    uint64_t seed;
    uint8_t wl = 0;             // Ray::wl of the camera rays
};

// Per-pixel body of render_tiles (src/renderer.rs:693-727).
template <class H>
static inline Color render_pixel_h(const SceneView& S, const Camera& cam, const SampleParams& P, uint32_t row, uint32_t col, SmallRng& rng, H& hk) {
    uint32_t i = row * P.width + col;
    float w_denom = (float)std::max<uint32_t>(P.width - 1, 1);
    float h_denom = (float)std::max<uint32_t>(P.height - 1, 1);
    uint32_t strata2 = P.strata * P.strata;
    float strata_f = (float)P.strata;
    uint32_t ray_y = P.height - 1 - row;
    float uj, vj;
    if (strata2 > 0) {
        uint32_t offset = (i * 0x9E3779B9u) % strata2;
        uint32_t s = (P.sample_idx + offset) % strata2;
        uint32_t sx = s % P.strata, sy = s / P.strata;
        uj = ((float)sx + rng.gen()) / strata_f;
        vj = ((float)sy + rng.gen()) / strata_f;
    } else {
        uj = rng.gen(); vj = rng.gen();
    }
    float u = ((float)col + uj) / w_denom;
    float v = ((float)ray_y + vj) / h_denom;
    Ray cam_ray = cam.get_ray(u, v, rng);
    if (P.use_frame_time) cam_ray.time = P.t0 + rng.gen() * (P.t1 - P.t0);
    cam_ray.wavelength = rng.gen_range(380.0f, P.wl ? 780.0f : 700.0f);
    cam_ray.wl = P.wl;
    return ray_color_h(cam_ray, S, P.bg_scale, (int)P.max_depth, rng, hk);
}
static inline Color render_pixel(const SceneView& S, const Camera& cam, const SampleParams& P, uint32_t row, uint32_t col, SmallRng& rng) {
    NoHook nh;
    return render_pixel_h(S, cam, P, row, col, rng, nh);
}

// The GPU path seeds one stream per pixel instead of one per row; the oracle
// offers that mode too so both sides can draw the same numbers.
static inline uint64_t pixel_seed(uint32_t pixel, uint32_t sample_idx, uint64_t seed) {
    return ((uint64_t)pixel * 0x9E3779B97F4A7C15ULL) ^ ((uint64_t)sample_idx * 6364136223846793005ULL)
           ^ (seed * 0xD1B54A32D192ED03ULL);
}

// mode 0: reference row streams (src/renderer.rs:688-691); mode 1: per-pixel streams.
static inline void render_tiles(float* out, const uint8_t* converged, const SceneView& S, const Camera& cam,
                                const SampleParams& P, int mode, int threads) {
    parallel_for(P.height, threads, [&](size_t row) {
        SmallRng rng = SmallRng::seed_from_u64(((uint64_t)row * 0x9E3779B97F4A7C15ULL) ^ ((uint64_t)P.sample_idx * 6364136223846793005ULL));
        for (uint32_t col = 0; col < P.width; col++) {
            size_t i = (size_t)row * P.width + col;
            if (converged && converged[i]) { out[i * 3] = out[i * 3 + 1] = out[i * 3 + 2] = 0.0f; continue; }
            if (mode == 1) rng = SmallRng::seed_from_u64(pixel_seed((uint32_t)i, P.sample_idx, P.seed));
            Color c = render_pixel(S, cam, P, (uint32_t)row, col, rng);
            out[i * 3] = c.x; out[i * 3 + 1] = c.y; out[i * 3 + 2] = c.z;
        }
    });
}

}  // namespace rc
