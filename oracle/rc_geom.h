// ORACLE — test infrastructure only (see rc_math.h).
//
// Geometry layer restated from rustracer:
//   HitRecord / Material / Hittable    src/hittable.rs:7-136
//   HittableList                      src/hittable.rs:138-225
//   Sphere                            src/sphere.rs:11-118
//   Quad                              src/quad.rs:8-119, make_box 201-215
//   Triangle                          src/triangle.rs:11-145
//   convex_hit / Diamond::hit          src/diamond.rs:203-274, micro_tilt src/hittable.rs:261-270
//   BvhTree (SAH QBVH, width 8)        src/bvh.rs:8-603 (AVX2 build: QBVH_WIDTH 8, depth 3)
//   UniformScale / Translate / Rotate  src/transform.rs:15-277
// Identification for parity: every Hittable carries `tag`; the world BvhTree
// writes its leaf's tag into HitRecord.obj_id, a nested BvhTree (mesh, box)
// into HitRecord.prim_id.
#pragma once
#include "rc_math.h"
#include <vector>
#include <memory>
#include <algorithm>
#include <cstring>

namespace rc {

struct Material;

struct HitRecord {
    Point3 p;
    Vec3 normal;
    const Material* mat = nullptr;
    float t = 0, u = 0, v = 0;
    bool front_face = false;
    Vec3 tangent;
    int obj_id = -1;
    int prim_id = 0;
    static HitRecord make(const Point3& p, float t, const Material* m, const Ray& r, const Vec3& outward) {
        HitRecord h;
        h.p = p; h.t = t; h.mat = m;
        h.front_face = r.direction.dot(outward) < 0.0f;
        h.normal = h.front_face ? outward : -outward;
        return h;
    }
};

struct Hittable {
    int tag = 0;
    virtual ~Hittable() {}
    virtual bool hit(const Ray& r, float t_min, float t_max, HitRecord& rec) const = 0;
    virtual bool bounding_box(Aabb& out) const = 0;
    // Hittable::photon_bounds (src/hittable.rs:105-109): the bounds that size
    // the caustic photon disk; proxy geometry (the infinite plane) has none.
    virtual bool photon_bounds(Aabb& out) const { return bounding_box(out); }
    virtual bool any_hit(const Ray& r, float t_min, float t_max) const {
        HitRecord tmp;
        return hit(r, t_min, t_max, tmp);
    }
    virtual float transmittance(const Ray& r, float t_min, float t_max) const {
        return any_hit(r, t_min, t_max) ? 0.0f : 1.0f;
    }
    virtual float light_area() const { return 0.0f; }
    virtual float pdf_value(const Point3&, const Vec3&, float) const { return 0.0f; }
    virtual Vec3 pdf_generate(const Point3&, SmallRng&, float) const { return Vec3(1, 0, 0); }
};
typedef std::shared_ptr<Hittable> HPtr;

bool material_is_emissive(const Material* m);

// ---- Sphere ----------------------------------------------------------------
struct Sphere : Hittable {
    Point3 center, center1;
    bool moving = false;
    float radius;
    const Material* mat;
    Sphere(Point3 c, float r, const Material* m) : center(c), radius(r), mat(m) {}
    Point3 center_at(float time) const {
        return moving ? center + time * (center1 - center) : center;
    }
    static void sphere_uv(const Vec3& p, float& u, float& v) {
        float theta = std::acos(rclamp(-p.y, -1.0f, 1.0f));
        float phi = std::atan2(-p.z, p.x) + PI;
        u = phi / (2.0f * PI);
        v = theta / PI;
    }
    float light_area() const override {
        return material_is_emissive(mat) ? 4.0f * PI * radius * radius : 0.0f;
    }
    bool hit(const Ray& r, float t_min, float t_max, HitRecord& rec) const override {
        Point3 c = center_at(r.time);
        Vec3 oc = r.origin - c;
        float a = r.direction.length_squared();
        if (!(a > 0.0f)) return false;
        float half_b = oc.dot(r.direction);
        float cc = oc.length_squared() - radius * radius;
        float disc = half_b * half_b - a * cc;
        if (disc < 0.0f) return false;
        float sqrtd = std::sqrt(disc);
        float root = (-half_b - sqrtd) / a;
        if (root < t_min || root > t_max) {
            root = (-half_b + sqrtd) / a;
            if (root < t_min || root > t_max) return false;
        }
        Point3 p = r.at(root);
        Vec3 outward = (p - c) / radius;
        float u, v;
        sphere_uv(outward, u, v);
        rec = HitRecord::make(p, root, mat, r, outward);
        rec.u = u; rec.v = v;
        return true;
    }
    bool any_hit(const Ray& r, float t_min, float t_max) const override {
        Point3 c = center_at(r.time);
        Vec3 oc = r.origin - c;
        float a = r.direction.length_squared();
        if (!(a > 0.0f)) return false;
        float half_b = oc.dot(r.direction);
        float cc = oc.length_squared() - radius * radius;
        float disc = half_b * half_b - a * cc;
        if (disc < 0.0f) return false;
        float sqrtd = std::sqrt(disc);
        float t1 = (-half_b - sqrtd) / a;
        if (t1 >= t_min && t1 <= t_max) return true;
        float t2 = (-half_b + sqrtd) / a;
        return t2 >= t_min && t2 <= t_max;
    }
    float pdf_value(const Point3& origin, const Vec3&, float time) const override {
        Vec3 to_center = center_at(time) - origin;
        float dist_sq = to_center.length_squared();
        float r2 = radius * radius;
        if (dist_sq <= r2) return 0.0f;
        float cos_theta_max = std::sqrt(1.0f - r2 / dist_sq);
        return 1.0f / (2.0f * PI * (1.0f - cos_theta_max));
    }
    Vec3 pdf_generate(const Point3& origin, SmallRng& rng, float time) const override {
        Vec3 to_center = center_at(time) - origin;
        float dist_sq = to_center.length_squared();
        float r2 = radius * radius;
        if (dist_sq <= r2) return to_center.unit();
        float cos_theta_max = std::sqrt(1.0f - r2 / dist_sq);
        float r1 = rng.gen();
        float r2r = rng.gen();
        float phi = 2.0f * PI * r1;
        float cos_theta = 1.0f - r2r * (1.0f - cos_theta_max);
        float sin_theta = std::sqrt(rmax(1.0f - cos_theta * cos_theta, 0.0f));
        Vec3 local(std::cos(phi) * sin_theta, std::sin(phi) * sin_theta, cos_theta);
        return Onb::from_w(to_center.unit()).local(local);
    }
    bool bounding_box(Aabb& out) const override {
        Vec3 rr(radius, radius, radius);
        Aabb b0(center - rr, center + rr);
        out = moving ? Aabb::surrounding(b0, Aabb(center1 - rr, center1 + rr)) : b0;
        return true;
    }
};

// ---- Quad ------------------------------------------------------------------
struct Quad : Hittable {
    Point3 q; Vec3 u, v, normal, w; float d, area;
    const Material* mat;
    Quad(Point3 q_, Vec3 u_, Vec3 v_, const Material* m) : q(q_), u(u_), v(v_), mat(m) {
        Vec3 n = u.cross(v);
        area = n.length();
        normal = n / area;
        d = normal.dot(q);
        w = n / n.dot(n);
    }
    bool hit(const Ray& r, float t_min, float t_max, HitRecord& rec) const override {
        float denom = normal.dot(r.direction);
        if (!(std::fabs(denom) >= 1e-8f)) return false;
        float t = (d - normal.dot(r.origin)) / denom;
        if (t < t_min || t > t_max) return false;
        Point3 p = r.at(t);
        Vec3 planar = p - q;
        float alpha = w.dot(planar.cross(v));
        float beta = w.dot(u.cross(planar));
        if (!(alpha >= 0.0f && alpha <= 1.0f) || !(beta >= 0.0f && beta <= 1.0f)) return false;
        rec = HitRecord::make(p, t, mat, r, normal);
        rec.u = alpha; rec.v = beta;
        return true;
    }
    bool any_hit(const Ray& r, float t_min, float t_max) const override {
        float denom = normal.dot(r.direction);
        if (!(std::fabs(denom) >= 1e-8f)) return false;
        float t = (d - normal.dot(r.origin)) / denom;
        if (t < t_min || t > t_max) return false;
        Point3 p = r.at(t);
        Vec3 planar = p - q;
        float alpha = w.dot(planar.cross(v));
        float beta = w.dot(u.cross(planar));
        return (alpha >= 0.0f && alpha <= 1.0f) && (beta >= 0.0f && beta <= 1.0f);
    }
    bool bounding_box(Aabb& out) const override {
        Point3 c[4] = {q, q + u, q + v, q + u + v};
        Point3 mn = c[0], mx = c[0];
        for (int i = 1; i < 4; i++) {
            mn = Point3(rmin(mn.x, c[i].x), rmin(mn.y, c[i].y), rmin(mn.z, c[i].z));
            mx = Point3(rmax(mx.x, c[i].x), rmax(mx.y, c[i].y), rmax(mx.z, c[i].z));
        }
        out = Aabb(mn, mx).pad();
        return true;
    }
    float pdf_value(const Point3& origin, const Vec3& dir, float) const override {
        Ray r = Ray::make(origin, dir);
        HitRecord rec;
        if (!hit(r, 0.001f, INF, rec)) return 0.0f;
        float dist2 = (rec.p - origin).length_squared();
        float cos_theta = std::fabs(dir.unit().dot(rec.normal));
        if (cos_theta < 1e-8f) return 0.0f;
        return dist2 / (cos_theta * area);
    }
    Vec3 pdf_generate(const Point3& origin, SmallRng& rng, float) const override {
        float r1 = rng.gen();
        float r2 = rng.gen();
        Point3 point = q + r1 * u + r2 * v;
        return (point - origin).unit();
    }
    float light_area() const override { return material_is_emissive(mat) ? area : 0.0f; }
};

// ---- Cylinder (src/cylinder.rs:124-224) ----------------------------------------------
static const float TAU_F = 6.28318530717958647692f;
static inline float rs_fract(float x) { return x - std::trunc(x); }

struct Cylinder : Hittable {
    Point3 center; float radius = 1.0f, height = 1.0f;
    const Material* mat = nullptr;
    bool hit(const Ray& r, float t_min, float t_max, HitRecord& rec) const override {
        Vec3 oc = r.origin - center;
        float dx = r.direction.x, dz = r.direction.z, dy = r.direction.y;
        float best_t = t_max;
        bool found = false;
        float a = dx * dx + dz * dz;
        if (a > 1e-8f) {
            float half_b = oc.x * dx + oc.z * dz;
            float c = oc.x * oc.x + oc.z * oc.z - radius * radius;
            float disc = half_b * half_b - a * c;
            if (disc >= 0.0f) {
                float sq = std::sqrt(disc);
                float roots[2] = {(-half_b - sq) / a, (-half_b + sq) / a};
                for (float root : roots) {
                    if (root < t_min || root > best_t) continue;
                    Point3 p = r.at(root);
                    float ly = p.y - center.y;
                    if (ly < 0.0f || ly > height) continue;
                    Vec3 outward((p.x - center.x) / radius, 0.0f, (p.z - center.z) / radius);
                    rec = HitRecord::make(p, root, mat, r, outward);
                    float phi = std::atan2(p.z - center.z, p.x - center.x);
                    rec.u = rs_fract(phi / TAU_F + 0.5f);
                    rec.v = ly / height;
                    best_t = root; found = true;
                }
            }
        }
        if (std::fabs(dy) > 1e-8f) {
            const float caps[2][2] = {{0.0f, -1.0f}, {height, 1.0f}};
            for (int k = 0; k < 2; k++) {
                float t = (caps[k][0] - oc.y) / dy;
                if (t < t_min || t > best_t) continue;
                Point3 p = r.at(t);
                float lx = p.x - center.x, lz = p.z - center.z;
                if (lx * lx + lz * lz > radius * radius) continue;
                rec = HitRecord::make(p, t, mat, r, Vec3(0.0f, caps[k][1], 0.0f));
                rec.u = (lx / radius + 1.0f) * 0.5f;
                rec.v = (lz / radius + 1.0f) * 0.5f;
                best_t = t; found = true;
            }
        }
        if (found) rec.prim_id = 0;
        return found;
    }
    bool bounding_box(Aabb& out) const override {
        Vec3 rv(radius, 0.0f, radius);
        out = Aabb(center - rv, center + rv + Vec3(0.0f, height, 0.0f));
        return true;
    }
};

// ---- Cone (src/cone.rs:11-129) ---------------------------------------------------------
struct Cone : Hittable {
    Point3 center; float radius = 1.0f, height = 1.0f;
    const Material* mat = nullptr;
    bool hit(const Ray& r, float t_min, float t_max, HitRecord& rec) const override {
        Vec3 oc = r.origin - center;
        float h = height;
        float k = (radius / h) * (radius / h);
        float dx = r.direction.x, dy = r.direction.y, dz = r.direction.z;
        float hy = h - oc.y;
        float best_t = t_max;
        bool found = false;
        float a = dx * dx + dz * dz - k * dy * dy;
        float half_b = oc.x * dx + oc.z * dz + k * hy * dy;
        float c = oc.x * oc.x + oc.z * oc.z - k * hy * hy;
        if (std::fabs(a) > 1e-8f) {
            float disc = half_b * half_b - a * c;
            if (disc >= 0.0f) {
                float sq = std::sqrt(disc);
                float roots[2] = {(-half_b - sq) / a, (-half_b + sq) / a};
                for (float root : roots) {
                    if (root < t_min || root > best_t) continue;
                    Point3 p = r.at(root);
                    float ly = p.y - center.y;
                    if (ly < 0.0f || ly > h) continue;
                    float lx = p.x - center.x, lz = p.z - center.z;
                    Vec3 grad(lx, k * (h - ly), lz);
                    Vec3 outward = grad.length_squared() > 1e-12f ? grad.unit() : Vec3(0.0f, 1.0f, 0.0f);
                    rec = HitRecord::make(p, root, mat, r, outward);
                    float phi = std::atan2(lz, lx);
                    rec.u = rs_fract(phi / TAU_F + 0.5f);
                    rec.v = ly / h;
                    best_t = root; found = true;
                }
            }
        }
        if (std::fabs(dy) > 1e-8f) {
            float t = -oc.y / dy;
            if (t >= t_min && t <= best_t) {
                Point3 p = r.at(t);
                float lx = p.x - center.x, lz = p.z - center.z;
                if (lx * lx + lz * lz <= radius * radius) {
                    rec = HitRecord::make(p, t, mat, r, Vec3(0.0f, -1.0f, 0.0f));
                    rec.u = (lx / radius + 1.0f) * 0.5f;
                    rec.v = (lz / radius + 1.0f) * 0.5f;
                    found = true;
                }
            }
        }
        if (found) rec.prim_id = 0;
        return found;
    }
    bool bounding_box(Aabb& out) const override {
        Vec3 rv(radius, 0.0f, radius);
        out = Aabb(center - rv, center + rv + Vec3(0.0f, height, 0.0f));
        return true;
    }
};

// ---- Disk (src/disk.rs:103-182) ----------------------------------------------------------
struct Disk : Hittable {
    Point3 center; Vec3 normal; float radius = 1.0f;
    const Material* mat = nullptr;
    Vec3 u_axis, v_axis; float area = 0.0f;
    void init(const Point3& c, const Vec3& n, float r, const Material* m) {
        center = c; normal = n.unit(); radius = r; mat = m;
        vonb(normal, u_axis, v_axis);
        area = PI * radius * radius;
    }
    bool hit(const Ray& r, float t_min, float t_max, HitRecord& rec) const override {
        float denom = normal.dot(r.direction);
        if (!(std::fabs(denom) >= 1e-8f)) return false;
        float t = normal.dot(center - r.origin) / denom;
        if (t < t_min || t > t_max) return false;
        Point3 p = r.at(t);
        Vec3 off = p - center;
        if (off.length_squared() > radius * radius) return false;
        rec = HitRecord::make(p, t, mat, r, normal);
        rec.u = (off.dot(u_axis) / radius + 1.0f) * 0.5f;
        rec.v = (off.dot(v_axis) / radius + 1.0f) * 0.5f;
        rec.prim_id = 0;
        return true;
    }
    bool bounding_box(Aabb& out) const override {
        Vec3 ext(radius * std::sqrt(rmax(1.0f - normal.x * normal.x, 0.0f)), radius * std::sqrt(rmax(1.0f - normal.y * normal.y, 0.0f)),
                 radius * std::sqrt(rmax(1.0f - normal.z * normal.z, 0.0f)));
        out = Aabb(center - ext, center + ext).pad();
        return true;
    }
    float light_area() const override { return material_is_emissive(mat) ? area : 0.0f; }
    float pdf_value(const Point3& origin, const Vec3& dir, float) const override {
        Ray r = Ray::make(origin, dir);
        HitRecord rec;
        if (!hit(r, 0.001f, INF, rec)) return 0.0f;
        float dist2 = (rec.p - origin).length_squared();
        float cos_theta = std::fabs(dir.unit().dot(rec.normal));
        if (cos_theta < 1e-8f || area <= 0.0f) return 0.0f;
        return dist2 / (cos_theta * area);
    }
    Vec3 pdf_generate(const Point3& origin, SmallRng& rng, float) const override {
        float rr = radius * std::sqrt(rng.gen());
        float theta = 2.0f * PI * rng.gen();
        Point3 p = center + rr * (std::cos(theta) * u_axis + std::sin(theta) * v_axis);
        return (p - origin).unit();
    }
};

// ---- InfinitePlane (src/plane.rs:14-107) -----------------------------------------------
struct InfinitePlane : Hittable {
    Point3 point; Vec3 normal; float wave_amplitude = 0.0f, wave_scale = 1.0f;
    const Material* mat = nullptr;
    Vec3 u_axis, v_axis;
    void init(const Point3& p, const Vec3& n, float amp, float sc, const Material* m) {
        point = p; normal = n.unit(); wave_amplitude = amp; wave_scale = sc; mat = m;
        vonb(normal, u_axis, v_axis);
    }
    bool hit(const Ray& r, float t_min, float t_max, HitRecord& rec) const override {
        float denom = normal.dot(r.direction);
        if (!(std::fabs(denom) >= 1e-8f)) return false;
        float t = normal.dot(point - r.origin) / denom;
        if (t < t_min || t > t_max) return false;
        Point3 p = r.at(t);
        Vec3 off = p - point;
        Vec3 sn = normal;
        if (wave_amplitude > 1e-6f) {
            float u = off.dot(u_axis), v = off.dot(v_axis);
            float s = wave_scale, a = wave_amplitude;
            float phi1 = s * u + 0.7f * s * v;
            float phi2 = 1.3f * s * v - 0.5f * s * u;
            float dh_du = a * s * std::cos(phi1) - 0.25f * a * s * std::cos(phi2);
            float dh_dv = 0.7f * a * s * std::cos(phi1) + 0.65f * a * s * std::cos(phi2);
            sn = (normal - dh_du * u_axis - dh_dv * v_axis).unit();
        }
        rec = HitRecord::make(p, t, mat, r, sn);
        rec.u = rem_euclid(off.dot(u_axis), 1.0f);
        rec.v = rem_euclid(off.dot(v_axis), 1.0f);
        rec.prim_id = 0;
        return true;
    }
    bool bounding_box(Aabb& out) const override {
        const float B = 1e5f, E = 1e-3f;
        float n[3] = {normal.x, normal.y, normal.z}, p[3] = {point.x, point.y, point.z}, lo[3], hi[3];
        for (int k = 0; k < 3; k++) {
            if (std::fabs(n[k]) > 0.999f) { lo[k] = p[k] - E; hi[k] = p[k] + E; } else { lo[k] = -B; hi[k] = B; }
        }
        out = Aabb(Point3(lo[0], lo[1], lo[2]), Point3(hi[0], hi[1], hi[2]));
        return true;
    }    bool photon_bounds(Aabb&) const override { return false; }   // src/plane.rs:102-106
};

// ---- Triangle ------------------------------------------------------------------
struct Triangle : Hittable {
    Point3 v0, v1, v2; Vec3 n0, n1, n2, t0, t1, t2;
    float uv0[2], uv1[2], uv2[2];
    const Material* mat; float area;
    Triangle(const float* v, const float* n, const float* t, const float* uv, const Material* m) : mat(m) {
        v0 = Vec3(v[0], v[1], v[2]); v1 = Vec3(v[3], v[4], v[5]); v2 = Vec3(v[6], v[7], v[8]);
        if (n) { n0 = Vec3(n[0], n[1], n[2]); n1 = Vec3(n[3], n[4], n[5]); n2 = Vec3(n[6], n[7], n[8]); }
        else { Vec3 fn = (v1 - v0).cross(v2 - v0).unit(); n0 = n1 = n2 = fn; }
        if (t) { t0 = Vec3(t[0], t[1], t[2]); t1 = Vec3(t[3], t[4], t[5]); t2 = Vec3(t[6], t[7], t[8]); }
        else { Vec3 ft = (v1 - v0).unit(); t0 = t1 = t2 = ft; }
        if (uv) { uv0[0] = uv[0]; uv0[1] = uv[1]; uv1[0] = uv[2]; uv1[1] = uv[3]; uv2[0] = uv[4]; uv2[1] = uv[5]; }
        else { uv0[0] = 0; uv0[1] = 0; uv1[0] = 1; uv1[1] = 0; uv2[0] = 0.5f; uv2[1] = 1; }
        area = 0.5f * (v1 - v0).cross(v2 - v0).length();
    }
    float light_area() const override { return material_is_emissive(mat) ? area : 0.0f; }
    float pdf_value(const Point3& origin, const Vec3& dir, float) const override {
        Ray r = Ray::make(origin, dir);
        HitRecord rec;
        if (!hit(r, 0.001f, INF, rec)) return 0.0f;
        float dist2 = (rec.p - origin).length_squared();
        float cos_theta = std::fabs(dir.unit().dot(rec.normal));
        if (cos_theta < 1e-8f || area <= 0.0f) return 0.0f;
        return dist2 / (cos_theta * area);
    }
    Vec3 pdf_generate(const Point3& origin, SmallRng& rng, float) const override {
        float r1 = rng.gen();
        float r2 = rng.gen();
        if (r1 + r2 > 1.0f) { r1 = 1.0f - r1; r2 = 1.0f - r2; }
        Point3 p = v0 + r1 * (v1 - v0) + r2 * (v2 - v0);
        return (p - origin).unit();
    }
    bool hit(const Ray& r, float t_min, float t_max, HitRecord& rec) const override {
        Vec3 e1 = v1 - v0, e2 = v2 - v0;
        Vec3 h = r.direction.cross(e2);
        float a = e1.dot(h);
        if (std::fabs(a) < 1e-8f) return false;
        float f = 1.0f / a;
        Vec3 s = r.origin - v0;
        float b1 = f * s.dot(h);
        if (!(b1 >= 0.0f && b1 <= 1.0f)) return false;
        Vec3 qv = s.cross(e1);
        float b2 = f * r.direction.dot(qv);
        if (b2 < 0.0f || b1 + b2 > 1.0f) return false;
        float t = f * e2.dot(qv);
        if (t < t_min || t > t_max) return false;
        float b0 = 1.0f - b1 - b2;
        Point3 p = r.at(t);
        Vec3 sn_raw = b0 * n0 + b1 * n1 + b2 * n2;
        Vec3 sn = sn_raw.length_squared() > 1e-12f ? sn_raw.unit() : e1.cross(e2).unit();
        float u = b0 * uv0[0] + b1 * uv1[0] + b2 * uv2[0];
        float v = b0 * uv0[1] + b1 * uv1[1] + b2 * uv2[1];
        Vec3 raw_t = b0 * t0 + b1 * t1 + b2 * t2;
        raw_t = raw_t - sn * raw_t.dot(sn);
        Vec3 tangent;
        if (raw_t.length_squared() > 1e-8f) tangent = raw_t.unit();
        else { Vec3 a_, b_; vonb(sn, a_, b_); tangent = a_; }
        rec = HitRecord::make(p, t, mat, r, sn);
        rec.u = u; rec.v = v; rec.tangent = tangent;
        return true;
    }
    bool any_hit(const Ray& r, float t_min, float t_max) const override {
        Vec3 e1 = v1 - v0, e2 = v2 - v0;
        Vec3 h = r.direction.cross(e2);
        float a = e1.dot(h);
        if (std::fabs(a) < 1e-8f) return false;
        float f = 1.0f / a;
        Vec3 s = r.origin - v0;
        float b1 = f * s.dot(h);
        if (!(b1 >= 0.0f && b1 <= 1.0f)) return false;
        Vec3 qv = s.cross(e1);
        float b2 = f * r.direction.dot(qv);
        if (b2 < 0.0f || b1 + b2 > 1.0f) return false;
        float t = f * e2.dot(qv);
        return t >= t_min && t <= t_max;
    }
    bool bounding_box(Aabb& out) const override {
        Point3 mn(rmin(rmin(v0.x, v1.x), v2.x), rmin(rmin(v0.y, v1.y), v2.y), rmin(rmin(v0.z, v1.z), v2.z));
        Point3 mx(rmax(rmax(v0.x, v1.x), v2.x), rmax(rmax(v0.y, v1.y), v2.y), rmax(rmax(v0.z, v1.z), v2.z));
        out = Aabb(mn, mx).pad();
        return true;
    }
};

// ---- Convex polyhedron (Diamond) ----------------------------------------------
struct Plane4 { Vec3 n; float d; };

static inline Vec3 micro_tilt(size_t idx, float amp, const Vec3& n) {
    if (amp <= 0.0f) return n;
    uint64_t h = ((uint64_t)idx * 0x9E3779B97F4A7C15ULL) ^ 0xD1B54A32D192ED03ULL;
    float a = ((float)(uint32_t)(h >> 32) / (float)UINT32_MAX) * 2.0f - 1.0f;
    float b = ((float)(uint32_t)h / (float)UINT32_MAX) * 2.0f - 1.0f;
    Vec3 up = std::fabs(n.x) < 0.9f ? Vec3(1, 0, 0) : Vec3(0, 1, 0);
    Vec3 t = n.cross(up).unit();
    Vec3 b2 = n.cross(t);
    return (n + t * (a * amp) + b2 * (b * amp)).unit();
}

// convex_hit (src/diamond.rs:203-253); returns the plane index of the normal.
static inline bool convex_hit(const Plane4* planes, int np, const Ray& r, float t_min, float t_max,
                              float& t_out, Vec3& n_out, int& idx_out) {
    float t_enter = t_min, t_exit = t_max;
    Vec3 enter_normal(0, 1, 0), exit_normal(0, 1, 0);
    int enter_idx = -1, exit_idx = -1;
    bool entry_found = false;
    for (int i = 0; i < np; i++) {
        const Vec3& n = planes[i].n; float d = planes[i].d;
        float nd = n.dot(r.direction);
        float t_face = (d - n.dot(r.origin)) / nd;
        if (std::fabs(nd) < 1e-8f) {
            if (n.dot(r.origin) > d) return false;
            continue;
        }
        if (nd < 0.0f) {
            if (t_face > t_enter) { t_enter = t_face; enter_normal = n; enter_idx = i; entry_found = true; }
        } else {
            if (t_face < t_exit) { t_exit = t_face; exit_normal = n; exit_idx = i; }
        }
        if (t_enter > t_exit) return false;
    }
    if (t_enter > t_exit) return false;
    if (t_enter == t_exit && t_enter < t_max) return false;
    if (entry_found) { t_out = t_enter; n_out = enter_normal; idx_out = enter_idx; return true; }
    if (t_exit > t_max) return false;
    t_out = t_exit; n_out = exit_normal; idx_out = exit_idx;
    return true;
}

struct Convex : Hittable {
    std::vector<Plane4> planes;
    Aabb bbox;
    const Material* mat;
    Point3 center; float crown_h, pavilion_h, girdle_h, polish;
    bool hit(const Ray& r, float t_min, float t_max, HitRecord& rec) const override {
        float t; Vec3 on; int idx;
        if (!convex_hit(planes.data(), (int)planes.size(), r, t_min, t_max, t, on, idx)) return false;
        Point3 p = r.at(t);
        Vec3 rel = p - center;
        float u = (std::atan2(-rel.z, rel.x) + PI) / (2.0f * PI);
        float v = rclamp((rel.y + 0.5f * girdle_h + pavilion_h) / (crown_h + pavilion_h + girdle_h), 0.0f, 1.0f);
        rec = HitRecord::make(p, t, mat, r, on);
        if (polish > 0.0f) {
            size_t facet = 0;
            for (size_t i = 0; i < planes.size(); i++) {
                const Vec3& n = planes[i].n;
                if (std::fabs(n.x - on.x) < 1e-6f && std::fabs(n.y - on.y) < 1e-6f && std::fabs(n.z - on.z) < 1e-6f) { facet = i; break; }
            }
            rec.normal = micro_tilt(facet, polish, rec.normal);
        }
        rec.u = u; rec.v = v;
        rec.prim_id = idx >= 0 ? idx : 0;     // bounding half-space (facet) index
        return true;
    }
    bool any_hit(const Ray& r, float t_min, float t_max) const override {
        float t; Vec3 on; int idx;
        return convex_hit(planes.data(), (int)planes.size(), r, t_min, t_max, t, on, idx);
    }
    bool bounding_box(Aabb& out) const override { out = bbox; return true; }
};

// ---- HittableList ------------------------------------------------------------
struct HittableList : Hittable {
    std::vector<HPtr> objects;
    bool hit(const Ray& r, float t_min, float t_max, HitRecord& rec) const override {
        float closest = t_max;
        bool found = false;
        for (auto& o : objects) {
            HitRecord tmp;
            if (o->hit(r, t_min, closest, tmp)) { closest = tmp.t; rec = tmp; found = true; }
        }
        return found;
    }
    bool any_hit(const Ray& r, float t_min, float t_max) const override {
        for (auto& o : objects) if (o->any_hit(r, t_min, t_max)) return true;
        return false;
    }
    float transmittance(const Ray& r, float t_min, float t_max) const override {
        float tr = 1.0f;
        for (auto& o : objects) {
            tr *= o->transmittance(r, t_min, t_max);
            if (tr <= 0.0f) return 0.0f;
        }
        return tr;
    }
    float light_area() const override {
        float s = 0.0f;
        for (auto& o : objects) s += o->light_area();
        return s;
    }
    bool bounding_box(Aabb& out) const override {
        if (objects.empty()) return false;
        bool have = false;
        for (auto& o : objects) {
            Aabb b;
            if (!o->bounding_box(b)) return false;
            out = have ? Aabb::surrounding(out, b) : b;
            have = true;
        }
        return true;
    }
    // src/hittable.rs:199-211: union of the children that have photon bounds
    bool photon_bounds(Aabb& out) const override {
        bool have = false;
        for (auto& o : objects) {
            Aabb b;
            if (!o->photon_bounds(b)) continue;
            out = have ? Aabb::surrounding(out, b) : b;
            have = true;
        }
        return have;
    }
    float pdf_value(const Point3& origin, const Vec3& dir, float time) const override {
        if (objects.empty()) return 0.0f;
        float weight = 1.0f / (float)objects.size();
        float s = 0.0f;
        for (auto& o : objects) s += weight * o->pdf_value(origin, dir, time);
        return s;
    }
    Vec3 pdf_generate(const Point3& origin, SmallRng& rng, float time) const override {
        if (objects.empty()) return Vec3(1, 0, 0);
        size_t idx = (size_t)rng.next_u32() % objects.size();
        return objects[idx]->pdf_generate(origin, rng, time);
    }
};

// ---- BvhTree -------------------------------------------------------------------
static const int QW = 8;      // QBVH_WIDTH on an AVX2 host (src/bvh.rs:10-20)
static const int QDEPTH = 3;  // log2(QW)
static const int NUM_BUCKETS = 12;

struct QbvhNode {
    float mn[3][QW], mx[3][QW];
    int32_t children[QW];
};

static inline float surface_area(const Aabb& b) {
    Vec3 d = b.max - b.min;
    return 2.0f * (d.x * d.y + d.y * d.z + d.z * d.x);
}

struct BvhTree : Hittable {
    std::vector<QbvhNode> nodes;
    std::vector<HPtr> objects;   // leaf order
    Aabb bbox;
    std::vector<float> light_prefix;
    float light_total = 0.0f;
    bool top = false;            // world tree writes obj_id, nested tree prim_id

    typedef std::pair<std::vector<HPtr>, std::vector<Aabb>> Group;

    static void sah_partition(const std::vector<HPtr>& objs, const std::vector<Aabb>& boxes, Group& L, Group& R) {
        size_t span = objs.size();
        float c_min[3] = {INF, INF, INF}, c_max[3] = {-INF, -INF, -INF};
        for (auto& b : boxes)
            for (int a = 0; a < 3; a++) {
                float c = (b.min[a] + b.max[a]) * 0.5f;
                c_min[a] = rmin(c_min[a], c);
                c_max[a] = rmax(c_max[a], c);
            }
        Aabb node_bbox = boxes[0];
        for (size_t i = 1; i < boxes.size(); i++) node_bbox = Aabb::surrounding(node_bbox, boxes[i]);
        float node_sa = surface_area(node_bbox);
        float best_cost = INF;
        int best_axis = 0, best_split = 0;
        for (int axis = 0; axis < 3; axis++) {
            float extent = c_max[axis] - c_min[axis];
            if (extent < 1e-6f) continue;
            bool bb_have[NUM_BUCKETS] = {false};
            Aabb bb[NUM_BUCKETS];
            uint32_t cnt[NUM_BUCKETS] = {0};
            for (auto& b : boxes) {
                float c = (b.min[axis] + b.max[axis]) * 0.5f;
                float t = (c - c_min[axis]) / extent;
                size_t bi = std::min<uint64_t>(f2usize((float)NUM_BUCKETS * t), NUM_BUCKETS - 1);
                cnt[bi]++;
                bb[bi] = bb_have[bi] ? Aabb::surrounding(bb[bi], b) : b;
                bb_have[bi] = true;
            }
            float left_sa[NUM_BUCKETS]; uint32_t left_count[NUM_BUCKETS];
            bool acc_have = false; Aabb acc; uint32_t acc_n = 0;
            for (int i = 0; i < NUM_BUCKETS; i++) {
                acc_n += cnt[i];
                if (bb_have[i]) { acc = acc_have ? Aabb::surrounding(acc, bb[i]) : bb[i]; acc_have = true; }
                left_count[i] = acc_n;
                left_sa[i] = acc_have ? surface_area(acc) : 0.0f;
            }
            acc_have = false; acc_n = 0;
            for (int split = NUM_BUCKETS - 2; split >= 0; split--) {
                acc_n += cnt[split + 1];
                if (bb_have[split + 1]) { acc = acc_have ? Aabb::surrounding(acc, bb[split + 1]) : bb[split + 1]; acc_have = true; }
                uint32_t nl = left_count[split], nr = acc_n;
                if (nl == 0 || nr == 0) continue;
                float right_sa = acc_have ? surface_area(acc) : 0.0f;
                float cost = 1.0f + (left_sa[split] * (float)nl + right_sa * (float)nr) / node_sa;
                if (cost < best_cost) { best_cost = cost; best_axis = axis; best_split = split; }
            }
        }
        float extent = c_max[best_axis] - c_min[best_axis];
        std::vector<size_t> bid(span);
        for (size_t i = 0; i < span; i++) {
            if (extent < 1e-6f) { bid[i] = 0; continue; }
            float c = (boxes[i].min[best_axis] + boxes[i].max[best_axis]) * 0.5f;
            float t = (c - c_min[best_axis]) / extent;
            bid[i] = std::min<uint64_t>(f2usize((float)NUM_BUCKETS * t), NUM_BUCKETS - 1);
        }
        std::vector<size_t> order(span);
        for (size_t i = 0; i < span; i++) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return bid[a] < bid[b]; });
        size_t mid_off = 0;
        while (mid_off < span && bid[order[mid_off]] <= (size_t)best_split) mid_off++;
        size_t mid = (mid_off == 0 || mid_off == span) ? span / 2 : mid_off;
        for (size_t k = 0; k < span; k++) {
            Group& g = k < mid ? L : R;
            g.first.push_back(objs[order[k]]);
            g.second.push_back(boxes[order[k]]);
        }
    }

    static void split_recursive(const std::vector<HPtr>& objs, const std::vector<Aabb>& boxes, int depth, std::vector<Group>& out) {
        if (depth == QDEPTH || objs.size() <= 1) { out.push_back(Group(objs, boxes)); return; }
        Group L, R;
        sah_partition(objs, boxes, L, R);
        split_recursive(L.first, L.second, depth + 1, out);
        split_recursive(R.first, R.second, depth + 1, out);
    }

    static std::vector<Group> split_n(const std::vector<HPtr>& objs, const std::vector<Aabb>& boxes) {
        std::vector<Group> groups;
        if (objs.size() <= (size_t)QW) {
            for (size_t i = 0; i < objs.size(); i++) groups.push_back(Group({objs[i]}, {boxes[i]}));
            return groups;
        }
        split_recursive(objs, boxes, 0, groups);
        return groups;
    }

    static QbvhNode node_from(const int32_t* children, const Aabb* bbs) {
        QbvhNode n;
        for (int c = 0; c < QW; c++) {
            for (int a = 0; a < 3; a++) { n.mn[a][c] = bbs[c].min[a]; n.mx[a][c] = bbs[c].max[a]; }
            n.children[c] = children[c];
        }
        return n;
    }

    int32_t build(const std::vector<HPtr>& objs, const std::vector<Aabb>& boxes, Aabb& out_bbox) {
        if (objs.size() == 1) {
            int32_t idx = (int32_t)objects.size();
            objects.push_back(objs[0]);
            out_bbox = boxes[0];
            return -(idx + 1);
        }
        std::vector<Group> groups = split_n(objs, boxes);
        size_t my_idx = nodes.size();
        QbvhNode sentinel;
        memset(&sentinel, 0, sizeof sentinel);
        for (int c = 0; c < QW; c++) sentinel.children[c] = INT32_MIN;
        nodes.push_back(sentinel);
        int32_t children[QW]; Aabb bbs[QW];
        for (int c = 0; c < QW; c++) { children[c] = INT32_MIN; bbs[c] = Aabb(); }
        bool have = false; Aabb nb;
        for (size_t i = 0; i < groups.size(); i++) {
            Aabb cb;
            int32_t cd = build(groups[i].first, groups[i].second, cb);
            children[i] = cd; bbs[i] = cb;
            nb = have ? Aabb::surrounding(nb, cb) : cb; have = true;
        }
        nodes[my_idx] = node_from(children, bbs);
        out_bbox = nb;
        return (int32_t)my_idx;
    }

    void from_list(const std::vector<HPtr>& objs) {
        has_photon_bbox = false;
        for (auto& o : objs) {
            Aabb b;
            if (!o->photon_bounds(b)) continue;
            photon_bbox = has_photon_bbox ? Aabb::surrounding(photon_bbox, b) : b;
            has_photon_bbox = true;
        }
        std::vector<Aabb> boxes(objs.size());
        for (size_t i = 0; i < objs.size(); i++) { Aabb b; if (objs[i]->bounding_box(b)) boxes[i] = b; }
        int32_t root = build(objs, boxes, bbox);
        if (root < 0) {
            int32_t ch[QW]; Aabb bbs[QW];
            for (int c = 0; c < QW; c++) { ch[c] = INT32_MIN; bbs[c] = Aabb(); }
            ch[0] = root; bbs[0] = boxes[0];
            nodes.push_back(node_from(ch, bbs));
        }
        float acc = 0.0f;
        for (auto& o : objects) { acc += o->light_area(); light_prefix.push_back(acc); }
        light_total = acc;
    }

    // test_children_avx2 (src/bvh.rs:330-364): minps/maxps return the second
    // operand when either is NaN; the hit compare is ordered (<=, false on NaN).
    static inline float mn_ps(float a, float b) { return a < b ? a : b; }
    static inline float mx_ps(float a, float b) { return a > b ? a : b; }
    static uint32_t test_children(const QbvhNode& node, const float ro[3], const float id[3], float t_min, float t_max, float t_near[QW]) {
        uint32_t mask = 0;
        for (int c = 0; c < QW; c++) {
            float t0 = t_min, t1 = t_max;
            for (int a = 0; a < 3; a++) {
                float da = (node.mn[a][c] - ro[a]) * id[a];
                float db = (node.mx[a][c] - ro[a]) * id[a];
                t0 = mx_ps(t0, mn_ps(da, db));
                t1 = mn_ps(t1, mx_ps(da, db));
            }
            t_near[c] = t0;
            if (t0 <= t1 && node.children[c] != INT32_MIN) mask |= 1u << c;
        }
        return mask;
    }

    bool hit_index(const Ray& r, float t_min, float t_max, HitRecord& best, size_t& best_idx) const {
        if (nodes.empty()) return false;
        std::vector<int32_t> stack; stack.reserve(128);
        stack.push_back(0);
        bool found = false;
        float closest = t_max;
        float ro[3] = {r.origin.x, r.origin.y, r.origin.z};
        float id[3] = {r.inv_dir.x, r.inv_dir.y, r.inv_dir.z};
        while (!stack.empty()) {
            int32_t entry = stack.back(); stack.pop_back();
            if (entry < 0) {
                size_t prim = (size_t)(-entry - 1);
                HitRecord rec;
                if (objects[prim]->hit(r, t_min, closest, rec)) {
                    closest = rec.t; best = rec; best_idx = prim; found = true;
                    if (top) best.obj_id = objects[prim]->tag; else best.prim_id = objects[prim]->tag;
                }
                continue;
            }
            const QbvhNode& node = nodes[(size_t)entry];
            float t_near[QW];
            uint32_t mask = test_children(node, ro, id, t_min, closest, t_near);
            if (!mask) continue;
            // far-to-near (descending t_near), equal keys keep child order
            float ht[QW]; int32_t hc[QW]; int n = 0;
            for (int c = 0; c < QW; c++) if (mask & (1u << c)) { ht[n] = t_near[c]; hc[n] = node.children[c]; n++; }
            for (int i = 1; i < n; i++) {
                float kt = ht[i]; int32_t kc = hc[i]; int j = i - 1;
                while (j >= 0 && total_gt(kt, ht[j])) { ht[j + 1] = ht[j]; hc[j + 1] = hc[j]; j--; }
                ht[j + 1] = kt; hc[j + 1] = kc;
            }
            for (int i = 0; i < n; i++) stack.push_back(hc[i]);
        }
        return found;
    }
    // f32::total_cmp ordering: b.0.total_cmp(&a.0) sorts descending.
    static inline int32_t total_key(float f) {
        int32_t i; memcpy(&i, &f, 4);
        return i ^ (int32_t)(((uint32_t)(i >> 31)) >> 1);
    }
    static inline bool total_gt(float a, float b) { return total_key(a) > total_key(b); }

    bool hit(const Ray& r, float t_min, float t_max, HitRecord& rec) const override {
        size_t idx;
        return hit_index(r, t_min, t_max, rec, idx);
    }
    bool any_hit(const Ray& r, float t_min, float t_max) const override {
        if (nodes.empty()) return false;
        std::vector<int32_t> stack; stack.reserve(128);
        stack.push_back(0);
        float ro[3] = {r.origin.x, r.origin.y, r.origin.z};
        float id[3] = {r.inv_dir.x, r.inv_dir.y, r.inv_dir.z};
        while (!stack.empty()) {
            int32_t entry = stack.back(); stack.pop_back();
            if (entry < 0) {
                if (objects[(size_t)(-entry - 1)]->any_hit(r, t_min, t_max)) return true;
                continue;
            }
            float t_near[QW];
            uint32_t mask = test_children(nodes[(size_t)entry], ro, id, t_min, t_max, t_near);
            for (int c = 0; c < QW; c++) if (mask & (1u << c)) stack.push_back(nodes[(size_t)entry].children[c]);
        }
        return false;
    }
    float transmittance(const Ray& r, float t_min, float t_max) const override {
        if (nodes.empty()) return 1.0f;
        std::vector<int32_t> stack; stack.reserve(128);
        stack.push_back(0);
        float tr = 1.0f;
        float ro[3] = {r.origin.x, r.origin.y, r.origin.z};
        float id[3] = {r.inv_dir.x, r.inv_dir.y, r.inv_dir.z};
        while (!stack.empty()) {
            int32_t entry = stack.back(); stack.pop_back();
            if (entry < 0) {
                tr *= objects[(size_t)(-entry - 1)]->transmittance(r, t_min, t_max);
                if (tr <= 0.0f) return 0.0f;
                continue;
            }
            float t_near[QW];
            uint32_t mask = test_children(nodes[(size_t)entry], ro, id, t_min, t_max, t_near);
            for (int c = 0; c < QW; c++) if (mask & (1u << c)) stack.push_back(nodes[(size_t)entry].children[c]);
        }
        return tr;
    }
    float light_area() const override { return light_total; }
    float pdf_value(const Point3& origin, const Vec3& dir, float time) const override {
        if (light_total <= 0.0f) return 0.0f;
        Ray r = Ray::make(origin, dir);
        HitRecord rec; size_t idx;
        if (!hit_index(r, 0.001f, INF, rec, idx)) return 0.0f;
        float w = idx == 0 ? light_prefix[0] : light_prefix[idx] - light_prefix[idx - 1];
        if (w <= 0.0f) return 0.0f;
        return (w / light_total) * objects[idx]->pdf_value(origin, dir, time);
    }
    Vec3 pdf_generate(const Point3& origin, SmallRng& rng, float time) const override {
        if (light_total <= 0.0f || light_prefix.empty()) return Vec3(1, 0, 0);
        float target = rng.gen() * light_total;
        size_t idx = 0;
        while (idx < light_prefix.size() && light_prefix[idx] < target) idx++;
        idx = std::min(idx, objects.size() - 1);
        return objects[idx]->pdf_generate(origin, rng, time);
    }
    bool bounding_box(Aabb& out) const override { if (nodes.empty()) return false; out = bbox; return true; }
    // src/bvh.rs:206-208, 604: union of the objects' photon bounds, in input order
    bool has_photon_bbox = false;
    Aabb photon_bbox;
    bool photon_bounds(Aabb& out) const override { if (!has_photon_bbox) return false; out = photon_bbox; return true; }
};

// ---- transforms ------------------------------------------------------------------
struct UniformScale : Hittable {
    HPtr object; float factor;
    Ray scaled(const Ray& r) const { float inv = 1.0f / factor; return Ray::scatter_from(inv * r.origin, inv * r.direction, r); }
    bool hit(const Ray& r, float t_min, float t_max, HitRecord& rec) const override {
        if (!object->hit(scaled(r), t_min, t_max, rec)) return false;
        rec.p = factor * rec.p;
        return true;
    }
    bool any_hit(const Ray& r, float t_min, float t_max) const override { return object->any_hit(scaled(r), t_min, t_max); }
    float transmittance(const Ray& r, float t_min, float t_max) const override {
        return std::pow(object->transmittance(scaled(r), t_min, t_max), factor);
    }
    float light_area() const override { return object->light_area() * factor * factor; }
    float pdf_value(const Point3& o, const Vec3& d, float time) const override { return object->pdf_value((1.0f / factor) * o, d, time); }
    Vec3 pdf_generate(const Point3& o, SmallRng& rng, float time) const override { return object->pdf_generate((1.0f / factor) * o, rng, time); }
    bool bounding_box(Aabb& out) const override {
        Aabb b; if (!object->bounding_box(b)) return false;
        out = Aabb(factor * b.min, factor * b.max); return true;
    }
    bool photon_bounds(Aabb& out) const override {      // src/transform.rs:76-80
        Aabb b; if (!object->photon_bounds(b)) return false;
        out = Aabb(factor * b.min, factor * b.max); return true;
    }
};

struct Translate : Hittable {
    HPtr object; Vec3 offset;
    Ray moved(const Ray& r) const { return Ray::scatter_from(r.origin - offset, r.direction, r); }
    bool hit(const Ray& r, float t_min, float t_max, HitRecord& rec) const override {
        if (!object->hit(moved(r), t_min, t_max, rec)) return false;
        rec.p += offset;
        return true;
    }
    bool any_hit(const Ray& r, float t_min, float t_max) const override { return object->any_hit(moved(r), t_min, t_max); }
    float transmittance(const Ray& r, float t_min, float t_max) const override { return object->transmittance(moved(r), t_min, t_max); }
    float light_area() const override { return object->light_area(); }
    float pdf_value(const Point3& o, const Vec3& d, float time) const override { return object->pdf_value(o - offset, d, time); }
    Vec3 pdf_generate(const Point3& o, SmallRng& rng, float time) const override { return object->pdf_generate(o - offset, rng, time); }
    bool bounding_box(Aabb& out) const override {
        Aabb b; if (!object->bounding_box(b)) return false;
        out = Aabb(b.min + offset, b.max + offset); return true;
    }
    bool photon_bounds(Aabb& out) const override {      // src/transform.rs:131-135
        Aabb b; if (!object->photon_bounds(b)) return false;
        out = Aabb(b.min + offset, b.max + offset); return true;
    }
};

struct Rotate : Hittable {
    HPtr object; float m[3][3]; Aabb bbox; bool has_bbox = false;
    Aabb photon_bbox; bool has_photon_bbox = false;      // src/transform.rs:146-148, 200
    Vec3 apply(const Vec3& v) const {
        return Vec3(m[0][0] * v.x + m[0][1] * v.y + m[0][2] * v.z,
                    m[1][0] * v.x + m[1][1] * v.y + m[1][2] * v.z,
                    m[2][0] * v.x + m[2][1] * v.y + m[2][2] * v.z);
    }
    Vec3 apply_t(const Vec3& v) const {
        return Vec3(m[0][0] * v.x + m[1][0] * v.y + m[2][0] * v.z,
                    m[0][1] * v.x + m[1][1] * v.y + m[2][1] * v.z,
                    m[0][2] * v.x + m[1][2] * v.y + m[2][2] * v.z);
    }
    Aabb rotated_aabb(const Aabb& bb) const {
        Point3 mn(INF, INF, INF), mx(-INF, -INF, -INF);
        for (int i = 0; i < 2; i++) for (int j = 0; j < 2; j++) for (int k = 0; k < 2; k++) {
            Vec3 c(i == 0 ? bb.min.x : bb.max.x, j == 0 ? bb.min.y : bb.max.y, k == 0 ? bb.min.z : bb.max.z);
            Vec3 r = apply(c);
            mn.x = rmin(mn.x, r.x); mx.x = rmax(mx.x, r.x);
            mn.y = rmin(mn.y, r.y); mx.y = rmax(mx.y, r.y);
            mn.z = rmin(mn.z, r.z); mx.z = rmax(mx.z, r.z);
        }
        return Aabb(mn, mx);
    }
    void init_bbox() {
        Aabb bb;
        has_bbox = object->bounding_box(bb);
        if (has_bbox) bbox = rotated_aabb(bb);
        has_photon_bbox = object->photon_bounds(bb);
        if (has_photon_bbox) photon_bbox = rotated_aabb(bb);
    }
    Ray rotated(const Ray& r) const { return Ray::scatter_from(apply_t(r.origin), apply_t(r.direction), r); }
    bool hit(const Ray& r, float t_min, float t_max, HitRecord& rec) const override {
        if (!object->hit(rotated(r), t_min, t_max, rec)) return false;
        rec.p = apply(rec.p); rec.normal = apply(rec.normal); rec.tangent = apply(rec.tangent);
        return true;
    }
    bool any_hit(const Ray& r, float t_min, float t_max) const override { return object->any_hit(rotated(r), t_min, t_max); }
    float transmittance(const Ray& r, float t_min, float t_max) const override { return object->transmittance(rotated(r), t_min, t_max); }
    float light_area() const override { return object->light_area(); }
    float pdf_value(const Point3& o, const Vec3& d, float time) const override { return object->pdf_value(apply_t(o), apply_t(d), time); }
    Vec3 pdf_generate(const Point3& o, SmallRng& rng, float time) const override { return apply(object->pdf_generate(apply_t(o), rng, time)); }
    bool bounding_box(Aabb& out) const override { if (!has_bbox) return false; out = bbox; return true; }
    bool photon_bounds(Aabb& out) const override { if (!has_photon_bbox) return false; out = photon_bbox; return true; }
};

}  // namespace rc
