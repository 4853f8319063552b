// ORACLE — test infrastructure only.  Nothing in the product (rustracer_amd/)
// includes, links or runs this code; only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg load liboracle_rtcore.so, and only as a checker.
//
// CPU restatement of rustracer's f32 math layer.  Every operation is written in
// the order the Rust source evaluates it and the file is compiled with
// -ffp-contract=off (Rust never fuses a*b+c), so results are bit-identical to a
// Rust build on the same libm for +,-,*,/,sqrt.
//   Vec3          src/vec3.rs:1-112
//   Ray           src/ray.rs:4-45
//   Aabb          src/aabb.rs:4-29
//   Onb           src/onb.rs:3-21
//   SmallRng      rand 0.8 SmallRng = Xoshiro256PlusPlus seeded by SplitMix64
//                 (third-party crate rand 0.8.x / rand_core 0.6.4, Cargo.lock;
//                 restated from the published xoshiro256++ and SplitMix64
//                 algorithms — parity of the exact streams is unpinned)
#pragma once
#include <cmath>
#include <cstdint>
#include <cfloat>
#include <limits>

namespace rc {

static const float PI = 3.14159265358979323846f;
static const float INF = std::numeric_limits<float>::infinity();

// Rust f32::max / f32::min ignore a NaN operand (IEEE maxNum), like fmaxf.
static inline float rmax(float a, float b) { return std::fmax(a, b); }
static inline float rmin(float a, float b) { return std::fmin(a, b); }
// Rust f32::clamp: NaN passes through.
static inline float rclamp(float x, float lo, float hi) {
    if (x < lo) x = lo;
    if (x > hi) x = hi;
    return x;
}
// Rust `x as usize` / `as i32` from f32: saturating, NaN -> 0.
static inline uint64_t f2usize(float x) {
    if (!(x > 0.0f)) return 0;
    if (x >= 18446744073709551616.0f) return UINT64_MAX;
    return (uint64_t)x;
}
static inline uint32_t f2u32(float x) {
    if (!(x > 0.0f)) return 0;
    if (x >= 4294967296.0f) return UINT32_MAX;
    return (uint32_t)x;
}
static inline int32_t f2i32(float x) {
    if (x != x) return 0;
    if (x <= -2147483648.0f) return INT32_MIN;
    if (x >= 2147483648.0f) return INT32_MAX;
    return (int32_t)x;
}
// Rust f32::rem_euclid
static inline float rem_euclid(float a, float b) {
    float r = std::fmod(a, b);
    return r < 0.0f ? r + std::fabs(b) : r;
}
// llvm.powi expansion for a small constant exponent (repeated squaring).
static inline float powi5(float a) {
    float a2 = a * a;
    float a4 = a2 * a2;
    return a * a4;
}

struct Vec3 {
    float x, y, z;
    Vec3() : x(0), y(0), z(0) {}
    Vec3(float x_, float y_, float z_) : x(x_), y(y_), z(z_) {}
    float length_squared() const { return x * x + y * y + z * z; }
    float length() const { return std::sqrt(length_squared()); }
    float dot(const Vec3& r) const { return x * r.x + y * r.y + z * r.z; }
    Vec3 cross(const Vec3& r) const {
        return Vec3(y * r.z - z * r.y, z * r.x - x * r.z, x * r.y - y * r.x);
    }
    Vec3 unit() const { float l = length(); return Vec3(x / l, y / l, z / l); }
    float luminance() const { return 0.2126f * x + 0.7152f * y + 0.0722f * z; }
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    bool near_zero() const {
        const float S = 1e-6f;
        return std::fabs(x) < S && std::fabs(y) < S && std::fabs(z) < S;
    }
};
typedef Vec3 Color;
typedef Vec3 Point3;

static inline Vec3 operator+(const Vec3& a, const Vec3& b) { return Vec3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline Vec3 operator-(const Vec3& a, const Vec3& b) { return Vec3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline Vec3 operator-(const Vec3& a) { return Vec3(-a.x, -a.y, -a.z); }
static inline Vec3 operator*(const Vec3& a, const Vec3& b) { return Vec3(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline Vec3 operator*(const Vec3& a, float t) { return Vec3(a.x * t, a.y * t, a.z * t); }
static inline Vec3 operator*(float t, const Vec3& a) { return Vec3(a.x * t, a.y * t, a.z * t); }
static inline Vec3 operator/(const Vec3& a, float t) { return Vec3(a.x / t, a.y / t, a.z / t); }
static inline Vec3& operator+=(Vec3& a, const Vec3& b) { a.x += b.x; a.y += b.y; a.z += b.z; return a; }
static inline Vec3& operator*=(Vec3& a, const Vec3& b) { a.x *= b.x; a.y *= b.y; a.z *= b.z; return a; }
static inline Vec3& operator*=(Vec3& a, float t) { a.x *= t; a.y *= t; a.z *= t; return a; }
static inline Vec3& operator/=(Vec3& a, float t) { a.x /= t; a.y /= t; a.z /= t; return a; }

// Vec3::onb (src/vec3.rs:37-42)
static inline void vonb(const Vec3& n, Vec3& u, Vec3& v) {
    Vec3 up = std::fabs(n.y) < 0.9f ? Vec3(0, 1, 0) : Vec3(1, 0, 0);
    u = n.cross(up).unit();
    v = n.cross(u);
}
// src/vec3.rs:49
static inline Vec3 reflect(const Vec3& v, const Vec3& n) { return v - 2.0f * v.dot(n) * n; }
// src/vec3.rs:56-62; returns false for the None case.
static inline bool refract(const Vec3& v, const Vec3& n, float eta, Vec3& out) {
    float cos_theta = rmin((-v).dot(n), 1.0f);
    Vec3 perp = eta * (v + cos_theta * n);
    float k = 1.0f - perp.length_squared();
    if (k < 0.0f) return false;
    out = perp + (-std::sqrt(k) * n);
    return true;
}

struct Ray {
    Point3 origin;
    Vec3 direction;
    Vec3 inv_dir;
    float time;
    float wavelength;
    bool spectral_weighted;
    // wavelength mode (extension): 0 the reference's [380, 700); 1 one
    // wavelength over [380, 780); 2 four hero wavelengths over [380, 780),
    // their tints applied per wavelength at each contribution (tint: 0 none,
    // 1 the CMF colour, 2 the fire tint)
    uint8_t wl = 0, tint = 0;
    Ray() : time(0), wavelength(550.0f), spectral_weighted(false) {}
    static Ray at_time(const Point3& o, const Vec3& d, float t) {
        Ray r; r.origin = o; r.direction = d;
        r.inv_dir = Vec3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
        r.time = t; r.wavelength = 550.0f; r.spectral_weighted = false; r.wl = 0; r.tint = 0;
        return r;
    }
    static Ray make(const Point3& o, const Vec3& d) { return at_time(o, d, 0.0f); }
    static Ray scatter_from(const Point3& o, const Vec3& d, const Ray& src) {
        Ray r; r.origin = o; r.direction = d;
        r.inv_dir = Vec3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
        r.time = src.time; r.wavelength = src.wavelength; r.spectral_weighted = src.spectral_weighted;
        r.wl = src.wl; r.tint = src.tint;
        return r;
    }
    Point3 at(float t) const { return origin + t * direction; }
};

struct Aabb {
    Point3 min, max;
    Aabb() {}
    Aabb(const Point3& a, const Point3& b) : min(a), max(b) {}
    static Aabb surrounding(const Aabb& a, const Aabb& b) {
        return Aabb(Point3(rmin(a.min.x, b.min.x), rmin(a.min.y, b.min.y), rmin(a.min.z, b.min.z)),
                    Point3(rmax(a.max.x, b.max.x), rmax(a.max.y, b.max.y), rmax(a.max.z, b.max.z)));
    }
    Aabb pad() const {
        const float D = 0.0001f;
        float px = max.x - min.x < D ? D : 0.0f;
        float py = max.y - min.y < D ? D : 0.0f;
        float pz = max.z - min.z < D ? D : 0.0f;
        return Aabb(Point3(min.x - px, min.y - py, min.z - pz), Point3(max.x + px, max.y + py, max.z + pz));
    }
};

struct Onb {
    Vec3 u, v, w;
    static Onb from_w(const Vec3& n) {
        Onb o;
        o.w = n.unit();
        Vec3 a = std::fabs(o.w.x) > 0.9f ? Vec3(0, 1, 0) : Vec3(1, 0, 0);
        o.v = o.w.cross(a).unit();
        o.u = o.w.cross(o.v);
        return o;
    }
    Vec3 local(const Vec3& a) const { return a.x * u + a.y * v + a.z * w; }
};

// ---- SmallRng -------------------------------------------------------------
static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

struct SmallRng {
    uint64_t s[4];
    static SmallRng seed_from_u64(uint64_t state) {
        SmallRng r;
        const uint64_t PHI = 0x9e3779b97f4a7c15ULL;
        for (int i = 0; i < 4; i++) {
            state += PHI;
            uint64_t z = state;
            z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
            z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
            z = z ^ (z >> 31);
            r.s[i] = z;
        }
        if ((r.s[0] | r.s[1] | r.s[2] | r.s[3]) == 0) return seed_from_u64(0);
        return r;
    }
    uint64_t next_u64() {
        uint64_t result = rotl(s[0] + s[3], 23) + s[0];
        uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return result;
    }
    uint32_t next_u32() { return (uint32_t)(next_u64() >> 32); }
    // Standard distribution for f32: 24 high bits of next_u32, [0, 1).
    float gen() { return (float)(next_u32() >> 8) * (1.0f / 16777216.0f); }
    // UniformFloat::sample_single for a half-open finite f32 range: draw a
    // [1,2) float from 23 random mantissa bits, shift to [0,1), scale, and
    // redraw on the rare rounding up to `high`.
    float gen_range(float low, float high) {
        float scale = high - low;
        for (;;) {
            uint32_t bits = (next_u32() >> 9) | 0x3f800000u;
            float v12; memcpy_f(&v12, bits);
            float v01 = v12 - 1.0f;
            float res = v01 * scale + low;
            if (res < high) return res;
        }
    }
    // UniformInt<u32>::sample_single for low..high (exclusive).
    uint32_t gen_range_u32(uint32_t low, uint32_t high) {
        uint32_t range = high - 1 - low + 1;
        if (range == 0) return next_u32();
        uint32_t zone = (range << __builtin_clz(range)) - 1;
        for (;;) {
            uint32_t v = next_u32();
            uint64_t m = (uint64_t)v * (uint64_t)range;
            uint32_t hi = (uint32_t)(m >> 32), lo = (uint32_t)m;
            if (lo <= zone) return low + hi;
        }
    }
    static void memcpy_f(float* f, uint32_t b) { union { uint32_t u; float f; } c; c.u = b; *f = c.f; }
};

// Vec3::random_in_unit_disk (src/vec3.rs:80-85)
static inline Vec3 random_in_unit_disk(SmallRng& rng) {
    for (;;) {
        float x = rng.gen_range(-1.0f, 1.0f);
        float y = rng.gen_range(-1.0f, 1.0f);
        Vec3 p(x, y, 0.0f);
        if (p.length_squared() < 1.0f) return p;
    }
}

}  // namespace rc
