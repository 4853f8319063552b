// Device-side f32 vector math, random streams and small helpers for the
// per-sample radiance kernels.  The library is compiled with
// -ffp-contract=off so every a*b+c rounds twice, as in the unfused f32 code
// the path is specified by (rustracer src/vec3.rs); +,-,*,/,sqrt are IEEE
// correctly rounded on gfx950 in this mode.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define RT_D __device__ __forceinline__
#define RT_HD __host__ __device__ __forceinline__

namespace rt {

static constexpr float PI = 3.14159265358979323846f;
static constexpr float FINF = __builtin_huge_valf();

struct V3 {
    float x, y, z;
};
RT_HD V3 mk(float x, float y, float z) { V3 v; v.x = x; v.y = y; v.z = z; return v; }
RT_HD V3 operator+(V3 a, V3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
RT_HD V3 operator-(V3 a, V3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
RT_HD V3 operator-(V3 a) { return mk(-a.x, -a.y, -a.z); }
RT_HD V3 operator*(V3 a, V3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
RT_HD V3 operator*(V3 a, float t) { return mk(a.x * t, a.y * t, a.z * t); }
RT_HD V3 operator*(float t, V3 a) { return mk(a.x * t, a.y * t, a.z * t); }
RT_HD V3 operator/(V3 a, float t) { return mk(a.x / t, a.y / t, a.z / t); }
RT_HD V3& operator+=(V3& a, V3 b) { a.x += b.x; a.y += b.y; a.z += b.z; return a; }
RT_HD V3& operator-=(V3& a, V3 b) { a.x -= b.x; a.y -= b.y; a.z -= b.z; return a; }
RT_HD V3& operator*=(V3& a, V3 b) { a.x *= b.x; a.y *= b.y; a.z *= b.z; return a; }
RT_HD V3& operator*=(V3& a, float t) { a.x *= t; a.y *= t; a.z *= t; return a; }
RT_HD V3& operator/=(V3& a, float t) { a.x /= t; a.y /= t; a.z /= t; return a; }
RT_HD float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
RT_HD float len2(V3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
RT_HD V3 cross(V3 a, V3 b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
RT_D float vlen(V3 a) { return sqrtf(len2(a)); }
RT_D V3 unit(V3 a) { float l = sqrtf(len2(a)); return mk(a.x / l, a.y / l, a.z / l); }
RT_HD float lum(V3 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }
RT_HD float comp(V3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

// maxNum / minNum: a NaN operand is ignored (same as fmaxf / fminf).
RT_D float fmx(float a, float b) { return fmaxf(a, b); }
RT_D float fmn(float a, float b) { return fminf(a, b); }
RT_D float fclamp(float x, float lo, float hi) { if (x < lo) x = lo; if (x > hi) x = hi; return x; }
// float -> unsigned index, saturating, NaN -> 0
RT_D uint32_t f2u(float x) { if (!(x > 0.0f)) return 0u; if (x >= 4294967296.0f) return 0xffffffffu; return (uint32_t)x; }
RT_D int32_t f2i(float x) {
    if (x != x) return 0;
    if (x <= -2147483648.0f) return INT32_MIN;
    if (x >= 2147483648.0f) return INT32_MAX;
    return (int32_t)x;
}
RT_D float rem1(float a) { float r = fmodf(a, 1.0f); return r < 0.0f ? r + 1.0f : r; }
RT_D float pow5(float a) { float a2 = a * a; float a4 = a2 * a2; return a * a4; }

// Tangent frame of a unit normal (Vec3::onb, src/vec3.rs:37-42).
RT_D void vonb(V3 n, V3& u, V3& v) {
    V3 up = fabsf(n.y) < 0.9f ? mk(0, 1, 0) : mk(1, 0, 0);
    u = unit(cross(n, up));
    v = cross(n, u);
}
RT_D V3 reflect(V3 v, V3 n) { return v - 2.0f * dot(v, n) * n; }
RT_D bool refract(V3 v, V3 n, float eta, V3& out) {
    float c = fmn(dot(-v, n), 1.0f);
    V3 perp = eta * (v + c * n);
    float k = 1.0f - len2(perp);
    if (k < 0.0f) return false;
    out = perp + (-sqrtf(k) * n);
    return true;
}

struct Onb { V3 u, v, w; };
RT_D Onb onb_from_w(V3 n) {
    Onb o;
    o.w = unit(n);
    V3 a = fabsf(o.w.x) > 0.9f ? mk(0, 1, 0) : mk(1, 0, 0);
    o.v = unit(cross(o.w, a));
    o.u = cross(o.w, o.v);
    return o;
}
RT_D V3 onb_local(const Onb& o, V3 a) { return a.x * o.u + a.y * o.v + a.z * o.w; }

// wl: 0 the reference's hero wavelength over [380, 700) nm; 1 one wavelength
// over [380, 780) (the wide CIE table); 2 four hero wavelengths over
// [380, 780), whose spectral tints are applied per wavelength when a
// contribution is made (tint: 0 none yet, 1 CMF colour, 2 fire tint).
struct Ray {
    V3 o, d, inv;
    float time, wavelength;
    bool spectral;
    uint8_t wl, tint;
};
RT_D Ray make_ray(V3 o, V3 d, float time) {
    Ray r; r.o = o; r.d = d; r.inv = mk(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    r.time = time; r.wavelength = 550.0f; r.spectral = false; r.wl = 0; r.tint = 0;
    return r;
}
RT_D Ray scatter_from(V3 o, V3 d, const Ray& src) {
    Ray r = make_ray(o, d, src.time);
    r.wavelength = src.wavelength; r.spectral = src.spectral; r.wl = src.wl; r.tint = src.tint;
    return r;
}
RT_D V3 at(const Ray& r, float t) { return r.o + t * r.d; }

// ---- xoshiro256++ stream, SplitMix64-seeded ---------------------------------
// The reference draws from rand's SmallRng (xoshiro256++, seed_from_u64 via
// SplitMix64).  One stream per (pixel, sample) here; f32 draws use the 24 high
// bits of next_u32 as rand's Standard distribution does.
RT_HD uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
struct Rng {
    uint64_t s0, s1, s2, s3;
};
RT_HD Rng rng_seed(uint64_t state) {
    Rng r;
    // an all-zero state is replaced by seed 0's (a loop, not a recursive
    // call, so the function inlines into every kernel)
    for (int pass = 0; pass < 2; pass++) {
        uint64_t sv[4];
        for (int i = 0; i < 4; i++) {
            state += 0x9e3779b97f4a7c15ULL;
            uint64_t z = state;
            z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
            z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
            sv[i] = z ^ (z >> 31);
        }
        r.s0 = sv[0]; r.s1 = sv[1]; r.s2 = sv[2]; r.s3 = sv[3];
        if ((r.s0 | r.s1 | r.s2 | r.s3) != 0) break;
        state = 0;
    }
    return r;
}
RT_HD uint64_t rng_u64(Rng& r) {
    uint64_t result = rotl64(r.s0 + r.s3, 23) + r.s0;
    uint64_t t = r.s1 << 17;
    r.s2 ^= r.s0; r.s3 ^= r.s1; r.s1 ^= r.s2; r.s0 ^= r.s3; r.s2 ^= t;
    r.s3 = rotl64(r.s3, 45);
    return result;
}
RT_HD uint32_t rng_u32(Rng& r) { return (uint32_t)(rng_u64(r) >> 32); }
RT_HD float rng_f(Rng& r) { return (float)(rng_u32(r) >> 8) * (1.0f / 16777216.0f); }
RT_D float rng_range(Rng& r, float lo, float hi) {
    float scale = hi - lo;
    for (;;) {
        float v12 = __uint_as_float((rng_u32(r) >> 9) | 0x3f800000u);
        float res = (v12 - 1.0f) * scale + lo;
        if (res < hi) return res;
    }
}
RT_D uint32_t rng_range_u32(Rng& r, uint32_t lo, uint32_t hi) {
    uint32_t range = hi - lo;
    if (range == 0) return rng_u32(r);
    uint32_t zone = (range << __builtin_clz(range)) - 1;
    for (;;) {
        uint32_t v = rng_u32(r);
        uint64_t m = (uint64_t)v * (uint64_t)range;
        if ((uint32_t)m <= zone) return lo + (uint32_t)(m >> 32);
    }
}
RT_D V3 random_in_unit_disk(Rng& r) {
    for (;;) {
        float x = rng_range(r, -1.0f, 1.0f);
        float y = rng_range(r, -1.0f, 1.0f);
        V3 p = mk(x, y, 0.0f);
        if (len2(p) < 1.0f) return p;
    }
}

// Stream key of one (pixel, sample): shared with the CPU checker's per-pixel mode.
RT_HD uint64_t pixel_seed(uint32_t pixel, uint32_t sample_idx, uint64_t seed) {
    return ((uint64_t)pixel * 0x9E3779B97F4A7C15ULL) ^ ((uint64_t)sample_idx * 6364136223846793005ULL)
           ^ (seed * 0xD1B54A32D192ED03ULL);
}

}  // namespace rt
