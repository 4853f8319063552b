// Per-pixel accumulation of sample passes: the firefly clamp and Welford
// statistics of accumulate_sample (src/main.rs:603-644) and the convergence
// test of mark_converged (src/main.rs:646-677), one pixel per call, in the
// reference's f32 operation order (unfused, component-wise Vec3 division).
#pragma once
#include "dev_math.h"
#include "../../include/rt_api.h"

namespace rt {

// Color::luminance (src/vec3.rs:33)
RT_D float accum_lum(V3 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }

// One sample into one pixel's state (sum a[3], count *n, Welford m2[3]).
// The caller has already skipped converged pixels.
RT_D void accumulate_pixel(float* a, uint32_t* n, float* m2, V3 sc, const rt_accum_opts& o) {
    if (!isfinite(sc.x) || !isfinite(sc.y) || !isfinite(sc.z)) return;
    uint32_t old_n = *n;
    V3 A = mk(a[0], a[1], a[2]);
    V3 old_mean = old_n > 0 ? A / (float)old_n : mk(0, 0, 0);
    float old_lum = accum_lum(old_mean);
    if (o.firefly_clamp > 0.0f && old_n >= o.firefly_min_samples) {
        float l = accum_lum(sc);
        float th = fmaxf(o.firefly_clamp * old_lum, o.firefly_abs_floor);
        if (l > th) sc = sc * (th / l);
    }
    A = A + sc;
    a[0] = A.x; a[1] = A.y; a[2] = A.z;
    *n = old_n + 1;
    V3 new_mean = A / (float)(old_n + 1);
    m2[0] = m2[0] + (sc.x - old_mean.x) * (sc.x - new_mean.x);
    m2[1] = m2[1] + (sc.y - old_mean.y) * (sc.y - new_mean.y);
    m2[2] = m2[2] + (sc.z - old_mean.z) * (sc.z - new_mean.z);
}

// Whether an unconverged pixel with n samples is converged now.
RT_D bool pixel_converged(const float* a, uint32_t n, const float* m2, float threshold) {
    if (n == 0) return false;
    float nf = (float)n;
    float nm1 = (float)(n - 1 > 1 ? n - 1 : 1);
    for (int c = 0; c < 3; c++) {
        float m = a[c] / nf;
        if (m < 1e-4f) continue;
        float variance = m2[c] / nm1;
        float std_err = sqrtf(variance / nf);
        if (std_err / m >= threshold) return false;
    }
    return true;
}

}  // namespace rt

namespace rt {

// ---- resolve: per-pixel mean, lens optics, tone map, sRGB encode (src/output.rs,
// src/main.rs:571-582), in the reference's f32 operation order ------------------
// aces_curve / aces_tonemap (src/output.rs:11-30): Narkowicz's fit between the
// sRGB -> AP1 and AP1 -> sRGB matrices.
RT_D float aces_curve(float x) {
    x = fmaxf(x, 0.0f);
    return (x * (2.51f * x + 0.03f)) / (x * (2.43f * x + 0.59f) + 0.14f);
}
RT_D V3 aces_tonemap(V3 c) {
    float r = c.x * 0.59719f + c.y * 0.35458f + c.z * 0.04823f;
    float g = c.x * 0.07600f + c.y * 0.90834f + c.z * 0.01566f;
    float b = c.x * 0.02840f + c.y * 0.13383f + c.z * 0.83777f;
    r = aces_curve(r); g = aces_curve(g); b = aces_curve(b);
    return mk(r * 1.60475f + g * -0.53108f + b * -0.07367f,
              r * -0.10208f + g * 1.10813f + b * -0.00605f,
              r * -0.00327f + g * -0.07276f + b * 1.07602f);
}
// agx_curve / agx_tonemap (src/output.rs:37-79): Blender AgX with the log2
// encoding over [-12.47393, +4.026069] EV, the degree-6 curve, the "Punchy"
// look (power 1.35, saturation 1.4), then the outset matrix.
RT_D float agx_curve(float x) {
    float x2 = x * x;
    float x4 = x2 * x2;
    return 15.5f * x4 * x2 - 40.14f * x4 * x + 31.96f * x4 - 6.868f * x2 * x + 0.4298f * x2 + 0.1191f * x - 0.00232f;
}
RT_D float agx_log(float v) { return fclamp((log2f(fmaxf(v, 1e-10f)) - -12.47393f) / 16.5f, 0.0f, 1.0f); }
RT_D V3 agx_tonemap(V3 c) {
    float r = c.x * 0.842479062253094f + c.y * 0.0784335999999992f + c.z * 0.0792237451477643f;
    float g = c.x * 0.0423282422610123f + c.y * 0.878468636469772f + c.z * 0.0791661274605434f;
    float b = c.x * 0.0423756549057051f + c.y * 0.0784336f + c.z * 0.879142973793104f;
    r = agx_curve(agx_log(r)); g = agx_curve(agx_log(g)); b = agx_curve(agx_log(b));
    float luma = accum_lum(mk(r, g, b));
    r = luma + 1.4f * (powf(r, 1.35f) - luma);
    g = luma + 1.4f * (powf(g, 1.35f) - luma);
    b = luma + 1.4f * (powf(b, 1.35f) - luma);
    return mk(r * 1.19687900512017f + g * -0.0980208811401368f + b * -0.0990297440797205f,
              r * -0.0528968517574562f + g * 1.15190312990417f + b * -0.0989611768448433f,
              r * -0.0529716355144234f + g * -0.0980434501171241f + b * 1.15107567775604f);
}
// srgb_encode and the 8-bit store of tone_map (src/output.rs:83-85, 103-111).
RT_D float srgb_encode(float x) { return x <= 0.0031308f ? 12.92f * x : 1.055f * powf(x, 1.0f / 2.4f) - 0.055f; }
RT_D uint8_t srgb_byte(float x) {
    float v = fclamp(srgb_encode(fmaxf(x, 0.0f)), 0.0f, 0.999f) * 256.0f;
    if (!(v > 0.0f)) return 0;
    return v >= 255.0f ? (uint8_t)255 : (uint8_t)v;
}

}  // namespace rt
