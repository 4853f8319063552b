// Device-side appearance and integrator of the radiance path: spectral
// weights, the material set, backgrounds (solid / gradient / Preetham /
// equirect env with 2-D CDF importance sampling), the caustic photon gather
// and ray_color.  Semantics follow rustracer src/material.rs, src/spectrum.rs,
// src/renderer.rs:23-585 and src/photon.rs:148-228; every branch draws random
// numbers in the reference's order.
#pragma once
#include "dev_scene.h"
#include "spectral_data.h"
#include "dev_walter.h"

namespace rt {

enum { M_LAMBERTIAN = 1, M_DIELECTRIC = 2, M_SPECTRAL_DIELECTRIC = 3, M_PBR = 4, M_DIFFUSE_LIGHT = 5,
       M_SPECTRAL_METAL = 6, M_ROUGH_DIELECTRIC = 7, M_ROUGH_SPECTRAL = 8, M_BLACKBODY = 9, M_SSS = 10, M_PEARL = 11,
       M_PHASE = 100 };
#define M_KINDS 11
enum { BGK_SOLID = 0, BGK_GRADIENT = 1, BGK_PHYSICAL = 2, BGK_ENV = 3, BGK_ENV_SPLIT = 4 };

// Initialised in place: every kernel translation unit holds its own copy.
__constant__ float c_cie_x[RT_CIE_WIDE_N] = {RT_CIE_X_DATA, RT_CIE_X_TAIL_DATA};
__constant__ float c_cie_y[RT_CIE_WIDE_N] = {RT_CIE_Y_DATA, RT_CIE_Y_TAIL_DATA};
__constant__ float c_cie_z[RT_CIE_WIDE_N] = {RT_CIE_Z_DATA, RT_CIE_Z_TAIL_DATA};
__constant__ float c_metal[6][RT_METAL_N] = {{RT_GOLD_N_DATA}, {RT_GOLD_K_DATA}, {RT_COPPER_N_DATA},
                                             {RT_COPPER_K_DATA}, {RT_SILVER_N_DATA}, {RT_SILVER_K_DATA}};

// ---- spectrum (src/spectrum.rs:64-277) ------------------------------------------
RT_HD void xyz_srgb(float x, float y, float z, float o[3]) {
    o[0] = fmaxf(3.2406f * x - 1.5372f * y - 0.4986f * z, 0.0f);
    o[1] = fmaxf(-0.9689f * x + 1.8758f * y + 0.0415f * z, 0.0f);
    o[2] = fmaxf(0.0557f * x - 0.2040f * y + 1.0570f * z, 0.0f);
}
// wide: the table through 780 nm (the wide wavelength modes), else the
// reference's, which stops (and clamps) at 700 nm
RT_D void cie_at(float lambda, float& x, float& y, float& z, bool wide = false) {
    const int n = wide ? RT_CIE_WIDE_N : RT_CIE_N;
    float t = fclamp((lambda - RT_CIE_START) / RT_CIE_STEP, 0.0f, (float)(n - 1));
    uint32_t i = min(f2u(t), (uint32_t)(n - 2));
    float fr = t - (float)i;
    x = c_cie_x[i] + fr * (c_cie_x[i + 1] - c_cie_x[i]);
    y = c_cie_y[i] + fr * (c_cie_y[i + 1] - c_cie_y[i]);
    z = c_cie_z[i] + fr * (c_cie_z[i + 1] - c_cie_z[i]);
}
RT_D V3 spectral_to_rgb(const DScene& S, float lambda, bool wide = false) {
    float x, y, z; cie_at(lambda, x, y, z, wide);
    float c[3]; xyz_srgb(x, y, z, c);
    const float* nm = wide ? S.norms_w : S.norms;
    return mk(c[0] * nm[0], c[1] * nm[1], c[2] * nm[2]);
}
RT_HD void fire_color(float x, float y, float z, float c[3]) {
    float w = fmaxf(powf(y, 0.5f), 0.08f);
    float rgb[3]; xyz_srgb(x, y, z, rgb);
    c[0] = rgb[0] * w; c[1] = rgb[1] * w; c[2] = rgb[2] * w;
    const float d = 0.5f;   // FIRE_DESAT default
    float l = 0.2126f * c[0] + 0.7152f * c[1] + 0.0722f * c[2];
    for (int k = 0; k < 3; k++) c[k] = c[k] * (1.0f - d) + l * d;
}
RT_D V3 spectral_fire_rgb(const DScene& S, float lambda, bool wide = false) {
    float x, y, z; cie_at(lambda, x, y, z, wide);
    float c[3]; fire_color(x, y, z, c);
    const float* nm = wide ? S.fire_norms_w : S.fire_norms;
    return mk(c[0] * nm[0], c[1] * nm[1], c[2] * nm[2]);
}
RT_D void metal_nk(int variant, float lambda, float& n, float& k) {
    const float* ns = c_metal[variant * 2];
    const float* ks = c_metal[variant * 2 + 1];
    float t = fclamp((lambda - RT_METAL_START) / RT_METAL_STEP, 0.0f, (float)(RT_METAL_N - 1));
    uint32_t i = min(f2u(t), (uint32_t)(RT_METAL_N - 2));
    float f = t - (float)i;
    n = ns[i] + f * (ns[i + 1] - ns[i]);
    k = ks[i] + f * (ks[i + 1] - ks[i]);
}
RT_D float fresnel_conductor(float ci, float n, float k) {
    float cos2 = ci * ci;
    float sin2 = fmx(1.0f - cos2, 0.0f);
    float n2 = n * n, k2 = k * k;
    float t0 = n2 - k2 - sin2;
    float ab = sqrtf(t0 * t0 + 4.0f * n2 * k2);
    float t1 = ab + cos2;
    float a = sqrtf(fmx(0.5f * (ab + t0), 0.0f));
    float t2 = 2.0f * ci * a;
    float rs = (t1 - t2) / fmx(t1 + t2, 1e-12f);
    float t3 = cos2 * ab + sin2 * sin2;
    float t4 = t2 * sin2;
    float rp = rs * (t3 - t4) / fmx(t3 + t4, 1e-12f);
    return fclamp(0.5f * (rp + rs), 0.0f, 1.0f);
}
RT_D float cauchy(float lambda, float b, float c) { float l = lambda * 1e-3f; return b + c / (l * l); }
RT_D float sellmeier(float lambda, const float* B, const float* C) {
    float l = lambda * 1e-3f; float l2 = l * l;
    return sqrtf(1.0f + B[0] * l2 / (l2 - C[0]) + B[1] * l2 / (l2 - C[1]) + B[2] * l2 / (l2 - C[2]));
}

// ---- dielectric / GGX helpers (src/material.rs:104-321) ------------------------------
RT_D float schlick(float c, float ref_idx) {
    float b = (1.0f - ref_idx) / (1.0f + ref_idx);
    float r0 = b * b; float u = 1.0f - c;
    return r0 + (1.0f - r0) * (u * u * u * u * u);
}
RT_D float fresnel_dielectric(float ci, float eta) {
    float s2i = fmx(1.0f - ci * ci, 0.0f);
    float s2t = eta * eta * s2i;
    if (s2t >= 1.0f) return 1.0f;
    float ct = sqrtf(1.0f - s2t);
    float rs = (eta * ci - ct) / (eta * ci + ct);
    float rp = (eta * ct - ci) / (eta * ct + ci);
    return (rs * rs + rp * rp) * 0.5f;
}
RT_D V3 dielectric_boundary(float ior, const Ray& r, const Rec& h, Rng& rng, bool& reflected) {
    float ratio = h.front ? 1.0f / ior : ior;
    V3 u = unit(r.d);
    float c = fmn(dot(-u, h.n), 1.0f);
    float s = sqrtf(1.0f - c * c);
    reflected = ratio * s > 1.0f || fresnel_dielectric(c, ratio) > rng_f(rng);
    if (reflected) return reflect(u, h.n);
    V3 out;
    if (refract(u, h.n, ratio, out)) return out;
    return reflect(u, h.n);
}
RT_D V3 schlick_color(float c, V3 f0) {
    float u = fmx(1.0f - c, 0.0f);
    float t = u * u * u * u * u;
    return f0 + (mk(1, 1, 1) - f0) * t;
}
RT_D float smith_g1(float c, float al) { float a2 = al * al, c2 = c * c; return 2.0f * c / (c + sqrtf(a2 + (1.0f - a2) * c2)); }
RT_D float smith_g1_aniso(float c, float tx, float ty, float ax, float ay) {
    float den = c + sqrtf(c * c + ax * ax * tx * tx + ay * ay * ty * ty);
    return fclamp(2.0f * c / fmx(den, 1e-6f), 0.0f, 1.0f);
}
RT_D float smith_lambda(float c, float al) { float a2 = al * al, c2 = c * c; return (sqrtf(a2 + (1.0f - a2) * c2) - c) / (2.0f * c); }
RT_D float smith_lambda_aniso(float c, float tx, float ty, float ax, float ay) {
    return (sqrtf(c * c + ax * ax * tx * tx + ay * ay * ty * ty) - c) / (2.0f * c);
}
RT_D float ggx_ndf(float ch, float al) { float a2 = al * al; float d = 1.0f + ch * ch * (a2 - 1.0f); return a2 / fmx(PI * d * d, 1e-12f); }
RT_D float ggx_ndf_aniso(V3 h, float ax, float ay) {
    float hx = h.x / ax, hy = h.y / ay;
    float d = hx * hx + hy * hy + h.z * h.z;
    return 1.0f / fmx(PI * ax * ay * d * d, 1e-12f);
}
RT_HD V3 vndf_sample(V3 wo, float ax, float ay, float xi1, float xi2) {
    V3 whr = mk(ax * wo.x, ay * wo.y, wo.z);
    float wl = sqrtf(len2(whr));
    V3 wh = mk(whr.x / wl, whr.y / wl, whr.z / wl);
    float len = sqrtf(wh.x * wh.x + wh.y * wh.y);
    V3 t1 = len > 1e-6f ? mk(-wh.y / len, wh.x / len, 0.0f) : mk(1, 0, 0);
    V3 t2 = cross(wh, t1);
    float a = 1.0f / (1.0f + wh.z);
    float r = sqrtf(xi1);
    float p1, p2;
    if (xi2 < a) { float phi = xi2 / a * PI; p1 = r * cosf(phi); p2 = r * sinf(phi); }
    else { float phi = PI + (xi2 - a) / (1.0f - a) * PI; p1 = r * cosf(phi); p2 = r * sinf(phi) * wh.z; }
    float p3 = sqrtf(fmaxf(1.0f - p1 * p1 - p2 * p2, 0.0f));
    V3 nh = t1 * p1 + t2 * p2 + wh * p3;
    V3 m = mk(ax * nh.x, ay * nh.y, fmaxf(nh.z, 0.0f));
    float ml = sqrtf(len2(m));
    return mk(m.x / ml, m.y / ml, m.z / ml);
}
RT_D float e_ss(const DScene& S, float mu, float al) {
    float x = fclamp(fclamp(mu, 0.0f, 1.0f) * 64.0f - 0.5f, 0.0f, 63.0f);
    float y = fclamp(fclamp(al, 0.0f, 1.0f) * 64.0f - 0.5f, 0.0f, 63.0f);
    uint32_t i0 = f2u(x), j0 = f2u(y);
    uint32_t i1 = min(i0 + 1, 63u), j1 = min(j0 + 1, 63u);
    float fx = x - (float)i0, fy = y - (float)j0;
    float e00 = S.elut[i0 * 64 + j0], e10 = S.elut[i1 * 64 + j0], e01 = S.elut[i0 * 64 + j1], e11 = S.elut[i1 * 64 + j1];
    return (e00 * (1.0f - fx) + e10 * fx) * (1.0f - fy) + (e01 * (1.0f - fx) + e11 * fx) * fy;
}
RT_D float energy_comp(const DScene& S, float mu, float al) { return fmn(1.0f / fmx(e_ss(S, mu, al), 0.05f), 10.0f); }

struct Scatter { V3 att; Ray ray; bool skip_pdf; };

// Texture::value (src/texture.rs:26-56).
RT_D V3 tex_value(const DScene& S, const DMaterial& m, float u, float v, V3 p) {
    V3 a = mk(m.p[0], m.p[1], m.p[2]);
    switch (m.tex) {
        case 1: {
            int32_t x = f2i(floorf(m.p[3] * p.x)), y = f2i(floorf(m.p[3] * p.y)), z = f2i(floorf(m.p[3] * p.z));
            return ((x + y + z) % 2 == 0) ? a : mk(m.p[4], m.p[5], m.p[6]);
        }
        case 2: {
            float uc = fclamp(u, 0.0f, 1.0f), vc = 1.0f - fclamp(v, 0.0f, 1.0f);
            uint32_t x = min(f2u(uc * (float)m.tw), m.tw - 1), y = min(f2u(vc * (float)m.th), m.th - 1);
            const uint8_t* px = S.tex_pix + m.toff + 3 * ((size_t)y * m.tw + x);
            return mk(S.srgb_lut[px[0]], S.srgb_lut[px[1]], S.srgb_lut[px[2]]);
        }
        case 3: {
            float t = 1.0f + sinf(m.p[3] * p.z + 10.0f * perlin_turb(S.grids + m.toff, p, 7));
            return mk(0.5f * t, 0.5f * t, 0.5f * t);
        }
        case 4: {
            float sc = m.p[3];
            float w = fclamp(2.0f * perlin_turb(S.grids + m.toff, mk(p.x * sc, p.y * sc, p.z * sc), 7), 0.0f, 1.0f);
            return a * (1.0f - w) + mk(m.p[4], m.p[5], m.p[6]) * w;
        }
        default: return a;
    }
}

// nacre_color (src/material.rs:1119-1136): the thin-film table (built on the
// host like NACRE_LUT, src/material.rs:1091-1117) at the optical path
// difference, blended toward grey at grazing angles.
RT_D V3 nacre_color(const DScene& S, float cos_theta, float film_ior, float th) {
    float sin_sq = fmx(1.0f - cos_theta * cos_theta, 0.0f);
    float cos_t = sqrtf(fmx(1.0f - sin_sq / (film_ior * film_ior), 0.0f));
    float opd = 2.0f * film_ior * th * cos_t;
    float t = fclamp(opd / 2500.0f, 0.0f, 1.0f) * 255.0f;
    uint32_t lo = f2u(t), hi = min(lo + 1u, 255u);
    V3 a = mk(S.nacre[3 * lo], S.nacre[3 * lo + 1], S.nacre[3 * lo + 2]);
    V3 b = mk(S.nacre[3 * hi], S.nacre[3 * hi + 1], S.nacre[3 * hi + 2]);
    V3 raw = a * (1.0f - (t - (float)lo)) + b * (t - (float)lo);
    float blend = powf(cos_theta, 0.5f);
    return raw * blend + mk(0.5f, 0.5f, 0.5f) * (1.0f - blend);
}
// oil_noise (src/material.rs:1138-1147)
RT_D float oil_noise(V3 p) {
    float a = sinf(p.x * 1.618f + p.z);
    float b = sinf(p.z * 1.618f - p.x);
    float c = sinf(p.x * 0.618f + p.y + p.z);
    return (a + b + c) * (1.0f / 3.0f);
}
// Texture-map fetches of PbrMaterial (src/material.rs:632-671): the 2x2
// bilinear neighbourhood, texel centres at half-integers, edges clamped.
struct Bilin { uint32_t i00, i10, i01, i11; float tx, ty; };
RT_D Bilin bilinear_at(uint32_t off, uint32_t w, uint32_t h, float u, float v) {
    u = fclamp(u, 0.0f, 1.0f);
    v = 1.0f - fclamp(v, 0.0f, 1.0f);
    float fx = fclamp(u * (float)w - 0.5f, 0.0f, (float)(w - 1));
    float fy = fclamp(v * (float)h - 0.5f, 0.0f, (float)(h - 1));
    uint32_t x0 = f2u(fx), y0 = f2u(fy);
    uint32_t x1 = min(x0 + 1u, w - 1u), y1 = min(y0 + 1u, h - 1u);
    Bilin b;
    b.i00 = off + 3u * (y0 * w + x0); b.i10 = off + 3u * (y0 * w + x1);
    b.i01 = off + 3u * (y1 * w + x0); b.i11 = off + 3u * (y1 * w + x1);
    b.tx = fx - (float)x0; b.ty = fy - (float)y0;
    return b;
}
RT_D V3 texel_srgb(const DScene& S, uint32_t i) { return mk(S.srgb_lut[S.tex_pix[i]], S.srgb_lut[S.tex_pix[i + 1]], S.srgb_lut[S.tex_pix[i + 2]]); }
RT_D V3 texel_rgb(const DScene& S, uint32_t i) {
    return mk((float)S.tex_pix[i] / 255.0f, (float)S.tex_pix[i + 1] / 255.0f, (float)S.tex_pix[i + 2] / 255.0f);
}
RT_D V3 sample_srgb(const DScene& S, const DMaterial& m, int k, float u, float v) {
    Bilin b = bilinear_at(m.map_off[k], m.map_w[k], m.map_h[k], u, v);
    V3 c0 = texel_srgb(S, b.i00) * (1.0f - b.tx) + texel_srgb(S, b.i10) * b.tx;
    V3 c1 = texel_srgb(S, b.i01) * (1.0f - b.tx) + texel_srgb(S, b.i11) * b.tx;
    return c0 * (1.0f - b.ty) + c1 * b.ty;
}
RT_D V3 sample_rgb(const DScene& S, const DMaterial& m, int k, float u, float v) {
    Bilin b = bilinear_at(m.map_off[k], m.map_w[k], m.map_h[k], u, v);
    V3 c0 = texel_rgb(S, b.i00) * (1.0f - b.tx) + texel_rgb(S, b.i10) * b.tx;
    V3 c1 = texel_rgb(S, b.i01) * (1.0f - b.tx) + texel_rgb(S, b.i11) * b.tx;
    return c0 * (1.0f - b.ty) + c1 * b.ty;
}
RT_D float sample_linear(const DScene& S, const DMaterial& m, int k, float u, float v) {
    Bilin b = bilinear_at(m.map_off[k], m.map_w[k], m.map_h[k], u, v);
    float v0 = (float)S.tex_pix[b.i00] / 255.0f * (1.0f - b.tx) + (float)S.tex_pix[b.i10] / 255.0f * b.tx;
    float v1 = (float)S.tex_pix[b.i01] / 255.0f * (1.0f - b.tx) + (float)S.tex_pix[b.i11] / 255.0f * b.tx;
    return v0 * (1.0f - b.ty) + v1 * b.ty;
}

// PbrMaterial parameter block: p0-2 albedo, p3 roughness, p4 metallic,
// p5 anisotropy, p6 angle, p7 clearcoat, p8 coat roughness, p9-11 emission,
// p12 strength, p13 sheen, p14 sheen_tint, p15 film thickness, p16 film ior;
// maps (DMaterial::map_*, width 0 = none): 0 albedo, 1 roughness,
// 2 metallic, 3 ambient occlusion, 4 tangent-space normal.
enum { MAP_ALBEDO = 0, MAP_ROUGH = 1, MAP_METAL = 2, MAP_AO = 3, MAP_NORMAL = 4 };
struct PbrP { V3 n, albedo; float rough, metal; };
// PbrMaterial::tbn_normal (src/material.rs:755-769): normal maps need the
// hit's tangent (meshes)
RT_D V3 pbr_normal(const DScene& S, const DMaterial& m, const Rec& h) {
    if (m.map_w[MAP_NORMAL] && len2(h.tangent) > 1e-6f) {
        V3 raw = sample_rgb(S, m, MAP_NORMAL, h.u, h.v);
        V3 ts = mk(raw.x * 2.0f - 1.0f, raw.y * 2.0f - 1.0f, raw.z * 2.0f - 1.0f);
        V3 ng = h.n, t = h.tangent;
        V3 b = cross(ng, t);
        V3 pert = t * ts.x + b * ts.y + ng * ts.z;
        return len2(pert) > 1e-12f ? unit(pert) : h.n;
    }
    return h.n;
}
RT_D V3 pbr_albedo(const DScene& S, const DMaterial& m, float u, float v) {
    V3 c = m.map_w[MAP_ALBEDO] ? sample_srgb(S, m, MAP_ALBEDO, u, v) : mk(m.p[0], m.p[1], m.p[2]);
    return m.map_w[MAP_AO] ? c * sample_linear(S, m, MAP_AO, u, v) : c;
}
// PbrMaterial::params (src/material.rs:771-782)
RT_D PbrP pbr_params(const DScene& S, const DMaterial& m, const Rec& h) {
    PbrP q;
    q.n = pbr_normal(S, m, h);
    q.albedo = pbr_albedo(S, m, h.u, h.v);
    q.rough = m.map_w[MAP_ROUGH] ? sample_linear(S, m, MAP_ROUGH, h.u, h.v) : m.p[3];
    q.metal = m.map_w[MAP_METAL] ? sample_linear(S, m, MAP_METAL, h.u, h.v) : m.p[4];
    return q;
}
RT_D V3 pbr_film(const DScene& S, const DMaterial& m, float cos_h) {
    return m.p[15] > 0.0f ? nacre_color(S, cos_h, m.p[16], m.p[15]) : mk(1, 1, 1);
}
RT_D void pbr_frame(const DMaterial& m, V3 n, V3& t, V3& b) {
    V3 t0, b0; vonb(n, t0, b0);
    float ca = cosf(m.p[6]), sa = sinf(m.p[6]);
    t = t0 * ca + b0 * sa;
    b = b0 * ca - t0 * sa;
}
RT_D float pbr_aspect(const DMaterial& m) { return sqrtf(fmx(1.0f - 0.9f * fclamp(m.p[5], 0.0f, 1.0f), 0.001f)); }

RT_D bool pbr_scatter(const DScene& S, const DMaterial& m, const PbrP& q, const Ray& r, const Rec& h, Rng& rng, Scatter& sr) {
    V3 n = q.n;
    V3 wo = unit(-r.d);
    float cos_o = dot(wo, n);
    if (cos_o <= 0.0f) return false;
    V3 albedo = q.albedo;
    float rough = q.rough, metallic = q.metal, clearcoat = m.p[7];
    float alpha = fmx(rough * rough, 1e-4f);
    V3 f0 = mk(0.04f, 0.04f, 0.04f) * (1.0f - metallic) + albedo * metallic;
    float p_coat = clearcoat > 1e-3f ? fmx(schlick(cos_o, 1.0f / 1.5f) * clearcoat, 0.12f * clearcoat) : 0.0f;
    V3 fa = schlick_color(cos_o, f0);
    float p_spec = metallic > 0.999f ? 1.0f : fclamp(lum(fa), 0.04f, 0.9f);
    if (rng_f(rng) < p_coat) {
        float ac = fmx(m.p[8] * m.p[8], 1e-4f);
        V3 t, b; vonb(n, t, b);
        float xi1 = rng_f(rng), xi2 = rng_f(rng);
        V3 wo_ts = mk(dot(wo, t), dot(wo, b), cos_o);
        V3 mts = vndf_sample(wo_ts, ac, ac, xi1, xi2);
        V3 mm = unit(t * mts.x + b * mts.y + n * mts.z);
        float voh = fmx(dot(wo, mm), 0.0f);
        V3 wi = unit(2.0f * voh * mm - wo);
        float cos_i = dot(wi, n);
        if (cos_i <= 0.0f) return false;
        float fc = schlick(voh, 1.0f / 1.5f) * clearcoat;
        float lo = smith_lambda(cos_o, ac), li = smith_lambda(cos_i, ac);
        float w = ((1.0f + lo) / (1.0f + lo + li)) * energy_comp(S, cos_o, ac);
        sr.att = pbr_film(S, m, voh) * (fc * w / p_coat);
        sr.ray = scatter_from(h.p, wi, r); sr.skip_pdf = true;
        return true;
    }
    if (rng_f(rng) < p_spec) {
        float xi1 = rng_f(rng), xi2 = rng_f(rng);
        float asp = pbr_aspect(m);
        float ax = alpha / asp, ay = alpha * asp;
        V3 t, b; pbr_frame(m, n, t, b);
        V3 wo_ts = mk(dot(wo, t), dot(wo, b), cos_o);
        V3 mts = vndf_sample(wo_ts, ax, ay, xi1, xi2);
        V3 mm = unit(t * mts.x + b * mts.y + n * mts.z);
        float voh = fmx(dot(wo, mm), 0.0f);
        V3 wi = unit(2.0f * voh * mm - wo);
        float cos_i = dot(wi, n);
        if (cos_i <= 0.0f) return false;
        V3 wi_ts = mk(dot(wi, t), dot(wi, b), cos_i);
        float lo = smith_lambda_aniso(cos_o, wo_ts.x, wo_ts.y, ax, ay);
        float li = smith_lambda_aniso(cos_i, wi_ts.x, wi_ts.y, ax, ay);
        float wg = ((1.0f + lo) / (1.0f + lo + li)) * energy_comp(S, cos_o, alpha);
        V3 f = schlick_color(voh, f0);
        sr.att = f * wg / (p_spec * (1.0f - p_coat));
        sr.ray = scatter_from(h.p, wi, r); sr.skip_pdf = true;
        return true;
    }
    V3 sheen_attn = mk(0, 0, 0);
    float sheen = m.p[13];
    if (sheen > 0.0f && metallic < 0.999f) {
        float l = lum(albedo);
        V3 ct = l > 1e-6f ? albedo / l : mk(1, 1, 1);
        V3 sc = mk(1, 1, 1) * (1.0f - m.p[14]) + ct * m.p[14];
        float fh = pow5(1.0f - cos_o);
        sheen_attn = sc * (sheen * (1.0f - metallic) * fh * PI);
    }
    sr.att = (albedo * (1.0f - metallic) + sheen_attn) / ((1.0f - p_spec) * (1.0f - p_coat));
    sr.ray = scatter_from(h.p, h.n, r); sr.skip_pdf = false;
    return true;
}

RT_D V3 pbr_brdf_cos(const DScene& S, const DMaterial& m, const PbrP& q, const Ray& r, const Rec& h, V3 wi) {
    V3 n = q.n, wo = unit(-r.d);
    float cos_o = dot(wo, n), cos_i = dot(wi, n);
    if (cos_o <= 0.0f || cos_i <= 0.0f) return mk(0, 0, 0);
    V3 hv = unit(wo + wi);
    float woh = fmx(dot(wo, hv), 0.0f), hn = fmx(dot(hv, n), 0.0f);
    V3 albedo = q.albedo;
    float metallic = q.metal, clearcoat = m.p[7];
    V3 f0 = mk(0.04f, 0.04f, 0.04f) * (1.0f - metallic) + albedo * metallic;
    V3 res = mk(0, 0, 0);
    if (clearcoat > 1e-3f) {
        float ac = fmx(m.p[8] * m.p[8], 1e-4f);
        float d = ggx_ndf(hn, ac);
        float g2 = 1.0f / (1.0f + smith_lambda(cos_o, ac) + smith_lambda(cos_i, ac));
        float fc = schlick(woh, 1.0f / 1.5f) * clearcoat;
        res += pbr_film(S, m, woh) * (d * fc * g2 / (4.0f * cos_o) * energy_comp(S, cos_o, ac));
    }
    float alpha = fmx(q.rough * q.rough, 1e-4f);
    float asp = pbr_aspect(m); float ax = alpha / asp, ay = alpha * asp;
    V3 t, b; pbr_frame(m, n, t, b);
    V3 hts = mk(dot(hv, t), dot(hv, b), hn), wots = mk(dot(wo, t), dot(wo, b), cos_o), wits = mk(dot(wi, t), dot(wi, b), cos_i);
    float d = ggx_ndf_aniso(hts, ax, ay);
    float g2 = 1.0f / (1.0f + smith_lambda_aniso(cos_o, wots.x, wots.y, ax, ay) + smith_lambda_aniso(cos_i, wits.x, wits.y, ax, ay));
    V3 f = schlick_color(woh, f0);
    res += f * (d * g2 / (4.0f * cos_o) * energy_comp(S, cos_o, alpha));
    return res;
}

RT_D float pbr_spec_pdf(const DMaterial& m, const PbrP& q, const Ray& r, const Rec& h, V3 wi) {
    V3 n = q.n, wo = unit(-r.d);
    float cos_o = dot(wo, n), cos_i = dot(wi, n);
    if (cos_o <= 0.0f || cos_i <= 0.0f) return 0.0f;
    V3 hv = unit(wo + wi);
    float hn = fmx(dot(hv, n), 0.0f);
    V3 albedo = q.albedo;
    float metallic = q.metal, clearcoat = m.p[7];
    V3 f0 = mk(0.04f, 0.04f, 0.04f) * (1.0f - metallic) + albedo * metallic;
    float p_coat = clearcoat > 1e-3f ? fmx(schlick(cos_o, 1.0f / 1.5f) * clearcoat, 0.12f * clearcoat) : 0.0f;
    V3 fa = schlick_color(cos_o, f0);
    float p_spec = metallic > 0.999f ? 1.0f : fclamp(lum(fa), 0.04f, 0.9f);
    float pdf = 0.0f;
    if (clearcoat > 1e-3f) {
        float ac = fmx(m.p[8] * m.p[8], 1e-4f);
        pdf += p_coat * ggx_ndf(hn, ac) * smith_g1(cos_o, ac) / (4.0f * cos_o);
    }
    if ((1.0f - p_coat) * p_spec > 1e-6f) {
        float alpha = fmx(q.rough * q.rough, 1e-4f);
        float asp = pbr_aspect(m); float ax = alpha / asp, ay = alpha * asp;
        V3 t, b; pbr_frame(m, n, t, b);
        V3 hts = mk(dot(hv, t), dot(hv, b), hn), wots = mk(dot(wo, t), dot(wo, b), cos_o);
        pdf += (1.0f - p_coat) * p_spec * ggx_ndf_aniso(hts, ax, ay) * smith_g1_aniso(cos_o, wots.x, wots.y, ax, ay) / (4.0f * cos_o);
    }
    return pdf;
}

RT_D bool smetal_scatter(const DScene& S, const DMaterial& m, const Ray& r, const Rec& h, Rng& rng, Scatter& sr) {
    V3 n = h.n, wo = unit(-r.d);
    float cos_o = dot(wo, n);
    if (cos_o <= 0.0f) return false;
    float rough = m.p[1];
    V3 dir; float cos_h;
    if (rough == 0.0f) { dir = reflect(unit(r.d), n); cos_h = cos_o; }
    else {
        float alpha = fmx(rough * rough, 1e-4f);
        V3 t, b; vonb(n, t, b);
        V3 wo_ts = mk(dot(wo, t), dot(wo, b), cos_o);
        float xi1 = rng_f(rng), xi2 = rng_f(rng);
        V3 mts = vndf_sample(wo_ts, alpha, alpha, xi1, xi2);
        V3 mm = unit(t * mts.x + b * mts.y + n * mts.z);
        float voh = fmx(dot(wo, mm), 0.0f);
        V3 wi = unit(2.0f * voh * mm - wo);
        if (dot(wi, n) <= 0.0f) return false;
        dir = wi; cos_h = voh;
    }
    if (dot(dir, n) <= 0.0f) return false;
    float in, ik; metal_nk((int)m.p[0], r.wavelength, in, ik);
    float fr = fresnel_conductor(cos_h, in, ik);
    float w = 1.0f;
    if (rough != 0.0f) {
        float alpha = fmx(rough * rough, 1e-4f);
        float lo = smith_lambda(cos_o, alpha), li = smith_lambda(fmx(dot(dir, n), 0.0f), alpha);
        w = ((1.0f + lo) / (1.0f + lo + li)) * energy_comp(S, cos_o, alpha);
    }
    float fw = fr * w;
    // four hero wavelengths: the tint is applied per wavelength when a contribution is made
    sr.att = (r.spectral || r.wl == 2) ? mk(fw, fw, fw) : spectral_to_rgb(S, r.wavelength, r.wl != 0) * fw;
    sr.ray = scatter_from(h.p, dir, r);
    if (!r.spectral) { sr.ray.spectral = true; sr.ray.tint = 1; }
    sr.skip_pdf = true;
    return true;
}
RT_D V3 smetal_brdf_cos(const DScene& S, const DMaterial& m, const Ray& r, const Rec& h, V3 wi) {
    float rough = m.p[1];
    if (rough == 0.0f) return mk(0, 0, 0);
    V3 n = h.n, wo = unit(-r.d);
    float cos_o = dot(wo, n), cos_i = dot(wi, n);
    if (cos_o <= 0.0f || cos_i <= 0.0f) return mk(0, 0, 0);
    V3 hv = unit(wo + wi);
    float woh = fmx(dot(wo, hv), 0.0f), hn = fmx(dot(hv, n), 0.0f);
    float alpha = fmx(rough * rough, 1e-4f);
    float d = ggx_ndf(hn, alpha);
    float g2 = 1.0f / (1.0f + smith_lambda(cos_o, alpha) + smith_lambda(cos_i, alpha));
    float in, ik; metal_nk((int)m.p[0], r.wavelength, in, ik);
    float fr = fresnel_conductor(woh, in, ik);
    float val = d * fr * g2 / (4.0f * cos_o) * energy_comp(S, cos_o, alpha);
    return (r.spectral || r.wl == 2) ? mk(val, val, val) : spectral_to_rgb(S, r.wavelength, r.wl != 0) * val;
}
RT_D float smetal_spec_pdf(const DMaterial& m, const Ray& r, const Rec& h, V3 wi) {
    float rough = m.p[1];
    if (rough == 0.0f) return 0.0f;
    V3 n = h.n, wo = unit(-r.d);
    float cos_o = dot(wo, n);
    if (cos_o <= 0.0f || dot(wi, n) <= 0.0f) return 0.0f;
    V3 hv = unit(wo + wi);
    float hn = fmx(dot(hv, n), 0.0f);
    float alpha = fmx(rough * rough, 1e-4f);
    return ggx_ndf(hn, alpha) * smith_g1(cos_o, alpha) / (4.0f * cos_o);
}

// Rough GGX dielectric (extension): VNDF microfacet, exact Fresnel, weight G1(wi).
// eta = IOR on the incident side over the other side; returns the new
// direction and whether it reflected.
RT_D bool rough_dielectric_core(float alpha, float eta, const Ray& r, const Rec& h, Rng& rng, V3& wi, bool& refl, float& g1) {
    V3 n = h.n, wo = unit(-r.d);
    float cos_o = dot(wo, n);
    if (cos_o <= 0.0f) return false;
    V3 t, b; vonb(n, t, b);
    float xi1 = rng_f(rng), xi2 = rng_f(rng);
    V3 wo_ts = mk(dot(wo, t), dot(wo, b), cos_o);
    V3 mts = vndf_sample(wo_ts, alpha, alpha, xi1, xi2);
    V3 mm = unit(t * mts.x + b * mts.y + n * mts.z);
    float com = fmn(dot(wo, mm), 1.0f);
    float fr = fresnel_dielectric(com, eta);
    refl = fr > rng_f(rng);
    if (!refl) { V3 out; if (refract(-wo, mm, eta, out)) wi = unit(out); else refl = true; }
    if (refl) wi = unit(reflect(-wo, mm));
    float ci = dot(wi, n);
    if (refl ? ci <= 0.0f : ci >= 0.0f) return false;
    g1 = smith_g1(fabsf(ci), alpha);
    return true;
}
RT_D bool rough_dielectric_scatter(const DMaterial& m, const Ray& r, const Rec& h, Rng& rng, Scatter& sr) {
    float alpha = fmx(m.p[1] * m.p[1], 1e-4f);
    float eta = h.front ? 1.0f / m.p[0] : m.p[0];
    V3 wi; bool refl; float g1;
    if (!rough_dielectric_core(alpha, eta, r, h, rng, wi, refl, g1)) return false;
    sr.att = mk(g1, g1, g1);
    sr.ray = scatter_from(h.p, wi, r); sr.skip_pdf = true;
    return true;
}

// Rough dispersive dielectric (extension): n(l) = n_d + s(l) - s(589.3 nm)
// with s the Sellmeier law of p1-6; GGX alpha per primitive group; the fire
// tint of the spectral dielectric on the first refraction.
static constexpr float RSD_LAMBDA_D = 589.3f;
RT_HD int rsd_alpha_field(int prim) { return prim <= 0 ? 7 : 8 + (prim - 1) % 16; }
RT_D float rsd_ior(const DMaterial& m, float lambda) {
    return m.p[0] + (sellmeier(lambda, m.p + 1, m.p + 4) - sellmeier(RSD_LAMBDA_D, m.p + 1, m.p + 4));
}
RT_D float rsd_alpha(const DMaterial& m, const Rec& h) { return fclamp(m.p[rsd_alpha_field(h.prim_id)], 1e-4f, 1.0f); }
RT_D bool rsd_scatter(const DScene& S, const DMaterial& m, const Ray& r, const Rec& h, Rng& rng, Scatter& sr) {
    float ior = rsd_ior(m, r.wavelength);
    float eta = h.front ? 1.0f / ior : ior;
    V3 wi; bool refl; float g1;
    if (!rough_dielectric_core(rsd_alpha(m, h), eta, r, h, rng, wi, refl, g1)) return false;
    V3 att = mk(g1, g1, g1);
    bool tinted = false;
    if (!r.spectral && !refl) { if (r.wl != 2) att *= spectral_fire_rgb(S, r.wavelength, r.wl != 0); tinted = true; }
    sr.ray = scatter_from(h.p, wi, r);
    if (tinted) { sr.ray.spectral = true; sr.ray.tint = 2; }
    sr.att = att; sr.skip_pdf = true;
    return true;
}

RT_D V3 hg_sample(V3 vin, float g, Rng& rng) {
    float xi1 = rng_f(rng), xi2 = rng_f(rng);
    float ct;
    if (fabsf(g) < 1e-3f) ct = 1.0f - 2.0f * xi1;
    else {
        float s = (1.0f - g * g) / fmx(1.0f - g + 2.0f * g * xi1, 1e-12f);
        ct = (1.0f + g * g - s * s) / (2.0f * g);
    }
    float st = sqrtf(fmx(1.0f - ct * ct, 0.0f));
    float phi = 2.0f * PI * xi2;
    V3 w = vin;
    V3 helper = fabsf(w.x) > 0.9f ? mk(0, 1, 0) : mk(1, 0, 0);
    V3 u = unit(cross(w, helper));
    V3 v = cross(w, u);
    return u * (st * cosf(phi)) + v * (st * sinf(phi)) + w * ct;
}

// SSSMaterial::scatter (src/material.rs:594-628): inside, an exponential
// free flight before the exit surface scatters with HG and Beer-Lambert up to
// that point; otherwise the dielectric boundary with absorption over the
// traversed segment (none on entry).
RT_D V3 beer_lambert(const float* sa, float t) { return mk(expf(-sa[0] * t), expf(-sa[1] * t), expf(-sa[2] * t)); }
RT_D bool sss_scatter(const DMaterial& m, const Ray& r, const Rec& h, Rng& rng, Scatter& sr) {
    V3 u = unit(r.d);
    float density = m.p[7];
    sr.skip_pdf = true;
    if (!h.front && density > 0.0f) {
        float path_length = vlen(h.p - r.o);
        float t_scat = -logf(fmx(rng_f(rng), 1e-9f)) / density;
        if (t_scat < path_length) {
            V3 ps = r.o + u * t_scat;
            V3 nd = hg_sample(u, m.p[8], rng);
            sr.att = mk(m.p[0], m.p[1], m.p[2]) * beer_lambert(m.p + 3, t_scat);
            sr.ray = scatter_from(ps, nd, r);
            return true;
        }
        bool refl;
        V3 d = dielectric_boundary(m.p[6], r, h, rng, refl);
        sr.att = beer_lambert(m.p + 3, path_length);
        sr.ray = scatter_from(h.p, d, r);
        return true;
    }
    bool refl;
    V3 d = dielectric_boundary(m.p[6], r, h, rng, refl);
    sr.att = mk(1, 1, 1);
    sr.ray = scatter_from(h.p, d, r);
    return true;
}

// Material kinds whose scatter is a delta / sampled-lobe event (skip_pdf):
// the wavefront's specular shade class.  MASK (a set of 1 << kind) lets a
// caller that knows the hit's class compile only those materials.
#define RT_MASK_ALL 0xffffffffu
#define RT_MASK_SPECULAR ((1u << M_DIELECTRIC) | (1u << M_SPECTRAL_DIELECTRIC) | (1u << M_ROUGH_DIELECTRIC) | \
                          (1u << M_ROUGH_SPECTRAL) | (1u << M_SSS))
#define RT_IN(MASK, K) (((MASK) >> (K)) & 1u)

// PearlMaterial (src/material.rs:1166-1285): p0-2 base colour, p3 ior,
// p4 film thickness, p5 orient strength, p6 film scale, p7 luster roughness.
RT_D float pearl_thickness(const DMaterial& m, V3 p) { return m.p[4] + oil_noise(p * m.p[6]) * 150.0f; }
RT_D bool pearl_scatter(const DScene& S, const DMaterial& m, const Ray& r, const Rec& h, Rng& rng, Scatter& sr) {
    V3 n = h.n, wo = unit(-r.d);
    float cos_theta = fclamp(dot(wo, n), 0.0f, 1.0f);
    float th = pearl_thickness(m, h.p);
    float ior = m.p[3];
    float f = fmx(schlick(cos_theta, 1.0f / ior), 0.10f);
    float cos_il = S.pearl_sun_on ? fclamp(dot(n, mk(S.pearl_sun[0], S.pearl_sun[1], S.pearl_sun[2])), 0.0f, 1.0f) : cos_theta;
    if (rng_f(rng) < f) {
        V3 t, b; vonb(n, t, b);
        float al = fmx(m.p[7] * m.p[7], 1e-4f);
        float xi1 = rng_f(rng), xi2 = rng_f(rng);
        V3 wo_ts = mk(dot(wo, t), dot(wo, b), cos_theta);
        V3 mts = vndf_sample(wo_ts, al, al, xi1, xi2);
        V3 mm = unit(t * mts.x + b * mts.y + n * mts.z);
        float voh = fmx(dot(wo, mm), 0.0f);
        V3 wi = unit(2.0f * voh * mm - wo);
        float cos_i = dot(wi, n);
        if (cos_i <= 0.0f) return false;
        float fh = schlick(voh, 1.0f / ior);
        float g1 = smith_g1(cos_i, al);
        sr.att = nacre_color(S, voh, ior, th) * (fh * g1 / f);
        sr.ray = scatter_from(h.p, wi, r); sr.skip_pdf = true;
        return true;
    }
    V3 orient = nacre_color(S, cos_il, ior, th);
    float s = m.p[5];
    V3 tinted = mk(m.p[0], m.p[1], m.p[2]) * (orient * s + mk(1, 1, 1) * (1.0f - s));
    sr.att = tinted * (1.0f / (1.0f - f));
    sr.ray = scatter_from(h.p, n, r); sr.skip_pdf = false;
    return true;
}
RT_D V3 pearl_brdf_cos(const DScene& S, const DMaterial& m, const Ray& r, const Rec& h, V3 wi) {
    V3 n = h.n, wo = unit(-r.d);
    float cos_o = fclamp(dot(wo, n), 0.0f, 1.0f), cos_i = dot(wi, n);
    if (cos_i <= 0.0f) return mk(0, 0, 0);
    V3 hv = unit(wo + wi);
    float woh = fmx(dot(wo, hv), 0.0f), hn = fmx(dot(hv, n), 0.0f);
    float al = fmx(m.p[7] * m.p[7], 1e-4f);
    float d = ggx_ndf(hn, al);
    float g1o = smith_g1(cos_o, al), g1i = smith_g1(cos_i, al);
    float fh = schlick(woh, 1.0f / m.p[3]);
    return nacre_color(S, woh, m.p[3], pearl_thickness(m, h.p)) * (d * fh * g1o * g1i / (4.0f * cos_o));
}
RT_D float pearl_spec_pdf(const DMaterial& m, const Ray& r, const Rec& h, V3 wi) {
    V3 n = h.n, wo = unit(-r.d);
    float cos_o = fclamp(dot(wo, n), 0.0f, 1.0f), cos_i = dot(wi, n);
    if (cos_i <= 0.0f) return 0.0f;
    V3 hv = unit(wo + wi);
    float hn = fmx(dot(hv, n), 0.0f);
    float f = fmx(schlick(cos_o, 1.0f / m.p[3]), 0.10f);
    float al = fmx(m.p[7] * m.p[7], 1e-4f);
    return f * ggx_ndf(hn, al) * smith_g1(cos_o, al) / (4.0f * cos_o);
}

// SpectralDielectric's IOR at lambda: the Sellmeier law when p5 != 0 (extension), else Cauchy.
RT_D float sdiel_ior(const DMaterial& m, float lambda) {
    return m.p[5] != 0.0f ? sellmeier(lambda, m.p + 6, m.p + 9) : cauchy(lambda, m.p[0], m.p[1]);
}

// Material::scatter dispatch.  Medium events scatter with the HG phase function.
template <uint32_t MASK = RT_MASK_ALL>
RT_D bool mat_scatter(const DScene& S, const Rec& h, const Ray& r, Rng& rng, Scatter& sr) {
    if (h.medium) {
        const DMedium& md = S.media[h.medium - 1];
        sr.att = mk(md.albedo[0], md.albedo[1], md.albedo[2]);
        sr.ray = scatter_from(h.p, hg_sample(unit(r.d), md.g, rng), r);
        sr.skip_pdf = true;
        return true;
    }
    const DMaterial& m = S.mats[h.mat];
    switch (m.kind) {
        case M_LAMBERTIAN:
            if (!RT_IN(MASK, M_LAMBERTIAN)) return false;
            sr.att = tex_value(S, m, h.u, h.v, h.p); sr.ray = scatter_from(h.p, h.n, r); sr.skip_pdf = false;
            return true;
        case M_DIELECTRIC: {
            if (!RT_IN(MASK, M_DIELECTRIC)) return false;
            bool refl;
            V3 d = dielectric_boundary(m.p[0], r, h, rng, refl);
            sr.att = mk(1, 1, 1); sr.ray = scatter_from(h.p, d, r); sr.skip_pdf = true;
            return true;
        }
        case M_SPECTRAL_DIELECTRIC: {
            if (!RT_IN(MASK, M_SPECTRAL_DIELECTRIC)) return false;
            float ior = sdiel_ior(m, r.wavelength);
            bool refl;
            V3 d = dielectric_boundary(ior, r, h, rng, refl);
            V3 att = mk(1, 1, 1);
            bool tinted = false;
            if (!r.spectral && !refl) { if (r.wl != 2) att = spectral_fire_rgb(S, r.wavelength, r.wl != 0); tinted = true; }
            if (!h.front) {
                float dd = h.t;
                att.x *= expf(-m.p[2] * dd); att.y *= expf(-m.p[3] * dd); att.z *= expf(-m.p[4] * dd);
            }
            sr.ray = scatter_from(h.p, d, r);
            if (tinted) { sr.ray.spectral = true; sr.ray.tint = 2; }
            sr.att = att; sr.skip_pdf = true;
            return true;
        }
        case M_PBR: return RT_IN(MASK, M_PBR) && pbr_scatter(S, m, pbr_params(S, m, h), r, h, rng, sr);
        case M_SPECTRAL_METAL: return RT_IN(MASK, M_SPECTRAL_METAL) && smetal_scatter(S, m, r, h, rng, sr);
        case M_ROUGH_DIELECTRIC: return RT_IN(MASK, M_ROUGH_DIELECTRIC) && rough_dielectric_scatter(m, r, h, rng, sr);
        case M_ROUGH_SPECTRAL: return RT_IN(MASK, M_ROUGH_SPECTRAL) && rsd_scatter(S, m, r, h, rng, sr);
        case M_SSS: return RT_IN(MASK, M_SSS) && sss_scatter(m, r, h, rng, sr);
        case M_PEARL: return RT_IN(MASK, M_PEARL) && pearl_scatter(S, m, r, h, rng, sr);
        default: return false;   // DiffuseLight
    }
}
// planck_raw (src/spectrum.rs:204-208): relative Planck radiance
RT_HD float planck_raw(float lambda_nm, float temp_k) {
    const float HC_OVER_K = 14388000.0f;
    float x = HC_OVER_K / (lambda_nm * temp_k);
    float l2 = lambda_nm * lambda_nm, l4 = l2 * l2;
    return 1.0f / ((lambda_nm * l4) * fmaxf(expf(x) - 1.0f, 1e-30f));
}
// Material::emitted_at for the ray that reached the emitter (its hero
// wavelength and whether its throughput already carries the spectral weight).
RT_D V3 mat_emitted(const DScene& S, const Rec& h, const Ray& r) {
    if (h.medium) return mk(0, 0, 0);
    const DMaterial& m = S.mats[h.mat];
    if (m.kind == M_DIFFUSE_LIGHT) return tex_value(S, m, h.u, h.v, h.p);
    if (m.kind == M_BLACKBODY) {
        // BlackbodyLight::emitted_at (src/material.rs:69-80)
        // the wide wavelength modes normalise over the table extended to 780 nm (p6-9)
        if (r.spectral) { float sc = planck_raw(r.wavelength, m.p[0]) * (r.wl != 0 ? m.p[6] : m.p[2]) * m.p[1]; return mk(sc, sc, sc); }
        return r.wl != 0 ? mk(m.p[7], m.p[8], m.p[9]) : mk(m.p[3], m.p[4], m.p[5]);
    }
    if (m.kind == M_PBR) return mk(m.p[9], m.p[10], m.p[11]) * m.p[12];
    return mk(0, 0, 0);
}
RT_D float mat_scattering_pdf(const Rec& h, V3 dir) { return fmx(dot(h.n, unit(dir)) / PI, 0.0f); }
// Material::shading_normal (src/hittable.rs:58-60): the PBR normal map's
// perturbed normal, the hit normal otherwise; the diffuse branch samples and
// evaluates its cosine lobe about it.
RT_D V3 mat_shading_normal(const DScene& S, const Rec& h) {
    if (h.medium) return h.n;
    const DMaterial& m = S.mats[h.mat];
    return m.kind == M_PBR ? pbr_normal(S, m, h) : h.n;
}
RT_D float scattering_pdf_n(V3 n, V3 dir) { return fmx(dot(n, unit(dir)) / PI, 0.0f); }
RT_D V3 mat_albedo_hint(const DScene& S, const Rec& h) {
    if (h.medium) return mk(1, 1, 1);
    const DMaterial& m = S.mats[h.mat];
    if (m.kind == M_LAMBERTIAN) return tex_value(S, m, h.u, h.v, h.p);
    if (m.kind == M_PBR) return pbr_albedo(S, m, h.u, h.v);
    if (m.kind == M_DIFFUSE_LIGHT) return tex_value(S, m, h.u, h.v, h.p);
    if (m.kind == M_BLACKBODY) return mk(m.p[3], m.p[4], m.p[5]);
    if (m.kind == M_SSS || m.kind == M_PEARL) return mk(m.p[0], m.p[1], m.p[2]);
    if (m.kind == M_SPECTRAL_METAL) {
        // SpectralMetal::new's avg_color (src/material.rs:435-443): the mean of
        // spectral_to_rgb(lambda) * F0(lambda) over the 13 table wavelengths
        V3 acc = mk(0, 0, 0);
        for (int i = 0; i < 13; i++) {
            float lam = 380.0f + (float)i * 25.0f;
            float n, k;
            metal_nk((int)m.p[0], lam, n, k);
            acc += spectral_to_rgb(S, lam) * fresnel_conductor(1.0f, n, k);
        }
        return acc / 13.0f;
    }
    return mk(1, 1, 1);
}
RT_D bool mat_receives_caustics(const DScene& S, const Rec& h) {
    if (h.medium) return false;
    const DMaterial& m = S.mats[h.mat];
    return m.kind == M_LAMBERTIAN || m.kind == M_PEARL || (m.kind == M_PBR && m.p[4] < 0.999f);
}
template <uint32_t MASK = RT_MASK_ALL>
RT_D V3 mat_spec_brdf_cos(const DScene& S, const Rec& h, const Ray& r, V3 wi) {
    if (h.medium) return mk(0, 0, 0);
    const DMaterial& m = S.mats[h.mat];
    if (RT_IN(MASK, M_PBR) && m.kind == M_PBR) return pbr_brdf_cos(S, m, pbr_params(S, m, h), r, h, wi);
    if (RT_IN(MASK, M_SPECTRAL_METAL) && m.kind == M_SPECTRAL_METAL) return smetal_brdf_cos(S, m, r, h, wi);
    if (RT_IN(MASK, M_PEARL) && m.kind == M_PEARL) return pearl_brdf_cos(S, m, r, h, wi);
    return mk(0, 0, 0);
}
template <uint32_t MASK = RT_MASK_ALL>
RT_D float mat_spec_pdf(const DScene& S, const Rec& h, const Ray& r, V3 wi) {
    if (h.medium) return 0.0f;
    const DMaterial& m = S.mats[h.mat];
    if (RT_IN(MASK, M_PBR) && m.kind == M_PBR) return pbr_spec_pdf(m, pbr_params(S, m, h), r, h, wi);
    if (RT_IN(MASK, M_SPECTRAL_METAL) && m.kind == M_SPECTRAL_METAL) return smetal_spec_pdf(m, r, h, wi);
    if (RT_IN(MASK, M_PEARL) && m.kind == M_PEARL) return pearl_spec_pdf(m, r, h, wi);
    return 0.0f;
}

// ---- background -------------------------------------------------------------------------
static constexpr float COS_SUN_MAX = 0.9999892f;

RT_D V3 preetham(V3 dir, V3 sun_dir, float turbidity) {
    V3 d = unit(dir), sun = unit(sun_dir);
    float cg = fclamp(dot(d, sun), -1.0f, 1.0f);
    float gamma = acosf(cg);
    float t = fclamp(turbidity, 1.7f, 20.0f);
    float cts = fmx(sun.y, 0.0175f);
    float ths = acosf(cts);
#define PEREZ(a, b, c, dp, e, ct, g, cgm) ((1.0f + (a) * expf((b) / (ct))) * (1.0f + (c) * expf((dp) * (g)) + (e) * (cgm) * (cgm)))
    float ay = 0.1787f * t - 1.4630f, by_ = -0.3554f * t + 0.4275f, cy = 0.0227f * t + 5.3251f, dy = 0.1206f * t - 2.5771f, ey = -0.0670f * t + 0.3703f;
    float ax = -0.0193f * t - 0.2592f, bx = -0.0665f * t + 0.0008f, cx = -0.0004f * t + 0.2125f, dx = -0.0641f * t - 0.8989f, ex = -0.0033f * t + 0.0452f;
    float acy = -0.0167f * t - 0.2608f, bcy = -0.0950f * t + 0.0092f, ccy = -0.0079f * t + 0.2102f, dcy = -0.0441f * t - 1.6537f, ecy = -0.0109f * t + 0.0529f;
    float fn_y = PEREZ(ay, by_, cy, dy, ey, 1.0f, ths, cts);
    float fn_x = PEREZ(ax, bx, cx, dx, ex, 1.0f, ths, cts);
    float fn_cy = PEREZ(acy, bcy, ccy, dcy, ecy, 1.0f, ths, cts);
    float yz = fmx((4.0453f * t - 4.9710f) * tanf((4.0f / 9.0f - t / 120.0f) * (PI - 2.0f * ths)) - 0.2155f * t + 2.4192f, 0.0f);
    if (cg > COS_SUN_MAX && sun.y > 0.0f && d.y > 0.0f) { float disc = yz * 111.0f; return mk(disc, disc * 0.95f, disc * 0.85f); }
    float ts = ths, ts2 = ts * ts, ts3 = ts2 * ts;
    float xz = t * t * (0.00166f * ts3 - 0.00375f * ts2 + 0.00209f * ts) + t * (-0.02903f * ts3 + 0.06377f * ts2 - 0.03202f * ts + 0.00394f)
             + (0.11693f * ts3 - 0.21196f * ts2 + 0.06052f * ts + 0.25886f);
    float yzc = t * t * (0.00275f * ts3 - 0.00610f * ts2 + 0.00317f * ts) + t * (-0.04214f * ts3 + 0.08970f * ts2 - 0.04153f * ts + 0.00516f)
              + (0.15346f * ts3 - 0.26756f * ts2 + 0.06670f * ts + 0.26688f);
    float ct = fmx(d.y, 0.001f);
    float fy = PEREZ(ay, by_, cy, dy, ey, ct, gamma, cg);
    float fx = PEREZ(ax, bx, cx, dx, ex, ct, gamma, cg);
    float fcy = PEREZ(acy, bcy, ccy, dcy, ecy, ct, gamma, cg);
#undef PEREZ
    float cap_y = fmx(yz * fy / fn_y, 0.0f);
    float cap_x = fclamp(xz * fx / fn_x, 0.001f, 0.998f);
    float cap_yc = fclamp(yzc * fcy / fn_cy, 0.001f, 0.998f);
    float l = cap_y * 0.05f;
    float X = cap_x / cap_yc * l, Y = l, Z = fmx((1.0f - cap_x - cap_yc) / cap_yc * l, 0.0f);
    V3 sky = mk(fmx(3.2406f * X - 1.5372f * Y - 0.4986f * Z, 0.0f), fmx(-0.9689f * X + 1.8758f * Y + 0.0415f * Z, 0.0f),
                fmx(0.0557f * X - 0.2040f * Y + 1.0570f * Z, 0.0f));
    if (d.y < 0.0f) { float tt = fmn(-d.y * 8.0f, 1.0f); return sky * (1.0f - tt) + mk(0.07f, 0.06f, 0.05f) * tt; }
    return sky;
}

RT_D V3 env_px(const DScene& S, int32_t ix, int32_t iy) {
    int32_t w = (int32_t)S.env_w, h = (int32_t)S.env_h;
    int32_t x = ((ix % w) + w) % w;
    int32_t y = iy < 0 ? 0 : (iy > h - 1 ? h - 1 : iy);
    const float* p = S.env_rgb + ((size_t)y * S.env_w + x) * 3;
    return mk(p[0], p[1], p[2]);
}
RT_D V3 env_unrot(const DScene& S, V3 d) { float s = sinf(S.env_rot), c = cosf(S.env_rot); return mk(d.x * c - d.z * s, d.y, d.x * s + d.z * c); }
RT_D V3 env_rot(const DScene& S, V3 d) { float s = sinf(S.env_rot), c = cosf(S.env_rot); return mk(d.x * c + d.z * s, d.y, -d.x * s + d.z * c); }
RT_D V3 env_sample(const DScene& S, V3 dir) {
    V3 d = env_unrot(S, unit(dir));
    float u = rem1(0.5f + atan2f(d.x, d.z) / (2.0f * PI));
    float v = acosf(fclamp(d.y, -1.0f, 1.0f)) / PI;
    float x = u * (float)S.env_w - 0.5f, y = v * (float)S.env_h - 0.5f;
    int32_t x0 = f2i(floorf(x)), y0 = f2i(floorf(y));
    float fx = x - (float)x0, fy = y - (float)y0;
    V3 c00 = env_px(S, x0, y0), c10 = env_px(S, x0 + 1, y0), c01 = env_px(S, x0, y0 + 1), c11 = env_px(S, x0 + 1, y0 + 1);
    return (c00 * (1.0f - fx) + c10 * fx) * (1.0f - fy) + (c01 * (1.0f - fx) + c11 * fx) * fy;
}
RT_D uint32_t partition_point(const float* a, uint32_t n, float xi) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) { uint32_t mid = lo + (hi - lo) / 2; if (a[mid] < xi) lo = mid + 1; else hi = mid; }
    return lo;
}
RT_D float env_pdf_at(const DScene& S, uint32_t i, uint32_t j) {
    const float* p = S.env_rgb + ((size_t)j * S.env_w + i) * 3;
    float l = lum(mk(p[0], p[1], p[2]));
    return l * (float)((size_t)S.env_w * S.env_h) * S.env_inv_total / (2.0f * PI * PI);
}
// cdf: the two CDF searches whatever the scene's sampler (rebuild_caustics
// draws through EnvMapData::sample_dir, src/scene.rs:58-62).
RT_D V3 env_sample_dir(const DScene& S, Rng& rng, float& pdf, bool cdf = false) {
    uint32_t i, j;
    if (S.env_sampling == 1 && !cdf) {
        // alias table over the W*H texels (Walker 1977, Vose 1991): the first
        // draw picks a column of the table (64-bit multiply-high, no modulo
        // bias), the second accepts it or takes its alias.  Same density and
        // the same number of draws as the two CDF searches.
        uint32_t n = S.env_w * S.env_h;
        uint32_t t = (uint32_t)__umul64hi(rng_u64(rng), (uint64_t)n);
        if (t >= n) t = n - 1;
        if (rng_f(rng) >= S.env_aprob[t]) t = S.env_alias[t];
        j = t / S.env_w;
        i = t - j * S.env_w;
    } else {
        j = min(partition_point(S.env_marg, S.env_h, rng_f(rng)), S.env_h - 1);
        i = min(partition_point(S.env_cond + (size_t)j * S.env_w, S.env_w, rng_f(rng)), S.env_w - 1);
    }
    float u = ((float)i + rng_f(rng)) / (float)S.env_w;
    float v = ((float)j + rng_f(rng)) / (float)S.env_h;
    float phi = 2.0f * PI * (u - 0.5f);
    float theta = PI * v;
    float st = sinf(theta);
    pdf = env_pdf_at(S, i, j);
    return env_rot(S, mk(st * sinf(phi), cosf(theta), st * cosf(phi)));
}
RT_D float env_dir_pdf(const DScene& S, V3 dir) {
    V3 d = env_unrot(S, unit(dir));
    float u = rem1(0.5f + atan2f(d.x, d.z) / (2.0f * PI));
    float v = acosf(fclamp(d.y, -1.0f, 1.0f)) / PI;
    uint32_t i = min(f2u(u * (float)S.env_w), S.env_w - 1);
    uint32_t j = min(f2u(v * (float)S.env_h), S.env_h - 1);
    return env_pdf_at(S, i, j);
}
RT_D bool has_env(const DScene& S) { return S.bg_kind == BGK_ENV || S.bg_kind == BGK_ENV_SPLIT; }
RT_D V3 bg_simple(const DScene& S, uint32_t k, V3 dir) {
    if (k == BGK_SOLID) return mk(S.bg_color[0], S.bg_color[1], S.bg_color[2]);
    float t = (fclamp(dir.y, -1.0f, 1.0f) + 1.0f) * 0.5f;
    return mk(S.bg_bottom[0], S.bg_bottom[1], S.bg_bottom[2]) * (1.0f - t) + mk(S.bg_top[0], S.bg_top[1], S.bg_top[2]) * t;
}
RT_D V3 bg_eval(const DScene& S, V3 dir) {
    switch (S.bg_kind) {
        case BGK_SOLID: case BGK_GRADIENT: return bg_simple(S, S.bg_kind, dir);
        case BGK_PHYSICAL: return preetham(dir, mk(S.sun_dir[0], S.sun_dir[1], S.sun_dir[2]), S.turbidity);
        case BGK_ENV: return env_sample(S, dir);
        default: return bg_simple(S, S.bg_visible, dir);
    }
}
RT_D V3 env_eval(const DScene& S, V3 dir) { return has_env(S) ? env_sample(S, dir) : mk(0, 0, 0); }
RT_D bool physical_sun(const DScene& S) { return S.bg_kind == BGK_PHYSICAL && S.sun_dir[1] > 0.0f; }
RT_D float sun_pdf_value(const DScene& S, V3 dir) {
    if (physical_sun(S) && dot(unit(dir), unit(mk(S.sun_dir[0], S.sun_dir[1], S.sun_dir[2]))) > COS_SUN_MAX)
        return 1.0f / (2.0f * PI * (1.0f - COS_SUN_MAX));
    return 0.0f;
}
RT_D float env_pdf_value(const DScene& S, V3 dir) { return has_env(S) ? env_dir_pdf(S, dir) : 0.0f; }
RT_D V3 sample_sun_cone(V3 axis, float ctm, Rng& rng) {
    float r1 = rng_f(rng), r2 = rng_f(rng);
    float ct = 1.0f - r1 * (1.0f - ctm);
    float st = sqrtf(fmx(1.0f - ct * ct, 0.0f));
    float phi = 2.0f * PI * r2;
    V3 up = fabsf(axis.x) < 0.999f ? mk(1, 0, 0) : mk(0, 1, 0);
    V3 t = unit(cross(axis, up));
    V3 b = cross(axis, t);
    return t * (st * cosf(phi)) + b * (st * sinf(phi)) + axis * ct;
}

// ---- light list sampling (HittableList::pdf_generate / pdf_value) -------------------------
RT_D float light_pdf_value(const DScene& S, const DObject& ob, V3 origin, V3 dir, float time, Stack& st) {
    // Translate/Rotate/Scale map the query into object space; the pdf of a
    // shape is computed through its own hit test, as Quad::pdf_value does.
    Ray wr = make_ray(origin, dir, time);
    Ray r = to_object(S, ob, wr);
    V3 o = r.o, d = r.d;
    if (ob.shape == SH_QUAD) {
        const DQuad& q = S.quads[ob.data_first];
        float t, a, b;
        if (!quad_test(q, r, 0.001f, FINF, t, a, b)) return 0.0f;
        V3 P = at(r, t);
        V3 n = mk(q.n[0], q.n[1], q.n[2]);
        V3 nn = dot(r.d, n) < 0.0f ? n : -n;
        float dist2 = len2(P - o);
        float ct = fabsf(dot(unit(d), nn));
        if (ct < 1e-8f) return 0.0f;
        return dist2 / (ct * q.area);
    }
    if (ob.shape == SH_SPHERE) {
        V3 c = mk(ob.p[0], ob.p[1], ob.p[2]);
        if (ob.p[7] != 0.0f) c = c + time * (mk(ob.p[4], ob.p[5], ob.p[6]) - c);
        V3 tc = c - o;
        float d2 = len2(tc), r2 = ob.p[3] * ob.p[3];
        if (d2 <= r2) return 0.0f;
        float ctm = sqrtf(1.0f - r2 / d2);
        return 1.0f / (2.0f * PI * (1.0f - ctm));
    }
    if (ob.shape == SH_DISK) {
        // Disk::pdf_value (src/disk.rs:130-137)
        float t;
        if (!disk_test(ob.p, r, 0.001f, FINF, t)) return 0.0f;
        float dist2 = len2(at(r, t) - o);
        float ct = fabsf(dot(unit(d), mk(ob.p[3], ob.p[4], ob.p[5])));
        if (ct < 1e-8f || ob.p[13] <= 0.0f) return 0.0f;
        return dist2 / (ct * ob.p[13]);
    }
    return 0.0f;
}
RT_D float lights_pdf_value(const DScene& S, V3 origin, V3 dir, float time, Stack& st) {
    if (S.num_lights == 0) return 0.0f;
    float w = 1.0f / (float)S.num_lights;
    float s = 0.0f;
    for (uint32_t i = 0; i < S.num_lights; i++) s += w * light_pdf_value(S, S.lights[i], origin, dir, time, st);
    return s;
}
RT_D V3 light_generate(const DScene& S, const DObject& ob, V3 origin, Rng& rng, float time) {
    // object-space origin, then the local direction back to world space
    Ray wr = make_ray(origin, mk(1, 0, 0), time);
    Ray r = to_object(S, ob, wr);
    V3 o = r.o, d;
    if (ob.shape == SH_QUAD) {
        const DQuad& q = S.quads[ob.data_first];
        float r1 = rng_f(rng), r2 = rng_f(rng);
        V3 pt = mk(q.q[0], q.q[1], q.q[2]) + r1 * mk(q.u[0], q.u[1], q.u[2]) + r2 * mk(q.v[0], q.v[1], q.v[2]);
        d = unit(pt - o);
    } else if (ob.shape == SH_SPHERE) {
        V3 c = mk(ob.p[0], ob.p[1], ob.p[2]);
        if (ob.p[7] != 0.0f) c = c + time * (mk(ob.p[4], ob.p[5], ob.p[6]) - c);
        V3 tc = c - o;
        float d2 = len2(tc), r2 = ob.p[3] * ob.p[3];
        if (d2 <= r2) d = unit(tc);
        else {
            float ctm = sqrtf(1.0f - r2 / d2);
            float r1 = rng_f(rng), r2r = rng_f(rng);
            float phi = 2.0f * PI * r1;
            float ct = 1.0f - r2r * (1.0f - ctm);
            float st = sqrtf(fmx(1.0f - ct * ct, 0.0f));
            d = onb_local(onb_from_w(unit(tc)), mk(cosf(phi) * st, sinf(phi) * st, ct));
        }
    } else if (ob.shape == SH_DISK) {
        // Disk::pdf_generate (src/disk.rs:139-145): r = R sqrt(u), theta = 2 pi v
        float rr = ob.p[6] * sqrtf(rng_f(rng));
        float th = 2.0f * PI * rng_f(rng);
        V3 pt = mk(ob.p[0], ob.p[1], ob.p[2]) + rr * (cosf(th) * mk(ob.p[7], ob.p[8], ob.p[9]) + sinf(th) * mk(ob.p[10], ob.p[11], ob.p[12]));
        d = unit(pt - o);
    } else d = mk(1, 0, 0);
    for (uint32_t k = 0; k < ob.xf_count; k++) {
        const DXform& x = S.xforms[ob.xf_first + k];
        if (x.kind == XF_ROTATE) d = rot_apply(x.v, d);
    }
    return d;
}
RT_D V3 lights_generate(const DScene& S, V3 origin, Rng& rng, float time) {
    if (S.num_lights == 0) return mk(1, 0, 0);
    uint32_t idx = rng_u32(rng) % S.num_lights;
    return light_generate(S, S.lights[idx], origin, rng, time);
}

// ---- photon gather (kd_gather, src/photon.rs:195-228), iterative --------------------------
// PhotonMap::irradiance (src/photon.rs:141-228): Epanechnikov-weighted sum of
// the photons within the gather radius whose normal faces the query normal,
// times 2/(pi^2 R^2).  Range search of the heap-ordered kd-tree without a
// stack: bit d of `pend` says the far child of the depth-d ancestor still
// needs a visit; when a descent leaves the tree, the deepest pending ancestor
// is recomputed from the index (in heap order the ancestor d levels up of
// node i is ((i + 1) >> d) - 1) and its far child entered.
// The Epanechnikov sum of kd_gather (src/photon.rs:195-228) over the
// heap-ordered tree, before normalisation; `count` (optional) receives the
// number of photons that passed the radius and hemisphere tests.
// nb (may be null): adds the bytes of the photon records the walk reads
// (16-B position + split axis and 16-B normal per node visited, 16-B power
// per photon summed), for the counting pass.
template <bool SUM>
RT_D V3 photon_walk(const DScene& S, V3 pos, V3 nrm, float r2, uint32_t& count, uint32_t* nb = nullptr) {
    uint32_t n = S.num_photons;
    uint32_t cnt = 0;
    V3 acc = mk(0, 0, 0);
    uint32_t node = 0, pend = 0;
    int depth = 0;
    for (;;) {
        if (node < n) {
            float4 P = S.ph_pos[node];
            float4 N = S.ph_nrm[node];
            if (nb) *nb += 32u;
            float dx = P.x - pos.x, dy = P.y - pos.y, dz = P.z - pos.z;
            float d2 = dx * dx + dy * dy + dz * dz;
            if (d2 < r2) {
                float dt = N.x * nrm.x + N.y * nrm.y + N.z * nrm.z;
                if (dt > 0.0f) {
                    if (SUM) { float4 W = S.ph_pow[node]; acc += mk(W.x, W.y, W.z) * (1.0f - d2 / r2); if (nb) *nb += 16u; }
                    cnt++;
                }
            }
            int axis = (int)P.w;
            float sq = comp(pos, axis) - (axis == 0 ? P.x : (axis == 1 ? P.y : P.z));
            if (sq * sq < r2) pend |= 1u << depth; else pend &= ~(1u << depth);
            node = sq <= 0.0f ? 2 * node + 1 : 2 * node + 2;
            depth++;
            continue;
        }
        if (pend == 0) break;
        int d = 31 - __clz(pend);
        pend &= ~(1u << d);
        uint32_t anc = ((node + 1) >> (depth - d)) - 1;      // heap order: the ancestor's index + 1 is a prefix of node + 1
        float4 P = S.ph_pos[anc];
        int axis = (int)P.w;
        float sq = comp(pos, axis) - (axis == 0 ? P.x : (axis == 1 ? P.y : P.z));
        node = sq <= 0.0f ? 2 * anc + 2 : 2 * anc + 1;
        depth = d + 1;
    }
    count = cnt;
    return acc;
}
RT_D V3 photon_gather(const DScene& S, V3 pos, V3 nrm, float r2, uint32_t* nb = nullptr) {
    uint32_t c;
    return photon_walk<true>(S, pos, nrm, r2, c, nb);
}

// k-NN gather (extension) over the bucket tree: a complete binary tree of
// pk_depth levels of splitting planes (heap order, {plane, axis bits}) whose
// leaves hold at most RT_PK_LEAF photons stored contiguously (position and
// normal in one 32-B record, power apart).  A leaf's photons load
// independently of one another, so a lane has them in flight together rather
// than one dependent load per photon as in the photon-per-node tree; the
// plane array is small and stays in cache.
#ifndef RT_PK_LEAF
#define RT_PK_LEAF 8
#endif

// Push d2 into this lane's max-heap of k floats (its LDS column, `stride`
// apart); once k are held, replace the root when d2 is smaller.  The heap is
// 4-ary (children of i: 4i+1 .. 4i+4): three levels for k = 64, and a
// level's child reads are independent, so a sift waits on half the LDS round
// trips a binary heap would.  Returns the new root (the largest held), which
// is the k-th smallest distance once k are held, whatever the heap's shape.
RT_D float knn_heap_add(lds_f32* heap, uint32_t stride, uint32_t k, uint32_t& cnt, float d2) {
    if (cnt < k) {                               // push, sift up
        uint32_t i = cnt++;
        float root = d2;
        while (i > 0) {
            uint32_t p = (i - 1) >> 2;
            float hp = heap[p * stride];
            if (hp >= d2) { root = p == 0 ? hp : heap[0]; break; }
            heap[i * stride] = hp;
            i = p;
        }
        heap[i * stride] = d2;
        return i == 0 ? d2 : root;
    }
    uint32_t i = 0;                              // replace the root, sift down
    float root = d2;
    for (;;) {
        const uint32_t c0 = 4 * i + 1;
        if (c0 >= k) break;
        uint32_t c = c0;
        float hc = heap[c0 * stride];
        float h1 = c0 + 1 < k ? heap[(c0 + 1) * stride] : -FINF;
        float h2 = c0 + 2 < k ? heap[(c0 + 2) * stride] : -FINF;
        float h3 = c0 + 3 < k ? heap[(c0 + 3) * stride] : -FINF;
        if (h1 > hc) { c = c0 + 1; hc = h1; }
        if (h2 > hc) { c = c0 + 2; hc = h2; }
        if (h3 > hc) { c = c0 + 3; hc = h3; }
        if (hc <= d2) break;
        if (i == 0) root = hc;
        heap[i * stride] = hc;
        i = c;
    }
    heap[i * stride] = d2;
    return root;
}

// Range walk of the bucket tree around `pos` over the photons on the query's
// hemisphere.  KNN = false: the Epanechnikov sum within r2 (the gather of
// src/photon.rs:195-228 in another order).  KNN = true: additionally keep the
// k smallest squared distances in the heap, the pruning radius shrinking to
// the heap's root once k are held (a far child queued under a larger radius
// is re-tested when it is popped); the sum then covers only the first k
// photons found, which is the whole gather when fewer than k lie within r2.
// KNN = false with stop > 0: give up once `stop` photons are counted (the
// caller then knows the query holds at least that many).  SUM = false (with
// KNN): no sum, for callers that already know k photons lie within r2.
// Stackless: bit d of `pend` marks the far child of the depth-d plane as
// still to visit; the ancestor d levels up of node i is ((i + 1) >> d) - 1.
// nb (may be null): adds the bytes the walk reads (8-B plane per split
// visited, 32-B record per photon of a visited leaf, 16-B power per photon
// summed), for the counting pass.
template <bool KNN, bool SUM = true>
RT_D V3 pk_walk(const DScene& S, V3 pos, V3 nrm, float r2, lds_f32* heap, uint32_t stride, uint32_t& count, float& bound_out,
                uint32_t stop = 0, uint32_t* nb = nullptr) {
    const uint32_t D = S.pk_depth, k = S.ph_knn;
    const uint32_t first_leaf = (1u << D) - 1u;
    float bound = r2;
    uint32_t cnt = 0;
    V3 acc = mk(0, 0, 0);
    uint32_t node = 0, pend = 0, depth = 0;
    for (;;) {
        while (depth < D) {                                  // descend to the near leaf
            float2 sp = S.pk_split[node];
            if (nb) *nb += 8u;
            float sq = comp(pos, __float_as_int(sp.y)) - sp.x;
            if (sq * sq < bound) pend |= 1u << depth; else pend &= ~(1u << depth);
            node = sq <= 0.0f ? 2 * node + 1 : 2 * node + 2;
            depth++;
        }
        uint32_t leaf = node - first_leaf;
        uint32_t b = S.pk_leaf[leaf], e = S.pk_leaf[leaf + 1];
        if (nb) *nb += 32u * (e - b);
        // the leaf's records (at most RT_PK_LEAF) all requested before the
        // first is used: one memory round trip per leaf, not per photon
        float4 PL[RT_PK_LEAF], NL[RT_PK_LEAF];
#pragma unroll
        for (uint32_t j = 0; j < RT_PK_LEAF; j++)
            if (b + j < e) { PL[j] = S.pk_rec[2 * (b + j)]; NL[j] = S.pk_rec[2 * (b + j) + 1]; }
#pragma unroll
        for (uint32_t j = 0; j < RT_PK_LEAF; j++) {
            if (b + j >= e) break;
            const uint32_t i = b + j;
            float4 P = PL[j], N = NL[j];
            float dx = P.x - pos.x, dy = P.y - pos.y, dz = P.z - pos.z;
            float d2 = dx * dx + dy * dy + dz * dz;
            if (d2 < bound && N.x * nrm.x + N.y * nrm.y + N.z * nrm.z > 0.0f) {
                if (!KNN || (SUM && cnt < k)) {
                    float4 W = S.pk_pow[i];
                    acc += mk(W.x, W.y, W.z) * (1.0f - d2 / bound);
                    if (nb) *nb += 16u;
                }
                if (KNN) {
                    float root = knn_heap_add(heap, stride, k, cnt, d2);
                    if (cnt == k) bound = root;
                } else {
                    cnt++;
                    if (stop && cnt >= stop) { count = cnt; bound_out = bound; return acc; }
                }
            }
        }
        bool more = false;                                   // pop the deepest pending far child
        while (pend) {
            uint32_t d = 31u - (uint32_t)__clz(pend);
            pend &= ~(1u << d);
            uint32_t anc = ((node + 1) >> (depth - d)) - 1;
            float2 sp = S.pk_split[anc];
            if (nb) *nb += 8u;
            float sq = comp(pos, __float_as_int(sp.y)) - sp.x;
            if (KNN && !(sq * sq < bound)) continue;
            node = sq <= 0.0f ? 2 * anc + 2 : 2 * anc + 1;
            depth = d + 1;
            more = true;
            break;
        }
        if (!more) break;
    }
    count = cnt;
    bound_out = bound;
    return acc;
}

// k-NN gather radius: the k-th smallest squared distance among the photons on
// the query's hemisphere closer than the gather radius.  Fewer than k (or a
// zero k-th distance): the gather radius, and `acc` / `count` < k carry that
// gather's sum.
RT_D float photon_knn_r2(const DScene& S, V3 pos, V3 nrm, lds_f32* heap, uint32_t stride, float bound0, V3& acc,
                         uint32_t& count, uint32_t* nb = nullptr) {
    float bound;
    acc = pk_walk<true>(S, pos, nrm, bound0, heap, stride, count, bound, 0, nb);
    if (count < S.ph_knn) return S.ph_r2;
    float r2 = heap[0];
    return r2 > 0.0f ? r2 : S.ph_r2;
}

// First pass of the wavefront's k-NN stage, without the heap: the gather
// within ph_r2, abandoned once k photons are counted.  True (irradiance in
// `e`) when fewer than k lie within the radius: the walk and its sum are the
// heap search's, whose radius cannot shrink before k are held.
RT_D bool photon_knn_sparse(const DScene& S, V3 pos, V3 nrm, V3& e) {
    uint32_t cnt;
    float unused;
    V3 acc = pk_walk<false>(S, pos, nrm, S.ph_r2, nullptr, 0, cnt, unused, S.ph_knn);
    if (cnt >= S.ph_knn) return false;
    e = acc * S.ph_norm;
    return true;
}

// k-NN irradiance: the heap search's radius, then the gather at that radius
// (with fewer than k photons the search already summed the gather).
RT_D V3 photon_knn_irradiance(const DScene& S, V3 pos, V3 nrm, lds_f32* heap, uint32_t stride, uint32_t* nb = nullptr) {
    V3 acc;
    uint32_t cnt;
    float r2 = photon_knn_r2(S, pos, nrm, heap, stride, S.ph_r2, acc, cnt, nb);
    if (cnt < S.ph_knn) return acc * S.ph_norm;
    float norm = r2 == S.ph_r2 ? S.ph_norm : 2.0f / (PI * PI * r2);
    float unused;
    return pk_walk<false>(S, pos, nrm, r2, nullptr, 0, cnt, unused, 0, nb) * norm;
}

// Second pass of the wavefront's k-NN stage, for a query that the first pass
// found to hold at least k photons within ph_r2, in two halves (two kernels:
// the heap search needs the LDS heap, the gather does not).  The heap search
// without the sum of its first k photons (never used then) gives the gather
// radius: the k-th distance, or ph_r2 should the search hold fewer than k
// after all or the k-th distance be 0 ...
RT_D float photon_knn_radius(const DScene& S, V3 pos, V3 nrm, lds_f32* heap, uint32_t stride) {
    uint32_t cnt;
    float unused;
    pk_walk<true, false>(S, pos, nrm, S.ph_r2, heap, stride, cnt, unused);
    if (cnt < S.ph_knn) return S.ph_r2;
    return heap[0] > 0.0f ? heap[0] : S.ph_r2;
}
// ... and the gather at that radius.  Together they equal
// photon_knn_irradiance: with fewer than k photons the gather within ph_r2
// is that function's sum (same walk, same order, the radius never shrinking).
RT_D V3 photon_knn_gather(const DScene& S, V3 pos, V3 nrm, float r2) {
    float norm = r2 == S.ph_r2 ? S.ph_norm : 2.0f / (PI * PI * r2);
    uint32_t cnt;
    float unused;
    return pk_walk<false>(S, pos, nrm, r2, nullptr, 0, cnt, unused) * norm;
}

// PhotonMap::irradiance (src/photon.rs:148-157).  A k-NN map is gathered here
// when the caller has an LDS heap (st.has_heap), else `knn_irr` holds the value
// the wavefront's k_wf_knn stage computed for this hit.
// The gather's photon-record bytes go to st.n_pbytes (counting pass).
RT_D V3 photon_irradiance(const DScene& S, V3 pos, V3 nrm, V3 knn_irr, Stack& st) {
    if (S.num_photons == 0) return mk(0, 0, 0);
    if (S.ph_knn) return st.has_heap ? photon_knn_irradiance(S, pos, nrm, st.heap, st.stride, &st.n_pbytes) : knn_irr;
    return photon_gather(S, pos, nrm, S.ph_r2, &st.n_pbytes) * S.ph_norm;
}

// ---- one bounce of ray_color (src/renderer.rs:371-585) ---------------------------------------
// The path loop is split at its trace calls so the same code serves the
// megakernel (which resolves shadow rays immediately) and the wavefront stages
// (which queue them).  Random numbers are drawn in the reference's order: a
// shadow ray's transmittance never draws, so deferring it leaves the stream
// unchanged.
struct PathState {
    Ray ray;
    V3 thr;
    float prev_mis, prev_sun_w;
    bool prev_specular;
    V3 knn_irr;          // k-NN photon irradiance of the current hit when computed ahead (wavefront)
};

// A deferred shadow query: `contrib` is added (times the transmittance) when
// the segment [0.001, t_max] toward `dir` is not blocked.
struct ShadowReq {
    V3 o, d;
    float t_max, time;
    V3 contrib;
};

RT_D void path_init(PathState& ps, const Ray& r) {
    ps.ray = r; ps.thr = mk(1, 1, 1); ps.prev_specular = true; ps.prev_mis = 1.0f; ps.prev_sun_w = 1.0f; ps.knn_irr = mk(0, 0, 0);
}

// ---- four hero wavelengths (extension) -------------------------------------------------
// Hero wavelength spectral sampling (Wilkie, Nawaz, Droske, Weidlich and
// Hanika, EGSR 2014): a camera path carries lambda_0 uniform over [380, 780)
// and the rotations lambda_k = lambda_0 + 100 k nm (mod the range), k = 1..3.
// Every decision is drawn at lambda_0 and every contribution is computed at
// lambda_0 without a spectral tint; it is then weighted by the balance
// heuristic over the four rotations,
//     sum_j (f_j / f_0) E_j T(lambda_j)  /  sum_j (p_j / p_0),
// where f_j / f_0 and p_j / p_0 are the products over the path's vertices of
// the BSDF-value and sampling-density ratios (q and r below; 1 at lambda_0),
// E_j / E_0 a blackbody emitter's ratio, and T the tint of the path's first
// spectral event (the CMF colour of a spectral metal, the fire tint of a
// dispersive refraction) evaluated per wavelength.  A dispersive specular
// refraction has density 0 at the other wavelengths: q and r drop to 0 there
// and the path goes on as a one-wavelength path.  Paths without a spectral
// event get weight 1, so they are the one-wavelength estimator's.  Next-event
// contributions are weighted on the path up to their vertex with the
// connection's own BSDF ratio; the next-event / BSDF MIS weights use
// lambda_0's densities, which at every spectral material that samples the
// sun (the spectral metal) do not depend on the wavelength.
struct Spec4 {
    float q[3], r[3];
    bool on;
};
RT_D void spec4_init(Spec4& s) {
    for (int k = 0; k < 3; k++) { s.q[k] = 1.0f; s.r[k] = 1.0f; }
    s.on = false;
}
RT_D float hero_lambda(float l0, int k) {
    float t = (l0 - 380.0f) + 100.0f * (float)k;
    if (t >= 400.0f) t -= 400.0f;
    return 380.0f + t;
}
RT_D V3 spec4_tint(const DScene& S, int kind, float lam) {
    if (kind == 1) return spectral_to_rgb(S, lam, true);
    if (kind == 2) return spectral_fire_rgb(S, lam, true);
    return mk(1, 1, 1);
}
// Weight of a contribution made at lambda_0 (see above): kind is the tint,
// b (may be null) the connection's own BSDF ratios f_k / f_0, bb_temp > 0 a
// blackbody emitter's temperature (its radiance ratio per wavelength).
// Paths of the other modes (a batch may mix them) get 1.
RT_D V3 spec4_weight(const DScene& S, const Ray& ray, const Spec4& s, int kind, const float* b, float bb_temp) {
    if (ray.wl != 2) return mk(1, 1, 1);
    if (!s.on && kind == 0 && b == nullptr && !(bb_temp > 0.0f)) return mk(1, 1, 1);
    float l0 = ray.wavelength;
    float e0 = bb_temp > 0.0f ? planck_raw(l0, bb_temp) : 1.0f;
    V3 num = spec4_tint(S, kind, l0);
    float den = 1.0f;
    for (int k = 0; k < 3; k++) {
        float lk = hero_lambda(l0, k + 1);
        float w = s.q[k] * (b ? b[k] : 1.0f);
        if (bb_temp > 0.0f) w *= planck_raw(lk, bb_temp) / e0;
        if (w != 0.0f) num += spec4_tint(S, kind, lk) * w;
        den += s.r[k];
    }
    return num / den;
}
// A spectral metal's BSDF ratio f_k / f_0 between wo and wi: the conductor
// Fresnel factor at the half vector (the rest does not depend on lambda).
RT_D void spec4_metal_ratio(const DMaterial& m, V3 wo, V3 wi, float l0, float b[3]) {
    float ch = fmx(dot(wo, unit(wo + wi)), 0.0f);
    float n0, k0; metal_nk((int)m.p[0], l0, n0, k0);
    float f0 = fresnel_conductor(ch, n0, k0);
    for (int k = 0; k < 3; k++) {
        float nk, kk; metal_nk((int)m.p[0], hero_lambda(l0, k + 1), nk, kk);
        b[k] = f0 > 0.0f ? fresnel_conductor(ch, nk, kk) / f0 : 0.0f;
    }
}
// q and r after the scatter from `ray` at h toward wi (spectral materials).
RT_D void spec4_scatter(const DScene& S, const Rec& h, const Ray& ray, V3 wi, Spec4& s) {
    if (h.medium || ray.wl != 2) return;
    const DMaterial& m = S.mats[h.mat];
    V3 wo = unit(-ray.d);
    float l0 = ray.wavelength;
    if (m.kind == M_SPECTRAL_METAL) {
        float b[3];
        spec4_metal_ratio(m, wo, wi, l0, b);
        for (int k = 0; k < 3; k++) s.q[k] *= b[k];
        s.on = true;
    } else if (m.kind == M_SPECTRAL_DIELECTRIC) {
        s.on = true;
        if (dot(wi, h.n) <= 0.0f) {                       // refraction: a delta at lambda_0 only
            for (int k = 0; k < 3; k++) { s.q[k] = 0.0f; s.r[k] = 0.0f; }
            return;
        }
        // reflection: probability 1 under total internal reflection, else F
        // (dielectric_boundary); the BSDF value over it is 1
        float c = fmn(dot(wo, h.n), 1.0f);
        float sn = sqrtf(1.0f - c * c);
        auto prefl = [&](float lam) {
            float ior = sdiel_ior(m, lam);
            float ratio = h.front ? 1.0f / ior : ior;
            return ratio * sn > 1.0f ? 1.0f : fresnel_dielectric(c, ratio);
        };
        float p0 = prefl(l0);
        for (int k = 0; k < 3; k++) {
            float f = p0 > 0.0f ? prefl(hero_lambda(l0, k + 1)) / p0 : 0.0f;
            s.q[k] *= f; s.r[k] *= f;
        }
    } else if (m.kind == M_ROUGH_SPECTRAL) {
        s.on = true;
        double al = (double)rsd_alpha(m, h);
        double n0 = (double)rsd_ior(m, l0);
        if (dot(wi, h.n) > 0.0f) {                        // reflection: only F depends on lambda
            float com = fmn(dot(wo, unit(wo + wi)), 1.0f);
            float f0 = fresnel_dielectric(com, h.front ? 1.0f / (float)n0 : (float)n0);
            for (int k = 0; k < 3; k++) {
                float nk = rsd_ior(m, hero_lambda(l0, k + 1));
                float f = f0 > 0.0f ? fresnel_dielectric(com, h.front ? 1.0f / nk : nk) / f0 : 0.0f;
                s.q[k] *= f; s.r[k] *= f;
            }
            return;
        }
        double lf0, lp0;
        double k0 = h.front ? n0 : 1.0 / n0;
        bool ok0 = rsd_logf<double>(wo, wi, h.n, k0, al, false, lf0) && rsd_logpdf_t<double>(wo, wi, h.n, k0, al, lp0);
        for (int k = 0; k < 3; k++) {
            double nk = (double)rsd_ior(m, hero_lambda(l0, k + 1));
            double kk = h.front ? nk : 1.0 / nk;
            double lf, lp;
            if (ok0 && rsd_logf<double>(wo, wi, h.n, kk, al, false, lf) && rsd_logpdf_t<double>(wo, wi, h.n, kk, al, lp)) {
                s.q[k] *= (float)exp(lf - lf0);
                s.r[k] *= (float)exp(lp - lp0);
            } else {
                s.q[k] = 0.0f; s.r[k] = 0.0f;
            }
        }
    }
}
RT_D float spec4_bb_temp(const DScene& S, const Rec& h, const Ray& ray) {
    if (h.medium || !ray.spectral) return 0.0f;
    const DMaterial& m = S.mats[h.mat];
    return m.kind == M_BLACKBODY ? m.p[0] : 0.0f;
}

// Where a vertex's shadow requests go.  ShadowList keeps them for the caller,
// sh[0..n) in the order they are filed; a hook sees each request with its
// filing index.  (The wavefront's shade stage files them straight into its
// shadow queue instead, so no request stays live past the point it is made.)
struct ShadowList {
    ShadowReq* sh;
    int& n;
    RT_D int index() const { return n; }
    RT_D void file(const ShadowReq& q) { sh[n++] = q; }
};

// Shade the hit (or miss) `id` of ps.ray at bounce `depth`.  Adds direct
// emission / background to `color`, files up to three shadow requests with
// `sink` and advances the path.  Returns false when the path ends.  H4: the
// path carries four hero wavelengths, their ratios in *s4.
template <class DH, uint32_t MASK = RT_MASK_ALL, bool H4 = false, class SINK>
RT_D bool bounce_s(const DScene& S, PathState& ps, const HitId& id, int depth, float bg_scale, bool use_photons, Rng& rng,
                   V3& color, SINK& sink, Stack& st, DH& dh, Spec4* s4 = nullptr) {
    Ray& ray = ps.ray;
    dh.segment(S, ray, id);
    const int ptint = ray.spectral ? ray.tint : 0;    // the tint the path carries into this vertex
    const bool lights_empty = S.num_lights == 0;
    const bool sun = physical_sun(S);
    const bool env = has_env(S);
    if (id.slot < 0) {
        float sp = sun_pdf_value(S, ray.d);
        float ep = env_pdf_value(S, ray.d);
        float ew = (sp + ep == 0.0f) ? 1.0f : (ps.prev_specular ? ps.prev_sun_w : ps.prev_mis);
        V3 bg = (depth == 0 || !env) ? bg_eval(S, ray.d) : env_eval(S, ray.d);
        if constexpr (H4) bg *= spec4_weight(S, ray, *s4, ptint, nullptr, 0.0f);
        color += dh.emit(ps.thr * bg * bg_scale * ew);
        return false;
    }
    Rec h = finalize(S, ray, id, 0.001f);
    float ew = (ps.prev_specular || lights_empty) ? 1.0f : ps.prev_mis;
    V3 em = mat_emitted(S, h, ray);
    if constexpr (H4) if (em.x != 0.0f || em.y != 0.0f || em.z != 0.0f) em *= spec4_weight(S, ray, *s4, ptint, nullptr, spec4_bb_temp(S, h, ray));
    color += dh.emit(ps.thr * em * ew);
    V3 w4 = mk(1, 1, 1);                                // this vertex's next-event weight (four wavelengths)
    if constexpr (H4) w4 = spec4_weight(S, ray, *s4, ptint, nullptr, 0.0f);
    Scatter sr;
    if (!mat_scatter<MASK>(S, h, ray, rng, sr)) return false;
    constexpr bool SPEC_ONLY = (MASK & ~RT_MASK_SPECULAR) == 0u;   // every scatter here is skip_pdf
    if (SPEC_ONLY || sr.skip_pdf) {
        ps.prev_sun_w = 1.0f;
        if (sun) {
            float p_sun = 1.0f / (2.0f * PI * (1.0f - COS_SUN_MAX));
            V3 su = unit(mk(S.sun_dir[0], S.sun_dir[1], S.sun_dir[2]));
            V3 ss = sample_sun_cone(su, COS_SUN_MAX, rng);
            V3 bc = mat_spec_brdf_cos<MASK>(S, h, ray, ss);
            if (bc.x > 0.0f || bc.y > 0.0f || bc.z > 0.0f) {
                float pm = mat_spec_pdf<MASK>(S, h, ray, ss);
                V3 sc = bg_eval(S, ss);
                if constexpr (H4) {
                    if (!h.medium && S.mats[h.mat].kind == M_SPECTRAL_METAL) {
                        float b[3];
                        spec4_metal_ratio(S.mats[h.mat], unit(-ray.d), ss, ray.wavelength, b);
                        sc *= spec4_weight(S, ray, *s4, ray.spectral ? ray.tint : 1, b, 0.0f);
                    } else {
                        sc *= w4;
                    }
                }
                float md = p_sun * p_sun + pm * pm;
                if (md > 0.0f) {
                    ShadowReq q;
                    q.o = h.p; q.d = ss; q.t_max = FINF; q.time = ray.time; q.contrib = ps.thr * bc * sc * p_sun / md;
                    dh.shadow(q, sink.index(), S, h, ray, false);
                    sink.file(q);
                }
            }
            float pmb = mat_spec_pdf<MASK>(S, h, ray, sr.ray.d);
            if (pmb > 0.0f) ps.prev_sun_w = (pmb * pmb) / fmx(pmb * pmb + p_sun * p_sun, 1e-8f);
        }
        dh.scatter(S, h, ray, sr.ray.d);
        if constexpr (H4) spec4_scatter(S, h, ray, sr.ray.d, *s4);
        ps.thr *= sr.att;
        ray = sr.ray;
        ps.prev_specular = true;
        ps.prev_mis = 1.0f;
    } else {
        const V3 sn = mat_shading_normal(S, h);
        if (use_photons) {
            V3 irr = photon_irradiance(S, h.p, h.n, ps.knn_irr, st);
            if (irr.x > 0.0f || irr.y > 0.0f || irr.z > 0.0f) color += dh.emit_vertex(ps.thr * mat_albedo_hint(S, h) * irr * w4, S, h, ray);
        }
        if (!lights_empty) {
            V3 ldir = lights_generate(S, h.p, rng, ray.time);
            Ray shadow = make_ray(h.p, ldir, ray.time);
            Rec lrec;
            if (lights_closest(S, shadow, 0.001f, FINF, lrec)) {
                float l_pdf = lights_pdf_value(S, h.p, ldir, ray.time, st);
                float brdf = scattering_pdf_n(sn, shadow.d);
                float sp = sun_pdf_value(S, ldir);
                float ep = env_pdf_value(S, ldir);
                float md = l_pdf * l_pdf + brdf * brdf + sp * sp + ep * ep;
                if (md > 0.0f && brdf > 0.0f) {
                    ShadowReq q;
                    q.o = h.p; q.d = ldir; q.t_max = lrec.t * (1.0f - 1e-4f); q.time = ray.time;
                    V3 le = mat_emitted(S, lrec, ray);
                    if constexpr (H4) le *= spec4_weight(S, ray, *s4, ptint, nullptr, spec4_bb_temp(S, lrec, ray));
                    q.contrib = ps.thr * sr.att * brdf * l_pdf * le / md;
                    dh.shadow(q, sink.index(), S, h, ray, true);
                    sink.file(q);
                }
            }
        }
        if (sun) {
            V3 su = unit(mk(S.sun_dir[0], S.sun_dir[1], S.sun_dir[2]));
            V3 ss = sample_sun_cone(su, COS_SUN_MAX, rng);
            float brdf = scattering_pdf_n(sn, ss);
            if (brdf > 0.0f) {
                float spdf = 1.0f / (2.0f * PI * (1.0f - COS_SUN_MAX));
                V3 sc = bg_eval(S, ss);
                float ap = lights_empty ? 0.0f : lights_pdf_value(S, h.p, ss, ray.time, st);
                float md = spdf * spdf + brdf * brdf + ap * ap;
                ShadowReq q;
                q.o = h.p; q.d = ss; q.t_max = FINF; q.time = ray.time;
                q.contrib = ps.thr * sr.att * brdf * spdf * sc * w4 / md;
                dh.shadow(q, sink.index(), S, h, ray, true);
                sink.file(q);
            }
        }
        if (env) {
            float epdf;
            V3 ed = env_sample_dir(S, rng, epdf);
            if (epdf > 0.0f) {
                float brdf = scattering_pdf_n(sn, ed);
                if (brdf > 0.0f) {
                    float ap = lights_empty ? 0.0f : lights_pdf_value(S, h.p, ed, ray.time, st);
                    V3 ec = env_eval(S, ed);
                    float md = epdf * epdf + brdf * brdf + ap * ap;
                    ShadowReq q;
                    q.o = h.p; q.d = ed; q.t_max = FINF; q.time = ray.time;
                    q.contrib = ps.thr * sr.att * brdf * epdf * ec * w4 / md;
                    dh.shadow(q, sink.index(), S, h, ray, true);
                    sink.file(q);
                }
            }
        }
        Onb uvw = onb_from_w(sn);
        float r1 = rng_f(rng), r2 = rng_f(rng);
        float phi = 2.0f * PI * r1;
        float sr2 = sqrtf(r2);
        V3 ind = onb_local(uvw, mk(cosf(phi) * sr2, sinf(phi) * sr2, sqrtf(1.0f - r2)));
        float pdf_val = fmx(dot(unit(ind), uvw.w) / PI, 0.0f);
        if (pdf_val <= 0.0f) return false;
        Ray sc = scatter_from(h.p, ind, ray);
        float scat_pdf = scattering_pdf_n(sn, sc.d);
        if (scat_pdf <= 0.0f) return false;
        if (dot(ind, h.n) <= 0.0f) return false;
        float a = lights_empty ? 0.0f : lights_pdf_value(S, h.p, ind, ray.time, st);
        float s = sun_pdf_value(S, ind);
        float e = env_pdf_value(S, ind);
        float nee = a * a + s * s + e * e;
        ps.prev_mis = (pdf_val * pdf_val) / fmx(pdf_val * pdf_val + nee, 1e-8f);
        dh.scatter(S, h, ray, sc.d);
        ps.thr *= sr.att * (scat_pdf / pdf_val);
        ray = sc;
        ps.prev_specular = false;
    }
    if (depth >= 2) {
        V3 eff = ps.thr;                                 // four wavelengths: with the path's spectral weight
        if constexpr (H4) eff *= spec4_weight(S, ray, *s4, ray.spectral ? ray.tint : 0, nullptr, 0.0f);
        float survive = fmx(fmx(eff.x, eff.y), eff.z);
        float p = fmn(survive, 1.0f);
        if (survive <= 0.0f || rng_f(rng) >= p) return false;
        ps.thr /= p;
    }
    return true;
}

// Hooks of bounce_t for the plain renderers: no derivative bookkeeping.
struct NoDiff {
    RT_D void segment(const DScene&, const Ray&, const HitId&) {}
    RT_D V3 emit(V3 c) { return c; }
    RT_D V3 emit_vertex(V3 c, const DScene&, const Rec&, const Ray&) { return c; }
    RT_D void shadow(ShadowReq&, int, const DScene&, const Rec&, const Ray&, bool) {}
    RT_D void scatter(const DScene&, const Rec&, const Ray&, V3) {}
};
// bounce_s with the requests kept in sh[0..nsh).
template <class DH, uint32_t MASK = RT_MASK_ALL, bool H4 = false>
RT_D bool bounce_t(const DScene& S, PathState& ps, const HitId& id, int depth, float bg_scale, bool use_photons, Rng& rng,
                   V3& color, ShadowReq* sh, int& nsh, Stack& st, DH& dh, Spec4* s4 = nullptr) {
    nsh = 0;
    ShadowList sink{sh, nsh};
    return bounce_s<DH, MASK, H4>(S, ps, id, depth, bg_scale, use_photons, rng, color, sink, st, dh, s4);
}
template <uint32_t MASK = RT_MASK_ALL, bool H4 = false>
RT_D bool bounce(const DScene& S, PathState& ps, const HitId& id, int depth, float bg_scale, bool use_photons, Rng& rng,
                 V3& color, ShadowReq* sh, int& nsh, Stack& st, Spec4* s4 = nullptr) {
    NoDiff nd;
    return bounce_t<NoDiff, MASK, H4>(S, ps, id, depth, bg_scale, use_photons, rng, color, sh, nsh, st, nd, s4);
}

// Megakernel form of ray_color: trace, shade, resolve shadow rays inline.
// H4: the ray carries four hero wavelengths (wl 2).
template <bool H4 = false>
RT_D V3 ray_color(const DScene& S, Ray ray, int max_depth, float bg_scale, bool use_photons, Rng& rng, Stack& st) {
    PathState ps;
    path_init(ps, ray);
    Spec4 s4;
    spec4_init(s4);
    V3 color = mk(0, 0, 0);
    ShadowReq sh[3];
    for (int depth = 0; depth < max_depth; depth++) {
        HitId id = trace_closest(S, ps.ray, 0.001f, FINF, st);
        int nsh = 0;
        bool cont = bounce<RT_MASK_ALL, H4>(S, ps, id, depth, bg_scale, use_photons, rng, color, sh, nsh, st, &s4);
        for (int k = 0; k < nsh; k++) {
            float tr = trace_transmittance(S, make_ray(sh[k].o, sh[k].d, sh[k].time), 0.001f, sh[k].t_max, st, false);
            if (tr > 0.0f) color += sh[k].contrib * tr;
        }
        if (!cont) break;
    }
    return color;
}

struct DCamera { float origin[3], lower_left[3], horizontal[3], vertical[3], u[3], v[3]; float lens_radius; uint32_t blades; };

RT_D Ray camera_ray(const DCamera& c, float s, float t, Rng& rng) {
    V3 rd;
    if (c.blades >= 3) {
        float n = (float)c.blades, rot = PI / n;
        float k = (float)rng_range_u32(rng, 0, c.blades);
        float a1 = 2.0f * PI * k / n + rot, a2 = 2.0f * PI * (k + 1.0f) / n + rot;
        V3 p1 = mk(cosf(a1), sinf(a1), 0.0f), p2 = mk(cosf(a2), sinf(a2), 0.0f);
        float r1 = rng_f(rng), r2 = rng_f(rng);
        if (r1 + r2 > 1.0f) { r1 = 1.0f - r1; r2 = 1.0f - r2; }
        rd = c.lens_radius * (p1 * r1 + p2 * r2);
    } else rd = c.lens_radius * random_in_unit_disk(rng);
    V3 cu = mk(c.u[0], c.u[1], c.u[2]), cv = mk(c.v[0], c.v[1], c.v[2]);
    V3 off = cu * rd.x + cv * rd.y;
    V3 org = mk(c.origin[0], c.origin[1], c.origin[2]);
    V3 dir = mk(c.lower_left[0], c.lower_left[1], c.lower_left[2]) + s * mk(c.horizontal[0], c.horizontal[1], c.horizontal[2])
             + t * mk(c.vertical[0], c.vertical[1], c.vertical[2]) - org - off;
    return make_ray(org + off, dir, rng_f(rng));
}

struct DParams {
    uint32_t sample_idx, strata, width, height, max_depth, use_frame_time, use_photons;
    uint32_t wl;                    // Ray::wl of the camera rays: 0 reference, 1 wide, 2 four hero wavelengths
    float bg_scale, t0, t1;
    uint64_t seed;
    unsigned long long* counters;   // counting pass: [0] node tests, [1] primitive tests, [2] rays, [3] samples
};

// Camera ray of one pixel sample: render_tiles body (src/renderer.rs:693-726).
RT_D Ray pixel_ray(const DCamera& cam, const DParams& P, uint32_t row, uint32_t col, uint32_t sample_idx, Rng& rng) {
    uint32_t i = row * P.width + col;
    float w_denom = (float)max(P.width - 1, 1u), h_denom = (float)max(P.height - 1, 1u);
    uint32_t s2 = P.strata * P.strata;
    uint32_t ray_y = P.height - 1 - row;
    float uj, vj;
    if (s2 > 0) {
        uint32_t off = (i * 0x9E3779B9u) % s2;
        uint32_t s = (sample_idx + off) % s2;
        uint32_t sx = s % P.strata, sy = s / P.strata;
        float sf = (float)P.strata;
        uj = ((float)sx + rng_f(rng)) / sf;
        vj = ((float)sy + rng_f(rng)) / sf;
    } else { uj = rng_f(rng); vj = rng_f(rng); }
    float u = ((float)col + uj) / w_denom;
    float v = ((float)ray_y + vj) / h_denom;
    Ray r = camera_ray(cam, u, v, rng);
    if (P.use_frame_time) r.time = P.t0 + rng_f(rng) * (P.t1 - P.t0);
    r.wavelength = rng_range(rng, 380.0f, P.wl != 0 ? 780.0f : 700.0f);
    r.wl = (uint8_t)P.wl;
    return r;
}

}  // namespace rt
