// ORACLE — test infrastructure only (see rtcore.cpp).
//
// Host checker of the material-parameter gradient (extension: the reference
// has no derivatives).  It differentiates the same estimator as the device
// adjoint — sampling decisions at the scene's parameters theta_s, BSDF values
// at theta — but by a different route:
//   * forward score accumulation instead of path replay: every radiance term
//     c added to a pixel contributes delta . (c * sum of dlog f of the
//     vertices whose BSDF value it carries), with the per-parameter prefix sum
//     carried along the path;
//   * dlog f by central differences of log f in double precision instead of
//     dual numbers in float.
// It also renders the detached estimator at theta != theta_s (BSDF ratios
// f_theta / f_theta_s per vertex) for finite differences with common random
// numbers.
//
// log f of the rough spectral dielectric is the Walter et al. (2007)
// microfacet BSDF (GGX D, separable Smith G1, exact unpolarised Fresnel),
// the density that the RoughSpectralDielectric sampler's weight G1(wi)
// implies; tests pin it by integrating f |cos| over the sphere against the
// sampler's mean weight.
#pragma once
#include "rc_render.h"
#include "../include/rt_api.h"
#include <cmath>
#include <vector>

namespace rc {

static inline double sellmeier_d(double lambda_nm, const float* B, const float* C) {
    double l = lambda_nm * 1e-3, l2 = l * l;
    return std::sqrt(1.0 + B[0] * l2 / (l2 - C[0]) + B[1] * l2 / (l2 - C[1]) + B[2] * l2 / (l2 - C[2]));
}
static inline double rsd_n(const float* p, double lambda) { return p[0] + sellmeier_d(lambda, p + 1, p + 4) - sellmeier_d(589.3, p + 1, p + 4); }
static inline int rsd_field(int prim) { return prim <= 0 ? 7 : 8 + (prim - 1) % 16; }
static inline double rsd_alpha_d(const float* p, int prim) { return std::min(std::max((double)p[rsd_field(prim)], 1e-4), 1.0); }

// log f at a rough spectral vertex for the parameter block p (24 floats).
static inline bool rsd_vertex_logf(const float* p, const HitRecord& rec, const Vec3& wo, const Vec3& wi, float lambda, double& lf) {
    double n = rsd_n(p, lambda);
    double k = rec.front_face ? n : 1.0 / n;
    bool refl = wi.dot(rec.normal) > 0.0f;
    return logf_rsd(d3(wo), d3(wi), d3(rec.normal), k, rsd_alpha_d(p, rec.prim_id), refl, lf);
}

// Constant media of one descriptor entry, for the sigma_t parameters: the
// medium objects built from it, and whether any of them sits under a transform.
struct MediumSet { std::vector<const ConstantMedium*> objs; bool transformed = false; bool constant = true; };

// Parameter map: fmap[24 * material + field] = parameter index or -1;
// mpar[k] = the parameter of media[k]'s density or -1.
struct ParamMap {
    std::vector<int8_t> fmap;
    std::vector<int> mpar;
    int n = 0;
    bool build(const std::vector<rt_material>& mats, const rt_param* prm, uint32_t np,
               const std::vector<MediumSet>* media = nullptr) {
        if (!prm || np == 0 || np > 64) return false;
        fmap.assign(24 * mats.size(), (int8_t)-1);
        mpar.assign(media ? media->size() : 0, -1);
        for (uint32_t i = 0; i < np; i++) {
            const rt_param& q = prm[i];
            if (q.material & RT_PARAM_MEDIUM) {
                uint32_t k = q.material & ~RT_PARAM_MEDIUM;
                if (!media || k >= media->size() || q.field != 0 || q.count != 1 || mpar[k] >= 0) return false;
                const MediumSet& ms = (*media)[k];
                if (ms.objs.empty() || ms.transformed || !ms.constant) return false;
                mpar[k] = (int)i;
                continue;
            }
            if (q.material >= mats.size() || q.count == 0 || q.field >= 24 || q.count > 24 - q.field) return false;
            const rt_material& m = mats[q.material];
            for (uint32_t f = q.field; f < q.field + q.count; f++) {
                bool ok = (m.kind == RT_MAT_LAMBERTIAN && m.tex == 0 && f < 3) || (m.kind == RT_MAT_ROUGH_SPECTRAL && (f <= 3 || f >= 7));
                if (!ok || fmap[24 * q.material + f] >= 0) return false;
                fmap[24 * q.material + f] = (int8_t)i;
            }
        }
        n = (int)np;
        return true;
    }
};

// dlog f / dtheta_j (per colour channel) at a vertex; out has one Color per parameter.
static inline void vertex_dlogf(const std::vector<rt_material>& mats, const ParamMap& pm, const HitRecord& rec,
                                const Ray& ray, const Vec3& wi, std::vector<Color>& out) {
    for (auto& c : out) c = Color();
    if (!rec.mat || rec.mat->id < 0) return;
    int mi = rec.mat->id;
    const rt_material& m = mats[mi];
    const int8_t* fm = &pm.fmap[24 * (size_t)mi];
    if (m.kind == RT_MAT_LAMBERTIAN) {
        if (m.tex != 0) return;
        for (int c = 0; c < 3; c++)
            if (fm[c] >= 0) {
                Color v;
                if (c == 0) v.x = 1.0f / m.p[0]; else if (c == 1) v.y = 1.0f / m.p[1]; else v.z = 1.0f / m.p[2];
                out[fm[c]] += v;
            }
        return;
    }
    if (m.kind != RT_MAT_ROUGH_SPECTRAL) return;
    Vec3 wo = (-ray.direction).unit();
    Vec3 wu = wi.unit();
    int fa = rsd_field(rec.prim_id);
    for (int f = 0; f < 24; f++) {
        if (fm[f] < 0) continue;
        if (f >= 7 && f != fa) continue;
        if (f >= 7 && !(m.p[f] > 1e-4f && m.p[f] < 1.0f)) continue;
        double h = f == 0 ? 1e-6 : (f <= 3 ? 1e-6 : 1e-5 * m.p[f]);
        float pp[24], pm_[24];
        for (int q = 0; q < 24; q++) pp[q] = pm_[q] = m.p[q];
        // the perturbation goes through float storage: take the realised step
        double lp, lm;
        float up = (float)(m.p[f] + h), dn = (float)(m.p[f] - h);
        pp[f] = up; pm_[f] = dn;
        if (!rsd_vertex_logf(pp, rec, wo, wu, ray.wavelength, lp) || !rsd_vertex_logf(pm_, rec, wo, wu, ray.wavelength, lm)) continue;
        double g = (lp - lm) / ((double)up - (double)dn);
        out[fm[f]] += Color((float)g, (float)g, (float)g);
    }
}

// f_theta / f_theta_s at a vertex (theta_s: the scene's table, theta: me).
static inline Color vertex_ratio(const std::vector<rt_material>& ms, const std::vector<rt_material>& me, const HitRecord& rec,
                                 const Ray& ray, const Vec3& wi) {
    if (!rec.mat || rec.mat->id < 0) return Color(1, 1, 1);
    int mi = rec.mat->id;
    const rt_material& a = ms[mi];
    const rt_material& b = me[mi];
    if (a.kind == RT_MAT_LAMBERTIAN && a.tex == 0) {
        return Color(a.p[0] != 0.0f ? b.p[0] / a.p[0] : 1.0f, a.p[1] != 0.0f ? b.p[1] / a.p[1] : 1.0f,
                     a.p[2] != 0.0f ? b.p[2] / a.p[2] : 1.0f);
    }
    if (a.kind == RT_MAT_ROUGH_SPECTRAL) {
        Vec3 wo = (-ray.direction).unit();
        double ls, le;
        if (!rsd_vertex_logf(a.p, rec, wo, wi.unit(), ray.wavelength, ls)) return Color(1, 1, 1);
        if (!rsd_vertex_logf(b.p, rec, wo, wi.unit(), ray.wavelength, le)) return Color(0, 0, 0);
        float r = (float)std::exp(le - ls);
        return Color(r, r, r);
    }
    return Color(1, 1, 1);
}

// A parameterised constant medium object: its density at theta_s (ds) and
// at theta (de), and its parameter index.
struct MedParam { const ConstantMedium* cm; int param; float ds, de; };
// In-medium length of the segment [t_min, t_end] (the medium's own interval
// logic), and whether the segment ended in one of its scattering events.
static inline float med_overlap(const MedParam& m, const Ray& r, float t_min, float t_end) {
    float len;
    return inside_length(*m.cm->boundary, r, t_min, t_end, len) ? len : 0.0f;
}
static inline std::vector<MedParam> med_params(const ParamMap& pm, const std::vector<MediumSet>& media, const float* de) {
    std::vector<MedParam> out;
    for (size_t k = 0; k < pm.mpar.size(); k++)
        if (pm.mpar[k] >= 0)
            for (const ConstantMedium* cm : media[k].objs) out.push_back(MedParam{cm, pm.mpar[k], cm->density, de ? de[pm.mpar[k]] : cm->density});
    return out;
}

// Forward score accumulation for one sample.  A traced segment adds
// dlog p / dsigma = -L (+ 1 / sigma when it ends in that medium's event) to
// every later term; a shadow ray adds -L to its own term.
struct ScoreHook {
    const std::vector<rt_material>* mats;
    const ParamMap* pm;
    const std::vector<MedParam>* med = nullptr;
    Color delta;
    std::vector<Color> prefix, tmp, shadow;
    bool shadow_on = false;
    double* grad;
    void init(const std::vector<rt_material>* m, const ParamMap* p, Color d, double* g, const std::vector<MedParam>* md) {
        mats = m; pm = p; delta = d; grad = g; med = md;
        prefix.assign(p->n, Color()); tmp.assign(p->n, Color()); shadow.assign(p->n, Color());
        shadow_on = false;
    }
    void acc(const Color& c, const std::vector<Color>* extra) {
        for (int j = 0; j < pm->n; j++) {
            Color s = prefix[j];
            if (extra) s += (*extra)[j];
            if (shadow_on) s += shadow[j];
            grad[j] += (double)delta.x * c.x * s.x + (double)delta.y * c.y * s.y + (double)delta.z * c.z * s.z;
        }
        shadow_on = false;
    }
    Color add(const Color& c) { acc(c, nullptr); return c; }
    Color add_v(const Color& c, const HitRecord& rec, const Ray& ray, const Vec3& wi) {
        vertex_dlogf(*mats, *pm, rec, ray, wi, tmp);
        acc(c, &tmp);
        return c;
    }
    void scatter(const HitRecord& rec, const Ray& ray, const Vec3& wi) {
        shadow_on = false;
        vertex_dlogf(*mats, *pm, rec, ray, wi, tmp);
        for (int j = 0; j < pm->n; j++) prefix[j] += tmp[j];
    }
    void segment(const Ray& r, bool hit, const HitRecord& rec) {
        shadow_on = false;
        if (!med) return;
        for (const MedParam& m : *med) {
            float s = -med_overlap(m, r, 0.001f, hit ? rec.t : INF);
            if (hit && rec.mat == &m.cm->phase) s += 1.0f / m.ds;
            prefix[m.param] += Color(s, s, s);
        }
    }
    void shadow_seg(const Ray& r, float t_max) {
        for (auto& c : shadow) c = Color();
        shadow_on = med && !med->empty();
        if (!shadow_on) return;
        for (const MedParam& m : *med) { float s = -med_overlap(m, r, 0.001f, t_max); shadow[m.param] += Color(s, s, s); }
    }
};

// Detached estimator: decisions at the scene's parameters, BSDF values and
// medium densities at theta.
struct RatioHook {
    const std::vector<rt_material>* ms;
    const std::vector<rt_material>* me;
    const std::vector<MedParam>* med = nullptr;
    Color ratio = Color(1, 1, 1);
    float shadow = 1.0f;
    Color add(const Color& c) { Color v = c * ratio * shadow; shadow = 1.0f; return v; }
    Color add_v(const Color& c, const HitRecord& rec, const Ray& ray, const Vec3& wi) {
        Color v = c * ratio * vertex_ratio(*ms, *me, rec, ray, wi) * shadow;
        shadow = 1.0f;
        return v;
    }
    void scatter(const HitRecord& rec, const Ray& ray, const Vec3& wi) { shadow = 1.0f; ratio *= vertex_ratio(*ms, *me, rec, ray, wi); }
    void segment(const Ray& r, bool hit, const HitRecord& rec) {
        shadow = 1.0f;
        if (!med) return;
        for (const MedParam& m : *med) {
            float f = std::exp(-(m.de - m.ds) * med_overlap(m, r, 0.001f, hit ? rec.t : INF));
            if (hit && rec.mat == &m.cm->phase) f *= m.de / m.ds;
            ratio *= f;
        }
    }
    void shadow_seg(const Ray& r, float t_max) {
        shadow = 1.0f;
        if (!med) return;
        for (const MedParam& m : *med) shadow *= std::exp(-(m.de - m.ds) * med_overlap(m, r, 0.001f, t_max));
    }
};

}  // namespace rc
