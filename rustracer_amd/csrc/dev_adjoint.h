// Derivatives of the radiance estimator with respect to material parameters
// (extension: the reference renders only; it has no derivatives).
//
// Estimator.  Every sampling decision of a path (directions, Fresnel and lobe
// choices, Russian roulette) is drawn with the scene's current parameters
// theta_s; every BSDF value is evaluated at theta.  A path's contribution is
// then  C(theta) = sum_j e_j * prod_{k<j} f_theta(v_k) |cos| / p_theta_s(v_k),
// which equals the plain renderer's at theta = theta_s and is a smooth
// function of theta for fixed random numbers: material parameters move no
// geometry and no decision.  Its derivative at theta = theta_s is
//   dC/dtheta = sum_k dlog f(v_k)/dtheta * (radiance reaching the camera
//               through vertex k's outgoing direction)
//             + sum_j e_j-local terms (NEE / photon gather at a vertex whose
//               own BSDF value depends on theta).
//
// Two hook sets plug into bounce_t:
//   DetachedHook  forward render at theta != theta_s (finite differences with
//                 common random numbers);
//   ReplayHook    path-replay backward: the sample's radiance L is computed
//                 first, then the path is traced again from the same stream
//                 and each vertex's dlog f is weighted by the radiance still to
//                 come (L minus everything collected before that vertex).
//
// Differentiable fields: Lambertian (solid) albedo p0-2 (dlog f_c/da_c =
// 1/a_c); rough spectral dielectric n_d, Sellmeier B1-3 and GGX alpha, with
// dlog f/dn and dlog f/dalpha from forward-mode dual numbers in double through the
// Walter et al. (2007) microfacet BSDF.
#pragma once

namespace rt {

// ---- forward-mode dual numbers -------------------------------------------------------------
// Value and derivative in R (float or double); scalar operands are float.
template <class R> struct DlT {
    R v, d;
    RT_HD DlT() : v(0), d(0) {}
    RT_HD DlT(R a, R b = 0) : v(a), d(b) {}
};
template <class R> RT_HD DlT<R> operator+(DlT<R> a, DlT<R> b) { return DlT<R>(a.v + b.v, a.d + b.d); }
template <class R> RT_HD DlT<R> operator-(DlT<R> a, DlT<R> b) { return DlT<R>(a.v - b.v, a.d - b.d); }
template <class R> RT_HD DlT<R> operator*(DlT<R> a, DlT<R> b) { return DlT<R>(a.v * b.v, a.d * b.v + a.v * b.d); }
template <class R> RT_HD DlT<R> operator/(DlT<R> a, DlT<R> b) { return DlT<R>(a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)); }
template <class R> RT_HD DlT<R> operator-(DlT<R> a) { return DlT<R>(-a.v, -a.d); }
template <class R> RT_HD DlT<R> operator+(DlT<R> a, float b) { return DlT<R>(a.v + (R)b, a.d); }
template <class R> RT_HD DlT<R> operator+(float a, DlT<R> b) { return DlT<R>((R)a + b.v, b.d); }
template <class R> RT_HD DlT<R> operator-(DlT<R> a, float b) { return DlT<R>(a.v - (R)b, a.d); }
template <class R> RT_HD DlT<R> operator-(float a, DlT<R> b) { return DlT<R>((R)a - b.v, -b.d); }
template <class R> RT_HD DlT<R> operator*(DlT<R> a, float b) { return DlT<R>(a.v * (R)b, a.d * (R)b); }
template <class R> RT_HD DlT<R> operator*(float a, DlT<R> b) { return DlT<R>((R)a * b.v, (R)a * b.d); }
template <class R> RT_HD DlT<R> operator/(DlT<R> a, float b) { return DlT<R>(a.v / (R)b, a.d / (R)b); }
template <class R> RT_HD DlT<R> operator/(float a, DlT<R> b) { return DlT<R>((R)a / b.v, -(R)a * b.d / (b.v * b.v)); }
template <class R> RT_HD R val(DlT<R> a) { return a.v; }
template <class R> RT_HD DlT<R> xsqrt(DlT<R> a) { R s = sqrt(a.v); return DlT<R>(s, s > 0 ? a.d * (R)0.5 / s : (R)0); }
template <class R> RT_HD DlT<R> xlog(DlT<R> a) { return DlT<R>(log(a.v), a.d / a.v); }
template <class R> RT_HD DlT<R> xabs(DlT<R> a) { return a.v < 0 ? -a : a; }
using Dd = DlT<double>;
// d n(lambda) / d B_j of the rough spectral dielectric's dispersion law.
RT_HD float rsd_dn_dB(const float* p, float lambda, int j) {
    auto term = [&](float lam) {
        float l = lam * 1e-3f, l2 = l * l;
        float s = sqrtf(1.0f + p[1] * l2 / (l2 - p[4]) + p[2] * l2 / (l2 - p[5]) + p[3] * l2 / (l2 - p[6]));
        return (l2 / (l2 - p[4 + j])) / (2.0f * s);
    };
    return term(lambda) - term(RSD_LAMBDA_D);
}

// n(lambda) in double: the detached ratio compares two parameter sets whose
// n differ by a few float ulps for a 1e-3 step of a Sellmeier coefficient.
RT_D double rsd_ior_d(const DMaterial& m, float lambda) {
    auto sell = [&](double lam) {
        double l = lam * 1e-3, l2 = l * l;
        return sqrt(1.0 + m.p[1] * l2 / (l2 - m.p[4]) + m.p[2] * l2 / (l2 - m.p[5]) + m.p[3] * l2 / (l2 - m.p[6]));
    };
    return (double)m.p[0] + (sell(lambda) - sell(RSD_LAMBDA_D));
}

// ---- per-vertex scores ---------------------------------------------------------------------
#define RT_SCORE_MAX 6
struct Score {
    int n;
    int p[RT_SCORE_MAX];
    V3 v[RT_SCORE_MAX];
};
RT_D void score_add(Score& sc, int param, V3 v) {
    for (int k = 0; k < sc.n; k++) if (sc.p[k] == param) { sc.v[k] += v; return; }
    if (sc.n < RT_SCORE_MAX) { sc.p[sc.n] = param; sc.v[sc.n] = v; sc.n++; }
}

// dlog f / dtheta at a vertex (wo toward the previous vertex, wi the new or
// light direction, both unit).  fmap[24 * mat + field] = parameter or -1.
RT_D void vertex_score(const DScene& S, const int8_t* fmap, const Rec& h, V3 wo, V3 wi, float lambda, Score& sc) {
    sc.n = 0;
    if (h.medium) return;
    const DMaterial& m = S.mats[h.mat];
    const int8_t* fm = fmap + 24 * (size_t)h.mat;
    if (m.kind == M_LAMBERTIAN) {
        if (m.tex != 0) return;
        if (fm[0] >= 0) score_add(sc, fm[0], mk(1.0f / m.p[0], 0.0f, 0.0f));
        if (fm[1] >= 0) score_add(sc, fm[1], mk(0.0f, 1.0f / m.p[1], 0.0f));
        if (fm[2] >= 0) score_add(sc, fm[2], mk(0.0f, 0.0f, 1.0f / m.p[2]));
    } else if (m.kind == M_ROUGH_SPECTRAL) {
        // dual numbers in double: dlog D / dalpha = 2 / alpha - 4 alpha c^2 / (1 + c^2 (alpha^2 - 1))
        // is a difference of two terms of order 1 / alpha whose sum over a
        // lobe nearly vanishes (D stays normalised), and in float the
        // cancelling denominator biases that small sum by tenths of a percent
        int fa = rsd_alpha_field(h.prim_id);
        double nl = rsd_ior_d(m, lambda);
        double k = h.front ? nl : 1.0 / nl, dk = h.front ? 1.0 : -1.0 / (nl * nl);
        double al = (double)fclamp(m.p[fa], 1e-4f, 1.0f);
        bool refl = dot(wi, h.n) > 0.0f;
        if (fm[0] >= 0 || fm[1] >= 0 || fm[2] >= 0 || fm[3] >= 0) {
            Dd lf;
            if (rsd_logf<Dd>(wo, wi, h.n, Dd(k, dk), Dd(al, 0.0), refl, lf)) {
                float gn = (float)lf.d;
                if (fm[0] >= 0) score_add(sc, fm[0], mk(gn, gn, gn));
                for (int j = 0; j < 3; j++)
                    if (fm[1 + j] >= 0) { float g = gn * rsd_dn_dB(m.p, lambda, j); score_add(sc, fm[1 + j], mk(g, g, g)); }
            }
        }
        if (fm[fa] >= 0 && m.p[fa] > 1e-4f && m.p[fa] < 1.0f) {
            Dd lf;
            if (rsd_logf<Dd>(wo, wi, h.n, Dd(k, 0.0), Dd(al, 1.0), refl, lf)) {
                float ga = (float)lf.d;
                score_add(sc, fm[fa], mk(ga, ga, ga));
            }
        }
    }
}

// f_theta / f_theta_s at a vertex: `me` holds the evaluation materials.  The
// rough spectral dielectric's logs are taken in double: at alpha ~ 0.005 the
// GGX denominator 1 + c^2 (alpha^2 - 1) cancels to a few float ulps, which
// would swamp the ratio of two nearby parameter sets.
RT_D V3 vertex_ratio(const DScene& S, const DMaterial* me, const Rec& h, V3 wo, V3 wi, float lambda) {
    if (h.medium) return mk(1, 1, 1);
    const DMaterial& ms = S.mats[h.mat];
    const DMaterial& mv = me[h.mat];
    if (ms.kind == M_LAMBERTIAN) {
        V3 a = tex_value(S, ms, h.u, h.v, h.p), b = tex_value(S, mv, h.u, h.v, h.p);
        return mk(a.x != 0.0f ? b.x / a.x : 1.0f, a.y != 0.0f ? b.y / a.y : 1.0f, a.z != 0.0f ? b.z / a.z : 1.0f);
    }
    if (ms.kind == M_ROUGH_SPECTRAL) {
        bool refl = dot(wi, h.n) > 0.0f;
        double ns = rsd_ior_d(ms, lambda), ne = rsd_ior_d(mv, lambda);
        double ls, le;
        bool vs = rsd_logf<double>(wo, wi, h.n, h.front ? ns : 1.0 / ns, (double)rsd_alpha(ms, h), refl, ls);
        bool ve = rsd_logf<double>(wo, wi, h.n, h.front ? ne : 1.0 / ne, (double)rsd_alpha(mv, h), refl, le);
        if (!vs) return mk(1, 1, 1);
        if (!ve) return mk(0, 0, 0);
        float r = (float)exp(le - ls);
        return mk(r, r, r);
    }
    return mk(1, 1, 1);
}

// ---- constant-medium densities ---------------------------------------------------------------
// A parameterised medium object: its leaf slot, parameter index, and density
// at theta_s (ds, the scene's) and at theta (de).  A traced segment has log
// density -sigma L (no event before its end, L the in-medium length) plus
// log sigma when it ends in that medium's event; a shadow ray's
// transmittance is exp(-sigma L).
struct MedPar { int32_t slot, param; float ds, de; };
RT_D float medium_overlap(const DScene& S, const DObject& ob, const Ray& wr, float t_min, float t_end) {
    Ray r = to_object(S, ob, wr);
    float a, b;
    if (!boundary_t(S, ob, r, -FINF, FINF, a)) return 0.0f;
    if (!boundary_t(S, ob, r, a + 0.0001f, FINF, b)) return 0.0f;
    float t1 = fmx(fmx(a, 0.0f), t_min), t2 = fmn(b, t_end);
    return t1 < t2 ? (t2 - t1) * vlen(r.d) : 0.0f;
}
RT_D float segment_length(const DScene& S, const MedPar& m, const Ray& r, const HitId& id) {
    return medium_overlap(S, S.objs[m.slot], r, 0.001f, id.slot < 0 ? FINF : id.t);
}
RT_D bool segment_event(const MedPar& m, const HitId& id) { return id.prim == -2 && id.slot == m.slot; }
RT_D float shadow_length(const DScene& S, const MedPar& m, const ShadowReq& q) {
    return medium_overlap(S, S.objs[m.slot], make_ray(q.o, q.d, q.time), 0.001f, q.t_max);
}

// ---- hooks -----------------------------------------------------------------------------------
struct DetachedHook {
    const DMaterial* me;
    const MedPar* mp;
    int nmp;
    V3 ratio;
    RT_D void segment(const DScene& S, const Ray& r, const HitId& id) {
        for (int k = 0; k < nmp; k++) {
            float f = expf(-(mp[k].de - mp[k].ds) * segment_length(S, mp[k], r, id));
            if (segment_event(mp[k], id)) f *= mp[k].de / mp[k].ds;
            ratio *= f;
        }
    }
    RT_D V3 emit(V3 c) { return c * ratio; }
    RT_D V3 emit_vertex(V3 c, const DScene& S, const Rec& h, const Ray& r) {
        V3 wo = unit(-r.d);
        return c * ratio * vertex_ratio(S, me, h, wo, wo, r.wavelength);
    }
    RT_D void shadow(ShadowReq& q, int, const DScene& S, const Rec& h, const Ray& r, bool vdep) {
        q.contrib *= ratio;
        if (vdep) q.contrib *= vertex_ratio(S, me, h, unit(-r.d), unit(q.d), r.wavelength);
        for (int k = 0; k < nmp; k++) q.contrib *= expf(-(mp[k].de - mp[k].ds) * shadow_length(S, mp[k], q));
    }
    RT_D void scatter(const DScene& S, const Rec& h, const Ray& r, V3 d) {
        ratio *= vertex_ratio(S, me, h, unit(-r.d), unit(d), r.wavelength);
    }
    RT_D void begin_vertex() {}
    RT_D void resolved(int, V3, const ShadowReq&, const DScene&) {}
    RT_D void end_vertex() {}
};

struct ReplayHook {
    const int8_t* fmap;
    const MedPar* mp;
    int nmp;
    float* g;             // per-block LDS accumulators, one per parameter
    V3 delta;             // dLoss / d(sample radiance)
    V3 rem;               // radiance of this sample not yet collected
    Score vs, pend;
    bool vs_ok, pend_ok;
    bool vdep[3];
    RT_D void add(const Score& sc, V3 c) {
        for (int k = 0; k < sc.n; k++) {
            float v = delta.x * c.x * sc.v[k].x + delta.y * c.y * sc.v[k].y + delta.z * c.z * sc.v[k].z;
            if (v != 0.0f && isfinite(v)) atomicAdd(g + sc.p[k], v);
        }
    }
    RT_D void add1(int p, float s, V3 c) {
        float v = (delta.x * c.x + delta.y * c.y + delta.z * c.z) * s;
        if (v != 0.0f && isfinite(v)) atomicAdd(g + p, v);
    }
    // the radiance still to come travelled this segment
    RT_D void segment(const DScene& S, const Ray& r, const HitId& id) {
        for (int k = 0; k < nmp; k++) {
            float s = -segment_length(S, mp[k], r, id);
            if (segment_event(mp[k], id)) s += 1.0f / mp[k].ds;
            add1(mp[k].param, s, rem);
        }
    }
    RT_D V3 emit(V3 c) { rem -= c; return c; }
    RT_D V3 emit_vertex(V3 c, const DScene& S, const Rec& h, const Ray& r) {
        rem -= c;
        Score sc;
        V3 wo = unit(-r.d);
        vertex_score(S, fmap, h, wo, wo, r.wavelength, sc);
        add(sc, c);
        return c;
    }
    RT_D void shadow(ShadowReq& q, int k, const DScene& S, const Rec& h, const Ray& r, bool vd) {
        vdep[k] = vd;
        if (vd && !vs_ok) { vertex_score(S, fmap, h, unit(-r.d), unit(q.d), r.wavelength, vs); vs_ok = true; }
    }
    RT_D void scatter(const DScene& S, const Rec& h, const Ray& r, V3 d) {
        vertex_score(S, fmap, h, unit(-r.d), unit(d), r.wavelength, pend);
        pend_ok = true;
    }
    RT_D void begin_vertex() { vs_ok = false; pend_ok = false; vdep[0] = vdep[1] = vdep[2] = false; }
    RT_D void resolved(int k, V3 c, const ShadowReq& q, const DScene& S) {
        rem -= c;
        if (vdep[k] && vs_ok) add(vs, c);
        for (int j = 0; j < nmp; j++) add1(mp[j].param, -shadow_length(S, mp[j], q), c);
    }
    // radiance still to come arrives through this vertex's sampled direction
    RT_D void end_vertex() { if (pend_ok) add(pend, rem); }
};

struct PlainHook : NoDiff {
    RT_D void begin_vertex() {}
    RT_D void resolved(int, V3, const ShadowReq&, const DScene&) {}
    RT_D void end_vertex() {}
};

// ray_color with hooks: the megakernel loop of ray_color (src/renderer.rs:371-585).
template <class DH>
RT_D V3 ray_color_t(const DScene& S, Ray ray, int max_depth, float bg_scale, bool use_photons, Rng& rng, Stack& st, DH& dh) {
    PathState ps;
    path_init(ps, ray);
    V3 color = mk(0, 0, 0);
    ShadowReq sh[3];
    for (int depth = 0; depth < max_depth; depth++) {
        HitId id = trace_closest(S, ps.ray, 0.001f, FINF, st);
        int nsh = 0;
        dh.begin_vertex();
        bool cont = bounce_t(S, ps, id, depth, bg_scale, use_photons, rng, color, sh, nsh, st, dh);
        for (int k = 0; k < nsh; k++) {
            float tr = trace_transmittance(S, make_ray(sh[k].o, sh[k].d, sh[k].time), 0.001f, sh[k].t_max, st, false);
            if (tr > 0.0f) { V3 c = sh[k].contrib * tr; color += c; dh.resolved(k, c, sh[k], S); }
        }
        dh.end_vertex();
        if (!cont) break;
    }
    return color;
}

}  // namespace rt
