// Walter et al. (2007) rough dielectric in log form (extension: the
// reference has no rough dispersive dielectric): the BSDF value log f and the
// sampler's density of a transmitted direction, over T = double or a dual
// number type, for the detached renderer and adjoint scores (ratios of two
// parameter sets) and for the four-wavelength mode (ratios of two
// wavelengths).  The same formulas as the sampler (fresnel_dielectric,
// smith_g1, ggx_ndf).
#pragma once

namespace rt {

RT_HD float val(float a) { return a; }
RT_HD float xsqrt(float a) { return sqrtf(a); }
RT_HD float xlog(float a) { return logf(a); }
RT_HD float xabs(float a) { return fabsf(a); }
RT_HD double val(double a) { return a; }
RT_HD double xsqrt(double a) { return sqrt(a); }
RT_HD double xlog(double a) { return log(a); }
RT_HD double xabs(double a) { return fabs(a); }

template <class T> RT_HD T t_fresnel(T ci, T eta) {
    T s2i = T(1.0f) - ci * ci;
    if (val(s2i) < 0.0f) s2i = T(0.0f);
    T s2t = eta * eta * s2i;
    if (val(s2t) >= 1.0f) return T(1.0f);
    T ct = xsqrt(T(1.0f) - s2t);
    T rs = (eta * ci - ct) / (eta * ci + ct);
    T rp = (eta * ct - ci) / (eta * ct + ci);
    return (rs * rs + rp * rp) * 0.5f;
}
template <class T> RT_HD T t_g1(float c, T al) { T a2 = al * al; return (2.0f * c) / (c + xsqrt(a2 + (T(1.0f) - a2) * (c * c))); }
template <class T> RT_HD T t_ggx(T ch, T al) { T a2 = al * al; T d = T(1.0f) + ch * ch * (a2 - 1.0f); return a2 / (PI * d * d); }

// log f(i, o) for unit i (toward the previous vertex), o (the new direction),
// shading normal n; k = IOR on o's side over IOR on i's side for a
// transmission (the Fresnel eta is 1/k in both cases).  False where f = 0.
template <class T>
RT_HD bool rsd_logf(V3 i, V3 o, V3 n, T k, T al, bool refl, T& lf) {
    float cin = dot(i, n), con = dot(o, n);
    float ain = fabsf(cin), aon = fabsf(con);
    if (!(ain > 0.0f) || !(aon > 0.0f)) return false;
    T eta = T(1.0f) / k;
    T G = t_g1(ain, al) * t_g1(aon, al);
    if (refl) {
        // half vector in T's precision: at alpha ~ 0.005, 1 - (h.n)^2 is a
        // few hundred float ulps
        T hx = T(i.x) + o.x, hy = T(i.y) + o.y, hz = T(i.z) + o.z;
        T hl = xsqrt(hx * hx + hy * hy + hz * hz);
        if (!(val(hl) > 0.0f)) return false;
        T ih = (hx * i.x + hy * i.y + hz * i.z) / hl;
        if (!(val(ih) * cin > 0.0f)) return false;
        if (val(ih) > 1.0f) ih = T(1.0f);
        T F = t_fresnel(ih, eta);
        T D = t_ggx((hx * n.x + hy * n.y + hz * n.z) / hl, al);
        if (!(val(F) > 0.0f) || !(val(D) > 0.0f) || !(val(G) > 0.0f)) return false;
        lf = xlog(F) + xlog(G) + xlog(D) - logf(4.0f * ain * aon);
        return true;
    }
    T hx = i.x + k * o.x, hy = i.y + k * o.y, hz = i.z + k * o.z;
    T len = xsqrt(hx * hx + hy * hy + hz * hz);
    if (!(val(len) > 0.0f)) return false;
    hx = hx / len; hy = hy / len; hz = hz / len;
    T hn = hx * n.x + hy * n.y + hz * n.z;
    if (val(hn) < 0.0f) { hx = -hx; hy = -hy; hz = -hz; hn = -hn; }
    T ih = hx * i.x + hy * i.y + hz * i.z;
    T oh = hx * o.x + hy * o.y + hz * o.z;
    // the microfacet must face both directions from their own sides
    if (!(val(ih) * cin > 0.0f) || !(val(oh) * con > 0.0f)) return false;
    T aih = xabs(ih);
    if (val(aih) > 1.0f) aih = T(1.0f);
    T tr = T(1.0f) - t_fresnel(aih, eta);
    T den = ih + k * oh;
    T D = t_ggx(hn, al);
    if (!(val(tr) > 0.0f) || !(val(D) > 0.0f) || !(val(G) > 0.0f) || val(ih) == 0.0f || val(oh) == 0.0f || val(den) == 0.0f)
        return false;
    lf = xlog(tr) + xlog(G) + xlog(D) + xlog(aih) + xlog(xabs(oh)) + 2.0f * xlog(k) - 2.0f * xlog(xabs(den)) - logf(ain * aon);
    return true;
}


// log of the sampler's density of the transmitted direction o (rough_dielectric_core:
// a visible microfacet m drawn from i, then refraction with probability 1 - F):
//   p(o) = (1 - F(|i.m|)) G1(i) |i.m| D(m) / |i.n|  *  k^2 |o.m| / (i.m + k o.m)^2,
// m the generalised half vector of (i, o), k as in rsd_logf.  The BSDF value
// over it is G1(o) / |o.n|, the sampler's weight.  False where p = 0.
template <class T>
RT_HD bool rsd_logpdf_t(V3 i, V3 o, V3 n, T k, T al, T& lp) {
    float cin = dot(i, n), con = dot(o, n);
    float ain = fabsf(cin);
    if (!(ain > 0.0f) || !(fabsf(con) > 0.0f)) return false;
    T eta = T(1.0f) / k;
    T hx = i.x + k * o.x, hy = i.y + k * o.y, hz = i.z + k * o.z;
    T len = xsqrt(hx * hx + hy * hy + hz * hz);
    if (!(val(len) > 0.0f)) return false;
    hx = hx / len; hy = hy / len; hz = hz / len;
    T hn = hx * n.x + hy * n.y + hz * n.z;
    if (val(hn) < 0.0f) { hx = -hx; hy = -hy; hz = -hz; hn = -hn; }
    T ih = hx * i.x + hy * i.y + hz * i.z;
    T oh = hx * o.x + hy * o.y + hz * o.z;
    if (!(val(ih) * cin > 0.0f) || !(val(oh) * con > 0.0f)) return false;
    T aih = xabs(ih);
    if (val(aih) > 1.0f) aih = T(1.0f);
    T tr = T(1.0f) - t_fresnel(aih, eta);
    T den = ih + k * oh;
    T D = t_ggx(hn, al);
    T g1 = t_g1(ain, al);
    if (!(val(tr) > 0.0f) || !(val(D) > 0.0f) || !(val(g1) > 0.0f) || val(den) == 0.0f) return false;
    lp = xlog(tr) + xlog(g1) + xlog(aih) + xlog(D) - logf(ain) + 2.0f * xlog(k) + xlog(xabs(oh)) - 2.0f * xlog(xabs(den));
    return true;
}

}  // namespace rt
