// ORACLE — test infrastructure only (see rtcore.cpp).
//
// Walter et al. (2007) rough dielectric in double precision (extension: the
// reference has no rough dispersive dielectric): log f of the microfacet BSDF
// (GGX D, separable Smith G1, exact unpolarised Fresnel) and log p of the
// RoughSpectralDielectric sampler's density for a transmitted direction.  The
// adjoint checker takes ratios of two parameter sets, the four-wavelength
// render ratios of two wavelengths.
#pragma once
#include "rc_math.h"
#include <cmath>
#include <algorithm>

namespace rc {

struct D3 { double x, y, z; };
static inline D3 d3(const Vec3& v) { return D3{v.x, v.y, v.z}; }
static inline double ddot(const D3& a, const D3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline D3 dnorm(const D3& a) { double l = std::sqrt(ddot(a, a)); return D3{a.x / l, a.y / l, a.z / l}; }

static inline double fresnel_d(double ci, double eta) {
    double s2i = std::max(1.0 - ci * ci, 0.0);
    double s2t = eta * eta * s2i;
    if (s2t >= 1.0) return 1.0;
    double ct = std::sqrt(1.0 - s2t);
    double rs = (eta * ci - ct) / (eta * ci + ct);
    double rp = (eta * ct - ci) / (eta * ct + ci);
    return 0.5 * (rs * rs + rp * rp);
}
static inline double g1_d(double c, double a) { double a2 = a * a; return 2.0 * c / (c + std::sqrt(a2 + (1.0 - a2) * c * c)); }
static inline double ggx_d(double ch, double a) { double a2 = a * a, d = 1.0 + ch * ch * (a2 - 1.0); return a2 / (M_PI * d * d); }

// log f(i, o): i toward the previous vertex, o the new direction, n the
// shading normal; k = IOR beyond the surface (o's side for a transmission)
// over IOR on i's side.  False where f = 0.
static inline bool logf_rsd(D3 i, D3 o, D3 n, double k, double a, bool refl, double& lf) {
    i = dnorm(i); o = dnorm(o);
    double ci = ddot(i, n), co = ddot(o, n);
    if (!(std::fabs(ci) > 0.0) || !(std::fabs(co) > 0.0)) return false;
    double G = g1_d(std::fabs(ci), a) * g1_d(std::fabs(co), a);
    double eta = 1.0 / k;
    if (refl) {
        D3 h = dnorm(D3{i.x + o.x, i.y + o.y, i.z + o.z});
        if (!(ddot(i, h) * ci > 0.0)) return false;
        double F = fresnel_d(std::min(ddot(i, h), 1.0), eta);
        double D = ggx_d(ddot(h, n), a);
        if (!(F > 0.0) || !(D > 0.0) || !(G > 0.0)) return false;
        lf = std::log(F * G * D / (4.0 * std::fabs(ci) * std::fabs(co)));
        return true;
    }
    D3 h = dnorm(D3{i.x + k * o.x, i.y + k * o.y, i.z + k * o.z});
    if (ddot(h, n) < 0.0) h = D3{-h.x, -h.y, -h.z};
    double ih = ddot(i, h), oh = ddot(o, h);
    // the microfacet must face both directions from their own sides
    if (!(ih * ci > 0.0) || !(oh * co > 0.0)) return false;
    double T = 1.0 - fresnel_d(std::min(std::fabs(ih), 1.0), eta);
    double D = ggx_d(ddot(h, n), a);
    double den = ih + k * oh;
    if (!(T > 0.0) || !(D > 0.0) || !(G > 0.0) || ih == 0.0 || oh == 0.0 || den == 0.0) return false;
    lf = std::log(std::fabs(ih) * std::fabs(oh) * k * k * T * G * D / (std::fabs(ci) * std::fabs(co) * den * den));
    return true;
}


// log of the sampler's density of the transmitted direction o: the visible
// normal m drawn from i (density G1(i) max(0, i.m) D(m) / (i.n)), refraction
// taken with probability 1 - F(|i.m|), and the change of variables from m to
// o, k^2 |o.m| / (i.m + k o.m)^2.  False where the density is 0.
static inline bool logpdf_rsd_t(D3 i, D3 o, D3 n, double k, double a, double& lp) {
    i = dnorm(i); o = dnorm(o);
    double ci = ddot(i, n), co = ddot(o, n);
    if (!(std::fabs(ci) > 0.0) || !(std::fabs(co) > 0.0)) return false;
    D3 h = dnorm(D3{i.x + k * o.x, i.y + k * o.y, i.z + k * o.z});
    if (ddot(h, n) < 0.0) h = D3{-h.x, -h.y, -h.z};
    double ih = ddot(i, h), oh = ddot(o, h);
    if (!(ih * ci > 0.0) || !(oh * co > 0.0)) return false;
    double T = 1.0 - fresnel_d(std::min(std::fabs(ih), 1.0), 1.0 / k);
    double D = ggx_d(ddot(h, n), a);
    double G1 = g1_d(std::fabs(ci), a);
    double den = ih + k * oh;
    if (!(T > 0.0) || !(D > 0.0) || !(G1 > 0.0) || den == 0.0) return false;
    lp = std::log(T * G1 * std::fabs(ih) * D / std::fabs(ci) * k * k * std::fabs(oh) / (den * den));
    return true;
}

}  // namespace rc
