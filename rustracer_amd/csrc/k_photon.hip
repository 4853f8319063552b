// Photon emission and tracing.
#include "rt_kern.h"

__global__ __launch_bounds__(BLOCK) void k_photon_trace(DScene S, DEmitter E, uint32_t n, float4* __restrict__ out, uint32_t* __restrict__ ok) {
    RT_STACK(st);
    uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    Rng rng = rng_seed(((uint64_t)i * 6364136223846793005ULL) ^ 0x9E3779B97F4A7C15ULL);
    V3 pos, dir;
    if (E.kind == 0) {
        float s = rng_f(rng), t = rng_f(rng);
        pos = mk(E.o[0], E.o[1], E.o[2]) + mk(E.u[0], E.u[1], E.u[2]) * s + mk(E.v[0], E.v[1], E.v[2]) * t;
        float ct = sqrtf(rng_f(rng));
        float sn = sqrtf(1.0f - ct * ct);
        float phi = 2.0f * PI * rng_f(rng);
        dir = sn * cosf(phi) * mk(E.ta[0], E.ta[1], E.ta[2]) + sn * sinf(phi) * mk(E.ba[0], E.ba[1], E.ba[2]) + ct * mk(E.n[0], E.n[1], E.n[2]);
    } else {
        float r = E.disk_r * sqrtf(rng_f(rng));
        float phi = 2.0f * PI * rng_f(rng);
        pos = mk(E.disk_c[0], E.disk_c[1], E.disk_c[2]) + mk(E.t[0], E.t[1], E.t[2]) * (r * cosf(phi)) + mk(E.b[0], E.b[1], E.b[2]) * (r * sinf(phi));
        dir = mk(E.sun_down[0], E.sun_down[1], E.sun_down[2]);
    }
    V3 power = mk(E.power[0], E.power[1], E.power[2]);
    uint32_t spec_depth = 0;
    float wl = rng_range(rng, 380.0f, E.wide ? 780.0f : 700.0f);
    uint32_t stored = 0;
    for (int d = 0; d < 12; d++) {
        Ray ray = make_ray(pos, dir, 0.0f);
        ray.wavelength = wl;
        ray.wl = E.wide ? 1 : 0;
        HitId id = trace_closest(S, ray, 0.001f, FINF, st);
        if (id.slot < 0) break;
        Rec h = finalize(S, ray, id, 0.001f);
        Scatter sr;
        if (!mat_scatter(S, h, ray, rng, sr)) break;
        if (sr.skip_pdf) {
            power *= sr.att;
            pos = sr.ray.o; dir = sr.ray.d;
            spec_depth++;
        } else {
            if (spec_depth > 0 && mat_receives_caustics(S, h)) {
                float l = lum(power);
                if (l > E.lum_clamp) power *= E.lum_clamp / l;
                out[3 * (size_t)i] = make_float4(h.p.x, h.p.y, h.p.z, 0.0f);
                out[3 * (size_t)i + 1] = make_float4(power.x, power.y, power.z, 0.0f);
                out[3 * (size_t)i + 2] = make_float4(h.n.x, h.n.y, h.n.z, 0.0f);
                stored = 1;
            }
            break;
        }
    }
    ok[i] = stored;
}

// The direction and colour of SceneData::rebuild_caustics' emitter
// (src/scene.rs:35-74), one thread: out[0..3) direction, out[3..6) colour,
// out[6] = 1 Physical sun, 2 env map, 0 neither (the caller's quad).
__global__ void k_caustic_emitter(DScene S, float* out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    V3 dir = mk(0, 0, 0), c = mk(0, 0, 0);
    float kind = 0.0f;
    if (S.bg_kind == BGK_PHYSICAL) {
        V3 sd = mk(S.sun_dir[0], S.sun_dir[1], S.sun_dir[2]);
        V3 perp = fabsf(sd.x) < 0.9f ? unit(mk(0.0f, -sd.z, sd.y)) : unit(mk(-sd.z, 0.0f, sd.x));
        V3 pd = unit(sd + perp * 0.04f);
        dir = sd;
        c = bg_eval(S, pd) * PI;
        kind = 1.0f;
    } else if (has_env(S)) {
        Rng rng = rng_seed(0xC0FFEEull);
        dir = mk(0.0f, 1.0f, 0.0f);
        for (int k = 0; k < 8; k++) {
            float pdf;
            dir = dir + env_sample_dir(S, rng, pdf, true);
        }
        dir = unit(dir);
        c = env_sample(S, dir) * PI;
        kind = 2.0f;
    }
    out[0] = dir.x; out[1] = dir.y; out[2] = dir.z;
    out[3] = c.x; out[4] = c.y; out[5] = c.z;
    out[6] = kind;
}
