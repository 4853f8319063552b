// Photon emission and tracing.
#include "rt_kern.h"

__global__ __launch_bounds__(BLOCK) void k_photon_trace(DScene S, DEmitter E, uint32_t n, float4* __restrict__ out, uint32_t* __restrict__ ok) {
    Stack st; init_stack(st);
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
    float wl = rng_range(rng, 380.0f, 700.0f);
    uint32_t stored = 0;
    for (int d = 0; d < 12; d++) {
        Ray ray = make_ray(pos, dir, 0.0f);
        ray.wavelength = wl;
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
