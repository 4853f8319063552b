// Small kernels of librtamd: trace batches, pixel-state bookkeeping, image resolve,
// the AOV pass, the E_ss table, the wavefront counter kernels and the photon
// irradiance queries.
#include "rt_kern.h"

__global__ __launch_bounds__(BLOCK) void k_trace_closest(DScene S, const rt_ray* __restrict__ rays, uint32_t n, rt_hit* __restrict__ hits) {
    RT_STACK(st);
    uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    rt_ray q = rays[i];
    Ray r = ray_from_c(q);
    HitId id = trace_closest(S, r, q.t_min, q.t_max, st);
    rt_hit h;
    memset(&h, 0, sizeof h);
    if (id.slot >= 0) {
        Rec rec = finalize(S, r, id, q.t_min);
        h.t = rec.t; h.p[0] = rec.p.x; h.p[1] = rec.p.y; h.p[2] = rec.p.z;
        h.n[0] = rec.n.x; h.n[1] = rec.n.y; h.n[2] = rec.n.z;
        h.u = rec.u; h.v = rec.v; h.obj_id = rec.obj_id; h.prim_id = rec.prim_id; h.front_face = rec.front ? 1u : 0u;
    } else { h.obj_id = -1; h.prim_id = -1; }
    hits[i] = h;
}

__global__ __launch_bounds__(BLOCK) void k_trace_tr(DScene S, const rt_ray* __restrict__ rays, uint32_t n, float* __restrict__ tr, uint8_t* __restrict__ occ) {
    RT_STACK(st);
    uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    rt_ray q = rays[i];
    Ray r = ray_from_c(q);
    if (occ) occ[i] = trace_transmittance(S, r, q.t_min, q.t_max, st, true) <= 0.0f ? 1 : 0;
    else tr[i] = trace_transmittance(S, r, q.t_min, q.t_max, st, false);
}

// accumulate_sample (src/main.rs:603-644), one thread per pixel.
__global__ __launch_bounds__(BLOCK) void k_accumulate(const float* __restrict__ sc, uint32_t npix, rt_accum_opts o,
                                                      float* acc, uint32_t* ns, float* m2, const uint8_t* conv) {
    uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= npix || conv[i]) return;
    accumulate_pixel(acc + 3 * (size_t)i, ns + i, m2 + 3 * (size_t)i, mk(sc[3 * (size_t)i], sc[3 * (size_t)i + 1], sc[3 * (size_t)i + 2]), o);
}

// mark_converged (src/main.rs:646-677): flag, then count every flagged pixel.
__global__ __launch_bounds__(BLOCK) void k_mark_converged(uint32_t npix, float threshold, const float* acc, const uint32_t* ns,
                                                          const float* m2, uint8_t* conv, uint32_t* count) {
    uint32_t local = 0;
    for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i < npix; i += gridDim.x * BLOCK) {
        bool c = conv[i] != 0;
        if (!c && pixel_converged(acc + 3 * (size_t)i, ns[i], m2 + 3 * (size_t)i, threshold)) { conv[i] = 1; c = true; }
        local += c ? 1u : 0u;
    }
    wave_count(local, count);
}

// Flagged pixels of a state handed to rt_render_progressive.
__global__ __launch_bounds__(BLOCK) void k_count_flags(uint32_t npix, const uint8_t* conv, uint32_t* count) {
    uint32_t local = 0;
    for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i < npix; i += gridDim.x * BLOCK) local += conv[i] != 0 ? 1u : 0u;
    wave_count(local, count);
}

// Mean radiance of pixel j: accum / max(n, 1) per component (src/main.rs:574-577).
__device__ __forceinline__ V3 resolve_raw(const float* acc, const uint32_t* ns, uint32_t spp, uint32_t j) {
    float n = (float)max(ns ? ns[j] : spp, 1u);
    return mk(acc[3 * (size_t)j] / n, acc[3 * (size_t)j + 1] / n, acc[3 * (size_t)j + 2] / n);
}
// rt_resolve_rgb8: apply_optics (src/output.rs:150-186) reading the means on
// the fly, then tone_map (src/output.rs:103-111); one thread per pixel.
__global__ __launch_bounds__(BLOCK) void k_resolve(const float* __restrict__ acc, const uint32_t* __restrict__ ns, uint32_t spp,
                                                   uint32_t w, uint32_t h, rt_resolve_opts o, uint8_t* __restrict__ out) {
    uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= w * h) return;
    uint32_t y = i / w, x = i - y * w;
    V3 c;
    if (o.vignette == 0.0f && o.distortion == 0.0f) {
        c = resolve_raw(acc, ns, spp, i);
    } else {
        float cx = ((float)w - 1.0f) * 0.5f, cy = ((float)h - 1.0f) * 0.5f;
        float rc = fmaxf(sqrtf(cx * cx + cy * cy), 1e-6f);
        float dx = ((float)x - cx) / rc, dy = ((float)y - cy) / rc;
        float r2 = dx * dx + dy * dy;
        float s = 1.0f - o.distortion * r2;
        float sx = fclamp(cx + dx * rc * s, 0.0f, (float)w - 1.0f);
        float sy = fclamp(cy + dy * rc * s, 0.0f, (float)h - 1.0f);
        uint32_t x0 = f2u(sx), y0 = f2u(sy);
        uint32_t x1 = min(x0 + 1, w - 1), y1 = min(y0 + 1, h - 1);
        float fx = sx - (float)x0, fy = sy - (float)y0;
        V3 c00 = resolve_raw(acc, ns, spp, y0 * w + x0), c10 = resolve_raw(acc, ns, spp, y0 * w + x1);
        V3 c01 = resolve_raw(acc, ns, spp, y1 * w + x0), c11 = resolve_raw(acc, ns, spp, y1 * w + x1);
        c = (c00 * (1.0f - fx) + c10 * fx) * (1.0f - fy) + (c01 * (1.0f - fx) + c11 * fx) * fy;
        c = c * fmaxf(1.0f - o.vignette * r2, 0.0f);
    }
    V3 m = o.tonemapper == RT_TONE_AGX ? agx_tonemap(c * o.exposure) : aces_tonemap(c * o.exposure);
    out[3 * (size_t)i] = srgb_byte(m.x);
    out[3 * (size_t)i + 1] = srgb_byte(m.y);
    out[3 * (size_t)i + 2] = srgb_byte(m.z);
}

// After pass `pass`: count it, and stop when adaptive and the converged count
// reaches convergence_pct % of the pixels (the reference's target_reached).
__global__ void k_progressive_step(uint32_t* ctl, uint32_t pass, uint32_t npix, rt_accum_opts o) {
    if (threadIdx.x != 0 || ctl[0]) return;
    ctl[2] = pass + 1;
    if (o.adaptive && (float)ctl[1] >= (float)npix * o.convergence_pct / 100.0f) ctl[0] = 1;
}

// render_aux_pass (src/renderer.rs:597-645): per pixel one ray through the
// pixel centre (its own stream, seeded i ^ 0x9E3779B97F4A7C15, for the lens
// and time draws); a hit gives the material's albedo hint clamped to [0, 1]
// and the hit normal, a miss the background colour and a zero normal.
__global__ __launch_bounds__(BLOCK) void k_aux(DScene S, DCamera cam, uint32_t w, uint32_t h, float* __restrict__ alb,
                                               float* __restrict__ nrm) {
    RT_STACK(st);
    uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= w * h) return;
    uint32_t row = i / w, col = i - row * w;
    Rng rng = rng_seed((uint64_t)i ^ 0x9E3779B97F4A7C15ull);
    float u = ((float)col + 0.5f) / (float)max(w - 1, 1u);
    float v = ((float)(h - 1 - row) + 0.5f) / (float)max(h - 1, 1u);
    Ray r = camera_ray(cam, u, v, rng);
    HitId id = trace_closest(S, r, 0.001f, FINF, st);
    V3 a, n;
    if (id.slot < 0) {
        a = bg_eval(S, r.d);
        n = mk(0, 0, 0);
    } else {
        Rec hr = finalize(S, r, id, 0.001f);
        V3 c = mat_albedo_hint(S, hr);
        a = mk(fclamp(c.x, 0.0f, 1.0f), fclamp(c.y, 0.0f, 1.0f), fclamp(c.z, 0.0f, 1.0f));
        n = hr.n;
    }
    alb[3 * (size_t)i] = a.x; alb[3 * (size_t)i + 1] = a.y; alb[3 * (size_t)i + 2] = a.z;
    nrm[3 * (size_t)i] = n.x; nrm[3 * (size_t)i + 1] = n.y; nrm[3 * (size_t)i + 2] = n.z;
}

// E_ss table (src/material.rs:206-243): one thread per (mu, alpha) texel.
__global__ void k_elut(float* out) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 64 * 64) return;
    uint32_t i = t / 64, j = t % 64;
    float mu = ((float)i + 0.5f) / 64.0f;
    float alpha = fmaxf(((float)j + 0.5f) / 64.0f, 1e-3f);
    Rng rng = rng_seed((uint64_t)(i * 64 + j) + 0x9E3779B9ULL);
    V3 wo = mk(0.0f, 0.0f, mu);
    float lo = smith_lambda(mu, alpha);
    double sum = 0.0;
    for (int k = 0; k < 64; k++)
        for (int l = 0; l < 64; l++) {
            float xi1 = ((float)k + rng_f(rng)) / 64.0f;
            float xi2 = ((float)l + rng_f(rng)) / 64.0f;
            V3 m = vndf_sample(wo, alpha, alpha, xi1, xi2);
            float voh = fmaxf(dot(wo, m), 0.0f);
            V3 wi = 2.0f * voh * m - wo;
            if (wi.z <= 0.0f) continue;
            float li = smith_lambda(wi.z, alpha);
            sum += (double)((1.0f + lo) / (1.0f + lo + li));
        }
    out[t] = (float)(sum / 4096.0);
}

// Zero the per-bounce counters (a kernel, not a memset, so graph replays are safe).
__global__ void k_wf_reset(uint32_t* counters, uint32_t active, const uint32_t* stop, int generated) {
    uint32_t i = threadIdx.x;
    if (generated) active = counters[WF_CNT_GEN];
    if (stop && *stop) active = 0;
    if (i < 32) counters[i] = (i == WF_CNT_ACTIVE) ? active : 0u;
    if (i < WF_NB) counters[WF_CNT_BKT(i)] = 0u;
}
__global__ void k_wf_gen_zero(uint32_t* counters) {
    if (threadIdx.x == 0) counters[WF_CNT_GEN] = 0u;
}
// After shade: survivors become the active set, everything else restarts at 0.
__global__ void k_wf_advance(uint32_t* counters) {
    uint32_t i = threadIdx.x;
    __shared__ uint32_t nxt;
    if (i == 0) nxt = counters[WF_CNT_NEXT];
    __syncthreads();
    if (i < 32) counters[i] = (i == WF_CNT_ACTIVE) ? nxt : 0u;
    if (i < WF_NB) counters[WF_CNT_BKT(i)] = 0u;
}

// counters[WF_CNT_FIRST..+1] hold the frame position of a graph-replayed
// chunk; they lie past the 32 per-bounce counters that reset/advance clear.
__global__ void k_wf_set_first(uint32_t* counters, uint64_t first) {
    if (threadIdx.x == 0) { counters[WF_CNT_FIRST] = (uint32_t)first; counters[WF_CNT_FIRST + 1] = (uint32_t)(first >> 32); }
}
__global__ void k_wf_bump_first(uint32_t* counters, uint32_t n) {
    if (threadIdx.x == 0) {
        uint64_t f = ((uint64_t)counters[WF_CNT_FIRST + 1] << 32) | counters[WF_CNT_FIRST];
        f += n;
        counters[WF_CNT_FIRST] = (uint32_t)f; counters[WF_CNT_FIRST + 1] = (uint32_t)(f >> 32);
    }
}

// The wavefront render loop's bookkeeping: K passes of per-sample radiance
// (lsamp, sample id k * npix + j for path j of pass k, which covered pixel
// tile_order_pixel(j)) into the pixel state, pass by pass per pixel as
// accumulate_sample does, with mark_converged after each pass when adaptive
// (K is then 1, so the stop test runs between passes).
__global__ __launch_bounds__(BLOCK) void k_accum_passes(const float* __restrict__ lsamp, uint32_t K, uint32_t w, uint32_t h,
                                                        uint32_t pass0, rt_accum_opts o, float* acc, uint32_t* ns, float* m2,
                                                        uint8_t* conv, uint32_t* ctl) {
    if (ctl[0]) return;
    uint32_t npix = w * h;
    uint32_t local = 0;
    for (uint32_t j = blockIdx.x * BLOCK + threadIdx.x; j < npix; j += gridDim.x * BLOCK) {
        uint32_t px = tile_order_pixel(j, w, h);
        float* a = acc + 3 * (size_t)px;
        float* v = m2 + 3 * (size_t)px;
        for (uint32_t k = 0; k < K && !conv[px]; k++) {
            const float* c = lsamp + 3 * ((size_t)k * npix + j);
            accumulate_pixel(a, ns + px, v, mk(c[0], c[1], c[2]), o);
            if (o.adaptive && pass0 + k + 1 >= o.min_samples && pixel_converged(a, ns[px], v, o.threshold)) {
                conv[px] = 1;
                local++;
            }
        }
    }
    wave_count(local, ctl + 1);
}

template <bool KNN>
__global__ __launch_bounds__(BLOCK) void k_photon_irr(DScene S, const float* __restrict__ pn, uint32_t n, float* __restrict__ out) {
    uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    Stack st;
    st.heap = KNN ? LdsHeap::column() : nullptr; st.has_heap = KNN; st.stride = BLOCK; st.n_pbytes = 0;
    V3 c = photon_irradiance(S, mk(pn[6 * i], pn[6 * i + 1], pn[6 * i + 2]), mk(pn[6 * i + 3], pn[6 * i + 4], pn[6 * i + 5]), mk(0, 0, 0), st);
    out[3 * i] = c.x; out[3 * i + 1] = c.y; out[3 * i + 2] = c.z;
}

template __global__ void k_photon_irr<true>(DScene, const float*, uint32_t, float*);
template __global__ void k_photon_irr<false>(DScene, const float*, uint32_t, float*);

// Fill n words with v (grid-stride): clears buffers without hipMemsetAsync,
// so a caller may capture the launch in a graph.
__global__ __launch_bounds__(BLOCK) void k_fill_u32(uint32_t* p, uint64_t n, uint32_t v) {
    for (uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (uint64_t)gridDim.x * BLOCK) p[i] = v;
}
