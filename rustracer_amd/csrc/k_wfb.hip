// Wavefront path replay (adjoint) stages.
#include "rt_kern.h"

__global__ __launch_bounds__(BLOCK) void k_wfb_init(WfQueues Q, WfAdjQ A, uint32_t n) {
    uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const WfPath& w = Q.paths[i];
    V3 L = mk(A.lsamp[3 * i], A.lsamp[3 * i + 1], A.lsamp[3 * i + 2]);
    bool ok = isfinite(L.x) && isfinite(L.y) && isfinite(L.z);
    size_t px = 3 * (size_t)w.pixel;
    WfAdj a;
    a.rem = ok ? make_float4(L.x, L.y, L.z, 0.0f) : make_float4(0, 0, 0, 0);
    a.delta = ok ? make_float4(A.dimg[px] * A.inv_spp, A.dimg[px + 1] * A.inv_spp, A.dimg[px + 2] * A.inv_spp, 0.0f)
                 : make_float4(0, 0, 0, 0);
    a.pn = 0; a.vn = 0;
    A.adj[i] = a;
}

struct WfReplayHook {
    const int8_t* fmap;
    const MedPar* mp;
    int nmp;
    float* g;
    V3 delta, rem;
    Score pend;
    bool pend_ok;
    Score vs;
    bool vs_ok;
    bool vdep[3];
    RT_D void add(const Score& sc, V3 c) {
        for (int k = 0; k < sc.n; k++) {
            float v = delta.x * c.x * sc.v[k].x + delta.y * c.y * sc.v[k].y + delta.z * c.z * sc.v[k].z;
            if (v != 0.0f && isfinite(v)) atomicAdd(g + sc.p[k], v);
        }
    }
    // the radiance still to come travelled this segment (medium densities)
    RT_D void segment(const DScene& S, const Ray& r, const HitId& id) {
        for (int k = 0; k < nmp; k++) {
            float s = -segment_length(S, mp[k], r, id);
            if (segment_event(mp[k], id)) s += 1.0f / mp[k].ds;
            float v = (delta.x * rem.x + delta.y * rem.y + delta.z * rem.z) * s;
            if (v != 0.0f && isfinite(v)) atomicAdd(g + mp[k].param, v);
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
};

__global__ __launch_bounds__(BLOCK) void k_wfb_shade(DScene S, DParams P, WfQueues Q, WfAdjQ A, int depth) {
    __shared__ float g[RT_MAX_PARAMS];
    for (uint32_t t = threadIdx.x; t < A.nparams; t += BLOCK) g[t] = 0.0f;
    __syncthreads();
    RT_STACK(st);
    uint32_t off[WF_NB + 1];
    off[0] = 0;
    for (int b = 0; b < WF_NB; b++) off[b + 1] = off[b] + Q.counters[WF_CNT_BKT(b)];
    uint32_t total = off[WF_NB];
    uint32_t stride = gridDim.x * BLOCK;
    for (uint32_t base = blockIdx.x * BLOCK; base < total; base += stride) {
        uint32_t i = base + threadIdx.x;
        bool valid = i < total;
        bool cont = false;
        int nsh = 0;
        ShadowReq sh[3];
        bool vdep[3] = {false, false, false};
        uint32_t pid = 0;
        if (valid) {
            int b = 0;
            while (i >= off[b + 1]) b++;
            pid = Q.buckets[(size_t)b * Q.cap + (i - off[b])];
            WfPath& w = Q.paths[pid];
            WfAdj& a = A.adj[pid];
            WfReplayHook rh;
            rh.fmap = A.fmap; rh.mp = A.mp; rh.nmp = A.nmp; rh.g = g;
            rh.delta = mk(a.delta.x, a.delta.y, a.delta.z);
            rh.rem = mk(a.rem.x, a.rem.y, a.rem.z);
            rh.pend_ok = false; rh.vs_ok = false;
            rh.vdep[0] = rh.vdep[1] = rh.vdep[2] = false;
            // what is left now arrived through the previous vertex's sampled direction
            if (a.pn > 0) {
                Score pv;
                pv.n = a.pn;
                for (int k = 0; k < a.pn; k++) { pv.p[k] = a.pp[k]; pv.v[k] = mk(a.pv[3 * k], a.pv[3 * k + 1], a.pv[3 * k + 2]); }
                rh.add(pv, rh.rem);
            }
            if (depth < (int)P.max_depth) {
                PathState ps; Rng rng;
                wf_load(w, ps, rng);
                HitId id; id.t = w.hit_t; id.slot = w.hit_slot; id.prim = w.hit_prim;
                V3 color = mk(0, 0, 0);
                cont = bounce_t(S, ps, id, depth, P.bg_scale, P.use_photons != 0, rng, color, sh, nsh, st, rh) &&
                       depth + 1 < (int)P.max_depth;
                if (cont) wf_store(w, ps, rng);
            }
            a.rem = make_float4(rh.rem.x, rh.rem.y, rh.rem.z, 0.0f);
            a.pn = (cont && rh.pend_ok) ? rh.pend.n : 0;
            for (int k = 0; k < a.pn; k++) {
                a.pp[k] = (int8_t)rh.pend.p[k];
                a.pv[3 * k] = rh.pend.v[k].x; a.pv[3 * k + 1] = rh.pend.v[k].y; a.pv[3 * k + 2] = rh.pend.v[k].z;
            }
            a.vn = rh.vs_ok ? min(rh.vs.n, 3) : 0;
            for (int k = 0; k < a.vn; k++) {
                a.vp[k] = (int8_t)rh.vs.p[k];
                a.vv[3 * k] = rh.vs.v[k].x; a.vv[3 * k + 1] = rh.vs.v[k].y; a.vv[3 * k + 2] = rh.vs.v[k].z;
            }
            for (int k = 0; k < 3; k++) vdep[k] = rh.vdep[k];
        }
        uint32_t nslot = wave_append(Q.counters + WF_CNT_NEXT, 0u, cont);
        if (cont) Q.next[nslot] = pid;
        for (int k = 0; k < 3; k++) {
            bool has = valid && k < nsh;
            uint32_t s = wave_append(Q.counters + WF_CNT_SHADOW, 0u, has);
            if (has) {
                WfShadow q;
                q.o_tmax = make_float4(sh[k].o.x, sh[k].o.y, sh[k].o.z, sh[k].t_max);
                q.d_time = make_float4(sh[k].d.x, sh[k].d.y, sh[k].d.z, sh[k].time);
                q.c_pixel = make_float4(sh[k].contrib.x, sh[k].contrib.y, sh[k].contrib.z, 0.0f);
                q.meta = make_uint4(0u, pid, vdep[k] ? 1u : 0u, 0u);
                Q.shadow[s] = q;
            }
        }
    }
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < A.nparams; t += BLOCK)
        if (g[t] != 0.0f) atomicAdd(A.grad + t, g[t]);
}

// Shadow rays of the replay: the collected radiance leaves the path's
// remainder; a vertex-dependent one (its value carries that vertex's BSDF)
// adds its own score term.
__global__ __launch_bounds__(BLOCK) void k_wfb_shadow(DScene S, WfQueues Q, WfAdjQ A) {
    __shared__ float g[RT_MAX_PARAMS];
    for (uint32_t t = threadIdx.x; t < A.nparams; t += BLOCK) g[t] = 0.0f;
    __syncthreads();
    RT_STACK(st);
    uint32_t count = Q.counters[WF_CNT_SHADOW];
    uint32_t stride = gridDim.x * BLOCK;
    for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i < count; i += stride) {
        WfShadow q = Q.shadow[i];
        Ray r = make_ray(mk(q.o_tmax.x, q.o_tmax.y, q.o_tmax.z), mk(q.d_time.x, q.d_time.y, q.d_time.z), q.d_time.w);
        float tr = trace_transmittance(S, r, 0.001f, q.o_tmax.w, st, false);
        if (tr > 0.0f) {
            WfAdj& a = A.adj[q.meta.y];
            V3 c = mk(q.c_pixel.x, q.c_pixel.y, q.c_pixel.z) * tr;
            atomicAdd(&a.rem.x, -c.x);
            atomicAdd(&a.rem.y, -c.y);
            atomicAdd(&a.rem.z, -c.z);
            if (q.meta.z) {
                V3 d = mk(a.delta.x, a.delta.y, a.delta.z);
                for (int k = 0; k < a.vn; k++) {
                    float v = d.x * c.x * a.vv[3 * k] + d.y * c.y * a.vv[3 * k + 1] + d.z * c.z * a.vv[3 * k + 2];
                    if (v != 0.0f && isfinite(v)) atomicAdd(g + a.vp[k], v);
                }
            }
            if (A.nmp) {
                ShadowReq sq;
                sq.o = r.o; sq.d = r.d; sq.t_max = q.o_tmax.w; sq.time = q.d_time.w;
                float dc = a.delta.x * c.x + a.delta.y * c.y + a.delta.z * c.z;
                for (int k = 0; k < A.nmp; k++) {
                    float v = -shadow_length(S, A.mp[k], sq) * dc;
                    if (v != 0.0f && isfinite(v)) atomicAdd(g + A.mp[k].param, v);
                }
            }
        }
    }
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < A.nparams; t += BLOCK)
        if (g[t] != 0.0f) atomicAdd(A.grad + t, g[t]);
}
