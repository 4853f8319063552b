// Wavefront stages that trace: camera-ray generation, extend (closest hit + material sort) and shadow.
#include "rt_kern.h"

template <bool BATCH>
__global__ __launch_bounds__(BLOCK) void k_wf_generate(DCamera cam0, DParams P0, WfQueues Q, uint64_t first, uint32_t n, WfBatch B) {
    uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    if (first == ~0ull) first = ((uint64_t)Q.counters[WF_CNT_FIRST + 1] << 32) | Q.counters[WF_CNT_FIRST];
    uint64_t g = first + i;
    uint32_t view = 0, fb_off = 0;
    DCamera cam = cam0;
    DParams P = P0;
    if (BATCH) {
        uint32_t lo = 0, hi = B.nviews;            // last view with views[v].first <= g
        while (hi - lo > 1) { uint32_t mid = (lo + hi) / 2; if (B.views[mid].first <= g) lo = mid; else hi = mid; }
        view = lo;
        const DView& V = B.views[view];
        cam = V.cam; P = V.P; fb_off = V.fb_off;
        g -= V.first;
    }
    uint32_t npix = P.width * P.height;
    uint32_t pixel = tile_order_pixel((uint32_t)(g % npix), P.width, P.height);
    uint32_t sidx = P.sample_idx + (uint32_t)(g / npix);
    uint32_t row = pixel / P.width, col = pixel - row * P.width;
    Rng rng = rng_seed(pixel_seed(pixel, sidx, P.seed));
    Ray r = pixel_ray(cam, P, row, col, sidx, rng);
    PathState ps;
    path_init(ps, r);
    WfPath w;
    wf_store(w, ps, rng);
    if (Q.h4) { Spec4 s4; spec4_init(s4); spec4_store(Q, i, s4); }
    w.pixel = fb_off + pixel; w.pad0 = view; w.hit_t = 0.0f; w.hit_slot = -1; w.hit_prim = -1; w.key = 0; w.sid = i;
    Q.paths[i] = w;
    if (Q.conv) {                         // compact: only unflagged pixels enter the active set
        bool live = Q.conv[pixel] == 0;
        uint32_t slot = wave_append(Q.counters + WF_CNT_GEN, 0u, live);
        if (live) Q.active[slot] = i;
    } else {
        Q.active[i] = i;
    }
}

// Bucket of a hit: per material kind (sort 1) or per shade class (sort 2:
// misses, media and the specular materials share k_wf_shade<..., 1>).
__device__ __forceinline__ uint32_t sort_key(const DScene& S, const HitId& id, uint32_t sort) {
    if (sort == 0) return 0u;
    if (id.slot < 0) return 0u;
    if (id.prim == -2) return sort == 1 ? (uint32_t)M_KINDS + 1u : 0u;
    uint32_t k = S.mats[S.objs[id.slot].material].kind;
    if (sort == 1) return k <= (uint32_t)M_KINDS ? k : 0u;
    return RT_IN(RT_MASK_SPECULAR, k) ? 0u : 1u;
}

__device__ __forceinline__ uint32_t extend_one(const DScene& S, WfPath& w, uint32_t sort, Stack& st) {
    float4 o = w.o_time, d = w.d_wl;
    Ray r = make_ray(mk(o.x, o.y, o.z), mk(d.x, d.y, d.z), o.w);
    r.wavelength = d.w; r.spectral = (w.flags & 1u) != 0;   // media seed their free-flight stream from the ray, wavelength included
    HitId id = trace_closest(S, r, 0.001f, FINF, st);
    w.hit_t = id.t; w.hit_slot = id.slot; w.hit_prim = id.prim;
    return sort_key(S, id, sort);
}



// OCC = minimum waves per SIMD the register allocator must allow (0: its own choice).
template <int OCC, bool BATCH, int BS>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(OCC > 0 ? OCC : 1, 8))) void k_wf_extend(DScene S, WfQueues Q, WfBatch B) {
    RT_STACK_BS(st, BS);
    uint32_t count = Q.counters[WF_CNT_ACTIVE];
    uint32_t stride = gridDim.x * BS;
    for (uint32_t base = blockIdx.x * BS; base < count; base += stride) {
        uint32_t i = base + threadIdx.x;
        bool valid = i < count;
        uint32_t pid = 0, key = 0;
        if (valid) {
            pid = Q.active[i];
            WfPath& w = Q.paths[pid];
            if (BATCH) key = extend_one(view_scene(B, w.pad0), w, Q.sort, st);
            else key = extend_one(S, w, Q.sort, st);
        }
        uint32_t slot = wave_append(Q.counters + WF_CNT_BKT(0), key, valid, WF_BKT_STRIDE);
        if (valid) Q.buckets[(size_t)key * Q.cap + slot] = pid;
    }
}


template <int OCC, bool BATCH, int BS>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(OCC > 0 ? OCC : 1, 8))) void k_wf_shadow(DScene S, WfQueues Q, WfBatch B) {
    RT_STACK_BS(st, BS);
    uint32_t count = Q.counters[WF_CNT_SHADOW];
    uint32_t stride = gridDim.x * BS;
    for (uint32_t i = blockIdx.x * BS + threadIdx.x; i < count; i += stride) {
        WfShadow q = Q.shadow[Q.shadow_idx[i]];
        Ray r = make_ray(mk(q.o_tmax.x, q.o_tmax.y, q.o_tmax.z), mk(q.d_time.x, q.d_time.y, q.d_time.z), q.d_time.w);
        float tr = BATCH ? trace_transmittance(view_scene(B, q.meta.x), r, 0.001f, q.o_tmax.w, st, false)
                         : trace_transmittance(S, r, 0.001f, q.o_tmax.w, st, false);
        if (tr > 0.0f) {
            float* dst = Q.lsamp ? Q.lsamp + 3 * (size_t)q.meta.y : Q.fb + 3 * (size_t)__float_as_uint(q.c_pixel.w);
            atomicAdd(dst, q.c_pixel.x * tr);
            atomicAdd(dst + 1, q.c_pixel.y * tr);
            atomicAdd(dst + 2, q.c_pixel.z * tr);
        }
    }
}

template __global__ void k_wf_generate<true>(DCamera, DParams, WfQueues, uint64_t, uint32_t, WfBatch);
template __global__ void k_wf_generate<false>(DCamera, DParams, WfQueues, uint64_t, uint32_t, WfBatch);
template __global__ void k_wf_extend<0, true, 64>(DScene, WfQueues, WfBatch);
template __global__ void k_wf_extend<0, true, 128>(DScene, WfQueues, WfBatch);
template __global__ void k_wf_extend<0, true, 256>(DScene, WfQueues, WfBatch);
template __global__ void k_wf_extend<0, false, 64>(DScene, WfQueues, WfBatch);
template __global__ void k_wf_extend<0, false, 128>(DScene, WfQueues, WfBatch);
template __global__ void k_wf_extend<0, false, 256>(DScene, WfQueues, WfBatch);
template __global__ void k_wf_extend<4, true, 64>(DScene, WfQueues, WfBatch);
template __global__ void k_wf_extend<4, true, 128>(DScene, WfQueues, WfBatch);
template __global__ void k_wf_extend<4, true, 256>(DScene, WfQueues, WfBatch);
template __global__ void k_wf_extend<4, false, 64>(DScene, WfQueues, WfBatch);
template __global__ void k_wf_extend<4, false, 128>(DScene, WfQueues, WfBatch);
template __global__ void k_wf_extend<4, false, 256>(DScene, WfQueues, WfBatch);
template __global__ void k_wf_extend<5, true, 64>(DScene, WfQueues, WfBatch);
template __global__ void k_wf_extend<5, true, 128>(DScene, WfQueues, WfBatch);
template __global__ void k_wf_extend<5, true, 256>(DScene, WfQueues, WfBatch);
template __global__ void k_wf_extend<5, false, 64>(DScene, WfQueues, WfBatch);
template __global__ void k_wf_extend<5, false, 128>(DScene, WfQueues, WfBatch);
template __global__ void k_wf_extend<5, false, 256>(DScene, WfQueues, WfBatch);
template __global__ void k_wf_shadow<0, true, 64>(DScene, WfQueues, WfBatch);
template __global__ void k_wf_shadow<0, true, 128>(DScene, WfQueues, WfBatch);
template __global__ void k_wf_shadow<0, true, 256>(DScene, WfQueues, WfBatch);
template __global__ void k_wf_shadow<0, false, 64>(DScene, WfQueues, WfBatch);
template __global__ void k_wf_shadow<0, false, 128>(DScene, WfQueues, WfBatch);
template __global__ void k_wf_shadow<0, false, 256>(DScene, WfQueues, WfBatch);
template __global__ void k_wf_shadow<5, true, 64>(DScene, WfQueues, WfBatch);
template __global__ void k_wf_shadow<5, true, 128>(DScene, WfQueues, WfBatch);
template __global__ void k_wf_shadow<5, true, 256>(DScene, WfQueues, WfBatch);
template __global__ void k_wf_shadow<5, false, 64>(DScene, WfQueues, WfBatch);
template __global__ void k_wf_shadow<5, false, 128>(DScene, WfQueues, WfBatch);
template __global__ void k_wf_shadow<5, false, 256>(DScene, WfQueues, WfBatch);
template __global__ void k_wf_shadow<6, true, 64>(DScene, WfQueues, WfBatch);
template __global__ void k_wf_shadow<6, true, 128>(DScene, WfQueues, WfBatch);
template __global__ void k_wf_shadow<6, true, 256>(DScene, WfQueues, WfBatch);
template __global__ void k_wf_shadow<6, false, 64>(DScene, WfQueues, WfBatch);
template __global__ void k_wf_shadow<6, false, 128>(DScene, WfQueues, WfBatch);
template __global__ void k_wf_shadow<6, false, 256>(DScene, WfQueues, WfBatch);
