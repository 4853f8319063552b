// Wavefront k-NN photon gather stage.
#include "rt_kern.h"

// k-NN photon maps: between extend and shade, the photon irradiance of every
// hit whose material can take the diffuse branch (Lambertian, PBR, Pearl),
// into the path record, so the shade kernel keeps its occupancy.  Two
// passes: k_wf_knn (no LDS, full occupancy) finishes every query with fewer
// than k photons within the gather radius and queues the rest; k_wf_knn_dense
// searches those: one query per lane with a heap in LDS when they are many,
// one query per wave when they are few (the walk's latency then spread over
// the wave's lanes).  Paths whose hit does not gather are
// left as they are.  The first pass leaves each dense query's point, normal
// and path id in the shadow queue's memory, which is idle between extend and
// shade (two float4 per dense slot: point and path id, normal and the path's
// view); the heap search reads the view and puts the radius it finds in its place.
template <bool BATCH>
__device__ __forceinline__ bool knn_query(const DScene& S, const WfPath& w, V3& pos, V3& nrm) {
    Ray r = make_ray(mk(w.o_time.x, w.o_time.y, w.o_time.z), mk(w.d_wl.x, w.d_wl.y, w.d_wl.z), w.o_time.w);
    r.wavelength = w.d_wl.w; r.spectral = (w.flags & 1u) != 0;
    HitId id; id.t = w.hit_t; id.slot = w.hit_slot; id.prim = w.hit_prim;
    Rec h = finalize(S, r, id, 0.001f);
    pos = h.p; nrm = h.n;
    return true;
}
__device__ __forceinline__ void knn_store(WfPath& w, V3 e) {
    w.pad1[0] = __float_as_uint(e.x); w.pad1[1] = __float_as_uint(e.y); w.pad1[2] = __float_as_uint(e.z);
}
template <bool BATCH>
__global__ __launch_bounds__(BLOCK) void k_wf_knn(DScene S0, DParams P0, WfQueues Q, WfBatch B) {
    uint32_t off[WF_NB + 1];                 // bucket starts, read once
    off[0] = 0;
    for (int b = 0; b < WF_NB; b++) off[b + 1] = off[b] + Q.counters[WF_CNT_BKT(b)];
    const uint32_t total = off[WF_NB];
    uint32_t stride = gridDim.x * BLOCK;
    for (uint32_t base0 = blockIdx.x * BLOCK; base0 < total; base0 += stride) {
        uint32_t i = base0 + threadIdx.x;
        bool dense = false;
        uint32_t pid = 0, view = 0;
        V3 pos = mk(0, 0, 0), nrm = mk(0, 0, 0);
        if (i < total) {
            uint32_t b = 0;
            while (i >= off[b + 1]) b++;
            pid = Q.buckets[(size_t)b * Q.cap + (i - off[b])];
            WfPath& w = Q.paths[pid];
            view = w.pad0;
            const DScene S = BATCH ? view_scene(B, view) : S0;
            bool photons = BATCH ? B.views[w.pad0].P.use_photons != 0 : P0.use_photons != 0;
            if (photons && knn_gathers(S, w)) {
                V3 e;
                knn_query<BATCH>(S, w, pos, nrm);
                if (photon_knn_sparse(S, pos, nrm, e)) knn_store(w, e);
                else dense = true;
            }
        }
        uint32_t slot = wave_append(Q.counters + WF_CNT_DENSE, 0u, dense);
        if (dense) {
            Q.dense[slot] = pid;
            float4* stash = reinterpret_cast<float4*>(Q.shadow);
            stash[2 * (size_t)slot] = make_float4(pos.x, pos.y, pos.z, __uint_as_float(pid));
            stash[2 * (size_t)slot + 1] = make_float4(nrm.x, nrm.y, nrm.z, __uint_as_float(BATCH ? view : 0u));
        }
    }
}
// ---- the dense pass: one query per wave ------------------------------------------------------
// k <= 64 = the wave width, so a wave can hold a query's k nearest photons
// one per lane, sorted across the lanes, instead of each lane keeping a heap
// of its own query in LDS.  The wave walks the bucket tree for its query in
// lockstep (the same near-first walk and the same pruning as pk_walk), takes
// up to 8 leaves (<= 64 photons) per step, one photon per lane, and merges
// the candidates closer than the current bound into the sorted set with
// bitonic networks over the lanes; the bound shrinks to the k-th entry as
// soon as k are held.  The gather at the k-th distance then sums exactly the
// entries of the set (every photon closer than the k-th distance is in it,
// the k-th itself weighs 0): one walk instead of two, no LDS, and a query's
// latency spread over 64 lanes instead of one.

// Compare-exchange of the entries of lanes (lane, lane ^ j): ascending order
// when `up`; equal distances stay where they are.
__device__ __forceinline__ void wave_cmpx(float& d, uint32_t& i, uint32_t lane, int j, bool up) {
    float od = __shfl_xor(d, j);
    uint32_t oi = (uint32_t)__shfl_xor((int)i, j);
    bool lower = (lane & (uint32_t)j) == 0u;
    bool take = (lower == up) ? (od < d) : (od > d);
    if (take) { d = od; i = oi; }
}
// Bitonic sort of the wave's 64 entries, ascending by lane.
__device__ __forceinline__ void wave_sort64(float& d, uint32_t& i, uint32_t lane) {
    for (int k = 2; k <= 64; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) wave_cmpx(d, i, lane, j, (lane & (uint32_t)k) == 0u);
}
// Keep the 64 smallest of two ascending sets (kd, ki) and (c, ci), ascending in (kd, ki).
__device__ __forceinline__ void wave_merge64(float& kd, uint32_t& ki, float c, uint32_t ci, uint32_t lane) {
    float rc = __shfl(c, 63 - (int)lane);
    uint32_t rci = (uint32_t)__shfl((int)ci, 63 - (int)lane);
    if (rc < kd) { kd = rc; ki = rci; }
    for (int j = 32; j > 0; j >>= 1) wave_cmpx(kd, ki, lane, j, true);
}
__device__ __forceinline__ float wave_sum(float v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// The k-NN irradiance of one query, computed by the whole wave (every lane
// passes the same pos / nrm and gets the result): photon_knn_irradiance's
// value, with the gather summed over the lanes.
__device__ V3 knn_wave(const DScene& S, V3 pos, V3 nrm, uint32_t lane) {
    const uint32_t D = S.pk_depth, k = S.ph_knn;
    const uint32_t first_leaf = (1u << D) - 1u;
    float kd = FINF;                  // this lane's entry of the sorted set (FINF: empty)
    uint32_t ki = 0;
    float bound = S.ph_r2;
    uint32_t node = 0, pend = 0, depth = 0;
    bool more = true;
    while (more) {
        uint32_t b = 0, e = 0;        // this lane's leaf of the step: lane / 8
        for (uint32_t nl = 0; nl < 8 && more; nl++) {
            while (depth < D) {                                  // descend to the near leaf
                float2 sp = S.pk_split[node];
                float sq = comp(pos, __float_as_int(sp.y)) - sp.x;
                if (sq * sq < bound) pend |= 1u << depth; else pend &= ~(1u << depth);
                node = sq <= 0.0f ? 2 * node + 1 : 2 * node + 2;
                depth++;
            }
            uint32_t leaf = node - first_leaf;
            if ((lane >> 3) == nl) { b = S.pk_leaf[leaf]; e = S.pk_leaf[leaf + 1]; }
            more = false;                                        // pop the deepest pending far child
            while (pend) {
                uint32_t d = 31u - (uint32_t)__clz(pend);
                pend &= ~(1u << d);
                uint32_t anc = ((node + 1) >> (depth - d)) - 1;
                float2 sp = S.pk_split[anc];
                float sq = comp(pos, __float_as_int(sp.y)) - sp.x;
                if (!(sq * sq < bound)) continue;
                node = sq <= 0.0f ? 2 * anc + 2 : 2 * anc + 1;
                depth = d + 1;
                more = true;
                break;
            }
        }
        uint32_t i = b + (lane & 7u);
        float c = FINF;
        uint32_t ci = 0;
        if (i < e) {
            float4 P = S.pk_rec[2 * i], N = S.pk_rec[2 * i + 1];
            float dx = P.x - pos.x, dy = P.y - pos.y, dz = P.z - pos.z;
            float d2 = dx * dx + dy * dy + dz * dz;
            if (d2 < bound && N.x * nrm.x + N.y * nrm.y + N.z * nrm.z > 0.0f) { c = d2; ci = i; }
        }
        if (__ballot(c < FINF) == 0ull) continue;
        wave_sort64(c, ci, lane);
        wave_merge64(kd, ki, c, ci, lane);
        float kth = __shfl(kd, (int)k - 1);
        if (kth < FINF) bound = kth;                             // k held: prune to the k-th distance
    }
    uint32_t held = (uint32_t)__popcll(__ballot(kd < FINF));
    float r2 = S.ph_r2, norm = S.ph_norm;
    if (held >= k) {
        float kth = __shfl(kd, (int)k - 1);
        if (!(kth > 0.0f)) {                                     // k photons at the query point itself
            uint32_t cnt;
            float unused;
            return pk_walk<false>(S, pos, nrm, S.ph_r2, nullptr, 0, cnt, unused) * S.ph_norm;
        }
        r2 = kth;
        norm = 2.0f / (PI * PI * r2);
    }
    V3 a = mk(0, 0, 0);
    if (lane < k && kd < r2) { float4 W = S.pk_pow[ki]; a = mk(W.x, W.y, W.z) * (1.0f - kd / r2); }
    return mk(wave_sum(a.x), wave_sum(a.y), wave_sum(a.z)) * norm;
}

// The dense pass is launched as three kernels back to back, each of which
// returns at once unless the dense count is in its regime.  More than
// knn_wave_max queries: k_wf_knn_dense runs each query's heap search (one
// query per lane, the heap in LDS) and leaves the radius it finds in the
// stash; k_wf_knn_gather then sums the gather at that radius with no LDS, at
// full occupancy (photon_knn_radius and photon_knn_gather: the walks and
// sums of a one-kernel search).  Otherwise k_wf_knn_wave takes one query per
// wave.
template <bool BATCH>
__global__ __launch_bounds__(BLOCK) void k_wf_knn_dense(DScene S0, WfQueues Q, WfBatch B) {
    const uint32_t total = Q.counters[WF_CNT_DENSE];
    if (total <= Q.knn_wave_max) return;
    lds_f32* heap = LdsHeap::column();
    float4* stash = reinterpret_cast<float4*>(Q.shadow);
    uint32_t stride = gridDim.x * BLOCK;
    for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i < total; i += stride) {
        float4 a = stash[2 * (size_t)i], n = stash[2 * (size_t)i + 1];
        const DScene S = BATCH ? view_scene(B, __float_as_uint(n.w)) : S0;   // the view, until the radius replaces it
        n.w = photon_knn_radius(S, mk(a.x, a.y, a.z), mk(n.x, n.y, n.z), heap, BLOCK);
        stash[2 * (size_t)i + 1] = n;
    }
}
template <bool BATCH>
__global__ __launch_bounds__(BLOCK) void k_wf_knn_gather(DScene S0, WfQueues Q, WfBatch B) {
    const uint32_t total = Q.counters[WF_CNT_DENSE];
    if (total <= Q.knn_wave_max) return;
    const float4* stash = reinterpret_cast<const float4*>(Q.shadow);
    uint32_t stride = gridDim.x * BLOCK;
    for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i < total; i += stride) {
        float4 a = stash[2 * (size_t)i], n = stash[2 * (size_t)i + 1];
        WfPath& w = Q.paths[__float_as_uint(a.w)];
        const DScene S = BATCH ? view_scene(B, w.pad0) : S0;
        knn_store(w, photon_knn_gather(S, mk(a.x, a.y, a.z), mk(n.x, n.y, n.z), n.w));
    }
}
template <bool BATCH>
__global__ __launch_bounds__(BLOCK) void k_wf_knn_wave(DScene S0, WfQueues Q, WfBatch B) {
    constexpr uint32_t WPB = BLOCK / 64;
    const uint32_t lane = __lane_id();
    const uint32_t total = Q.counters[WF_CNT_DENSE];
    if (total > Q.knn_wave_max) return;
    const uint32_t waves = gridDim.x * WPB;
    for (uint32_t q = blockIdx.x * WPB + threadIdx.x / 64; q < total; q += waves) {
        WfPath& w = Q.paths[Q.dense[q]];
        const DScene S = BATCH ? view_scene(B, w.pad0) : S0;
        V3 pos, nrm;
        knn_query<BATCH>(S, w, pos, nrm);
        V3 e = knn_wave(S, pos, nrm, lane);
        if (lane == 0) knn_store(w, e);
    }
}

template __global__ void k_wf_knn<true>(DScene, DParams, WfQueues, WfBatch);
template __global__ void k_wf_knn_dense<true>(DScene, WfQueues, WfBatch);
template __global__ void k_wf_knn_gather<true>(DScene, WfQueues, WfBatch);
template __global__ void k_wf_knn_wave<true>(DScene, WfQueues, WfBatch);
template __global__ void k_wf_knn<false>(DScene, DParams, WfQueues, WfBatch);
template __global__ void k_wf_knn_dense<false>(DScene, WfQueues, WfBatch);
template __global__ void k_wf_knn_gather<false>(DScene, WfQueues, WfBatch);
template __global__ void k_wf_knn_wave<false>(DScene, WfQueues, WfBatch);
