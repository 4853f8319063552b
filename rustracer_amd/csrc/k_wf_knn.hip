// Wavefront k-NN photon gather stage.
#include "rt_kern.h"

// k-NN photon maps: between extend and shade, the photon irradiance of every
// hit whose material can take the diffuse branch (Lambertian, PBR, Pearl),
// into the path record, so the shade kernel keeps its occupancy.  Two
// passes: k_wf_knn (no LDS, full occupancy) finishes every query with fewer
// than k photons within the gather radius and queues the rest; k_wf_knn_dense
// runs the heap search (32 KB of LDS per block) on those only, so its lanes
// all do the same kind of work.  Paths whose hit does not gather are left as
// they are.
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
    uint32_t total = 0;
    for (int b = 0; b < WF_NB; b++) total += Q.counters[WF_CNT_BUCKET + b];
    uint32_t stride = gridDim.x * BLOCK;
    for (uint32_t base0 = blockIdx.x * BLOCK; base0 < total; base0 += stride) {
        uint32_t i = base0 + threadIdx.x;
        bool dense = false;
        uint32_t pid = 0;
        if (i < total) {
            uint32_t b = 0, base = 0;
            while (i >= base + Q.counters[WF_CNT_BUCKET + b]) { base += Q.counters[WF_CNT_BUCKET + b]; b++; }
            pid = Q.buckets[(size_t)b * Q.cap + (i - base)];
            WfPath& w = Q.paths[pid];
            const DScene& S = BATCH ? B.scenes[w.pad0] : S0;
            bool photons = BATCH ? B.views[w.pad0].P.use_photons != 0 : P0.use_photons != 0;
            if (photons && knn_gathers(S, w)) {
                V3 pos, nrm, e;
                knn_query<BATCH>(S, w, pos, nrm);
                if (photon_knn_sparse(S, pos, nrm, e)) knn_store(w, e);
                else dense = true;
            }
        }
        uint32_t slot = wave_append(Q.counters + WF_CNT_DENSE, 0u, dense);
        if (dense) Q.dense[slot] = pid;
    }
}
template <bool BATCH>
__global__ __launch_bounds__(BLOCK) void k_wf_knn_dense(DScene S0, WfQueues Q, WfBatch B) {
    float* heap = LdsHeap::column();
    uint32_t total = Q.counters[WF_CNT_DENSE];
    uint32_t stride = gridDim.x * BLOCK;
    for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i < total; i += stride) {
        WfPath& w = Q.paths[Q.dense[i]];
        const DScene& S = BATCH ? B.scenes[w.pad0] : S0;
        V3 pos, nrm;
        knn_query<BATCH>(S, w, pos, nrm);
        knn_store(w, photon_knn_dense(S, pos, nrm, heap, BLOCK));
    }
}

template __global__ void k_wf_knn<true>(DScene, DParams, WfQueues, WfBatch);
template __global__ void k_wf_knn_dense<true>(DScene, WfQueues, WfBatch);
template __global__ void k_wf_knn<false>(DScene, DParams, WfQueues, WfBatch);
template __global__ void k_wf_knn_dense<false>(DScene, WfQueues, WfBatch);
