// Declarations of the kernels for the host code.  Each kernel is defined (and
// its template arguments instantiated) in one k_*.hip translation unit; the
// host launches it through the device stub that unit exports.
#pragma once
#include "rt_kern.h"

// k_misc.hip
__global__ void k_trace_closest(DScene S, const rt_ray* rays, uint32_t n, rt_hit* hits);
__global__ void k_trace_tr(DScene S, const rt_ray* rays, uint32_t n, float* tr, uint8_t* occ);
__global__ void k_accumulate(const float* sc, uint32_t npix, rt_accum_opts o, float* acc, uint32_t* ns, float* m2, const uint8_t* conv);
__global__ void k_mark_converged(uint32_t npix, float threshold, const float* acc, const uint32_t* ns, const float* m2, uint8_t* conv,
                                 uint32_t* count);
__global__ void k_count_flags(uint32_t npix, const uint8_t* conv, uint32_t* count);
__global__ void k_resolve(const float* acc, const uint32_t* ns, uint32_t spp, uint32_t w, uint32_t h, rt_resolve_opts o, uint8_t* out);
__global__ void k_progressive_step(uint32_t* ctl, uint32_t pass, uint32_t npix, rt_accum_opts o);
__global__ void k_aux(DScene S, DCamera cam, uint32_t w, uint32_t h, float* alb, float* nrm);
__global__ void k_elut(float* out);
__global__ void k_wf_reset(uint32_t* counters, uint32_t active, const uint32_t* stop, int generated);
__global__ void k_wf_gen_zero(uint32_t* counters);
__global__ void k_wf_advance(uint32_t* counters);
__global__ void k_wf_set_first(uint32_t* counters, uint64_t first);
__global__ void k_wf_bump_first(uint32_t* counters, uint32_t n);
__global__ void k_accum_passes(const float* lsamp, uint32_t K, uint32_t w, uint32_t h, uint32_t pass0, rt_accum_opts o, float* acc,
                               uint32_t* ns, float* m2, uint8_t* conv, uint32_t* ctl);
template <bool KNN> __global__ void k_photon_irr(DScene S, const float* pn, uint32_t n, float* out);
__global__ void k_fill_u32(uint32_t* p, uint64_t n, uint32_t v);

// k_render.hip, k_render_o2.hip (k_render.inc), k_progressive.hip
template <int OCC, bool KNN, bool H4> __global__ void k_render(DScene S, DCamera cam, DParams P, const uint8_t* conv, float* out, uint32_t spp, int accumulate);
template <bool KNN, bool H4> __global__ void k_progressive(DScene S, DCamera cam, DParams P, rt_accum_opts o, uint32_t pass, float* acc,
                                                  uint32_t* ns, float* m2, uint8_t* conv, uint32_t* ctl);

// k_adjoint.hip
__global__ void k_detached(DScene S, DCamera cam, DParams P, const DMaterial* me, const MedPar* mp, int nmp, float* out, uint32_t spp);
__global__ void k_backward(DScene S, DCamera cam, DParams P, const int8_t* fmap, const MedPar* mp, int nmp, uint32_t nparams,
                           const float* dimg, float inv_spp, uint32_t spp, float* grad);

// k_photon.hip
__global__ void k_photon_trace(DScene S, DEmitter E, uint32_t n, float4* out, uint32_t* ok);
__global__ void k_caustic_emitter(DScene S, float* out);

// k_wf_trace.hip, k_wf_knn.hip, k_wf_shade_*.hip
template <bool BATCH> __global__ void k_wf_generate(DCamera cam0, DParams P0, WfQueues Q, uint64_t first, uint32_t n, WfBatch B);
template <int OCC, bool BATCH, int BS> __global__ void k_wf_extend(DScene S, WfQueues Q, WfBatch B);
template <int OCC, bool BATCH, int BS> __global__ void k_wf_shadow(DScene S, WfQueues Q, WfBatch B);
template <bool BATCH> __global__ void k_wf_knn(DScene S0, DParams P0, WfQueues Q, WfBatch B);
template <bool BATCH> __global__ void k_wf_knn_dense(DScene S0, WfQueues Q, WfBatch B);
template <bool BATCH> __global__ void k_wf_knn_gather(DScene S0, WfQueues Q, WfBatch B);
template <bool BATCH> __global__ void k_wf_knn_wave(DScene S0, WfQueues Q, WfBatch B);
template <bool BATCH, int CLASS, bool H4> __global__ void k_wf_shade(DScene S, DParams P, WfQueues Q, int depth, WfBatch B);

// k_wfb.hip
__global__ void k_wfb_init(WfQueues Q, WfAdjQ A, uint32_t n);
__global__ void k_wfb_shade(DScene S, DParams P, WfQueues Q, WfAdjQ A, int depth);
__global__ void k_wfb_shadow(DScene S, WfQueues Q, WfAdjQ A);
