// The plain per-pixel megakernel.
#include "rt_kern.h"

// Plain megakernel: one thread per pixel, `spp` consecutive sample passes,
// radiance summed in registers and added to (or stored into) `out`.  H4:
// four hero wavelengths per camera ray (an instantiation of its own, so the
// reference path keeps its register budget).
template <bool KNN, bool H4>
__global__ __launch_bounds__(BLOCK) void k_render(DScene S, DCamera cam, DParams P, const uint8_t* __restrict__ conv,
                                                  float* __restrict__ out, uint32_t spp, int accumulate) {
    RT_STACK(st);
    if (KNN) { st.heap = LdsHeap::column(); st.has_heap = true; }
    uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    uint32_t npix = P.width * P.height;
    if (i >= npix) return;
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    if (!(conv && conv[i])) {
        uint32_t row = i / P.width, col = i - row * P.width;
        for (uint32_t k = 0; k < spp; k++) {
            uint32_t sidx = P.sample_idx + k;
            Rng rng = rng_seed(pixel_seed(i, sidx, P.seed));
            Ray r = pixel_ray(cam, P, row, col, sidx, rng);
            V3 c = ray_color<H4>(S, r, (int)P.max_depth, P.bg_scale, P.use_photons != 0, rng, st);
            sx += c.x; sy += c.y; sz += c.z;
        }
    }
    if (accumulate) { out[3 * i] += sx; out[3 * i + 1] += sy; out[3 * i + 2] += sz; }
    else { out[3 * i] = sx; out[3 * i + 1] = sy; out[3 * i + 2] = sz; }
    if (P.counters) {
        atomicAdd(P.counters + 0, (unsigned long long)st.n_nodes);
        atomicAdd(P.counters + 1, (unsigned long long)st.n_prims);
        atomicAdd(P.counters + 2, (unsigned long long)st.n_rays);
        atomicAdd(P.counters + 3, (unsigned long long)((conv && conv[i]) ? 0u : spp));
        atomicAdd(P.counters + 4, (unsigned long long)st.s_nodes);
        atomicAdd(P.counters + 5, (unsigned long long)st.s_prims);
        atomicAdd(P.counters + 6, (unsigned long long)st.s_rays);
    }
}

template __global__ void k_render<true, false>(DScene, DCamera, DParams, const uint8_t*, float*, uint32_t, int);
template __global__ void k_render<false, false>(DScene, DCamera, DParams, const uint8_t*, float*, uint32_t, int);
template __global__ void k_render<true, true>(DScene, DCamera, DParams, const uint8_t*, float*, uint32_t, int);
template __global__ void k_render<false, true>(DScene, DCamera, DParams, const uint8_t*, float*, uint32_t, int);
