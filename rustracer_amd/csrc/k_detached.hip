// Megakernel adjoint, forward half: the detached-sampling render.
#include "rt_kern.h"

// Detached-sampling forward render: the k_render frame with BSDF values
// evaluated on the materials `me` while every decision follows the scene's.
__global__ __launch_bounds__(BLOCK) void k_detached(DScene S, DCamera cam, DParams P, const DMaterial* __restrict__ me,
                                                    const MedPar* __restrict__ mp, int nmp, float* __restrict__ out, uint32_t spp) {
    RT_STACK(st);
    uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    uint32_t npix = P.width * P.height;
    if (i >= npix) return;
    uint32_t row = i / P.width, col = i - row * P.width;
    V3 sum = mk(0, 0, 0);
    for (uint32_t k = 0; k < spp; k++) {
        uint32_t sidx = P.sample_idx + k;
        Rng rng = rng_seed(pixel_seed(i, sidx, P.seed));
        Ray r = pixel_ray(cam, P, row, col, sidx, rng);
        DetachedHook dh;
        dh.me = me; dh.mp = mp; dh.nmp = nmp; dh.ratio = mk(1, 1, 1);
        sum += ray_color_t(S, r, (int)P.max_depth, P.bg_scale, P.use_photons != 0, rng, st, dh);
    }
    out[3 * i] += sum.x; out[3 * i + 1] += sum.y; out[3 * i + 2] += sum.z;
}

