// Megakernel adjoint, backward half: the path-replay gradient pass.
#include "rt_kern.h"

// Path-replay adjoint: per sample, the radiance L from its stream, then the
// same path again distributing L over the vertices' dlog f (see
// dev_adjoint.h).  Gradients gather in LDS per block, then one global atomic
// per parameter and block.
__global__ __launch_bounds__(BLOCK) void k_backward(DScene S, DCamera cam, DParams P, const int8_t* __restrict__ fmap,
                                                    const MedPar* __restrict__ mp, int nmp,
                                                    uint32_t nparams, const float* __restrict__ dimg, float inv_spp,
                                                    uint32_t spp, float* __restrict__ grad) {
    __shared__ float g[RT_MAX_PARAMS];
    for (uint32_t t = threadIdx.x; t < nparams; t += BLOCK) g[t] = 0.0f;
    __syncthreads();
    RT_STACK(st);
    uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    uint32_t npix = P.width * P.height;
    if (i < npix) {
        uint32_t row = i / P.width, col = i - row * P.width;
        V3 delta = mk(dimg[3 * i], dimg[3 * i + 1], dimg[3 * i + 2]) * inv_spp;
        if (delta.x != 0.0f || delta.y != 0.0f || delta.z != 0.0f) {
            for (uint32_t k = 0; k < spp; k++) {
                uint32_t sidx = P.sample_idx + k;
                uint64_t seed = pixel_seed(i, sidx, P.seed);
                Rng rng = rng_seed(seed);
                Ray r = pixel_ray(cam, P, row, col, sidx, rng);
                PlainHook ph;
                V3 L = ray_color_t(S, r, (int)P.max_depth, P.bg_scale, P.use_photons != 0, rng, st, ph);
                if (!isfinite(L.x) || !isfinite(L.y) || !isfinite(L.z)) continue;
                rng = rng_seed(seed);
                r = pixel_ray(cam, P, row, col, sidx, rng);
                ReplayHook rh;
                rh.fmap = fmap; rh.mp = mp; rh.nmp = nmp; rh.g = g; rh.delta = delta; rh.rem = L;
                ray_color_t(S, r, (int)P.max_depth, P.bg_scale, P.use_photons != 0, rng, st, rh);
            }
        }
    }
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < nparams; t += BLOCK)
        if (g[t] != 0.0f) atomicAdd(grad + t, g[t]);
}
